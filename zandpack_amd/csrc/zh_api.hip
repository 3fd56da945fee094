// zandhip C ABI: problem handle, device-resident DOP54 attempt, explicit RHS.
//
// Device layout (per k block, batched over NK on the GEMM z axis):
//   slots 0..6 : k_1..k_7 (h-scaled stage derivatives)          sigma | Psi | Omega
//   slot  7/8  : y0 / y_new (swapped on acceptance; DOP54's 7th stage input
//                 is y0 + sum b5_j k_j, so the committed state is materialised
//                 by the last stage combination and no separate commit pass runs)
//   slot  9    : stage input for sigma and Omega (stage input Psi lives in Ap)
//   Ap  (M, n+nc)  = [ psi_stage | dGP*xi on coupled orbitals ]     A of the Psi GEMM
//   Bp  (n+nc, n)  = [ H0^T ; hbar * sigma_stage[:, coupled]^T ]      B of the Psi GEMM
//   Aomg (M, 2nc)  = [ psi_stage[:, c] | dGP*xi_c ]                   A of the Omega GEMM
//   Bomg (M, 2nc)  = [ dGM*Ixi_c | conj psi_stage[:, c] ]             B^T of the Omega GEMM
// Psi/Omega storage is complex128 (ZH_FP64) or complex64 (ZH_MIXED); sigma,
// the sigma-path GEMMs (H sigma, Pi), the error norm and the controller stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/zandhip.h"
#include "zh_gemm.hpp"
#include "zh_kernels.hpp"

namespace zh {

static thread_local std::string g_err;

struct Error {
  int code;
  std::string msg;
};

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      throw Error{ZH_EHIP, std::string(#x) + ": " + hipGetErrorString(e_) + " (line " + \
                              std::to_string(__LINE__) + ")"};                            \
  } while (0)

static void require(bool ok, const std::string& msg, int code = ZH_EINVAL) {
  if (!ok) throw Error{code, msg};
}

static Tableau dop54() {
  Tableau t;
  std::memset(&t, 0, sizeof(t));
  const double c[7] = {0.0, 1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
  for (int i = 0; i < 7; ++i) t.c[i] = c[i];
  t.a[1][0] = 1.0 / 5;
  t.a[2][0] = 3.0 / 40; t.a[2][1] = 9.0 / 40;
  t.a[3][0] = 44.0 / 45; t.a[3][1] = -56.0 / 15; t.a[3][2] = 32.0 / 9;
  t.a[4][0] = 19372.0 / 6561; t.a[4][1] = -25360.0 / 2187; t.a[4][2] = 64448.0 / 6561;
  t.a[4][3] = -212.0 / 729;
  t.a[5][0] = 9017.0 / 3168; t.a[5][1] = -355.0 / 33; t.a[5][2] = 46732.0 / 5247;
  t.a[5][3] = 49.0 / 176; t.a[5][4] = -5103.0 / 18656;
  t.a[6][0] = 35.0 / 384; t.a[6][1] = 0.0; t.a[6][2] = 500.0 / 1113; t.a[6][3] = 125.0 / 192;
  t.a[6][4] = -2187.0 / 6784; t.a[6][5] = 11.0 / 84;
  const double b5[7] = {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84, 0.0};
  const double b4[7] = {5179.0 / 57600, 0.0, 7571.0 / 16695, 393.0 / 640, -92097.0 / 339200,
                        187.0 / 2100, 1.0 / 40};
  for (int j = 0; j < 7; ++j) t.ct[j] = b5[j] - b4[j];
  return t;
}

// Kernel classes timed with HIP events when profiling is on.
enum KClass { KC_PARAMS, KC_COMBINE, KC_GATHER, KC_GEMM_PI, KC_CURRENTS, KC_GEMM_HSIG, KC_SIGSYM,
              KC_GEMM_OMGXI, KC_GEMM_PSI, KC_GEMM_OMEGA, KC_TE, KC_CTRL, KC_N };
static const char* kc_names[KC_N] = {"stage_params", "combine",   "gather",     "gemm_pi",
                                      "currents",     "gemm_hsig", "sigma_sym",  "gemm_omgxi",
                                      "gemm_psi",     "gemm_omega", "te_reduce", "controller"};

struct EngineBase {
  virtual ~EngineBase() {}
  virtual void set_constants(const double* H0, const double* xi, const double* Ixi,
                             const double* glp, const double* dgp, const double* dgm,
                             const double* xpp, const double* xpm) = 0;
  virtual void set_hubbard(double U, const double* n0) = 0;
  virtual void set_pulse(const PulseDesc& p) = 0;
  virtual void state_put(const double* sig, const double* psi, const double* omg) = 0;
  virtual void state_get(double* sig, double* psi, double* omg) = 0;
  virtual void sigma_get(double* sig) = 0;
  virtual void rhs(const double* Ht, const double* delta, const double* sig, const double* psi,
                   const double* omg, double* dsig, double* dpsi, double* domg, double* J) = 0;
  virtual void attempt(zh_attempt_info* info, double* J) = 0;
  bool use_graph = true;
  bool profile = false;
  double kc_ms[KC_N] = {0};
  long kc_count[KC_N] = {0};
  double attempt_ms = 0.0;
  long attempt_count = 0;
};

template <typename R>  // storage real type of Psi / Omega and compute type of their GEMMs
struct Engine : EngineBase {
  typedef typename Cplx<R>::type cs;
  int NK, n, NL, NX, NOT, M, Ma, nc = 0;
  double hbar;
  int device;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  Tableau tab = dop54();
  PulseDesc pulse;
  double U = 0.0;
  bool have_consts = false, have_state = false, h_dirty = false;
  std::vector<double> H0h;  // host copy of H0 (re-installed after an explicit rhs call)

  // device buffers
  int *cidx = nullptr, *cpos = nullptr, *lead = nullptr;
  double2 *Hcur = nullptr, *Ixic = nullptr, *Pic = nullptr, *G = nullptr, *dvec = nullptr;
  double2 *glp = nullptr, *chip = nullptr, *chim = nullptr;
  double* n0 = nullptr;
  cs *Bp = nullptr, *Ap = nullptr, *Aomg = nullptr, *Bomg = nullptr, *xic = nullptr, *W = nullptr;
  double2* sig[10] = {nullptr};
  cs* psi[10] = {nullptr};
  cs* omg[10] = {nullptr};
  double* te_part = nullptr;
  Ctrl* ctrl = nullptr;
  StageParams* sp = nullptr;
  double* Jd = nullptr;
  double* Jpart = nullptr;
  int iy0 = 7, iyn = 8;
  static constexpr int kPiParts = 8;  // split-K of Π_c = ψᵀ Ixi_c (K = M)
  static constexpr int kWParts = 2;   // split-K of W = Ω ξ_c (K = M)
  // pinned host mailboxes
  Ctrl* ctrl_h = nullptr;
  double* J_h = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipEvent_t ev[2 * KC_N + 2];

  template <typename T> T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess)
      throw Error{ZH_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    allocs.push_back(p);
    HIPCHK(hipMemsetAsync(p, 0, count * sizeof(T), stream));
    return (T*)p;
  }

  Engine(const zh_dims& d)
      : NK(d.NK), n(d.n), NL(d.NL), NX(d.NX), NOT(d.NOT), hbar(d.hbar), device(d.device) {
    M = NL * NX * NOT;
    Ma = NX * NOT;
    std::memset(&pulse, 0, sizeof(pulse));
    pulse.hbar = hbar;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc((void**)&ctrl_h, sizeof(Ctrl), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&J_h, sizeof(double) * NL * NK, hipHostMallocDefault));
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    const size_t nn = (size_t)NK * n * n, mn = (size_t)NK * M * n, mm = (size_t)NK * M * M;
    for (int s = 0; s < 10; ++s) {
      sig[s] = dalloc<double2>(nn);
      if (s != 9) psi[s] = dalloc<cs>(mn);
      omg[s] = dalloc<cs>(mm);
    }
    lead = dalloc<int>(M);
    Hcur = dalloc<double2>(nn);
    G = dalloc<double2>(nn);
    dvec = dalloc<double2>((size_t)NK * n);
    glp = dalloc<double2>((size_t)NK * M);
    chip = dalloc<double2>((size_t)NK * M);
    chim = dalloc<double2>((size_t)NK * M);
    n0 = dalloc<double>((size_t)NK * n);
    te_part = dalloc<double>(3 * kTEBlocks);
    ctrl = dalloc<Ctrl>(1);
    sp = dalloc<StageParams>(7);
    Jd = dalloc<double>((size_t)NL * NK);
    Jpart = dalloc<double>((size_t)NL * NK * kCurChunks);
    std::vector<int> ld(M);
    for (int m = 0; m < M; ++m) ld[m] = m / Ma;
    HIPCHK(hipMemcpyAsync(lead, ld.data(), sizeof(int) * M, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  ~Engine() override {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graph)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (ctrl_h) (void)hipHostFree(ctrl_h);
    if (J_h) (void)hipHostFree(J_h);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graphs() {
    for (auto& g : graph)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  }

  // ---------------------------------------------------------------- set-up
  void set_constants(const double* H0, const double* xi_, const double* Ixi_, const double* glp_,
                     const double* dgp_, const double* dgm_, const double* xpp_,
                     const double* xpm_) override {
    HIPCHK(hipSetDevice(device));
    drop_graphs();
    const std::complex<double>* xi = (const std::complex<double>*)xi_;
    const std::complex<double>* Ixi = (const std::complex<double>*)Ixi_;
    const std::complex<double>* dgp = (const std::complex<double>*)dgp_;
    const std::complex<double>* dgm = (const std::complex<double>*)dgm_;
    // coupled orbitals: where any xi or Ixi entry is non-zero (the reference
    // prunes orbitals the same way, Zandpack/mpi/zand:589-602)
    std::vector<int> cp(n, -1), ci;
    for (int o = 0; o < n; ++o) {
      bool nz = false;
      for (long bm = 0; bm < (long)NK * M && !nz; ++bm)
        nz = (xi[bm * n + o] != 0.0) || (Ixi[bm * n + o] != 0.0);
      if (nz) { cp[o] = (int)ci.size(); ci.push_back(o); }
    }
    require(!have_consts || (int)ci.size() == nc, "coupled-orbital set changed; create a new problem");
    if (!have_consts) {
      nc = (int)ci.size();
      const size_t mc = (size_t)NK * M * nc;
      cidx = dalloc<int>(nc);
      cpos = dalloc<int>(n);
      Ixic = dalloc<double2>(mc);
      Pic = dalloc<double2>((size_t)NK * kPiParts * n * nc);
      xic = dalloc<cs>(mc);
      W = dalloc<cs>(mc * kWParts);
      Ap = dalloc<cs>((size_t)NK * M * (n + nc));
      Bp = dalloc<cs>((size_t)NK * (n + nc) * n);
      Aomg = dalloc<cs>((size_t)NK * M * 2 * nc);
      Bomg = dalloc<cs>((size_t)NK * M * 2 * nc);
    }
    HIPCHK(hipMemcpyAsync(cidx, ci.data(), sizeof(int) * std::max(nc, 1) * (nc > 0), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(cpos, cp.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
    const size_t mc = (size_t)NK * M * nc;
    std::vector<double2> ixic(mc);
    std::vector<cs> xc(mc), apr(mc), aor(mc), bol(mc);
    for (long bm = 0; bm < (long)NK * M; ++bm) {
      for (int c = 0; c < nc; ++c) {
        const std::complex<double> x = xi[bm * n + ci[c]], ix = Ixi[bm * n + ci[c]];
        ixic[bm * nc + c] = make_double2(ix.real(), ix.imag());
        xc[bm * nc + c] = from_d2<cs>(make_double2(x.real(), x.imag()));
        const std::complex<double> px = dgp[bm] * x, mx = dgm[bm] * ix;
        apr[bm * nc + c] = from_d2<cs>(make_double2(px.real(), px.imag()));
        bol[bm * nc + c] = from_d2<cs>(make_double2(mx.real(), mx.imag()));
      }
    }
    if (nc > 0) {
      HIPCHK(hipMemcpyAsync(Ixic, ixic.data(), sizeof(double2) * mc, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(xic, xc.data(), sizeof(cs) * mc, hipMemcpyHostToDevice, stream));
      // right block of Ap, right block of Aomg, left block of Bomg (constant columns)
      HIPCHK(hipMemcpy2DAsync(Ap + n, sizeof(cs) * (n + nc), apr.data(), sizeof(cs) * nc,
                              sizeof(cs) * nc, (size_t)NK * M, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpy2DAsync(Aomg + nc, sizeof(cs) * 2 * nc, apr.data(), sizeof(cs) * nc,
                              sizeof(cs) * nc, (size_t)NK * M, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpy2DAsync(Bomg, sizeof(cs) * 2 * nc, bol.data(), sizeof(cs) * nc,
                              sizeof(cs) * nc, (size_t)NK * M, hipMemcpyHostToDevice, stream));
    }
    HIPCHK(hipMemcpyAsync(glp, glp_, sizeof(double2) * NK * M, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(chip, xpp_, sizeof(double2) * NK * M, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(chim, xpm_, sizeof(double2) * NK * M, hipMemcpyHostToDevice, stream));
    H0h.assign(H0, H0 + 2 * (size_t)NK * n * n);
    install_h(H0);
    HIPCHK(hipStreamSynchronize(stream));
    h_dirty = false;
    have_consts = true;
  }

  void install_h(const double* H) {
    HIPCHK(hipMemcpyAsync(Hcur, H, sizeof(double2) * NK * n * n, hipMemcpyHostToDevice, stream));
    dim3 grid((n + 31) / 32, (n + 31) / 32, NK);
    hipLaunchKernelGGL(k_transpose_h<cs>, grid, dim3(32, 8), 0, stream, n,
                       (long)(n + nc) * n, (const double2*)Hcur, Bp);
    HIPCHK(hipGetLastError());
  }

  void set_hubbard(double U_, const double* n0_) override {
    require(U_ == 0.0 || NK == 2, "the Hubbard model couples two spin blocks: NK must be 2");
    drop_graphs();
    U = U_;
    if (U != 0.0) {
      HIPCHK(hipMemcpyAsync(n0, n0_, sizeof(double) * NK * n, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
  }

  void set_pulse(const PulseDesc& p) override {
    drop_graphs();
    pulse = p;
    pulse.hbar = hbar;
  }

  // ------------------------------------------------------------- transfers
  // host complex128 (rows x cols dense) -> device storage C (leading dim ld)
  template <typename C>
  void upload(C* dst, long ld, const double* src, long rows, long cols) {
    if (sizeof(C) == sizeof(double2)) {
      HIPCHK(hipMemcpy2DAsync(dst, sizeof(C) * ld, src, sizeof(double2) * cols,
                              sizeof(double2) * cols, rows, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    } else {
      std::vector<C> tmp((size_t)rows * cols);
      for (size_t i = 0; i < tmp.size(); ++i)
        tmp[i] = from_d2<C>(make_double2(src[2 * i], src[2 * i + 1]));
      HIPCHK(hipMemcpy2DAsync(dst, sizeof(C) * ld, tmp.data(), sizeof(C) * cols, sizeof(C) * cols,
                              rows, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
  }
  template <typename C>
  void download(double* dst, const C* src, size_t count) {
    if (sizeof(C) == sizeof(double2)) {
      HIPCHK(hipMemcpyAsync(dst, src, sizeof(double2) * count, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    } else {
      std::vector<C> tmp(count);
      HIPCHK(hipMemcpyAsync(tmp.data(), src, sizeof(C) * count, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      for (size_t i = 0; i < count; ++i) {
        dst[2 * i] = (double)tmp[i].x;
        dst[2 * i + 1] = (double)tmp[i].y;
      }
    }
  }

  void state_put(const double* s, const double* p, const double* o) override {
    HIPCHK(hipSetDevice(device));
    upload(sig[iy0], n, s, (long)NK * n, n);
    upload(psi[iy0], n, p, (long)NK * M, n);
    upload(omg[iy0], M, o, (long)NK * M, M);
    have_state = true;
  }
  void state_get(double* s, double* p, double* o) override {
    require(have_state, "no resident state", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    if (s) download(s, sig[iy0], (size_t)NK * n * n);
    if (p) download(p, psi[iy0], (size_t)NK * M * n);
    if (o) download(o, omg[iy0], (size_t)NK * M * M);
  }
  void sigma_get(double* s) override { state_get(s, nullptr, nullptr); }

  // ---------------------------------------------------------------- launches
  void mark(int kc, bool start) {
    if (profile) HIPCHK(hipEventRecord(ev[2 * kc + (start ? 0 : 1)], stream));
  }
  void accum(int kc) {
    if (!profile) return;
    float ms = 0.f;
    HIPCHK(hipEventSynchronize(ev[2 * kc + 1]));
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * kc], ev[2 * kc + 1]));
    kc_ms[kc] += ms;
    kc_count[kc] += 1;
  }
  // launch wrapper: brackets one kernel class with events when profiling
  template <typename F> void timed(int kc, F&& f) {
    mark(kc, true);
    f();
    mark(kc, false);
    if (profile) accum(kc);
  }

  void combine(int s, bool first, const double2* sig_dst, cs* omg_dst, bool last) {
    const int nt = first ? 0 : s;
    auto fill = [&](auto& g, auto* const* ks) {
      g.nterms = nt;
      for (int j = 0; j < nt; ++j) { g.k[j] = ks[j]; g.a[j] = tab.a[s][j]; }
    };
    const int blocks = 2048;
    timed(KC_COMBINE, [&] {
      if (!first) {
        CombineArgs<double2> gs{};
        gs.y0 = sig[iy0];
        fill(gs, sig);
        gs.d1 = (double2*)sig_dst; gs.ld1 = n; gs.d2 = nullptr; gs.ld2 = 0;
        gs.rows = (long)NK * n; gs.cols = n;
        hipLaunchKernelGGL(k_combine<double2>, dim3(blocks), dim3(256), 0, stream, gs);
        CombineArgs<cs> go{};
        go.y0 = omg[iy0];
        fill(go, omg);
        go.d1 = omg_dst; go.ld1 = M; go.d2 = nullptr; go.ld2 = 0;
        go.rows = (long)NK * M; go.cols = M;
        hipLaunchKernelGGL(k_combine<cs>, dim3(blocks), dim3(256), 0, stream, go);
      }
      CombineArgs<cs> gp{};
      gp.y0 = psi[iy0];
      fill(gp, psi);
      gp.d1 = Ap; gp.ld1 = n + nc;
      gp.d2 = last ? psi[iyn] : nullptr; gp.ld2 = n;
      gp.rows = (long)NK * M; gp.cols = n;
      hipLaunchKernelGGL(k_combine<cs>, dim3(blocks), dim3(256), 0, stream, gp);
    });
    HIPCHK(hipGetLastError());
  }

  // Right-hand side of stage s from (sig_in, Ap, om_in) into slot kslot.
  void rhs_stage(int s, const double2* sig_in, const cs* om_in, int kslot, bool currents,
                 bool hub) {
    const StageParams* sps = sp + s;
    const double ih = 1.0 / hbar;
    timed(KC_GATHER, [&] {
      GatherArgs<cs> g;
      g.NK = NK; g.n = n; g.M = M; g.nc = nc;
      g.Ap = Ap; g.ldAp = n + nc; g.sAp = (long)M * (n + nc);
      g.Aomg = Aomg; g.Bomg = Bomg; g.Bp = Bp; g.sig = sig_in; g.cidx = cidx;
      g.dvec = dvec; g.n0 = n0; g.U = U; g.hbar = hbar; g.hubbard = hub && U != 0.0;
      hipLaunchKernelGGL(k_gather<cs>, dim3(512, 3), dim3(256), 0, stream, g);
    });
    HIPCHK(hipGetLastError());
    if (nc > 0) {
      timed(KC_GEMM_PI, [&] {
        // Π_c = ψᵀ Ixi_c  (n x M) (M x nc), FP64
        GemmArgs<R, double> g{n, nc, M, Ap, n + nc, (long)M * (n + nc), Ixic, nc, (long)M * nc,
                              kPiParts};
        EpiStore<double2> e{Pic, nc, (long)n * nc, 1.0, 0.0};
        HIPCHK((launch_zgemm<double, R, double, true, false, 64, 32, 16>(g, NK, e, stream)));
      });
    }
    if (currents) {
      // with no coupled orbitals the reductions run over empty ranges and give J = 0
      timed(KC_CURRENTS, [&] {
        hipLaunchKernelGGL(k_currents<cs>, dim3(kCurChunks, NL, NK), dim3(256), 0, stream, NK, M,
                           Ma, nc, (const cs*)Aomg, (const double2*)Ixic, Jpart);
        hipLaunchKernelGGL(k_currents_final, dim3(1), dim3(((NL * NK + 63) / 64) * 64), 0, stream,
                           NL * NK, (const double*)Jpart, ih * ih, Jd);
      });
      HIPCHK(hipGetLastError());
    }
    timed(KC_GEMM_HSIG, [&] {
      GemmArgs<double, double> g{n, n, n, Hcur, n, (long)n * n, sig_in, n, (long)n * n};
      EpiSigma e{G, sig_in, dvec, Pic, cpos, n, nc, ih, kPiParts};
      HIPCHK((launch_zgemm<double, double, double, false, false, 64, 64, 16>(g, NK, e, stream)));
    });
    timed(KC_SIGSYM, [&] {
      dim3 grid((n + 31) / 32, (n + 31) / 32, NK);
      hipLaunchKernelGGL(k_sigma_sym, grid, dim3(32, 8), 0, stream, n, (const double2*)G,
                         sig[kslot], sps, ih);
    });
    HIPCHK(hipGetLastError());
    if (nc > 0) {
      timed(KC_GEMM_OMGXI, [&] {
        // W = Ω ξ_c  (M x M)(M x nc)
        GemmArgs<R, R> g{M, nc, M, om_in, M, (long)M * M, xic, nc, (long)M * nc, kWParts};
        EpiStore<cs> e{W, nc, (long)M * nc, 1.0, 0.0};
        HIPCHK((launch_zgemm<R, R, R, false, false, 64, 32, 16>(g, NK, e, stream)));
      });
    }
    timed(KC_GEMM_PSI, [&] {
      // [ψ | ΔPξ_c] [H0ᵀ ; ħ σ_cᵀ]  (M x (n+nc)) ((n+nc) x n)
      GemmArgs<R, R> g{M, n, n + nc, Ap, n + nc, (long)M * (n + nc), Bp, n, (long)(n + nc) * n};
      EpiPsi<cs> e{psi[kslot], Ap, n + nc, xic, W, glp, chip, dvec, lead, cpos, sps, M, n, nc,
                   ih, ih * ih, kWParts};
      HIPCHK((launch_zgemm<R, R, R, false, false, 64, 64, 16>(g, NK, e, stream)));
    });
    timed(KC_GEMM_OMEGA, [&] {
      // [ψ_c | ΔPξ_c] [ΔM Ixi_c | ψ_c*]ᵀ  (M x 2nc)(2nc x M), outer-subtraction epilogue
      EpiOmega<cs> e{omg[kslot], om_in, chip, chim, lead, sps, M, ih};
      if (nc > 0) {
        GemmArgs<R, R> g{M, M, 2 * nc, Aomg, 2 * nc, (long)M * 2 * nc, Bomg, 2 * nc,
                         (long)M * 2 * nc};
        HIPCHK((launch_zgemm<R, R, R, false, true, 64, 64, 16>(g, NK, e, stream)));
      } else {
        GemmArgs<R, R> g{M, M, 0, Aomg, 1, 0, Bomg, 1, 0};
        HIPCHK((launch_zgemm<R, R, R, false, true, 64, 64, 16>(g, NK, e, stream)));
      }
    });
  }

  void te_and_control() {
    timed(KC_TE, [&] {
      TEArgs<double2> ts{};
      TEArgs<cs> tp{}, to{};
      for (int j = 0; j < 7; ++j) {
        ts.k[j] = sig[j]; tp.k[j] = psi[j]; to.k[j] = omg[j];
        ts.ct[j] = tp.ct[j] = to.ct[j] = tab.ct[j];
      }
      ts.count = (long)NK * n * n;
      tp.count = (long)NK * M * n;
      to.count = (long)NK * M * M;
      hipLaunchKernelGGL(k_te<double2>, dim3(kTEBlocks), dim3(256), 0, stream, ts, te_part);
      hipLaunchKernelGGL(k_te<cs>, dim3(kTEBlocks), dim3(256), 0, stream, tp, te_part + kTEBlocks);
      hipLaunchKernelGGL(k_te<cs>, dim3(kTEBlocks), dim3(256), 0, stream, to,
                         te_part + 2 * kTEBlocks);
    });
    HIPCHK(hipGetLastError());
    timed(KC_CTRL, [&] {
      hipLaunchKernelGGL(k_controller, dim3(1), dim3(256), 0, stream, (const double*)te_part,
                         3 * kTEBlocks, ctrl);
    });
    HIPCHK(hipGetLastError());
  }

  // The seven stages, error norm and controller of one attempt (graph body).
  void attempt_body() {
    timed(KC_PARAMS, [&] {
      hipLaunchKernelGGL(k_stage_params, dim3(1), dim3(1), 0, stream, (const Ctrl*)ctrl, pulse,
                         tab, NL, sp);
    });
    HIPCHK(hipGetLastError());
    for (int s = 0; s < 7; ++s) {
      const bool last = (s == 6);
      double2* sdst = last ? sig[iyn] : sig[9];
      cs* odst = last ? omg[iyn] : omg[9];
      combine(s, s == 0, sdst, odst, last);
      const double2* sin_ = (s == 0) ? sig[iy0] : sdst;
      const cs* oin = (s == 0) ? omg[iy0] : odst;
      rhs_stage(s, sin_, oin, s, s == 0, true);
    }
    te_and_control();
  }

  void ensure_h0() {
    if (h_dirty) {
      drop_graphs();
      install_h(H0h.data());
      HIPCHK(hipStreamSynchronize(stream));
      h_dirty = false;
    }
  }

  void attempt(zh_attempt_info* info, double* J) override {
    require(have_consts, "constants not set", ZH_ESTATE);
    require(have_state, "no resident state", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    ensure_h0();
    ctrl_h->t0 = info->t0;
    ctrl_h->h = info->h;
    ctrl_h->eps = info->eps;
    ctrl_h->force = info->force_accept ? 1 : 0;
    HIPCHK(hipMemcpyAsync(ctrl, ctrl_h, sizeof(Ctrl), hipMemcpyHostToDevice, stream));
    HIPCHK(hipEventRecord(ev[2 * KC_N], stream));
    const int par = (iy0 == 7) ? 0 : 1;
    if (use_graph && !profile) {
      if (!graph[par]) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        attempt_body();
        HIPCHK(hipStreamEndCapture(stream, &g));
        HIPCHK(hipGraphInstantiate(&graph[par], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
      }
      HIPCHK(hipGraphLaunch(graph[par], stream));
    } else {
      attempt_body();
    }
    HIPCHK(hipEventRecord(ev[2 * KC_N + 1], stream));
    HIPCHK(hipMemcpyAsync(ctrl_h, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(J_h, Jd, sizeof(double) * NL * NK, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * KC_N], ev[2 * KC_N + 1]));
    attempt_ms += ms;
    attempt_count += 1;
    info->te = ctrl_h->te;
    info->h_next = ctrl_h->h_next;
    info->accepted = ctrl_h->accept;
    if (J) std::memcpy(J, J_h, sizeof(double) * NL * NK);
    if (info->accepted) std::swap(iy0, iyn);
  }

  void rhs(const double* Ht, const double* delta, const double* s, const double* p,
           const double* o, double* ds, double* dp, double* dom, double* J) override {
    require(have_consts, "constants not set", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    install_h(Ht);
    h_dirty = true;
    StageParams st;
    std::memset(&st, 0, sizeof(st));
    st.ts = 0.0;
    st.h = 1.0;
    for (int a = 0; a < NL; ++a) st.delta[a] = delta[a];
    HIPCHK(hipMemcpyAsync(sp, &st, sizeof(st), hipMemcpyHostToDevice, stream));
    upload(sig[9], n, s, (long)NK * n, n);
    upload(Ap, n + nc, p, (long)NK * M, n);
    upload(omg[9], M, o, (long)NK * M, M);
    rhs_stage(0, sig[9], omg[9], 0, J != nullptr, false);
    HIPCHK(hipStreamSynchronize(stream));
    if (ds) download(ds, sig[0], (size_t)NK * n * n);
    if (dp) download(dp, psi[0], (size_t)NK * M * n);
    if (dom) download(dom, omg[0], (size_t)NK * M * M);
    if (J) {
      HIPCHK(hipMemcpy(J, Jd, sizeof(double) * NL * NK, hipMemcpyDeviceToHost));
    }
  }
};

// ----------------------------------------------------------- GEMM self-test
template <typename R, bool TA, bool TB>
static void selftest_impl(int M, int N, int K, const double* A, const double* B, double* C) {
  typedef typename Cplx<R>::type cs;
  const size_t na = (size_t)M * K, nb = (size_t)K * N, ncn = (size_t)M * N;
  std::vector<cs> ha(na), hb(nb);
  for (size_t i = 0; i < na; ++i) ha[i] = from_d2<cs>(make_double2(A[2 * i], A[2 * i + 1]));
  for (size_t i = 0; i < nb; ++i) hb[i] = from_d2<cs>(make_double2(B[2 * i], B[2 * i + 1]));
  cs *da = nullptr, *db = nullptr;
  double2* dc = nullptr;
  HIPCHK(hipMalloc(&da, sizeof(cs) * std::max<size_t>(na, 1)));
  HIPCHK(hipMalloc(&db, sizeof(cs) * std::max<size_t>(nb, 1)));
  HIPCHK(hipMalloc(&dc, sizeof(double2) * std::max<size_t>(ncn, 1)));
  HIPCHK(hipMemcpy(da, ha.data(), sizeof(cs) * na, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(db, hb.data(), sizeof(cs) * nb, hipMemcpyHostToDevice));
  GemmArgs<R, R> g{M, N, K, da, TA ? M : K, 0, db, TB ? K : N, 0};
  EpiStore<double2> e{dc, N, 0, 1.0, 0.0};
  HIPCHK((launch_zgemm<R, R, R, TA, TB, 64, 64, 16>(g, 1, e, nullptr)));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(C, dc, sizeof(double2) * ncn, hipMemcpyDeviceToHost));
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dc);
}

}  // namespace zh

using namespace zh;

struct zh_problem {
  std::unique_ptr<EngineBase> eng;
  zh_dims dims;
};

template <typename F>
static int zh_try(F&& f) {
  try {
    f();
    return ZH_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return ZH_EINVAL;
  }
}

extern "C" {

int zh_abi_version(void) { return ZH_ABI_VERSION; }
const char* zh_last_error(void) { return g_err.c_str(); }

int zh_problem_create(const zh_dims* d, zh_problem** out) {
  return zh_try([&] {
    require(d && out, "null argument");
    require(d->NK > 0 && d->n > 0 && d->NL > 0 && d->NX > 0 && d->NOT > 0, "dimensions must be positive");
    require(d->NL <= ZH_MAX_LEADS, "too many leads");
    require(d->precision == ZH_FP64 || d->precision == ZH_MIXED, "unknown precision");
    require(d->hbar > 0, "hbar must be positive");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    require(d->device >= 0 && d->device < ndev, "no such HIP device");
    auto* p = new zh_problem();
    p->dims = *d;
    try {
      if (d->precision == ZH_FP64) p->eng.reset(new Engine<double>(*d));
      else p->eng.reset(new Engine<float>(*d));
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

void zh_problem_destroy(zh_problem* p) { delete p; }

int zh_set_constants(zh_problem* p, const double* H0, const double* xi, const double* Ixi,
                     const double* GL_P, const double* dGP, const double* dGM, const double* Xpp,
                     const double* Xpm) {
  return zh_try([&] {
    require(p && H0 && xi && Ixi && GL_P && dGP && dGM && Xpp && Xpm, "null argument");
    p->eng->set_constants(H0, xi, Ixi, GL_P, dGP, dGM, Xpp, Xpm);
  });
}

int zh_set_hubbard(zh_problem* p, double U, const double* n0) {
  return zh_try([&] {
    require(p && (U == 0.0 || n0), "null argument");
    p->eng->set_hubbard(U, n0);
  });
}

int zh_set_pulse(zh_problem* p, int kind, double amp, double p1, double p2, const double* w) {
  return zh_try([&] {
    require(p && w, "null argument");
    require(kind >= ZH_PULSE_NONE && kind <= ZH_PULSE_BOX, "unknown pulse kind");
    require(!((kind == ZH_PULSE_GAUSS || kind == ZH_PULSE_BOX) && p2 == 0.0), "zero pulse width");
    require(!(kind == ZH_PULSE_RAMP && p1 == 0.0), "zero ramp time");
    PulseDesc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = kind; d.amp = amp; d.p1 = p1; d.p2 = p2;
    for (int a = 0; a < p->dims.NL; ++a) d.w[a] = w[a];
    p->eng->set_pulse(d);
  });
}

int zh_state_put(zh_problem* p, const double* sig, const double* psi, const double* omg) {
  return zh_try([&] {
    require(p && sig && psi && omg, "null argument");
    p->eng->state_put(sig, psi, omg);
  });
}

int zh_state_get(zh_problem* p, double* sig, double* psi, double* omg) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->state_get(sig, psi, omg);
  });
}

int zh_sigma_get(zh_problem* p, double* sig) {
  return zh_try([&] {
    require(p && sig, "null argument");
    p->eng->sigma_get(sig);
  });
}

int zh_rhs(zh_problem* p, const double* Ht, const double* delta, const double* sig,
           const double* psi, const double* omg, double* dsig, double* dpsi, double* domg,
           double* J) {
  return zh_try([&] {
    require(p && Ht && delta && sig && psi && omg, "null argument");
    p->eng->rhs(Ht, delta, sig, psi, omg, dsig, dpsi, domg, J);
  });
}

int zh_attempt(zh_problem* p, zh_attempt_info* info, double* J) {
  return zh_try([&] {
    require(p && info, "null argument");
    require(info->h > 0 && std::isfinite(info->h), "step size must be positive and finite");
    p->eng->attempt(info, J);
    if (!std::isfinite(info->te)) throw Error{ZH_ENONFINITE, "non-finite error estimate"};
  });
}

int zh_propagate(zh_problem* p, double t0, double t1, double h0, double eps, int time_update,
                 long max_attempts, long cap, double* times, double* J, double* h_used,
                 long* n_steps, long* n_rejected, double* h_final) {
  return zh_try([&] {
    require(p && n_steps && n_rejected, "null argument");
    require(h0 > 0 && eps > 0, "h0 and eps must be positive");
    const int NLK = p->dims.NL * p->dims.NK;
    std::vector<double> Jtmp(NLK);
    double t = t0, h = h0;
    long steps = 0, rej = 0, att = 0;
    *n_steps = 0;
    *n_rejected = 0;
    while (t < t1) {
      if (att >= max_attempts) {
        *n_steps = steps; *n_rejected = rej; if (h_final) *h_final = h;
        throw Error{ZH_ELIMIT, "attempt limit reached"};
      }
      zh_attempt_info info;
      std::memset(&info, 0, sizeof(info));
      info.t0 = t; info.h = h; info.eps = eps;
      p->eng->attempt(&info, Jtmp.data());
      ++att;
      if (!std::isfinite(info.te) || !std::isfinite(info.h_next)) {
        *n_steps = steps; *n_rejected = rej; if (h_final) *h_final = h;
        throw Error{ZH_ENONFINITE, "non-finite error estimate at t=" + std::to_string(t)};
      }
      const double h_old = h;
      h = info.h_next;
      if (!info.accepted) { ++rej; continue; }
      t += (time_update == ZH_TIME_REFERENCE) ? h : h_old;
      if (steps < cap) {
        if (times) times[steps] = t;
        if (h_used) h_used[steps] = h_old;
        if (J) std::memcpy(J + steps * NLK, Jtmp.data(), sizeof(double) * NLK);
      }
      ++steps;
    }
    *n_steps = steps;
    *n_rejected = rej;
    if (h_final) *h_final = h;
  });
}

int zh_set_graph(zh_problem* p, int enable) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->use_graph = enable != 0;
  });
}

int zh_set_profile(zh_problem* p, int enable) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->profile = enable != 0;
  });
}

int zh_timing(zh_problem* p, int reset, int cap, char* names, double* ms, long* counts) {
  if (!p) { g_err = "null argument"; return ZH_EINVAL; }
  EngineBase* e = p->eng.get();
  int k = 0;
  if (cap > 0 && names && ms && counts) {
    // entry 0: whole attempt; then kernel classes
    std::snprintf(names, 32, "%s", "attempt");
    ms[0] = e->attempt_count ? e->attempt_ms / e->attempt_count : 0.0;
    counts[0] = e->attempt_count;
    k = 1;
    for (int c = 0; c < KC_N && k < cap; ++c, ++k) {
      std::snprintf(names + 32 * k, 32, "%s", kc_names[c]);
      ms[k] = e->kc_count[c] ? e->kc_ms[c] / e->kc_count[c] : 0.0;
      counts[k] = e->kc_count[c];
    }
  }
  if (reset) {
    for (int c = 0; c < KC_N; ++c) { e->kc_ms[c] = 0; e->kc_count[c] = 0; }
    e->attempt_ms = 0; e->attempt_count = 0;
  }
  return k;
}

int zh_selftest_gemm(int precision, int M, int N, int K, int transA, int transB, const double* A,
                     const double* B, double* C) {
  return zh_try([&] {
    require(A && B && C && M > 0 && N > 0 && K >= 0, "bad argument");
    if (precision == ZH_FP64) {
      if (!transA && !transB) selftest_impl<double, false, false>(M, N, K, A, B, C);
      else if (!transA && transB) selftest_impl<double, false, true>(M, N, K, A, B, C);
      else if (transA && !transB) selftest_impl<double, true, false>(M, N, K, A, B, C);
      else selftest_impl<double, true, true>(M, N, K, A, B, C);
    } else {
      if (!transA && !transB) selftest_impl<float, false, false>(M, N, K, A, B, C);
      else if (!transA && transB) selftest_impl<float, false, true>(M, N, K, A, B, C);
      else if (transA && !transB) selftest_impl<float, true, false>(M, N, K, A, B, C);
      else selftest_impl<float, true, true>(M, N, K, A, B, C);
    }
  });
}

}  // extern "C"
