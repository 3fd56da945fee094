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
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <array>
#include <string>
#include <vector>

#include "../../include/zandhip.h"
#include "zh_gemm.hpp"
#include "zh_kernels.hpp"
#include "zh_scf.hpp"
#include "zh_omega.hpp"
#include "zh_omgxi.hpp"
#include "zh_psi.hpp"

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
              KC_GEMM_OMGXI, KC_GEMM_PSI, KC_GEMM_OMEGA, KC_TE, KC_CTRL, KC_WSUM, KC_N };
static const char* kc_names[KC_N] = {"stage_params", "combine",   "gather",     "gemm_pi",
                                      "currents",     "gemm_hsig", "sigma_sym",  "gemm_omgxi",
                                      "gemm_psi",     "gemm_omega", "te_reduce", "controller", "wsum"};

struct EngineBase {
  virtual ~EngineBase() {}
  virtual void set_constants(const double* H0, const double* xi, const double* Ixi,
                             const double* glp, const double* dgp, const double* dgm,
                             const double* xpp, const double* xpm) = 0;
  virtual void set_hubbard(double U, const double* n0) = 0;
  virtual void set_drive(int nt, const int* kinds, const double* amps, const double* par,
                         const int* nnz, const int* ks, const int* rows, const int* cols,
                         const double* vals) = 0;
  virtual void set_pulse(const PulseDesc& p) = 0;
  virtual void set_instance_pulse(int g, const PulseDesc& p) = 0;
  virtual int instances() const = 0;
  virtual void set_pulse_spline(int nknots, const double* x, const double* c) = 0;
  virtual void set_dissipator(const DissDesc& d, const double* sig_eq) = 0;
  virtual void set_overlap(const double* Sinv) = 0;
  virtual void set_tolerance(double rtol, double atol) = 0;
  virtual void pi_get(int hermitian, double* Q) = 0;
  virtual void state_put(const double* sig, const double* psi, const double* omg) = 0;
  virtual void state_get(double* sig, double* psi, double* omg) = 0;
  virtual void sigma_get(double* sig) = 0;
  virtual void steady_state(double t0, double tol, int maxiter, const zh_precond* pc,
                            int* iters, double* resid) = 0;
  virtual void rhs(const double* Ht, const double* delta, const double* sig, const double* psi,
                   const double* omg, double* dsig, double* dpsi, double* domg, double* J) = 0;
  // finite_next: an accepted attempt whose next step size is not finite (TE = 0 or
  // non-finite) is not committed and raises ZH_ENONFINITE
  virtual void attempt(zh_attempt_info* info, double* J, bool finite_next = false) = 0;
  virtual void launch(const zh_attempt_info* info) = 0;
  virtual void collect(zh_attempt_info* info, double* J, bool finite_next = false) = 0;
  virtual int ready() = 0;
  virtual long bytes() const = 0;
  virtual void autotune(int force, int* out) = 0;
  virtual void set_tiles(const int* sel) = 0;
  virtual void set_layout(int mode) = 0;
  virtual int layout_info(int* out) = 0;
  virtual void set_sparse(int mode) = 0;
  virtual void hamiltonian_info(int* out) = 0;
  virtual void set_fusion(int mode) = 0;
  virtual void set_antihermitian(int mode) = 0;
  virtual int antihermitian_info() const = 0;
  virtual void set_omgxi(int mode) = 0;
  virtual void set_psi_kernel(int mode) = 0;
  virtual int psi_kernel_info() const = 0;
  virtual void set_fsal(int mode) = 0;
  virtual void fsal_info(long* taken, int* on) const = 0;
  virtual void omgxi_info(int* out) const = 0;
  bool use_graph = true;
  bool profile = false;
  double kc_ms[KC_N] = {0};
  long kc_count[KC_N] = {0};
  double attempt_ms = 0.0;
  long attempt_count = 0;
};

// Candidate tiles per product (index 0 is the default).  The autotuner
// (zh_autotune) times them on the problem's own resident data and keeps the
// fastest per product; choices are cached per (precision, size bucket).
enum TuneSlot { TS_PSI, TS_HSIG, TS_OMEGA, TS_OMGXI, TS_PI, TS_N };
constexpr int kCands = 3;
template <typename R> struct Cands;
template <> struct Cands<double> {
  // FP64 tiles carry occupancy hints (waves per SIMD): without them the compiler keeps
  // the accumulators in AGPRs beside ~120-190 VGPRs and runs 1-2 waves per SIMD,
  // too few to hide the epilogue's and the next tile's memory latency.
  typedef TileList<Tile<64, 64, 8, 2, false, 3>, Tile<64, 64, 16, 2, false, 3>,
                   Tile<64, 32, 16, 2, false, 4>> psi;
  typedef TileList<Tile<64, 64, 8, 2, false, 3>, Tile<64, 64, 16, 2, false, 3>,
                   Tile<64, 64, 16, 2, false, 2>> psiR;
  typedef TileList<Tile<64, 64, 8, 2, false, 3>, Tile<64, 32, 8, 2, false, 4>,
                   Tile<32, 32, 8, 2, false, 4>> omega;
  typedef TileList<Tile<64, 48, 16, 4, false, 3>, Tile<64, 32, 16, 2, false, 3>,
                   Tile<64, 32, 32, 2, false, 3>> omgxi;
  // X ξ_c of an anti-Hermitian X held as its triangle (mirrored first K segment)
  typedef TileList<Tile<64, 32, 16, 2, false, 3>, Tile<64, 48, 16, 4, false, 2>,
                   Tile<64, 32, 8, 2, false, 3>> omgxiT;
};
template <> struct Cands<float> {
  typedef TileList<Tile<64, 64, 32, 2>, Tile<64, 64, 16, 2>, Tile<128, 64, 16, 2>> psi;
  typedef TileList<Tile<64, 64, 32, 2>, Tile<64, 64, 16, 2>, Tile<128, 128, 16, 2>> psiR;
  typedef TileList<Tile<64, 64, 8, 2>, Tile<64, 32, 8, 2>, Tile<32, 32, 8, 2>> omega;
  typedef TileList<Tile<64, 48, 16, 4>, Tile<64, 48, 32, 4>, Tile<64, 32, 16, 2>> omgxi;
  typedef TileList<Tile<64, 48, 16, 4>, Tile<64, 32, 16, 2>, Tile<64, 32, 8, 2>> omgxiT;
};
// σ-path products are FP64 in both precision modes
typedef TileList<Tile<64, 32, 16, 2, false, 3>, Tile<64, 64, 8, 2, false, 3>,
                 Tile<64, 64, 16, 2, false, 3>> CandHsig;
typedef TileList<Tile<64, 32, 16, 2, false, 3>, Tile<64, 64, 16, 2, false, 3>,
                 Tile<64, 64, 8, 2, false, 4>> CandHsigR;
typedef TileList<Tile<64, 48, 16, 4, false, 3>, Tile<64, 32, 16, 2, false, 3>,
                 Tile<32, 32, 16, 2, false, 4>> CandPi;

struct TuneKey {
  int prec, nb, mb, cb, hr, tri;
  bool operator<(const TuneKey& o) const {
    return std::tie(prec, nb, mb, cb, hr, tri) < std::tie(o.prec, o.nb, o.mb, o.cb, o.hr, o.tri);
  }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, std::array<int, TS_N + 2>> g_tune_cache;  // + fused Ω, streaming Ωξ
static int pow2_bucket(int v) {
  int b = 1;
  while (b < v) b <<= 1;
  return b;
}

// Scratch device buffers of one call, freed when it returns (also on error).
struct DevBuf {
  std::vector<void*> p;
  template <typename T> T* get(size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
      throw Error{ZH_ENOMEM, "hipMalloc failed for a scratch buffer"};
    p.push_back(q);
    return (T*)q;
  }
  ~DevBuf() {
    for (void* q : p) (void)hipFree(q);
  }
};

template <typename R>  // storage real type of Psi / Omega and compute type of their GEMMs
struct Engine : EngineBase {
  typedef typename Cplx<R>::type cs;
  static constexpr int kTile = 64;    // Ω GEMM / Ωξ GEMM row and column tile
  static constexpr int kPiParts = 8;  // split-K of Π_c = ψᵀ Ixi_c (K = modes)
  static constexpr int kWParts = 4;   // split-K of W = Ω ξ_c (K = modes)
  static constexpr int kEllMax = 64;  // most entries per row of the row-compressed H
  // complex FP64 products on the matrix cores in the three-multiplication form (three
  // real MFMAs per tile product); the FP32 mixed mode keeps four, since the form's
  // larger rounding reaches its error norm
  static constexpr int kRM3 = std::is_same<R, double>::value ? 3 : 0;
  static constexpr int kCGroup = 16;  // coupled-orbital column group (lead-aligned)
  int NK, n, NL, NX, NOT;
  // Instances: NI independent problems of KI k blocks each stacked on the k axis (NK = NI
  // KI), sharing the device mode order and every launch; each has its own control block
  // (t, h, eps), pulse, stage parameters, error norm and acceptance.
  int NI = 1, KI = 1;
  int M;        // modes per k in the reference layout, NL*NX*NOT
  int Md = 0;   // modes per k on the device (active block padded to kTile)
  int ML = 0;   // active block [0, ML); passive modes in [ML, Md)
  int nc = 0;   // coupled-orbital columns (lead groups padded to kCGroup when NL > 1)
  int nc_real = 0;  // coupled orbitals
  double hbar;
  int device;
  hipStream_t stream = nullptr;
  // Within a stage the Ω branch (its GEMM), the Ψ branch (Ωξ, Ψ right-hand side) and
  // the σ branch (Π, currents, Hσ) are independent once the stage input is gathered:
  // they run on three streams (fork/join events, also inside the captured graph), so
  // MFMA-heavy and HBM-heavy kernels share the CUs.  All streams at default priority.
  // the Ω and Ψ branches of a stage (they take turns, rhs_stage)
  hipStream_t aux[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  std::vector<void*> allocs;
  Tableau tab = dop54();
  std::vector<PulseDesc> pulse;  // per instance
  PulseDesc* pulse_d = nullptr;  // device copies read by k_stage_params
  bool pulse_dirty = true;
  DissDesc diss{ZH_DISS_NONE, 0.0, 0.0, 1.0};
  double2* sig_eq = nullptr;  // σ_eq of the dissipator (allocated on first use)
  // Drive terms H(t) = H0 + Σ_i f_i(t) M_i: the distinct entries (k, row, col) of the M_i,
  // each with its H0 value and its terms' coefficients (CSR), written into the H
  // operands every stage (k_drive_update)
  DriveDesc drive{};
  std::vector<int3> de_h;                 // (k, row, col)
  std::vector<int> de_beg_h, de_term_h;
  std::vector<double2> de_coef_h;
  std::vector<int> de_slot_h;
  int *de_k = nullptr, *de_r = nullptr, *de_c = nullptr, *de_slot = nullptr, *de_beg = nullptr,
      *de_term = nullptr;
  double2 *de_h0 = nullptr, *de_coef = nullptr;
  size_t de_cap = 0, dt_cap = 0;
  // Overlap with a σ- or t-dependent H (Hubbard or drive): S⁻¹H0 and S⁻¹M_i (dense),
  // S⁻¹H(t, σ) formed per stage; the epilogues then see no separate Hubbard diagonal
  double2 *SH0 = nullptr, *SM = nullptr, *dzero = nullptr;
  int sm_cap = -1;
  bool model_dirty = true;
  bool ovl_model() const { return overlap && (U != 0.0 || drive.nt > 0); }
  double U = 0.0;
  bool have_consts = false, have_state = false, h_dirty = false;
  bool h_real = false;     // the Hamiltonian in Hcur / Bp has no imaginary part
  bool xi_real = false;    // ξ and Ixi on the coupled orbitals are real
  bool pruned = false;     // active-first device mode order in use
  bool ff = false;         // passive x passive block of Ω is zero and skipped
  bool ff_dirty = false;   // that block may hold stale values in some slot
  int layout_req = 0;      // 0 auto, 1 reference mode order, 2 active-first (pruned)
  int sparse_req = 0;      // 0 auto, 1 dense H products, 2 row-compressed H products
  bool h_sparse = false;   // the H products run as gathers over the row-compressed H
  int ell_R = 0;           // entries per row of the row-compressed H
  int hpsi_rb = 0;         // mode rows per workgroup of the sparse ψHᵀ kernel
  // one-pass sparse Ψ right-hand side (zh_psi.hpp): per 64-orbital tile the orbitals its
  // H rows reach, [x, x + y); used when the widest fits kPsiFusedWMax
  int2* psi_span = nullptr;
  int psi_wmax = 0;
  int psi_req = 0;         // 0 auto (one pass when it fits), 1 split, 2 one pass
  bool psi_fused() const {
    return h_sparse && !overlap && psi_wmax > 0 && psi_wmax <= kPsiFusedWMax && psi_req != 1;
  }
  std::vector<double> H0h; // host copy of H0 (re-installed after an explicit rhs call)
  std::vector<int> perm;   // device mode -> reference mode (-1: inert padding mode)
  // Non-orthogonal basis (the nozand form): S⁻¹ multiplies H in both H products and
  // carries the constant and Ωξ terms of dΨ and Π of dσ from the coupled orbitals to
  // every orbital.  H products then run dense with the complex S⁻¹H.
  bool overlap = false;
  double2* Sdev = nullptr;    // S⁻¹ (NK, n, n)
  double2* Sc = nullptr;      // S⁻¹[c, :] (NK, nc, n): B operand of Π_c S⁻¹[c, :]
  cs* Bs = nullptr;           // S⁻¹[:, c]ᵀ (NK, nc, n): B operand of the Ψ term
  double2* Pis = nullptr;     // Π_c with its split-K partials summed (NK, n, nc)
  double2* Qs = nullptr;      // Π_c S⁻¹[c, :] (NK, n, n)
  cs* Yov = nullptr;          // ħ GL_P ξ_c + Ω_s ξ_c / ħ (NK, Md, nc)
  int* cpos_none = nullptr;   // (n) all -1: the coupled-column epilogue terms off
  std::vector<int> cidx_h; // coupled orbitals in device column order
  std::vector<int4> lrange_h;  // per lead: device modes in the active and passive blocks

  // device buffers
  int *cidx = nullptr, *cpos = nullptr, *lead = nullptr;
  int2* t1 = nullptr;      // per Ω column tile: term1 K range
  int2* t2 = nullptr;      // per Ω row tile: term2 K range (relative to nc)
  int2* txi = nullptr;     // per Ω column tile: ξ_c column range of its modes' leads
  int* gmask = nullptr;    // per 16-column group of the coupled order: lead mask
  int4* lrange = nullptr;  // per lead: device rows in the active and passive blocks
  double2 *Hcur = nullptr, *Ixic = nullptr, *Pic = nullptr, *G = nullptr, *dvec = nullptr;
  double* Hre = nullptr;  // Re H (A operand of Hσ when H is real)
  int* ell_col = nullptr;     // row-compressed H: (NK, kEllMax, n) column indices
  double2* ell_val = nullptr; // and values
  std::vector<int> ell_col_h;
  std::vector<double2> ell_val_h;
  R* HreT = nullptr;      // Re Hᵀ (B operand of ψHᵀ when H is real)
  double2 *glp = nullptr, *chip = nullptr, *chim = nullptr;
  double2 *uvec = nullptr, *vvec = nullptr;  // per-stage (h/ħ)(χ± + Δ) of the Ω epilogue
  double* n0 = nullptr;
  cs *Bp = nullptr, *Ap = nullptr, *Aomg = nullptr, *Bomg = nullptr, *xic = nullptr, *W = nullptr;
  // the Ψ stage input: two buffers, stage s of an attempt in apb(s) (Ap points at the
  // current one; Ap0 outside attempts), so that a stage's Ψ kernel can form the next
  // stage's input while its own is still read
  cs *Ap0 = nullptr, *Ap1 = nullptr;
  cs* apb(int s) const { return (s & 1) ? Ap1 : Ap0; }
  bool psi_nx = false;  // attempt body being issued: the one-pass Ψ kernel of stage s >= 1
                        // forms stage s + 1's Ψ input
  // attempt body being issued: the last stage's input pass also leaves the error norm's
  // weighted sum of k_1 .. k_6 (E6) in the free k_2 slot, so that the error-norm pass
  // reads E6 and k_7 only (σ and Ψ; for complex64 Ψ, E6 rounded to complex64 moved the
  // norm no further from the oracle's than the six slots do, over 36 attempts of four
  // problems at steps 0.005-0.02)
  bool e6_sig = false, e6_psi = false;
  // with E6 the last stage's σ / Ψ derivative k_7 goes straight over the k_1 slot (no
  // later pass reads k_1), as the next attempt's k_1 (first same as last)
  bool k7_sig0 = false, k7_psi0 = false;
  bool e6_ok() const {
    bool ok = tab.a[6][1] == 0.0 && tab.ct[1] == 0.0 && tab.ct[6] != 0.0;
    for (int j = 0; j < 6; ++j) ok = ok && ((tab.a[6][j] != 0.0) == (tab.ct[j] != 0.0));
    return ok;
  }
  cs* W2 = nullptr;  // second buffer of the Ωξ partials: stage s's product goes to wpart(s)
  cs* wpart(int s) const { return (s & 1) ? W2 : W; }
  bool xi_early = false;  // attempt body being issued: each stage's product issued early
  // attempt body being issued: the last stage issues the σ and Ψ error-norm partials on
  // its σ and Ψ branches' streams (te_launch), the attempt ends with the controller alone
  bool te_stage = false;
  cs* Pred[7] = {nullptr};  // y0 ξ_c and k_j ξ_c of the current attempt (Ω_s ξ_c by linearity)
  cs* Wst = nullptr;        // Ω_s ξ_c of the current stage
  // Fused Ω stage kernel (mixed precision): each stage's Ω kernel also writes k_s ξ_c
  // (the last stage y_new ξ_c) as per-chunk partials, double-buffered by stage parity;
  // PredY[par] holds y0 ξ_c (par: which of the two y0 / y_new slots is y0), PredY[1-par]
  // receives y_new ξ_c, so an accepted attempt leaves the next one's y0 ξ_c in place.
  static constexpr int kFusedNG = 8;            // ξ_c column groups of 16 the kernel carries
  static constexpr int kFusedChunk[3] = {512, 256, 768};
  int fusion_req = 0;       // 0 auto (the autotuner's choice; unfused until tuned), 1 off, 2 on
  bool fused_tuned = false; // the autotuner found the fused kernel faster
  cs* Wpart[2] = {nullptr, nullptr};
  cs* PredY[2] = {nullptr, nullptr};
  int wpart_max = 0;        // chunks the partial buffers hold
  bool y0xi_valid = false;  // PredY[par] is y0 ξ_c of the resident state
  bool fused_used = false;  // the attempt in flight ran the fused kernel
  // Streaming Ωξ kernel (zh_omgxi.hpp): column tiles of one lead each, built with the
  // layout; used when the autotuner timed it faster than the tiled GEMM (or forced).
  int xi_req = 0;          // 0 auto, 1 tiled GEMM, 2 streaming kernel
  bool xi_tuned = false;   // the autotuner found the streaming kernel faster
  bool xi_ok = false;      // the layout allows it
  int xi_ng = 1;           // 16-column groups per column tile
  std::vector<int4> xi_items;
  bool xi_stream() const { return xi_ok && (xi_req == 2 || (xi_req == 0 && xi_tuned)); }
  bool fusable() const {
    return std::is_same<R, float>::value && nc > 0 && nc <= 16 * kFusedNG && !overlap &&
           !tri_on && (fusion_req == 2 || (fusion_req == 0 && fused_tuned));
  }
  // Anti-Hermitian Ω.  dΩ_ij/dt = ΔM_j (ψ_i·Ixi_j) + ΔP_i (ξ_i·ψ_j*) + i(u_i - v_j) Ω_ij
  // maps Ω_ij -> -conj(Ω_ji) onto itself when every mode has ΔM_j Ixi_j = -conj(ΔP_j ξ_j)
  // and χ⁻_j = conj(χ⁺_j) (Lorentzian modes: Ixi = ξ*, ΔP = ΔM imaginary; Fermi-pole
  // modes: ΔP = ΔM = 0), whatever ψ and H are.  From an anti-Hermitian start (zero, or
  // psinought's Ω₀) Ω then stays anti-Hermitian, and only its block upper triangle (64 x
  // 64 blocks) is computed and stored: half of the Ω stage kernel's bytes and products.
  int anti_req = 0;         // 0 auto, 1 off, 2 on
  bool anti_ok = false;     // the constants have the symmetry
  bool state_anti = false;  // the resident Ω is anti-Hermitian
  bool tri_on = false;      // the resident Ω is held as its block upper triangle
  static constexpr double kAntiTol = 1e-13;  // relative, for both checks
  // Chunk layout of the fused kernel for chunk width C.
  OmegaFusedArgs fused_args(int C) const {
    OmegaFusedArgs a{};
    a.Md = Md; a.K = 2 * nc; a.nc = nc; a.C = C;
    a.ML = ML;
    a.SA = (ML + C - 1) / C;
    a.SP = (Md - ML + C - 1) / C;
    a.nbands = (Md + kFusedBM - 1) / kFusedBM;
    a.nfull = ff ? ML / kFusedBM : a.nbands;
    return a;
  }
  // error-norm partials written by the last stage's Ω kernel: one per workgroup
  int te_omega_blocks() const {
    if (fused_used) return fused_items(fused_args(kFusedChunk[tsel[TS_OMEGA]])) * NK;
    int bm = 0, bn = 0;
    tile_dims(tsel[TS_OMEGA], typename Cands<R>::omega(), bm, bn);
    return ((Md + bm - 1) / bm) * ((Md + bn - 1) / bn) * NK;
  }
  double2* sig[10] = {nullptr};
  cs* psi[10] = {nullptr};
  cs* omg[10] = {nullptr};
  double* te_part = nullptr;
  Ctrl* ctrl = nullptr;
  StageParams* sp = nullptr;
  double* Jd = nullptr;
  double* Jpart = nullptr;
  // first same as last (the attempt's k_1 is the previous accepted attempt's k_7): per
  // instance the device's decision (skip1, written by k_stage_params) and the currents
  // the last stage formed from y_new (Jn); on the host, which instances may offer it
  // at their next attempt and the step that made their k_7
  int* skip1 = nullptr;
  double* Jn = nullptr;
  std::vector<char> fsal_ready;
  std::vector<double> fsal_hprev;
  // where a next attempt may start and still take k_7: t0 + h (the last stage's time)
  // always, t0 + h_next when the controller found the stage parameters equal there
  std::vector<double> fsal_t7, fsal_tn;
  std::vector<char> fsal_hn;
  bool omit1_pending = false;  // the attempt in flight leaves out its first stage
  int fsal_req = 0;       // 0 on, 1 off
  long fsal_taken = 0;    // instance attempts that reused k_7
  const int* skip_now = nullptr;  // skip1 while the first stage of an attempt is issued
  int iy0 = 7, iyn = 8;
  int tsel[TS_N] = {0, 0, 0, 0, 0};  // chosen candidate tile per product
  // pinned host mailboxes
  Ctrl* ctrl_h = nullptr;
  double* J_h = nullptr;
  // by the parity of the y0 / y_new slots, and with / without the first stage
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev[2 * KC_N + 2];

  long nbytes = 0;
  bool pending = false;  // an attempt was launched and not collected yet
  std::vector<zh_attempt_info> pend_info;  // per instance

  template <typename T> T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess)
      throw Error{ZH_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    allocs.push_back(p);
    nbytes += (long)(count * sizeof(T));
    HIPCHK(hipMemsetAsync(p, 0, count * sizeof(T), stream));
    return (T*)p;
  }

  Engine(const zh_dims& d, int instances = 1)
      : NK(d.NK * instances), n(d.n), NL(d.NL), NX(d.NX), NOT(d.NOT), NI(instances), KI(d.NK),
        hbar(d.hbar), device(d.device) {
    M = NL * NX * NOT;
    PulseDesc p0;
    std::memset(&p0, 0, sizeof(p0));
    p0.hbar = hbar;
    pulse.assign(NI, p0);
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& a : aux) HIPCHK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    for (auto& e : ev_join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    HIPCHK(hipHostMalloc((void**)&ctrl_h, sizeof(Ctrl) * NI, hipHostMallocDefault));
    std::memset(ctrl_h, 0, sizeof(Ctrl) * NI);
    for (int g = 0; g < NI; ++g) ctrl_h[g].norm_scale = 1.0;
    HIPCHK(hipHostMalloc((void**)&J_h, sizeof(double) * NL * NK, hipHostMallocDefault));
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    ctrl = dalloc<Ctrl>(NI);
    sp = dalloc<StageParams>(7 * (size_t)NI);  // stage-major: sp[s * NI + g]
    pulse_d = dalloc<PulseDesc>(NI);
    Jd = dalloc<double>((size_t)NL * NK);
    Jpart = dalloc<double>((size_t)NL * NK * kCurChunks);
    skip1 = dalloc<int>(NI);
    Jn = dalloc<double>((size_t)NL * NK);
    fsal_ready.assign(NI, 0);
    fsal_hprev.assign(NI, 0.0);
    fsal_t7.assign(NI, 0.0);
    fsal_tn.assign(NI, 0.0);
    fsal_hn.assign(NI, 0);
    HIPCHK(hipStreamSynchronize(stream));
  }

  ~Engine() override {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : graph)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (auto& e : ev_join)
      if (e) (void)hipEventDestroy(e);

    for (auto& a : aux)
      if (a) (void)hipStreamDestroy(a);

    for (void* p : allocs) (void)hipFree(p);
    if (ctrl_h) (void)hipHostFree(ctrl_h);
    if (J_h) (void)hipHostFree(J_h);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graphs() {
    for (auto& g : graph)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    fsal_reset();  // whatever changed, the next attempt forms its own k_1
  }
  void fsal_reset() { std::fill(fsal_ready.begin(), fsal_ready.end(), 0); }
  // First same as last applies to the unfused attempts: the last stage stores k_7 over
  // the k_1 slot (Ω: in its epilogue; σ, Ψ: in the error-norm pass) and forms the
  // currents at y_new.  DOP54: the last row of the tableau is its weights, c_7 = 1.
  bool fsal_on() const { return fsal_req == 0 && !fused_used && tab.c[6] == 1.0; }

  // ---------------------------------------------------------------- set-up
  // Device mode order and coupled-orbital order, decided once from the constants.
  void plan_layout(const std::complex<double>* xi, const std::complex<double>* Ixi,
                   const std::complex<double>* dgp, const std::complex<double>* dgm) {
    // coupled orbitals (the reference's pruning, Zandpack/mpi/zand:589-602), grouped by
    // the first lead that couples them so each lead's orbitals form a short K range
    std::vector<int> first_lead(n, -1);
    for (int b = 0; b < NK; ++b)
      for (int m = 0; m < M; ++m) {
        const int a = m / (NX * NOT);
        for (int o = 0; o < n; ++o) {
          const long off = ((long)b * M + m) * n + o;
          if ((xi[off] != 0.0 || Ixi[off] != 0.0) && (first_lead[o] < 0 || a < first_lead[o]))
            first_lead[o] = a;
        }
      }
    // With several leads each lead's group is padded (cidx -1, zero columns) to a
    // multiple of 16, so that 16-aligned column tiles of the ξ products hold one lead
    // and run over that lead's modes only.
    cidx_h.clear();
    nc_real = 0;
    for (int a = 0; a < NL; ++a) {
      for (int o = 0; o < n; ++o)
        if (first_lead[o] == a) { cidx_h.push_back(o); ++nc_real; }
      if (NL > 1)
        while (cidx_h.size() % kCGroup) cidx_h.push_back(-1);
    }
    nc = (int)cidx_h.size();
    // active modes: ΔP or ΔM non-zero for some k (Fermi slots have both zero)
    std::vector<char> active(M, 0);
    for (int b = 0; b < NK; ++b)
      for (int m = 0; m < M; ++m)
        if (dgp[(long)b * M + m] != 0.0 || dgm[(long)b * M + m] != 0.0) active[m] = 1;
    int na = 0;
    for (int m = 0; m < M; ++m) na += active[m];
    const int np = M - na;
    const int ml_pad = ((na + kTile - 1) / kTile) * kTile;
    bool use = false;
    if (layout_req == 2) use = np > 0;
    else if (layout_req == 0) use = np > 0 && (ml_pad + np) <= M + M / 10;
    perm.clear();
    if (use) {
      for (int m = 0; m < M; ++m)
        if (active[m]) perm.push_back(m);
      while ((int)perm.size() < ml_pad) perm.push_back(-1);
      ML = ml_pad;
      for (int m = 0; m < M; ++m)
        if (!active[m]) perm.push_back(m);
    } else {
      for (int m = 0; m < M; ++m) perm.push_back(m);
      ML = M;
    }
    Md = (int)perm.size();
    pruned = use;
  }

  void allocate() {
    const size_t nn = (size_t)NK * n * n, mn = (size_t)NK * Md * n, mm = (size_t)NK * Md * Md;
    for (int s = 0; s < 10; ++s) {
      sig[s] = dalloc<double2>(nn);
      if (s != 9) psi[s] = dalloc<cs>(mn);
      if (s != 6) omg[s] = dalloc<cs>(mm);  // k_7 of Ω is consumed in place (error norm)
    }
    lead = dalloc<int>(Md);
    Hcur = dalloc<double2>(nn);
    Hre = dalloc<double>(nn);
    ell_col = dalloc<int>((size_t)NK * kEllMax * n);
    ell_val = dalloc<double2>((size_t)NK * kEllMax * n);
    HreT = dalloc<R>(nn);
    G = dalloc<double2>(nn);
    dvec = dalloc<double2>((size_t)NK * n);
    dzero = dalloc<double2>((size_t)NK * n);
    glp = dalloc<double2>((size_t)NK * Md);
    chip = dalloc<double2>((size_t)NK * Md);
    chim = dalloc<double2>((size_t)NK * Md);
    uvec = dalloc<double2>((size_t)NK * Md);
    vvec = dalloc<double2>((size_t)NK * Md);
    n0 = dalloc<double>((size_t)NK * n);
    const size_t mc = (size_t)NK * Md * nc;
    cidx = dalloc<int>(nc);
    cpos = dalloc<int>(n);
    t1 = dalloc<int2>((Md + kTile - 1) / kTile);
    t2 = dalloc<int2>((Md + kTile - 1) / kTile);
    txi = dalloc<int2>((Md + kTile - 1) / kTile);
    gmask = dalloc<int>((nc + kCGroup - 1) / kCGroup + 1);
    lrange = dalloc<int4>(NL);
    Ixic = dalloc<double2>(mc);
    Pic = dalloc<double2>((size_t)NK * kPiParts * n * nc);
    xic = dalloc<cs>(mc);
    W = dalloc<cs>(mc * kWParts);
    W2 = dalloc<cs>(mc * kWParts);
    for (auto& p : Pred) p = dalloc<cs>(mc);
    Wst = dalloc<cs>(mc);
    for (auto& q : PredY) q = dalloc<cs>(mc);
    if (std::is_same<R, float>::value && nc > 0 && nc <= 16 * kFusedNG) {
      // the narrowest chunk gives the most partials
      const int C = kFusedChunk[1];
      wpart_max = (Md + C - 1) / C + 1;
      for (auto& w : Wpart) w = dalloc<cs>(mc * wpart_max);
    }
    // room for the error-norm partials of the finest Ω tile candidate (32 x 32)
    const long t32 = (Md + 31) / 32;
    te_part = dalloc<double>((size_t)2 * kTEBlocks * NI + t32 * t32 * NK);
    Ap = Ap0 = dalloc<cs>((size_t)NK * Md * (n + nc));
    Ap1 = dalloc<cs>((size_t)NK * Md * (n + nc));
    psi_span = dalloc<int2>((n + 63) / 64);
    Bp = dalloc<cs>((size_t)NK * (n + nc) * n);
    Aomg = dalloc<cs>((size_t)NK * Md * 2 * nc);
    Bomg = dalloc<cs>((size_t)NK * Md * 2 * nc);
  }

  void set_constants(const double* H0, const double* xi_, const double* Ixi_, const double* glp_,
                     const double* dgp_, const double* dgm_, const double* xpp_,
                     const double* xpm_) override {
    HIPCHK(hipSetDevice(device));
    drop_graphs();
    typedef std::complex<double> cd;
    const cd* xi = (const cd*)xi_;
    const cd* Ixi = (const cd*)Ixi_;
    const cd* dgp = (const cd*)dgp_;
    const cd* dgm = (const cd*)dgm_;
    const cd* glpc = (const cd*)glp_;
    const cd* xpp = (const cd*)xpp_;
    const cd* xpm = (const cd*)xpm_;
    if (!have_consts) {
      plan_layout(xi, Ixi, dgp, dgm);
      allocate();
    } else {
      const int nc0 = nc, Md0 = Md, ML0 = ML;
      std::vector<int> perm0 = perm, cidx0 = cidx_h;
      plan_layout(xi, Ixi, dgp, dgm);
      require(nc == nc0 && Md == Md0 && ML == ML0 && perm == perm0 && cidx_h == cidx0,
              "coupling structure changed; create a new problem");
    }
    std::vector<int> cp(n, -1);
    for (int c = 0; c < nc; ++c)
      if (cidx_h[c] >= 0) cp[cidx_h[c]] = c;
    // per lead: range of its orbitals in the coupled order (term1 / term2 K range), and
    // per coupled column the leads that couple it
    std::vector<int> llo(NL, nc), lhi(NL, 0), cmask(nc, 0);
    for (int b = 0; b < NK; ++b)
      for (int m = 0; m < M; ++m) {
        const int a = m / (NX * NOT);
        for (int c = 0; c < nc; ++c) {
          if (cidx_h[c] < 0) continue;
          const long off = ((long)b * M + m) * n + cidx_h[c];
          if (xi[off] != 0.0 || Ixi[off] != 0.0) {
            llo[a] = std::min(llo[a], c);
            lhi[a] = std::max(lhi[a], c + 1);
            cmask[c] |= 1 << a;
          }
        }
      }
    const int ntile = (Md + kTile - 1) / kTile;
    std::vector<int2> tr(ntile);
    std::vector<int> ld(Md, 0);
    for (int d = 0; d < Md; ++d) ld[d] = perm[d] >= 0 ? perm[d] / (NX * NOT) : 0;
    for (int t = 0; t < ntile; ++t) {
      int lo = nc, hi = 0;
      for (int d = t * kTile; d < std::min(Md, (t + 1) * kTile); ++d) {
        const int m = perm[d];
        if (m < 0) continue;
        bool nzm = false;
        for (int b = 0; b < NK; ++b) nzm = nzm || dgm[(long)b * M + m] != 0.0;
        if (!nzm) continue;  // term1 vanishes on this column
        lo = std::min(lo, llo[ld[d]]);
        hi = std::max(hi, lhi[ld[d]]);
      }
      tr[t] = make_int2(lo < hi ? lo : 0, lo < hi ? hi : 0);
    }
    // per row tile: term2 K range (ΔP ξ_c vanishes outside the row leads' orbitals)
    // per column tile: ξ_c column range of the leads of its modes (the Ωξ product's
    // contraction over the tile's columns touches only those ξ_c columns)
    std::vector<int2> trx(ntile);
    for (int t = 0; t < ntile; ++t) {
      int lo = nc, hi = 0;
      for (int d = t * kTile; d < std::min(Md, (t + 1) * kTile); ++d) {
        if (perm[d] < 0) continue;
        lo = std::min(lo, llo[ld[d]]);
        hi = std::max(hi, lhi[ld[d]]);
      }
      trx[t] = make_int2(lo < hi ? lo : 0, lo < hi ? hi : 0);
    }
    std::vector<int2> tr2(ntile);
    for (int t = 0; t < ntile; ++t) {
      int lo = nc, hi = 0;
      for (int d = t * kTile; d < std::min(Md, (t + 1) * kTile); ++d) {
        const int m = perm[d];
        if (m < 0) continue;
        bool nzp = false;
        for (int b = 0; b < NK; ++b) nzp = nzp || dgp[(long)b * M + m] != 0.0;
        if (!nzp) continue;  // term2 vanishes on this row
        lo = std::min(lo, llo[ld[d]]);
        hi = std::max(hi, lhi[ld[d]]);
      }
      tr2[t] = make_int2(lo < hi ? lo : 0, lo < hi ? hi : 0);
    }
    // per 16-column group of the coupled order: mask of the leads that couple it; per
    // lead: its device rows in the active block [x, y) and in the passive block [z, w)
    const int ngroup = (nc + kCGroup - 1) / kCGroup;
    std::vector<int> gm(std::max(ngroup, 1), 0);
    for (int c = 0; c < nc; ++c) gm[c / kCGroup] |= cmask[c];
    std::vector<int4> lr(NL);
    for (int a = 0; a < NL; ++a) {
      int alo = Md, ahi = 0, plo = Md, phi = 0;
      for (int d = 0; d < Md; ++d) {
        if (perm[d] < 0 || ld[d] != a) continue;
        if (d < ML) { alo = std::min(alo, d); ahi = std::max(ahi, d + 1); }
        else { plo = std::min(plo, d); phi = std::max(phi, d + 1); }
      }
      lr[a] = make_int4(alo < ahi ? alo : 0, alo < ahi ? ahi : 0, plo < phi ? plo : 0,
                        plo < phi ? phi : 0);
    }
    // column tiles of the streaming Ωξ kernel: runs of up to 4 groups of one lead (a
    // group coupled by no lead gets a tile of zeros).  Its K chunks may reach past a
    // lead's modes to the chunk boundary, which stays inside the active or the passive
    // block when the active block is 64-aligned (or all of the modes).
    xi_items.clear();
    xi_ng = 1;
    xi_ok = nc > 0 && NL <= kXiMaxLeads && Md % 2 == 0 && nc % 2 == 0 &&
            (ML % kTile == 0 || ML == Md) && (long)Md * Md < (1L << 31) &&
            (long)Md * nc < (1L << 31);
    for (int g = 0; xi_ok && g < ngroup;) {
      const int mk = gm[g];
      if (mk & (mk - 1)) { xi_ok = false; break; }
      int e = g + 1;
      while (e < ngroup && gm[e] == mk && e - g < 4) ++e;
      xi_items.push_back(make_int4(kCGroup * g, kCGroup * (e - g), mk ? __builtin_ffs(mk) - 1 : -1, 0));
      xi_ng = std::max(xi_ng, e - g);
      g = e;
    }
    if ((int)xi_items.size() > kXiMaxItems) xi_ok = false;
    const size_t mc = (size_t)NK * Md * nc;
    std::vector<double2> ixic(mc), g1((size_t)NK * Md), x1((size_t)NK * Md), x2((size_t)NK * Md);
    std::vector<cs> xc(mc), apr(mc), bol(mc);
    for (int b = 0; b < NK; ++b)
      for (int d = 0; d < Md; ++d) {
        const int m = perm[d];
        const long bd = (long)b * Md + d;
        if (m < 0) {  // inert padding mode
          g1[bd] = x1[bd] = x2[bd] = make_double2(0, 0);
          for (int c = 0; c < nc; ++c) {
            ixic[bd * nc + c] = make_double2(0, 0);
            xc[bd * nc + c] = apr[bd * nc + c] = bol[bd * nc + c] = from_d2<cs>(make_double2(0, 0));
          }
          continue;
        }
        const long bm = (long)b * M + m;
        g1[bd] = make_double2(glpc[bm].real(), glpc[bm].imag());
        x1[bd] = make_double2(xpp[bm].real(), xpp[bm].imag());
        x2[bd] = make_double2(xpm[bm].real(), xpm[bm].imag());
        for (int c = 0; c < nc; ++c) {
          const bool pad = cidx_h[c] < 0;
          const cd x = pad ? cd(0.0) : xi[bm * n + cidx_h[c]];
          const cd ix = pad ? cd(0.0) : Ixi[bm * n + cidx_h[c]];
          ixic[bd * nc + c] = make_double2(ix.real(), ix.imag());
          xc[bd * nc + c] = from_d2<cs>(make_double2(x.real(), x.imag()));
          const cd px = dgp[bm] * x, mx = dgm[bm] * ix;
          apr[bd * nc + c] = from_d2<cs>(make_double2(px.real(), px.imag()));
          bol[bd * nc + c] = from_d2<cs>(make_double2(mx.real(), mx.imag()));
        }
      }
    xi_real = true;
    for (size_t i = 0; i < mc && xi_real; ++i)
      xi_real = ixic[i].y == 0.0 && (double)to_d2(xc[i]).y == 0.0;
    if (nc > 0) {
      HIPCHK(hipMemcpyAsync(cidx, cidx_h.data(), sizeof(int) * nc, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(Ixic, ixic.data(), sizeof(double2) * mc, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(xic, xc.data(), sizeof(cs) * mc, hipMemcpyHostToDevice, stream));
      // constant column blocks of Ap (right), Aomg (right) and Bomg (left): one
      // contiguous upload into the Ωξ partials buffer (free outside an attempt), then
      // strided device copies (a strided copy from pageable host memory goes row by row)
      HIPCHK(hipMemcpyAsync(W, apr.data(), sizeof(cs) * mc, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(W + mc, bol.data(), sizeof(cs) * mc, hipMemcpyHostToDevice, stream));
      for (cs* a : {Ap0, Ap1})
        HIPCHK(hipMemcpy2DAsync(a + n, sizeof(cs) * (n + nc), W, sizeof(cs) * nc,
                                sizeof(cs) * nc, (size_t)NK * Md, hipMemcpyDeviceToDevice,
                                stream));
      HIPCHK(hipMemcpy2DAsync(Aomg + nc, sizeof(cs) * 2 * nc, W, sizeof(cs) * nc,
                              sizeof(cs) * nc, (size_t)NK * Md, hipMemcpyDeviceToDevice, stream));
      HIPCHK(hipMemcpy2DAsync(Bomg, sizeof(cs) * 2 * nc, W + mc, sizeof(cs) * nc,
                              sizeof(cs) * nc, (size_t)NK * Md, hipMemcpyDeviceToDevice, stream));
    }
    HIPCHK(hipMemcpyAsync(cpos, cp.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(t1, tr.data(), sizeof(int2) * ntile, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(t2, tr2.data(), sizeof(int2) * ntile, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(txi, trx.data(), sizeof(int2) * ntile, hipMemcpyHostToDevice, stream));
    y0xi_valid = false;
    HIPCHK(hipMemcpyAsync(gmask, gm.data(), sizeof(int) * gm.size(), hipMemcpyHostToDevice,
                          stream));
    HIPCHK(hipMemcpyAsync(lrange, lr.data(), sizeof(int4) * NL, hipMemcpyHostToDevice, stream));
    lrange_h = lr;
    HIPCHK(hipMemcpyAsync(lead, ld.data(), sizeof(int) * Md, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(glp, g1.data(), sizeof(double2) * NK * Md, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(chip, x1.data(), sizeof(double2) * NK * Md, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(chim, x2.data(), sizeof(double2) * NK * Md, hipMemcpyHostToDevice, stream));
    H0h.assign(H0, H0 + 2 * (size_t)NK * n * n);
    model_dirty = true;
    install_h(H0);
    HIPCHK(hipStreamSynchronize(stream));
    h_dirty = false;
    anti_ok = constants_antihermitian((long)NK * M, n, xi, Ixi, dgp, dgm, xpp, xpm);
    have_consts = true;
    update_tri();
  }

  // Whether every mode has ΔM_j Ixi_j = -conj(ΔP_j ξ_j) and χ⁻_j = conj(χ⁺_j) (to
  // kAntiTol of the largest term): the constants keep Ω anti-Hermitian.
  static bool constants_antihermitian(long nm, int n, const std::complex<double>* xi,
                                      const std::complex<double>* Ixi,
                                      const std::complex<double>* dgp,
                                      const std::complex<double>* dgm,
                                      const std::complex<double>* xpp,
                                      const std::complex<double>* xpm) {
    double scale = 0.0, err = 0.0, xs = 0.0, xe = 0.0;
    for (long m = 0; m < nm; ++m) {
      xs = std::max(xs, std::abs(xpp[m]));
      xe = std::max(xe, std::abs(xpm[m] - std::conj(xpp[m])));
      if (dgp[m] == 0.0 && dgm[m] == 0.0) continue;
      for (int o = 0; o < n; ++o) {
        const std::complex<double> a = dgm[m] * Ixi[m * n + o], b = -std::conj(dgp[m] * xi[m * n + o]);
        scale = std::max(scale, std::max(std::abs(a), std::abs(b)));
        err = std::max(err, std::abs(a - b));
      }
    }
    return err <= kAntiTol * scale && xe <= kAntiTol * xs;
  }
  // Whether Ω (NK, M, M, reference layout, complex128) is anti-Hermitian to kAntiTol of
  // its largest entry (64 x 64 blocks, for the cache).
  bool omega_antihermitian(const double* o) const {
    typedef std::complex<double> cd;
    const cd* w = (const cd*)o;
    double scale = 0.0, err = 0.0;
    for (int b = 0; b < NK; ++b)
      for (int i0 = 0; i0 < M; i0 += 64)
        for (int j0 = i0; j0 < M; j0 += 64)
          for (int i = i0; i < std::min(M, i0 + 64); ++i)
            for (int j = std::max(j0, i); j < std::min(M, j0 + 64); ++j) {
              const cd a = w[((size_t)b * M + i) * M + j], c = w[((size_t)b * M + j) * M + i];
              scale = std::max(scale, std::max(std::abs(a), std::abs(c)));
              err = std::max(err, std::abs(a + std::conj(c)));
            }
    return err <= kAntiTol * scale;
  }
  // Hold Ω as its block upper triangle when the constants, the request and the resident
  // state allow it; leaving the mode fills the resident lower triangle back in.
  void update_tri() {
    const bool want = anti_ok && anti_req != 1 && state_anti;
    if (want == tri_on) return;
    if (tri_on && have_state) mirror_lower(omg[iy0]);
    tri_on = want;
    y0xi_valid = false;
    drop_graphs();
  }
  void mirror_lower(cs* X) {
    HIPCHK(hipSetDevice(device));
    hipLaunchKernelGGL(k_mirror_lower<cs>, dim3(2048), dim3(256), 0, stream, NK, Md, X);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
  }
  void set_antihermitian(int mode) override {
    require(mode >= 0 && mode <= 2, "anti-Hermitian Omega: 0 auto, 1 off, 2 on");
    if (mode == 2) {
      require(have_consts, "set the constants first", ZH_ESTATE);
      require(anti_ok, "anti-Hermitian Omega: the constants do not keep Omega anti-Hermitian "
                       "(every mode needs dGM_j Ixi_j = -conj(dGP_j xi_j) and "
                       "Xpm_j = conj(Xpp_j))");
      require(!have_state || state_anti,
              "anti-Hermitian Omega: the resident Omega is not anti-Hermitian");
    }
    anti_req = mode;
    update_tri();
  }
  int antihermitian_info() const override {
    return (tri_on ? 1 : 0) | (anti_ok ? 2 : 0) | (state_anti ? 4 : 0);
  }

  // Mode rows per workgroup of k_psi_ell: those rows live in LDS (at most 32 KiB,
  // or one row of up to 64 KiB); 0 when a row does not fit.
  int hpsi_rows() const {
    const size_t row = (size_t)n * sizeof(cs);
    for (int rb : {4, 2, 1})
      if (rb * row <= 32768) return rb;
    return row <= 65536 ? 1 : 0;
  }

  // S⁻¹H (FP64 MFMA on the device), read back for the host-side analysis below.
  std::vector<double> apply_sinv(const double* H) {
    const long nn = (long)n * n;
    HIPCHK(hipMemcpyAsync(Hcur, H, sizeof(double2) * NK * nn, hipMemcpyHostToDevice, stream));
    GemmArgs<double, double> g{n, n, n, Sdev, n, nn, Hcur, n, nn};
    EpiStore<double2> e{Qs, n, nn, 1.0, 0.0};
    HIPCHK((launch_sel<double, double, double, false, false, 3>(tsel[TS_HSIG], g, NK, e, stream,
                                                                KFull(), CandHsig())));
    std::vector<double> out(2 * (size_t)NK * nn);
    HIPCHK(hipMemcpyAsync(out.data(), Qs, sizeof(double2) * NK * nn, hipMemcpyDeviceToHost,
                          stream));
    HIPCHK(hipStreamSynchronize(stream));
    return out;
  }

  // Installs H (host, (NK, n, n)) into the Hamiltonian operands.  With `model` (H0 of
  // the attempts) the drive's entries are part of the structure: a row-compressed H holds
  // them even where H0 is zero, complex drive values make H complex, and the entries'
  // H0 values and compressed-row slots go to the device for k_drive_update.
  void install_h(const double* H, bool model = true) {
    std::vector<double> SH;
    if (overlap) {
      SH = apply_sinv(H);
      H = SH.data();
    }
    const bool with_drive = model && drive.nt > 0 && !overlap;
    // drive entries per (k, row), for the compressed rows
    std::map<long, std::vector<int>> dcols;
    if (with_drive)
      for (const int3& e : de_h) dcols[(long)e.x * n + e.y].push_back(e.z);
    // a real Hamiltonian halves the MFMA work of the Hσ and ψHᵀ products
    bool real = !overlap;
    for (size_t i = 0; i < (size_t)NK * n * n && real; ++i) real = H[2 * i + 1] == 0.0;
    if (with_drive)
      for (const double2& c : de_coef_h) real = real && c.y == 0.0;
    // entries per row: with few, both H products run as gathers (row-compressed H)
    auto row_cols = [&](long r) {
      std::vector<int> cols;
      const double* row = H + 2 * r * n;
      for (int o = 0; o < n; ++o)
        if (row[2 * o] != 0.0 || row[2 * o + 1] != 0.0) cols.push_back(o);
      auto it = dcols.find(r);
      if (it != dcols.end()) {
        cols.insert(cols.end(), it->second.begin(), it->second.end());
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      }
      return cols;
    };
    int rmax = 0;
    for (long r = 0; r < (long)NK * n; ++r) rmax = std::max(rmax, (int)row_cols(r).size());
    const int rb = hpsi_rows();
    const bool fits = rmax <= kEllMax && rb > 0 && !overlap;
    bool sparse = false;
    if (sparse_req == 2) {
      require(!overlap, "row-compressed H products are not available with an overlap");
      require(fits, "row-compressed H needs at most " + std::to_string(kEllMax) +
                        " non-zeros per row and a Psi row of at most 64 KiB (rows hold " +
                        std::to_string(rmax) + ")");
      sparse = true;
    } else if (sparse_req == 0) {
      sparse = fits && 8 * rmax <= n;
    }
    const int R_ = sparse ? std::max(rmax, 1) : 0;
    if (real != h_real || sparse != h_sparse || R_ != ell_R) drop_graphs();
    h_real = real;
    h_sparse = sparse;
    ell_R = R_;
    hpsi_rb = rb;
    HIPCHK(hipMemcpyAsync(Hcur, H, sizeof(double2) * NK * n * n, hipMemcpyHostToDevice, stream));
    if (sparse) {
      ell_col_h.assign((size_t)NK * R_ * n, 0);
      ell_val_h.assign((size_t)NK * R_ * n, make_double2(0.0, 0.0));
      for (int b = 0; b < NK; ++b)
        for (int o = 0; o < n; ++o) {
          const double* row = H + 2 * ((long)b * n + o) * n;
          const std::vector<int> cols = row_cols((long)b * n + o);
          int r = 0;
          for (int c : cols) {
            const size_t at = ((size_t)b * R_ + r) * n + o;
            ell_col_h[at] = c;
            ell_val_h[at] = make_double2(row[2 * c], row[2 * c + 1]);
            ++r;
          }
          for (; r < R_; ++r) ell_col_h[((size_t)b * R_ + r) * n + o] = o;  // zero padding
        }
      HIPCHK(hipMemcpyAsync(ell_col, ell_col_h.data(), sizeof(int) * ell_col_h.size(),
                            hipMemcpyHostToDevice, stream));
      // per 64-orbital tile: the orbitals the tile's rows of H reach (and the tile itself)
      const int nt = (n + 63) / 64;
      std::vector<int2> spn(nt);
      psi_wmax = 0;
      for (int ct = 0; ct < nt; ++ct) {
        int lo = ct * 64, hi = std::min(n, ct * 64 + 64) - 1;
        for (int b = 0; b < NK; ++b)
          for (int r = 0; r < R_; ++r)
            for (int o = ct * 64; o < std::min(n, ct * 64 + 64); ++o) {
              const int c = ell_col_h[((size_t)b * R_ + r) * n + o];
              lo = std::min(lo, c);
              hi = std::max(hi, c);
            }
        spn[ct] = make_int2(lo, hi - lo + 1);
        psi_wmax = std::max(psi_wmax, hi - lo + 1);
      }
      HIPCHK(hipMemcpyAsync(psi_span, spn.data(), sizeof(int2) * nt, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));  // spn is read asynchronously
      HIPCHK(hipMemcpyAsync(ell_val, ell_val_h.data(), sizeof(double2) * ell_val_h.size(),
                            hipMemcpyHostToDevice, stream));
    }
    if (with_drive) {
      // the entries' H0 values and compressed-row slots
      const size_t E = de_h.size();
      std::vector<double2> h0(E);
      de_slot_h.assign(E, -1);
      for (size_t e = 0; e < E; ++e) {
        const int3 q = de_h[e];
        const double* v = H + 2 * (((long)q.x * n + q.y) * n + q.z);
        h0[e] = make_double2(v[0], v[1]);
        if (sparse)
          for (int r = 0; r < R_; ++r)
            if (ell_col_h[((size_t)q.x * R_ + r) * n + q.y] == q.z) { de_slot_h[e] = r; break; }
      }
      HIPCHK(hipMemcpyAsync(de_h0, h0.data(), sizeof(double2) * E, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(de_slot, de_slot_h.data(), sizeof(int) * E, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));  // h0 and de_slot_h are read asynchronously
    }
    dim3 grid((n + 31) / 32, (n + 31) / 32, NK);
    hipLaunchKernelGGL((k_transpose_h<cs, R>), grid, dim3(32, 8), 0, stream, n,
                       (long)(n + nc) * n, (const double2*)Hcur, Bp, real ? Hre : nullptr,
                       real ? HreT : nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));  // the host copies above are read asynchronously
  }

  // Drive terms: the distinct entries of the M_i with their terms' coefficients.
  void set_drive(int nt, const int* kinds, const double* amps, const double* par, const int* nnz,
                 const int* ks, const int* rows, const int* cols, const double* vals) override {
    require(have_consts, "set the constants first", ZH_ESTATE);
    require(nt >= 0 && nt <= ZH_MAX_DRIVE, "at most ZH_MAX_DRIVE drive terms");
    HIPCHK(hipSetDevice(device));
    DriveDesc d{};
    d.nt = nt;
    std::map<std::tuple<int, int, int>, std::vector<std::pair<int, double2>>> ent;
    long off = 0;
    for (int i = 0; i < nt; ++i) {
      require(kinds[i] >= ZH_WAVE_CONST && kinds[i] <= ZH_WAVE_GAUSS_SIN, "unknown drive waveform");
      require(!(kinds[i] == ZH_WAVE_SIN_FERMI && par[4 * i + 3] == 0.0), "zero switch-off width");
      require(!(kinds[i] == ZH_WAVE_GAUSS_SIN && par[4 * i] == 0.0), "zero envelope width");
      require(nnz[i] >= 0, "negative entry count");
      d.kind[i] = kinds[i];
      d.amp[i] = amps[i];
      for (int j = 0; j < 4; ++j) d.par[i][j] = par[4 * i + j];
      for (int q = 0; q < nnz[i]; ++q, ++off) {
        require(ks[off] >= 0 && ks[off] < NK && rows[off] >= 0 && rows[off] < n && cols[off] >= 0 &&
                    cols[off] < n, "drive entry out of range");
        ent[std::make_tuple(ks[off], rows[off], cols[off])].push_back(
            {i, make_double2(vals[2 * off], vals[2 * off + 1])});
      }
    }
    drop_graphs();
    de_h.clear(); de_beg_h.assign(1, 0); de_term_h.clear(); de_coef_h.clear();
    for (auto& kv : ent) {
      de_h.push_back(make_int3(std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)));
      for (auto& tc : kv.second) { de_term_h.push_back(tc.first); de_coef_h.push_back(tc.second); }
      de_beg_h.push_back((int)de_term_h.size());
    }
    const size_t E = de_h.size(), T = de_term_h.size();
    if (E > de_cap) {
      de_k = dalloc<int>(E); de_r = dalloc<int>(E); de_c = dalloc<int>(E); de_slot = dalloc<int>(E);
      de_h0 = dalloc<double2>(E); de_beg = dalloc<int>(E + 1);
      de_cap = E;
    }
    if (T > dt_cap) {
      de_term = dalloc<int>(T); de_coef = dalloc<double2>(T);
      dt_cap = T;
    }
    if (E > 0) {
      std::vector<int> kk(E), rr(E), cc(E);
      for (size_t e = 0; e < E; ++e) { kk[e] = de_h[e].x; rr[e] = de_h[e].y; cc[e] = de_h[e].z; }
      HIPCHK(hipMemcpyAsync(de_k, kk.data(), sizeof(int) * E, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(de_r, rr.data(), sizeof(int) * E, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(de_c, cc.data(), sizeof(int) * E, hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(de_beg, de_beg_h.data(), sizeof(int) * (E + 1), hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipMemcpyAsync(de_term, de_term_h.data(), sizeof(int) * T, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipMemcpyAsync(de_coef, de_coef_h.data(), sizeof(double2) * T, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    drive = d;
    if (E == 0) drive.nt = 0;
    model_dirty = true;
    install_h(H0h.data());
    h_dirty = false;
  }

  // S⁻¹H0 and S⁻¹M_i of the overlap form with a σ- or t-dependent H (before an attempt).
  void prepare_model() {
    if (!model_dirty) return;
    model_dirty = false;
    if (!ovl_model()) return;
    const long nn = (long)n * n;
    if (!SH0) SH0 = dalloc<double2>((size_t)NK * nn);
    if (drive.nt > sm_cap) {
      SM = dalloc<double2>((size_t)std::max(drive.nt, 1) * NK * nn);
      sm_cap = drive.nt;
    }
    // S⁻¹ X for X = H0 and each dense M_i (FP64 MFMA)
    auto sinv_times = [&](const double* X, double2* out) {
      HIPCHK(hipMemcpyAsync(G, X, sizeof(double2) * NK * nn, hipMemcpyHostToDevice, stream));
      GemmArgs<double, double> g{n, n, n, Sdev, n, nn, G, n, nn};
      EpiStore<double2> e{out, n, nn, 1.0, 0.0};
      HIPCHK((launch_zgemm<double, double, double, false, false, Tile<64, 64, 16, 2, false, 3>, 3>(
          g, NK, e, stream)));
      HIPCHK(hipStreamSynchronize(stream));
    };
    sinv_times(H0h.data(), SH0);
    std::vector<double> M((size_t)2 * NK * nn);
    for (int i = 0; i < drive.nt; ++i) {
      std::fill(M.begin(), M.end(), 0.0);
      for (size_t e = 0; e < de_h.size(); ++e)
        for (int q = de_beg_h[e]; q < de_beg_h[e + 1]; ++q)
          if (de_term_h[q] == i) {
            const size_t at = ((size_t)de_h[e].x * n + de_h[e].y) * n + de_h[e].z;
            M[2 * at] = de_coef_h[q].x;
            M[2 * at + 1] = de_coef_h[q].y;
          }
      sinv_times(M.data(), SM + (size_t)i * NK * nn);
    }
  }

  // H(t_s, σ_s) of stage sps into the Hamiltonian operands (attempts and the steady
  // state; the explicit zh_rhs takes the caller's H): drive entries, or with an overlap
  // the dense S⁻¹H(t, σ) with the stage's Hubbard diagonal.
  void apply_model(const StageParams* sps, hipStream_t st) {
    if (ovl_model()) {
      hipLaunchKernelGGL((k_overlap_model<cs>), dim3(2048), dim3(256), 0, st, NK, n, nc, drive.nt,
                         (const double2*)SH0, (const double2*)SM, (const double2*)Sdev,
                         (const double2*)(U != 0.0 ? dvec : nullptr), sps, KI, Hcur, Bp);
      HIPCHK(hipGetLastError());
    } else if (drive.nt > 0) {
      DriveArgs<cs, R> a{(int)de_h.size(), n, nc, ell_R, de_k, de_r, de_c, de_slot, de_h0,
                         de_beg, de_term, de_coef, Hcur, Bp, h_real ? Hre : nullptr,
                         h_real ? HreT : nullptr, h_sparse ? ell_val : nullptr, sps, KI};
      hipLaunchKernelGGL((k_drive_update<cs, R>), dim3(std::min<int>(1024, (a.E + 255) / 256)),
                         dim3(256), 0, st, a);
      HIPCHK(hipGetLastError());
    }
  }
  // The Hubbard diagonal the σ and Ψ epilogues apply (with an overlap it is part of
  // S⁻¹H(t, σ) instead).
  const double2* dvec_epi() const { return ovl_model() ? dzero : dvec; }

  Ell ell() const { return Ell{ell_col, ell_val, ell_R, n}; }

  // The [ΔPξ_c][ħ σ_cᵀ] part of the Ψ right-hand side: rows from g0_rows() on have
  // ΔP = 0 (passive block of the active-first order), so their tiles are skipped.
  int g0_rows() const { return pruned ? ML : Md; }
  KSegRowLead kseg_g0() const {
    return KSegRowLead{t2, ML, pruned ? 1 : 0, (Md + kTile - 1) / kTile};
  }

  void set_hubbard(double U_, const double* n0_) override {
    require(U_ == 0.0 || KI == 2,
            "the Hubbard model couples two spin blocks: NK must be 2 (per instance)");
    require(have_consts, "set the constants first", ZH_ESTATE);
    drop_graphs();
    U = U_;
    model_dirty = true;
    if (U != 0.0) {
      HIPCHK(hipMemcpyAsync(n0, n0_, sizeof(double) * NK * n, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
  }

  // Π_α of the resident state per lead (the stage-0 Π of the next attempt,
  // Zandpack/mpi/zand:363-366 Q_np; nozand:484-486 PI_np): Q (NK, NL, n, n) complex128
  // = (Π_α + Π_α†)/ħ (hermitian) or Π_α/ħ, Π_α[o, o'] = Σ_{m in α} ψ_m[o] Ixi_m[o'].
  void pi_get(int hermitian, double* Q) override {
    require(have_consts && have_state, "Pi needs constants and a state", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    const long nn = (long)n * n;
    std::memset(Q, 0, sizeof(double) * 2 * (size_t)NK * NL * nn);
    if (nc == 0) return;
    std::vector<double2> part((size_t)NK * n * nc);
    for (int a = 0; a < NL; ++a) {
      const int4 r = lrange_h[a];
      GemmArgs<R, double> g{n, nc, Md, psi[iy0], n, (long)Md * n, Ixic, nc, (long)Md * nc};
      EpiStore<double2> e{Pic, nc, (long)n * nc, 1.0, 0.0};
      HIPCHK((launch_zgemm<double, R, double, true, false, Tile<64, 32, 16, 2, false, 3>, 3>(
          g, NK, e, stream, KSegFixed{r.x, r.y, r.z, r.w})));
      HIPCHK(hipMemcpyAsync(part.data(), Pic, sizeof(double2) * part.size(),
                            hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      const double ih = 1.0 / hbar;
      for (int b = 0; b < NK; ++b) {
        double* q = Q + 2 * ((size_t)b * NL + a) * nn;
        for (int o = 0; o < n; ++o)
          for (int c = 0; c < nc; ++c) {
            const int oc = cidx_h[c];
            if (oc < 0) continue;
            const double2 v = part[((size_t)b * n + o) * nc + c];
            q[2 * ((size_t)o * n + oc)] += v.x * ih;
            q[2 * ((size_t)o * n + oc) + 1] += v.y * ih;
            if (hermitian) {
              q[2 * ((size_t)oc * n + o)] += v.x * ih;
              q[2 * ((size_t)oc * n + o) + 1] -= v.y * ih;
            }
          }
      }
    }
  }

  // S⁻¹ of the non-orthogonal form (null: back to the orthogonal form).
  // Scaled error norm: TE = sqrt(mean |err/(atol + rtol max(|y0|, |y1|))|²) over the
  // reference-layout entries of σ, Ψ and Ω; (0, 0) restores zand's absolute norm.
  double tol_rtol = 0.0, tol_atol = 0.0;
  bool tol_scaled = false;
  void set_tolerance(double rtol, double atol) override {
    require(rtol >= 0.0 && atol >= 0.0 && std::isfinite(rtol) && std::isfinite(atol),
            "tolerances must be finite and non-negative");
    require(rtol == 0.0 || atol > 0.0, "the scaled norm needs atol > 0 (zero entries)");
    drop_graphs();
    tol_rtol = rtol;
    tol_atol = atol;
    tol_scaled = atol > 0.0;
  }
  double norm_scale() const {
    if (!tol_scaled) return 1.0;
    const double cnt = (double)KI * ((double)n * n + (double)M * n + (double)M * M);
    return 1.0 / cnt;
  }

  void set_overlap(const double* S) override {
    HIPCHK(hipSetDevice(device));
    require(have_consts, "set the constants first", ZH_ESTATE);
    drop_graphs();
    model_dirty = true;
    if (!S) {
      overlap = false;
      install_h(H0h.data());
      h_dirty = false;
      return;
    }
    const long nn = (long)n * n;
    if (!Sdev) {
      Sdev = dalloc<double2>((size_t)NK * nn);
      Sc = dalloc<double2>((size_t)NK * nc * n);
      Bs = dalloc<cs>((size_t)NK * nc * n);
      Pis = dalloc<double2>((size_t)NK * n * nc);
      Qs = dalloc<double2>((size_t)NK * nn);
      Yov = dalloc<cs>((size_t)NK * Md * nc);
      cpos_none = dalloc<int>(n);
      std::vector<int> none(n, -1);
      HIPCHK(hipMemcpyAsync(cpos_none, none.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    std::vector<double2> sc((size_t)NK * nc * n), bs_d((size_t)NK * nc * n);
    const double2* Sv = (const double2*)S;
    for (int b = 0; b < NK; ++b)
      for (int c = 0; c < nc; ++c) {
        const int o = cidx_h[c];
        for (int j = 0; j < n; ++j) {
          const size_t at = ((size_t)b * nc + c) * n + j;
          sc[at] = o < 0 ? make_double2(0, 0) : Sv[(size_t)b * nn + (size_t)o * n + j];
          bs_d[at] = o < 0 ? make_double2(0, 0) : Sv[(size_t)b * nn + (size_t)j * n + o];
        }
      }
    std::vector<cs> bs(bs_d.size());
    for (size_t i = 0; i < bs.size(); ++i) bs[i] = from_d2<cs>(bs_d[i]);
    HIPCHK(hipMemcpyAsync(Sdev, S, sizeof(double2) * NK * nn, hipMemcpyHostToDevice, stream));
    if (nc > 0) {
      HIPCHK(hipMemcpyAsync(Sc, sc.data(), sizeof(double2) * sc.size(), hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipMemcpyAsync(Bs, bs.data(), sizeof(cs) * bs.size(), hipMemcpyHostToDevice, stream));
    }
    overlap = true;
    install_h(H0h.data());  // synchronises: the host vectors above are read before return
    h_dirty = false;
  }

  void set_dissipator(const DissDesc& d, const double* seq) override {
    require(have_consts, "set the constants first", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    drop_graphs();
    if (d.kind != ZH_DISS_NONE) {
      if (!sig_eq) sig_eq = dalloc<double2>((size_t)NK * n * n);
      HIPCHK(hipMemcpyAsync(sig_eq, seq, sizeof(double2) * NK * n * n, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    diss = d;
  }

  void set_pulse(const PulseDesc& p) override {
    for (int g = 0; g < NI; ++g) set_instance_pulse(g, p);
  }
  int instances() const override { return NI; }
  void set_instance_pulse(int g, const PulseDesc& p) override {
    require(g >= 0 && g < NI, "instance index out of range");
    require(p.kind != ZH_PULSE_SPLINE || spl_n >= 2, "set the pulse spline first", ZH_ESTATE);
    pulse[g] = p;
    pulse[g].hbar = hbar;
    pulse[g].sx = spl_x;
    pulse[g].sc = spl_c;
    pulse[g].ns = spl_n;
    pulse_dirty = true;
    fsal_ready[g] = 0;  // its last k_7 was formed under the old pulse
  }
  // the pulses into pulse_d before a launch that reads them (stream order)
  void upload_pulses() {
    if (!pulse_dirty) return;
    HIPCHK(hipMemcpyAsync(pulse_d, pulse.data(), sizeof(PulseDesc) * NI, hipMemcpyHostToDevice,
                          stream));
    HIPCHK(hipStreamSynchronize(stream));  // pulse is host memory that may change
    pulse_dirty = false;
  }

  // knots and coefficients of ZH_PULSE_SPLINE (device copies; replaced on each call)
  double *spl_x = nullptr, *spl_c = nullptr;
  int spl_n = 0, spl_cap = 0;
  void set_pulse_spline(int nk, const double* x, const double* c) override {
    HIPCHK(hipSetDevice(device));
    drop_graphs();
    if (nk > spl_cap) {
      spl_x = dalloc<double>(nk);
      spl_c = dalloc<double>(4 * (size_t)(nk - 1));
      spl_cap = nk;
    }
    HIPCHK(hipMemcpyAsync(spl_x, x, sizeof(double) * nk, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(spl_c, c, sizeof(double) * 4 * (nk - 1), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    spl_n = nk;
    for (auto& q : pulse) {
      q.sx = spl_x;
      q.sc = spl_c;
      q.ns = spl_n;
    }
    pulse_dirty = true;
  }

  // ------------------------------------------------------------- transfers
  // host complex128 sigma-like (rows x cols dense) -> device storage C (leading dim ld)
  // (dense rows only: every caller passes ld == cols, and a strided copy from
  // pageable host memory would go row by row)
  template <typename C>
  void upload(C* dst, long ld, const double* src, long rows, long cols) {
    require(ld == cols, "upload: rows must be dense");
    if (sizeof(C) == sizeof(double2)) {
      HIPCHK(hipMemcpyAsync(dst, src, sizeof(double2) * rows * cols, hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));
    } else {
      std::vector<C> tmp((size_t)rows * cols);
      for (size_t i = 0; i < tmp.size(); ++i)
        tmp[i] = from_d2<C>(make_double2(src[2 * i], src[2 * i + 1]));
      HIPCHK(hipMemcpyAsync(dst, tmp.data(), sizeof(C) * tmp.size(), hipMemcpyHostToDevice,
                            stream));
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
  // Ψ (NK, M, n) reference rows -> device rows (leading dim ld), permuted; with
  // ld > n the rows go through `stage` (NK Md n elements) and a strided device copy
  void upload_psi(cs* dst, long ld, const double* src, cs* stage = nullptr) {
    std::vector<cs> tmp((size_t)NK * Md * n);
    for (int b = 0; b < NK; ++b)
      for (int d = 0; d < Md; ++d) {
        cs* row = tmp.data() + ((size_t)b * Md + d) * n;
        const int m = perm[d];
        for (int o = 0; o < n; ++o) {
          if (m < 0) { row[o] = from_d2<cs>(make_double2(0, 0)); continue; }
          const double* s = src + 2 * (((size_t)b * M + m) * n + o);
          row[o] = from_d2<cs>(make_double2(s[0], s[1]));
        }
      }
    if (ld == n) {
      HIPCHK(hipMemcpyAsync(dst, tmp.data(), sizeof(cs) * tmp.size(), hipMemcpyHostToDevice,
                            stream));
    } else {
      require(stage != nullptr, "upload_psi: strided rows need a staging buffer");
      HIPCHK(hipMemcpyAsync(stage, tmp.data(), sizeof(cs) * tmp.size(), hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipMemcpy2DAsync(dst, sizeof(cs) * ld, stage, sizeof(cs) * n, sizeof(cs) * n,
                              (size_t)NK * Md, hipMemcpyDeviceToDevice, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
  }
  void download_psi(double* dst, const cs* src) {
    std::vector<cs> tmp((size_t)NK * Md * n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, sizeof(cs) * tmp.size(), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int b = 0; b < NK; ++b)
      for (int d = 0; d < Md; ++d) {
        const int m = perm[d];
        if (m < 0) continue;
        for (int o = 0; o < n; ++o) {
          const cs v = tmp[((size_t)b * Md + d) * n + o];
          double* t = dst + 2 * (((size_t)b * M + m) * n + o);
          t[0] = (double)v.x;
          t[1] = (double)v.y;
        }
      }
  }
  // Ω (NK, M, M) reference -> device (NK, Md, Md), both axes permuted.
  // Returns whether the passive x passive block is exactly zero.
  bool upload_omega(cs* dst, const double* src) {
    std::vector<cs> tmp((size_t)NK * Md * Md);
    bool ffz = true;
    for (int b = 0; b < NK; ++b)
      for (int d1 = 0; d1 < Md; ++d1) {
        cs* row = tmp.data() + ((size_t)b * Md + d1) * Md;
        const int m1 = perm[d1];
        for (int d2 = 0; d2 < Md; ++d2) {
          const int m2 = perm[d2];
          if (m1 < 0 || m2 < 0) { row[d2] = from_d2<cs>(make_double2(0, 0)); continue; }
          const double* s = src + 2 * (((size_t)b * M + m1) * M + m2);
          row[d2] = from_d2<cs>(make_double2(s[0], s[1]));
          if (d1 >= ML && d2 >= ML && (s[0] != 0.0 || s[1] != 0.0)) ffz = false;
        }
      }
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), sizeof(cs) * tmp.size(), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return ffz;
  }
  // With Ω held as its block upper triangle the rest is filled in, Ω_ij = -conj(Ω_ji).
  void download_omega(double* dst, const cs* src) {
    std::vector<cs> tmp((size_t)NK * Md * Md);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, sizeof(cs) * tmp.size(), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int b = 0; b < NK; ++b)
      for (int d1 = 0; d1 < Md; ++d1) {
        const int m1 = perm[d1];
        if (m1 < 0) continue;
        for (int d2 = 0; d2 < Md; ++d2) {
          const int m2 = perm[d2];
          if (m2 < 0) continue;
          double* t = dst + 2 * (((size_t)b * M + m1) * M + m2);
          if (tri_on && (d1 >> 6) > (d2 >> 6)) {
            const cs v = tmp[((size_t)b * Md + d2) * Md + d1];
            t[0] = -(double)v.x;
            t[1] = (double)v.y;
          } else {
            const cs v = tmp[((size_t)b * Md + d1) * Md + d2];
            t[0] = (double)v.x;
            t[1] = (double)v.y;
          }
        }
      }
  }
  // zero the passive x passive block of every Ω slot
  void clear_ff() {
    if (ML >= Md) return;
    y0xi_valid = false;
    for (int s = 0; s < 10; ++s)
      for (int b = 0; b < NK && omg[s]; ++b)
        HIPCHK(hipMemset2DAsync(omg[s] + ((size_t)b * Md + ML) * Md + ML, sizeof(cs) * Md, 0,
                                sizeof(cs) * (Md - ML), Md - ML, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void set_ff(bool on) {
    on = on && pruned;
    if (on != ff) drop_graphs();
    if (on && (!ff || ff_dirty)) { clear_ff(); ff_dirty = false; }
    ff = on;
  }

  void state_put(const double* s, const double* p, const double* o) override {
    require(have_consts, "set the constants first", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    fsal_reset();
    // a null Ω is zero, anti-Hermitian
    const bool anti_in = o ? (anti_ok && omega_antihermitian(o)) : true;
    require(anti_req != 2 || anti_in,
            "anti-Hermitian Omega is on and the Omega given is not anti-Hermitian");
    if (tri_on) {  // the resident state is replaced in full
      tri_on = false;
      drop_graphs();
    }
    upload(sig[iy0], n, s, (long)NK * n, n);
    // a null Ψ or Ω is zero: cleared on the device instead of uploaded
    if (p) {
      upload_psi(psi[iy0], n, p);
    } else {
      HIPCHK(hipMemsetAsync(psi[iy0], 0, sizeof(cs) * (size_t)NK * Md * n, stream));
    }
    bool ffz = true;
    if (o) {
      ffz = upload_omega(omg[iy0], o);
    } else {
      HIPCHK(hipMemsetAsync(omg[iy0], 0, sizeof(cs) * (size_t)NK * Md * Md, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    if (!ffz) ff_dirty = true;
    set_ff(ffz);
    have_state = true;
    y0xi_valid = false;
    state_anti = anti_in;
    update_tri();
  }
  void state_get(double* s, double* p, double* o) override {
    require(have_state, "no resident state", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    if (s) download(s, sig[iy0], (size_t)NK * n * n);
    if (p) download_psi(p, psi[iy0]);
    if (o) download_omega(o, omg[iy0]);
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

  void set_layout(int mode) override {
    require(!have_consts, "the layout is fixed once the constants are set", ZH_ESTATE);
    require(mode >= 0 && mode <= 2, "layout: 0 auto, 1 reference order, 2 active-first");
    layout_req = mode;
  }
  int layout_info(int* out) override {
    out[0] = pruned ? 1 : 0; out[1] = ff ? 1 : 0; out[2] = Md; out[3] = ML; out[4] = nc_real;
    return 5;
  }
  void set_sparse(int mode) override {
    require(mode >= 0 && mode <= 2, "H products: 0 auto, 1 dense, 2 row-compressed");
    const int old = sparse_req;
    sparse_req = mode;
    if (have_consts) {
      try {
        drop_graphs();
        install_h(H0h.data());
        h_dirty = false;
      } catch (...) {
        sparse_req = old;
        install_h(H0h.data());
        throw;
      }
    }
  }
  void set_fusion(int mode) override {
    require(mode >= 0 && mode <= 2, "fusion: 0 auto, 1 off, 2 on");
    fusion_req = mode;
    drop_graphs();
  }
  void set_omgxi(int mode) override {
    require(mode >= 0 && mode <= 2, "Omega xi_c kernel: 0 auto, 1 tiled GEMM, 2 streaming");
    xi_req = mode;
    drop_graphs();
  }
  void set_psi_kernel(int mode) override {
    require(mode >= 0 && mode <= 2, "Psi kernel: 0 auto, 1 split, 2 one pass");
    if (mode == 2)
      require(!h_sparse || (psi_wmax > 0 && psi_wmax <= kPsiFusedWMax),
              "one-pass Psi kernel: the rows of H reach more than " +
                  std::to_string(kPsiFusedWMax) + " orbitals per 64-orbital tile");
    psi_req = mode;
    drop_graphs();
  }
  int psi_kernel_info() const override { return psi_fused() ? 1 : 0; }
  void set_fsal(int mode) override {
    require(mode == 0 || mode == 1, "first same as last: 0 on, 1 off");
    fsal_req = mode;
    drop_graphs();
  }
  void fsal_info(long* taken, int* on) const override {
    if (taken) *taken = fsal_taken;
    if (on) *on = fsal_on() ? 1 : 0;
  }
  void omgxi_info(int* out) const override {
    out[0] = xi_stream() && !fusable() ? 1 : 0;
    out[1] = xi_ok ? 1 : 0;
  }
  void hamiltonian_info(int* out) override {
    out[0] = h_real ? 1 : 0; out[1] = h_sparse ? 1 : 0; out[2] = ell_R;
  }


  // Stage inputs of σ and Ψ (Ω's is formed inside its own right-hand side kernel);
  // terms with a zero coefficient are left out.
  // do_sig / do_psi: the σ and the Ψ part (the attempts issue σ's for stage s on the main
  // stream right after stage s - 1's σ branch, ahead of that stage's join)
  void combine(int s, bool first, double2* sig_dst, bool last, bool do_sig = true,
               bool do_psi = true) {
    // one workgroup per row up to 4096 (at n = 1000, M = 4000 the concurrent attempts ran
    // 0.6% faster with 4096 than with 2048 for the complex64 Ψ rows; 8192 and 16384 the same)
    auto blocks = [](long rows) { return (int)std::max(1L, std::min(4096L, rows)); };
    int nt = 0, slot[6];
    double coef[6];
    if (!first)
      for (int j = 0; j < s; ++j)
        if (tab.a[s][j] != 0.0) { slot[nt] = j; coef[nt] = tab.a[s][j]; ++nt; }
    // the k_1 term (slot 0) is scaled by the instance's r1 (first same as last), and
    // the first stage's copy skips instances whose first stage is skipped
    const int k1 = (nt > 0 && slot[0] == 0) ? 0 : -1;
    const StageParams* sps = sp + s * NI;
    timed(KC_COMBINE, [&] {
      if (!first && do_sig) {
        CombineArgs<double2> gs{};
        gs.y0 = sig[iy0];
        gs.nterms = nt;
        for (int j = 0; j < nt; ++j) { gs.k[j] = sig[slot[j]]; gs.a[j] = coef[j]; }
        gs.d1 = sig_dst; gs.d2 = nullptr;
        gs.sp = sps; gs.ki = KI; gs.k1 = k1; gs.rows_b = n;
        if (s == 6 && e6_sig) {
          gs.e_out = sig[1];
          for (int j = 0; j < nt; ++j) gs.ect[j] = tab.ct[slot[j]];
        }
        MapDense ms{NK * n, n, n, n, n};
        launch_combine(gs, ms, blocks((long)NK * n), stream);
      }
      if (!do_psi) return;
      CombineArgs<cs> gp{};
      gp.y0 = psi[iy0];
      gp.nterms = nt;
      for (int j = 0; j < nt; ++j) { gp.k[j] = psi[slot[j]]; gp.a[j] = coef[j]; }
      gp.d1 = Ap;
      gp.d2 = last ? psi[iyn] : nullptr;
      gp.sp = sps; gp.ki = KI; gp.k1 = k1; gp.rows_b = Md;
      if (s == 6 && e6_psi) {
        gp.e_out = psi[1];
        for (int j = 0; j < nt; ++j) gp.ect[j] = tab.ct[slot[j]];
      }
      MapDense mp{NK * Md, n, n, n + nc, n};
      launch_combine(gp, mp, blocks((long)NK * Md), stream);
    });
    HIPCHK(hipGetLastError());
  }

  // Stage inputs on the coupled orbitals, the Hubbard diagonal and the per-mode
  // shifts of the Ω outer term (k_gather), on the main stream.
  void gather(const StageParams* sps, const double2* sig_in, bool hub) {
    GatherArgs<cs> g;
    g.NK = NK; g.n = n; g.M = Md; g.nc = nc;
    g.Ap = Ap; g.ldAp = n + nc; g.sAp = (long)Md * (n + nc);
    g.Aomg = Aomg; g.Bomg = Bomg; g.Bp = Bp; g.sig = sig_in; g.cidx = cidx;
    g.dvec = dvec; g.n0 = n0; g.U = U; g.hbar = hbar; g.hubbard = hub && U != 0.0;
    g.sp = sps; g.ki = KI; g.chip = chip; g.chim = chim; g.lead = lead; g.u = uvec; g.v = vvec;
    // one element per thread of the largest job up to 4096 workgroups (at n = 1000,
    // M = 4000: 1.3% more steps/s in both precisions than 512 workgroups)
    const long most = std::max(std::max((long)NK * Md * nc, (long)NK * nc * n),
                               (long)NK * std::max(Md, n));
    const int gx = (int)std::max(1L, std::min(4096L, (most + 255) / 256));
    hipLaunchKernelGGL(k_gather<cs>, dim3(gx, 4), dim3(256), 0, stream, g);
    HIPCHK(hipGetLastError());
  }

  // Ω side of one stage: stage input y0 + Σ a_j k_j, where Ω_s goes (ynew, the
  // last stage), whether k_s is stored and whether the error norm is accumulated.
  struct OmegaStage {
    const cs* y0 = nullptr;
    int nterms = 0;
    const cs* k[6] = {nullptr};
    double a[6] = {0};
    double ct[6] = {0};
    double ct_self = 0.0;
    cs* ynew = nullptr;
    bool store_k = true;
    bool te = false;
    // fused Ωξ (attempts in mixed precision): the partials buffer this stage's Ω kernel
    // writes, and the slot parity of y0 (PredY[par] = y0 ξ_c)
    cs* wpart = nullptr;
    int par = 0;
    // attempts: y0 ξ_c is resident in PredY[par] and the last stage leaves y_new ξ_c in
    // PredY[1 - par]; otherwise (a single evaluation) the first stage forms y0 ξ_c
    bool y0xi_resident = false;
    // attempts: the stage's parameters (r1 of the k_1 term, k[k1]) and where k_s goes
    // when not to its own slot (the last stage's k_7 over the k_1 slot)
    const StageParams* sps = nullptr;
    int k1 = -1;
    cs* kout = nullptr;
  };

  // Ωξ product of stage t (X = the k of stage t - 1, omg[t - 1]) into wpart(t), on st:
  // the stream of the kernel that wrote X, which stage t's Ψ branch then continues
  void issue_xi(int t, hipStream_t st) {
    timed(KC_GEMM_OMGXI, [&] { omgxi_run<kRM3>(omg[t - 1], st, true, wpart(t)); });
  }

  // Right-hand side of stage s from (sig_in, Ap, Ω stage os) into slot kslot.
  void rhs_stage(int s, const double2* sig_in, const OmegaStage& os, int kslot, double* jout,
                 bool hub, int sig_next = -1) {
    const StageParams* sps = sp + s * NI;
    timed(KC_GATHER, [&] {
      gather(sps, sig_in, hub);
      if (hub) apply_model(sps, stream);  // H(t_s, σ_s) of the attempts
    });
    // the three branches (serial on the main stream when profiling, so that the
    // per-class event timings stay meaningful).  The two side streams take turns: stage
    // s's Ω kernel runs on aux[s & 1] and is followed there by the next stage's Ωξ
    // product of its k (early, overlapping the rest of this stage and the next one's
    // stage combination and gathers), which that stage's Ψ branch continues; stage s's
    // Ψ branch runs on the other one, after this stage's product.  Three streams in all,
    // one hardware queue each.
    const bool conc = !profile;
    hipStream_t s_psi = stream, s_omg = stream;
    if (conc) {
      s_omg = aux[s & 1];
      s_psi = aux[(s + 1) & 1];
      HIPCHK(hipEventRecord(ev_fork, stream));
      HIPCHK(hipStreamWaitEvent(s_omg, ev_fork, 0));
      HIPCHK(hipStreamWaitEvent(s_psi, ev_fork, 0));
    }
    stage_omega(os, kslot, s_omg);
    if (conc) HIPCHK(hipEventRecord(ev_join[0], s_omg));  // the join does not wait for
    if (xi_early && s + 1 < 7) issue_xi(s + 1, s_omg);    // the early product
    stage_psi(s, os, (s == 6 && k7_psi0) ? 0 : kslot, s_psi);
    if (te_stage && s == 6) te_launch(false, true, s_psi);
    if (conc) HIPCHK(hipEventRecord(ev_join[1], s_psi));
    stage_sigma(s, sig_in, (s == 6 && k7_sig0) ? 0 : kslot, jout, stream);
    if (te_stage && s == 6) te_launch(true, false, stream);
    // the σ stage input of the next stage, as soon as this stage's σ derivative is there
    // (this stage's other branches do not read σ's stage input)
    if (sig_next > 0)
      combine(sig_next, false, sig_next == 6 ? sig[iyn] : sig[9], sig_next == 6, true, false);
    if (conc) {
      HIPCHK(hipStreamWaitEvent(stream, ev_join[0], 0));
      HIPCHK(hipStreamWaitEvent(stream, ev_join[1], 0));
    }
  }

  // σ branch: Π_c, currents (stage 0), Hσ with the σ right-hand side.
  // jout: the currents of the stage input go there (the first stage: J of the attempt;
  // the last: Jn, the next attempt's when it reuses k_7), or none
  void stage_sigma(int s, const double2* sig_in, int kslot, double* jout, hipStream_t st) {
    const StageParams* sps = sp + s * NI;
    const double ih = 1.0 / hbar;
    if (nc > 0) {
      timed(KC_GEMM_PI, [&] {
        // Π_c = ψᵀ Ixi_c  (n x Md) (Md x nc), FP64, split over K
        GemmArgs<R, double> g{n, nc, Md, Ap, n + nc, (long)Md * (n + nc), Ixic, nc,
                              (long)Md * nc, kPiParts};
        g.skip = skip_now; g.skip_ki = KI;
        EpiStore<double2> e{Pic, nc, (long)n * nc, 1.0, 0.0};
        KSegLead ks{gmask, lrange, (nc + kCGroup - 1) / kCGroup, ML, 0};
        // FP64 MFMAs: three real products per complex tile product
        HIPCHK((launch_sel<double, R, double, true, false, 3>(tsel[TS_PI], g, NK, e, st, ks,
                                                              CandPi())));
      });
    }
    const bool ovl = overlap && nc > 0;
    if (ovl) {
      timed(KC_GEMM_PI, [&] {
        // Q = Π_c S⁻¹[c, :]  (n x nc)(nc x n), FP64  (Zandpack/mpi/nozand:489-497)
        hipLaunchKernelGGL(k_sum_parts, dim3(512), dim3(256), 0, st, NK, kPiParts,
                           (long)n * nc, (const double2*)Pic, Pis);
        HIPCHK(hipGetLastError());
        GemmArgs<double, double> g{n, n, nc, Pis, nc, (long)n * nc, Sc, n, (long)nc * n};
        EpiStore<double2> e{Qs, n, (long)n * n, 1.0, 0.0};
        HIPCHK((launch_sel<double, double, double, false, false, 3>(tsel[TS_HSIG], g, NK, e, st,
                                                                    KFull(), CandHsig())));
      });
    }
    if (jout) {
      // with no coupled orbitals the reductions run over empty ranges and give J = 0; an
      // instance whose first stage is skipped takes Jn
      timed(KC_CURRENTS, [&] {
        hipLaunchKernelGGL(k_currents<cs>, dim3(kCurChunks, NL, NK), dim3(256), 0, st, NK, Md,
                           nc, (const int*)lead, (const cs*)Ap, (long)(n + nc),
                           (long)Md * (n + nc), (const int*)cidx, (const double2*)Ixic, Jpart,
                           sps, KI);
        hipLaunchKernelGGL(k_currents_final, dim3(1), dim3(((NL * NK + 63) / 64) * 64), 0, st,
                           NL * NK, (const double*)Jpart, ih * ih, jout, NK, sps, KI,
                           (const double*)Jn);
      });
      HIPCHK(hipGetLastError());
    }
    timed(KC_GEMM_HSIG, [&] {
      GemmArgs<double, double> g{n, n, n, Hcur, n, (long)n * n, sig_in, n, (long)n * n};
      g.skip = skip_now; g.skip_ki = KI;
      EpiSigma e{G, sig_in, dvec_epi(), Pic, overlap ? cpos_none : cpos, n, nc, ih, kPiParts,
                 ovl ? Qs : nullptr, skip_now, KI};
      if (h_sparse) {
        constexpr int rb = 1;  // rows per workgroup: more, shorter workgroups keep more
                               // gathers in flight (C3: 8 -> 1 rows, 50 -> 27 us)
        const dim3 grid((n + 255) / 256, (n + rb - 1) / rb, NK);
        if (ell_R == 4)
          hipLaunchKernelGGL((k_hsig_ell<rb, 4, EpiSigma>), grid, dim3(256), 0, st, sig_in, ell(), e);
        else if (ell_R == 8)
          hipLaunchKernelGGL((k_hsig_ell<rb, 8, EpiSigma>), grid, dim3(256), 0, st, sig_in, ell(), e);
        else
          hipLaunchKernelGGL((k_hsig_ell<rb, 0, EpiSigma>), grid, dim3(256), 0, st, sig_in, ell(), e);
        HIPCHK(hipGetLastError());
      } else if (h_real) {  // real H: the A operand is Re H
        g.Areal = Hre; g.ldar = n; g.sAr = (long)n * n;
        HIPCHK((launch_sel<double, double, double, false, false, 1>(tsel[TS_HSIG], g, NK, e, st,
                                                                    KFull(), CandHsigR())));
      } else {
        HIPCHK((launch_sel<double, double, double, false, false, 3>(tsel[TS_HSIG], g, NK, e, st,
                                                                    KFull(), CandHsig())));
      }
    });
    timed(KC_SIGSYM, [&] {
      dim3 grid((n + 31) / 32, (n + 31) / 32, NK);
      hipLaunchKernelGGL(k_sigma_sym, grid, dim3(32, 8), 0, st, n, (const double2*)G,
                         sig[kslot], sps, KI, ih, sig_in,
                         (const double2*)(diss.kind != ZH_DISS_NONE ? sig_eq : nullptr));
    });
    HIPCHK(hipGetLastError());
  }

  // Ψ branch: Ω_s ξ_c by linearity, then the Ψ right-hand side.
  void stage_psi(int s, const OmegaStage& os, int kslot, hipStream_t st) {
    const StageParams* sps = sp + s * NI;
    const double ih = 1.0 / hbar;
    cs* Wc = Wst;  // Ω_s ξ_c of this stage
    if (nc > 0 && os.wpart) {
      // fused: y0 ξ_c is resident and the earlier stages' Ω kernels left k_j ξ_c as
      // per-chunk partials; stage 0 needs y0 ξ_c alone
      if (s == 0) {
        Wc = PredY[os.par];
      } else {
        timed(KC_WSUM, [&] {
          const OmegaFusedArgs fa = fused_args(kFusedChunk[tsel[TS_OMEGA]]);
          WSumArgs<cs> w{};
          w.part = Wpart[(s - 1) % 2];
          w.pred[0] = PredY[os.par];
          for (int j = 1; j < 7; ++j) w.pred[j] = Pred[j];
          w.W = Wst;
          for (int j = 0; j < 6; ++j) w.a[j] = (j < s) ? tab.a[s][j] : 0.0;
          w.s = s; w.parts = fa.SA + fa.SP; w.NK = NK; w.M = Md; w.nc = nc;
          w.parts_pass = ff ? fa.SA : fa.SA + fa.SP;
          w.ML = ML;
          hipLaunchKernelGGL(k_wsum<cs>, dim3(1024), dim3(256), 0, st, w);
          HIPCHK(hipGetLastError());
        });
      }
    } else if (nc > 0 && s == 0 && os.y0xi_resident) {
      // y0 ξ_c is resident (PredY[par]): the previous attempt's last stage left it (y_new
      // ξ_c after an accepted step), or it was formed before the attempt
      Wc = PredY[os.par];
    } else if (nc > 0) {
      // X ξ_c  (Md x Md)(Md x nc) for X = y0 (a single evaluation's first stage) or the
      // previous stage's k, then Ω_s ξ_c = y0 ξ_c + Σ_j a_sj k_j ξ_c: Ω is streamed once
      // per stage and no stage-input copy of it is needed.  The last stage's Ω_7 ξ_c is
      // y_new ξ_c and goes to PredY[1 - par], the next attempt's y0 ξ_c if this one is
      // accepted.
      const bool res = os.y0xi_resident;
      WSumArgs<cs> w{};
      if (xi_early && s >= 1) {
        // issued as soon as the previous stage's Ω kernel wrote its k, earlier on this
        // stream (issue_xi)
        w.part = wpart(s);
      } else {
        timed(KC_GEMM_OMGXI, [&] { omgxi_run<kRM3>((s == 0) ? os.y0 : omg[s - 1], st); });
        w.part = W;
      }
      for (int j = 0; j < 7; ++j) w.pred[j] = Pred[j];
      if (res) w.pred[0] = PredY[os.par];
      w.W = (res && s == 6) ? PredY[1 - os.par] : Wst;
      for (int j = 0; j < 6; ++j) w.a[j] = (j < s) ? tab.a[s][j] : 0.0;
      w.s = s; w.parts = kWParts; w.NK = NK; w.M = Md; w.nc = nc;
      w.sp = os.sps; w.ki = KI;
      timed(KC_WSUM, [&] {
        hipLaunchKernelGGL(k_wsum<cs>, dim3(1024), dim3(256), 0, st, w);
        HIPCHK(hipGetLastError());
      });
      Wc = w.W;
    }
    const bool ovl = overlap && nc > 0;
    if (ovl) {
      timed(KC_GEMM_PSI, [&] {
        // (ħ GL_P ξ_c + Ω_s ξ_c/ħ) S⁻¹[:, c]ᵀ into the k slot  (Md x nc)(nc x n); the
        // Ψ epilogue adds it to ψ(S⁻¹H)ᵀ + ħ ΔPξ_c σ_cᵀ  (Zandpack/mpi/nozand:1003, 1208-1210)
        const long total = (long)NK * Md * nc;
        hipLaunchKernelGGL(k_overlap_y<cs>, dim3(1024), dim3(256), 0, st, total, nc,
                           (const double2*)glp, (const cs*)xic, (const cs*)Wc, Yov, hbar);
        HIPCHK(hipGetLastError());
        GemmArgs<R, R> g{Md, n, nc, Yov, nc, (long)Md * nc, Bs, n, (long)nc * n};
        g.skip = skip_now; g.skip_ki = KI;
        EpiStore<cs> e{psi[kslot], n, (long)Md * n, 1.0, 0.0};
        HIPCHK((launch_sel<R, R, R, false, false, kRM3>(tsel[TS_PSI], g, NK, e, st, KFull(),
                                                     typename Cands<R>::psi())));
      });
    }
    timed(KC_GEMM_PSI, [&] {
      if (psi_fused() && !ovl) {
        // one pass: stage-Ψ tile rows in LDS, ħ ΔP(σξ) on the VALU, gather, epilogue
        PsiFusedArgs<cs> f{};
        f.out = psi[kslot];
        f.Ap = Ap; f.sAp = (long)Md * (n + nc); f.ldAp = n + nc;
        f.Bc = Bp + (long)n * n; f.sBc = (long)(n + nc) * n;
        f.xic = xic; f.W = Wc; f.glp = glp; f.chip = chip; f.dvec = dvec_epi();
        f.lead = lead; f.cpos = cpos; f.t2 = t2; f.span = psi_span;
        f.col = ell_col; f.val = ell_val; f.R = ell_R;
        f.sp = sps; f.ki = KI;
        f.Md = Md; f.n = n; f.nc = nc; f.acc0_rows = nc > 0 ? g0_rows() : 0;
        f.inv_hbar = ih; f.inv_hbar2 = ih * ih;
        if (psi_nx && s >= 1 && s + 1 < 7) {
          // stage s + 1's input y0 + Σ_j a_{s+1,j} k_j, this stage's k from registers
          const int t = s + 1;
          f.nx_out = apb(t);
          f.nx_y = (t == 6) ? psi[iyn] : nullptr;
          f.nx_y0 = psi[iy0];
          for (int j = 0; j < t; ++j)
            if (tab.a[t][j] != 0.0) {
              require(f.nx_n < 6, "stage combination: too many terms");
              if (j == s) f.nx_self = f.nx_n;
              if (j == 0) f.nx_k1 = f.nx_n;
              f.nx_k[f.nx_n] = psi[j];
              f.nx_a[f.nx_n] = tab.a[t][j];
              f.nx_ct[f.nx_n] = tab.ct[j];
              ++f.nx_n;
            }
          if (t == 6 && e6_psi) f.nx_e = psi[1];
        }
        constexpr int TR = 16;  // mode rows per workgroup (C3: 16 beat 8, 32 and 64)
        constexpr bool XL = kPsiXL<cs>;
        const size_t lds = (size_t)TR * (psi_wmax + (XL ? 64 : 0)) * sizeof(cs);
        const dim3 grid((n + 63) / 64, (Md + TR - 1) / TR, NK);
        auto go = [&](auto nxt) {
          constexpr bool NXT = decltype(nxt)::value;
          if (ell_R == 4)
            hipLaunchKernelGGL((k_psi_fused<cs, TR, 4, XL, NXT>), grid, dim3(256), lds, st, f);
          else if (ell_R == 8)
            hipLaunchKernelGGL((k_psi_fused<cs, TR, 8, XL, NXT>), grid, dim3(256), lds, st, f);
          else
            hipLaunchKernelGGL((k_psi_fused<cs, TR, 0, XL, NXT>), grid, dim3(256), lds, st, f);
        };
        if (f.nx_out) go(std::true_type());
        else go(std::false_type());
        HIPCHK(hipGetLastError());
      } else if (h_sparse) {
        // [ΔPξ_c] [ħ σ_cᵀ] into the k slot (MFMA, K = nc), then the gather ψHᵀ with
        // the Ψ epilogue adding it
        EpiStore<cs> e0{psi[kslot], n, (long)Md * n, 1.0, 0.0};
        if (nc > 0) {
          GemmArgs<R, R> g0{Md, n, nc, Ap + n, n + nc, (long)Md * (n + nc), Bp + (long)n * n, n,
                            (long)(n + nc) * n};
          g0.skip = skip_now; g0.skip_ki = KI;
          HIPCHK((launch_sel<R, R, R, false, false, kRM3>(tsel[TS_PSI], g0, NK, e0, st,
                                                          kseg_g0(),
                                                          typename Cands<R>::psi())));
        }
        EpiPsi<cs> e{psi[kslot], nc > 0 ? psi[kslot] : nullptr, Ap, n + nc, xic, Wc, glp, chip,
                     dvec_epi(), lead, cpos, sps, Md, n, nc, ih, ih * ih, 1, g0_rows(), KI};
        const size_t lds = (size_t)hpsi_rb * n * sizeof(cs);
        const long sAp = (long)Md * (n + nc);
        const dim3 grid((Md + hpsi_rb - 1) / hpsi_rb, NK);
        if (hpsi_rb == 4)
          hipLaunchKernelGGL((k_psi_ell<cs, 4>), grid, dim3(256), lds, st, e, sAp, ell());
        else if (hpsi_rb == 2)
          hipLaunchKernelGGL((k_psi_ell<cs, 2>), grid, dim3(256), lds, st, e, sAp, ell());
        else
          hipLaunchKernelGGL((k_psi_ell<cs, 1>), grid, dim3(256), lds, st, e, sAp, ell());
        HIPCHK(hipGetLastError());
      } else if (h_real) {
        // [ΔPξ_c] [ħ σ_cᵀ] into the k slot, then ψ Re(H)ᵀ with the epilogue adding it
        EpiStore<cs> e0{psi[kslot], n, (long)Md * n, 1.0, 0.0};
        if (nc > 0) {
          GemmArgs<R, R> g0{Md, n, nc, Ap + n, n + nc, (long)Md * (n + nc), Bp + (long)n * n, n,
                            (long)(n + nc) * n};
          g0.skip = skip_now; g0.skip_ki = KI;
          HIPCHK((launch_sel<R, R, R, false, false, kRM3>(tsel[TS_PSI], g0, NK, e0, st,
                                                          kseg_g0(),
                                                          typename Cands<R>::psi())));
        }
        GemmArgs<R, R> g{Md, n, n, Ap, n + nc, (long)Md * (n + nc), Bp, n, (long)(n + nc) * n};
        g.Breal = HreT; g.ldbr = n; g.sBr = (long)n * n;
        g.skip = skip_now; g.skip_ki = KI;
        EpiPsi<cs> e{psi[kslot], nc > 0 ? psi[kslot] : nullptr, Ap, n + nc, xic, Wc, glp, chip,
                     dvec_epi(), lead, cpos, sps, Md, n, nc, ih, ih * ih, 1, g0_rows(), KI};
        HIPCHK((launch_sel<R, R, R, false, false, 2>(tsel[TS_PSI], g, NK, e, st, KFull(),
                                                     typename Cands<R>::psiR())));
      } else {
        // [ψ | ΔPξ_c] [H0ᵀ ; ħ σ_cᵀ]  (Md x (n+nc)) ((n+nc) x n); with an overlap
        // H0 is S⁻¹H and the coupled-column terms come in through acc0
        GemmArgs<R, R> g{Md, n, n + nc, Ap, n + nc, (long)Md * (n + nc), Bp, n,
                         (long)(n + nc) * n};
        g.skip = skip_now; g.skip_ki = KI;
        EpiPsi<cs> e{psi[kslot], ovl ? psi[kslot] : nullptr, Ap, n + nc, xic, Wc, glp, chip,
                     dvec_epi(), lead, overlap ? cpos_none : cpos, sps, Md, n, nc, ih, ih * ih, 1, Md,
                     KI};
        HIPCHK((launch_sel<R, R, R, false, false, kRM3>(tsel[TS_PSI], g, NK, e, st, KFull(),
                                                     typename Cands<R>::psi())));
      }
    });
  }

  // Ω branch: sources, stage input, outer term and (last stage) the error norm.
  // X ξ_c (Md x Md)(Md x nc) into the split-K partials W (X = y0 or a k slot).  With Ω
  // held as its block upper triangle each row block also runs the mirrored part,
  // -X[J < I, I]^H ξ_c, in the same K loop (KSegTri).
  template <int NG> void xi_launch(const cs* X, hipStream_t st, cs* part) {
    XiArgs<cs> a{};
    a.X = X; a.sX = (long)Md * Md;
    a.xi = xic; a.sXi = (long)Md * nc;
    a.part = part; a.sPart = (long)Md * nc;
    a.Md = Md; a.nc = nc; a.parts = kWParts; a.tri = tri_on ? 1 : 0; a.ML = ML; a.ff = ff ? 1 : 0;
    a.nitem = (int)xi_items.size();
    for (int i = 0; i < a.nitem; ++i) a.item[i] = xi_items[i];
    for (int l = 0; l < NL; ++l) a.lr[l] = lrange_h[l];
    const dim3 grid((Md + 63) / 64, a.nitem, NK * kWParts);
    // 4 chunk buffers (at most 64 KB of LDS per workgroup), 2 workgroups per CU (both
    // element types).  At n = 1000, M = 4000 the product alone is faster with 3 buffers
    // at 3 per CU (complex64 117 against 134 µs, complex128 187 against 212), but in the
    // attempts, where it shares the CUs with the HBM-bound kernels of the other branches,
    // 2 per CU gave 1.2% (mixed) and 0.6% (FP64) more steps/s; 1 per CU with 4 or 6
    // buffers lost (FP64 225 against 245 steps/s)
    hipLaunchKernelGGL((k_xi_stream<cs, NG, 4, 2>), grid, dim3(256), 0, st, a);
    HIPCHK(hipGetLastError());
  }
  // stream: the attempts' products may take the streaming kernel (the steady-state
  // operator keeps the four-multiplication GEMM)
  // part: the partials' buffer (W unless given)
  template <int RM> void omgxi_run(const cs* X, hipStream_t st, bool stream = true,
                                   cs* part = nullptr) {
    if (!part) part = W;
    if (stream && xi_stream()) {
      require(xi_ng >= 1 && xi_ng <= 4 && (int)lrange_h.size() == NL && NL <= kXiMaxLeads,
              "streaming Omega xi_c kernel: bad column tiles");
      if (xi_ng == 1) xi_launch<1>(X, st, part);
      else if (xi_ng == 2) xi_launch<2>(X, st, part);
      else if (xi_ng == 3) xi_launch<3>(X, st, part);
      else xi_launch<4>(X, st, part);
      return;
    }
    GemmArgs<R, R> g{Md, nc, Md, X, Md, (long)Md * Md, xic, nc, (long)Md * nc, kWParts};
    EpiStore<cs> e{part, nc, (long)Md * nc, 1.0, 0.0};
    KSegLead ks{gmask, lrange, (nc + kCGroup - 1) / kCGroup, ML, ff ? 1 : 0};
    if (!tri_on) {
      HIPCHK((launch_sel<R, R, R, false, false, RM>(tsel[TS_OMGXI], g, NK, e, st, ks,
                                                    typename Cands<R>::omgxi())));
      return;
    }
    int bm = 0, bn = 0;
    tile_dims(tsel[TS_OMGXI], typename Cands<R>::omgxiT(), bm, bn);
    require(bm == 64, "anti-Hermitian Omega: the X xi_c tiles must be 64 rows");
    HIPCHK((launch_sel<R, R, R, false, false, RM>(tsel[TS_OMGXI], g, NK, e, st,
                                                  KSegTri<KSegLead>{ks},
                                                  typename Cands<R>::omgxiT())));
  }

  void stage_omega(const OmegaStage& os, int kslot, hipStream_t st) {
    if (os.wpart) {
      stage_omega_fused(os, kslot, st);
      return;
    }
    stage_omega_gemm(os, kslot, st);
  }

  void stage_omega_gemm(const OmegaStage& os, int kslot, hipStream_t st) {
    timed(KC_GEMM_OMEGA, [&] {
      // [ψ_c | ΔPξ_c] [ΔM Ixi_c | ψ_c*]ᵀ  (Md x 2nc)(2nc x Md), with the stage input
      // formed, the outer-subtraction term applied and (last stage) the error norm
      // accumulated in the epilogue
      KSegOmega ks{t1, t2, nc, ML, ff ? 1 : 0, (Md + kTile - 1) / kTile};
      ks.tri = tri_on ? 1 : 0;
      GemmArgs<R, R> g{Md, Md, 2 * nc, Aomg, 2 * nc, (long)Md * 2 * nc, Bomg, 2 * nc,
                       (long)Md * 2 * nc};
      g.xcd_order = 0;  // the skipped passive block would unbalance contiguous XCD bands
      g.skip = skip_now; g.skip_ki = KI;
      auto fill = [&](auto& e) {
        e.tri = tri_on ? 1 : 0;
        e.out = os.kout ? os.kout : (os.store_k ? omg[kslot] : nullptr);
        e.sp = os.sps; e.kper = KI; e.k1 = os.k1;
        e.y0 = os.y0;
        e.nterms = os.nterms;
        for (int j = 0; j < 6; ++j) { e.k[j] = os.k[j]; e.a[j] = os.a[j]; e.ct[j] = os.ct[j]; }
        e.ynew = os.ynew;
        e.u = uvec; e.v = vvec; e.M = Md;
        e.ct_self = os.ct_self;
        e.te_part = te_part + 2 * kTEBlocks * NI;
        e.rtol = tol_rtol; e.atol = tol_atol; e.scaled = tol_scaled ? 1 : 0;
      };
      if (os.te) {
        EpiOmegaRK<cs, true> e{};
        fill(e);
        HIPCHK((launch_sel<R, R, R, false, true, 0>(tsel[TS_OMEGA], g, NK, e, st, ks,
                                                    typename Cands<R>::omega())));
      } else {
        EpiOmegaRK<cs, false> e{};
        fill(e);
        HIPCHK((launch_sel<R, R, R, false, true, 0>(tsel[TS_OMEGA], g, NK, e, st, ks,
                                                    typename Cands<R>::omega())));
      }
    });
  }

  // Ω branch with the Ωξ product fused (mixed precision, attempts): the kernel of
  // zh_omega.hpp on chunks of kFusedChunk[tsel[TS_OMEGA]] columns.
  template <typename E> void launch_fused(const OmegaFusedArgs& fa, const E& e, hipStream_t st) {
    KSegOmega ks{t1, t2, nc, ML, ff ? 1 : 0, (Md + kTile - 1) / kTile};
    const dim3 grid(fused_items(fa), NK);
    const int ng = (nc + 15) / 16;
    if (ng <= 2) hipLaunchKernelGGL((k_omega_fused<2, E>), grid, dim3(kThreads), 0, st, fa, e, ks);
    else if (ng <= 4) hipLaunchKernelGGL((k_omega_fused<4, E>), grid, dim3(kThreads), 0, st, fa, e, ks);
    else if (ng <= 6) hipLaunchKernelGGL((k_omega_fused<6, E>), grid, dim3(kThreads), 0, st, fa, e, ks);
    else hipLaunchKernelGGL((k_omega_fused<8, E>), grid, dim3(kThreads), 0, st, fa, e, ks);
    HIPCHK(hipGetLastError());
  }
  void stage_omega_fused(const OmegaStage& os, int kslot, hipStream_t st) {
    if constexpr (std::is_same<R, float>::value) {
      timed(KC_GEMM_OMEGA, [&] {
        OmegaFusedArgs fa = fused_args(kFusedChunk[tsel[TS_OMEGA]]);
        require(fa.SA + fa.SP <= wpart_max, "fused Omega kernel: partial buffer too small");
        const long t32 = (Md + 31) / 32;
        require((long)fused_items(fa) * NK <= t32 * t32 * NK,
                "fused Omega kernel: more error-norm partials than te_part holds");
        fa.A = Aomg; fa.B = Bomg; fa.xic = xic; fa.txi = txi; fa.wpart = os.wpart;
        auto fill = [&](auto& e) {
          e.out = os.store_k ? omg[kslot] : nullptr;
          e.y0 = os.y0;
          e.nterms = os.nterms;
          for (int j = 0; j < 6; ++j) { e.k[j] = os.k[j]; e.a[j] = os.a[j]; e.ct[j] = os.ct[j]; }
          e.ynew = os.ynew;
          e.u = uvec; e.v = vvec; e.M = Md;
          e.ct_self = os.ct_self;
          e.te_part = te_part + 2 * kTEBlocks * NI;
          e.rtol = tol_rtol; e.atol = tol_atol; e.scaled = tol_scaled ? 1 : 0;
        };
        if (os.te) {
          EpiOmegaRK<cs, true> e{};
          fill(e);
          launch_fused(fa, e, st);
        } else {
          EpiOmegaRK<cs, false> e{};
          fill(e);
          launch_fused(fa, e, st);
        }
      });
    } else {
      require(false, "the fused Omega kernel is FP32 only");
    }
  }

  // y0 ξ_c of the resident state into PredY[par] (the first fused attempt from a new
  // state): the Ωξ product of the unfused path on y0.
  void compute_y0xi(int par, hipStream_t st) {
    omgxi_run<kRM3>(omg[iy0], st);
    WSumArgs<cs> w{};
    w.part = W;
    w.pred[0] = PredY[par];
    for (int j = 1; j < 7; ++j) w.pred[j] = Pred[j];
    w.W = Wst;
    w.s = 0; w.parts = kWParts; w.NK = NK; w.M = Md; w.nc = nc;
    hipLaunchKernelGGL(k_wsum<cs>, dim3(1024), dim3(256), 0, st, w);
    HIPCHK(hipGetLastError());
  }

  // Error-norm partials of σ (do_sig) and Ψ (do_psi) on st.  The attempts issue each on
  // its branch's stream right after the last stage's σ / Ψ derivative, so that they
  // overlap the last Ω kernel instead of following it.
  void te_launch(bool do_sig, bool do_psi, hipStream_t st) {
    timed(KC_TE, [&] {
      // slots with a non-zero error weight (b5 - b4 vanishes for k_2)
      // (the Ω part is accumulated by the last stage's Ω kernel)
      TEArgs<double2> ts{};
      TEArgs<cs> tp{};
      int nw = 0;
      int last_w = -1;
      for (int j = 0; j < 7; ++j) {
        if (tab.ct[j] == 0.0) continue;
        if (j == 0) ts.k1 = tp.k1 = nw;
        ts.k[nw] = sig[j]; tp.k[nw] = psi[j];
        ts.ct[nw] = tp.ct[nw] = tab.ct[j];
        last_w = j;
        ++nw;
      }
      require(nw == 6, "error weights: expected six non-zero entries");
      ts.sp = tp.sp = sp;  // the first stage's entries: r1 per instance
      if (fsal_on()) {
        // k_7 (the last weighted slot) over the k_1 slot: the next attempt's k_1
        require(last_w == 6 && ts.k1 == 0, "first same as last: k_7 must carry an error weight");
        ts.kdst = sig[0]; tp.kdst = psi[0];
      }
      // E6 (weighted sum of k_1 .. k_6, r1 applied) in the k_2 slot, then k_7
      auto two = [&](auto& t, auto* e6, auto* k7) {
        t.k[0] = e6; t.k[1] = k7;
        t.ct[0] = 1.0; t.ct[1] = tab.ct[6];
        t.k1 = -1;
      };
      if (e6_sig) two(ts, sig[1], sig[k7_sig0 ? 0 : 6]);
      if (e6_psi) two(tp, psi[1], psi[k7_psi0 ? 0 : 6]);
      if (k7_sig0) ts.kdst = nullptr;  // k_7 is there already
      if (k7_psi0) tp.kdst = nullptr;
      if (tol_scaled) {
        ts.y0 = sig[iy0]; ts.y1 = sig[iyn]; tp.y0 = psi[iy0]; tp.y1 = psi[iyn];
        ts.rtol = tp.rtol = tol_rtol; ts.atol = tp.atol = tol_atol;
        ts.scaled = tp.scaled = 1;
      }
      // per instance (blockIdx.y): KI k blocks of rows, kTEBlocks partials each
      MapDense ms{KI * n, n, n, n, n};
      MapDense mp{KI * Md, n, n, n, n};
      const dim3 grid(kTEBlocks, NI);
      if (do_sig) {
        if (e6_sig)
          hipLaunchKernelGGL((k_te<double2, MapDense, 2>), grid, dim3(256), 0, st, ts, ms,
                             te_part, (long)KI * n * n);
        else
          hipLaunchKernelGGL((k_te<double2, MapDense, 6>), grid, dim3(256), 0, st, ts, ms,
                             te_part, (long)KI * n * n);
      }
      if (!do_psi) return;
      bool done = false;
      if constexpr (sizeof(cs) == 8) {
        if (mp.pairs()) {
          if (e6_psi)
            hipLaunchKernelGGL((k_te_pairs<cs, MapDense, 2>), grid, dim3(256), 0, st, tp, mp,
                               te_part + (size_t)kTEBlocks * NI, (long)KI * Md * n);
          else
            hipLaunchKernelGGL((k_te_pairs<cs, MapDense, 6>), grid, dim3(256), 0, st, tp, mp,
                               te_part + (size_t)kTEBlocks * NI, (long)KI * Md * n);
          done = true;
        }
      }
      if (!done) {
        if (e6_psi)
          hipLaunchKernelGGL((k_te<cs, MapDense, 2>), grid, dim3(256), 0, st, tp, mp,
                             te_part + (size_t)kTEBlocks * NI, (long)KI * Md * n);
        else
          hipLaunchKernelGGL((k_te<cs, MapDense, 6>), grid, dim3(256), 0, st, tp, mp,
                             te_part + (size_t)kTEBlocks * NI, (long)KI * Md * n);
      }
    });
    HIPCHK(hipGetLastError());
  }

  // The controller over the σ, Ψ and Ω error-norm partials, on the main stream.
  void control() {
    timed(KC_CTRL, [&] {
      // the Ω kernels' partials are one per workgroup, batch-major: KI blocks of each
      // instance are contiguous
      const int no = te_omega_blocks();
      require(no % NI == 0, "error-norm partials do not split over the instances");
      hipLaunchKernelGGL(k_controller, dim3(NI), dim3(kCtrlThreads), 0, stream,
                         (const double*)te_part, kTEBlocks, no / NI, NI, ctrl,
                         (const PulseDesc*)pulse_d, diss, drive, NL,
                         (const StageParams*)(sp + 6 * NI));
    });
    HIPCHK(hipGetLastError());
  }

  // The seven stages, error norm and controller of one attempt (graph body).
  // omit1: the first stage is left out (every instance takes k_7, first same as last)
  void attempt_body(bool omit1) {
    timed(KC_PARAMS, [&] {
      hipLaunchKernelGGL(k_stage_params, dim3(NI), dim3(64), 0, stream, ctrl,
                         (const PulseDesc*)pulse_d,
                         diss, drive, tab, NL, NI, sp, skip1, NK, (const double*)Jn, Jd);
    });
    HIPCHK(hipGetLastError());
    const bool fs = fsal_on();
    // each stage's Ωξ product issued as soon as its X is written (unfused attempts)
    xi_early = !fused_used && nc > 0;
    struct Off {  // also on error
      bool& f;
      ~Off() { f = false; }
    } off{xi_early};
    if (xi_early && omit1) {
      // X = k_1 (the last attempt's k_7) is there; stage 1's Ψ branch runs on aux[0]
      hipStream_t st = stream;
      if (!profile) {
        st = aux[0];
        HIPCHK(hipEventRecord(ev_fork, stream));
        HIPCHK(hipStreamWaitEvent(st, ev_fork, 0));
      }
      issue_xi(1, st);
    }
    const int s0 = omit1 ? 1 : 0;
    // the one-pass Ψ kernel of stage s >= 1 forms stage s + 1's Ψ input (complex128; at
    // n = 1000, M = 4000: FP64 226 -> 235 steps/s, while the complex64 kernel, shorter and
    // more latency-bound, lost more than the separate combination cost: 349 -> 333)
    psi_nx = psi_fused() && !(overlap && nc > 0) && sizeof(cs) == 16;
    e6_sig = e6_ok();
    e6_psi = e6_sig;
    k7_sig0 = fs && e6_sig;
    k7_psi0 = fs && e6_psi;
    struct Back {  // Ap back to its buffer outside attempts, also on error
      Engine& e;
      ~Back() {
        e.Ap = e.Ap0;
        e.psi_nx = e.e6_sig = e.e6_psi = e.k7_sig0 = e.k7_psi0 = e.te_stage = false;
      }
    } back{*this};
    te_stage = true;
    for (int s = s0; s < 7; ++s) {
      const bool last = (s == 6);
      double2* sdst = last ? sig[iyn] : sig[9];
      Ap = apb(s);
      // σ's part of the stage combination was issued by the stage before (rhs_stage),
      // and Ψ's too from stage 2 on with the one-pass Ψ kernel
      const bool psi_done = psi_nx && s - 1 >= 1 && s - 1 >= s0;
      combine(s, s == 0, sdst, last, s == s0, !psi_done);
      const double2* sin_ = (s == 0) ? sig[iy0] : sdst;
      OmegaStage os;
      os.y0 = omg[iy0];
      os.sps = sp + s * NI;
      for (int j = 0; j < s; ++j)
        if (tab.a[s][j] != 0.0) {
          if (j == 0) os.k1 = os.nterms;
          os.k[os.nterms] = omg[j];
          os.a[os.nterms] = tab.a[s][j];
          os.ct[os.nterms] = tab.ct[j];
          ++os.nterms;
        }
      if (last) {
        // y_new = Ω_7; the error norm is accumulated in place; k_7 goes over the k_1 slot
        // (each element once its k_1 is read) as the next attempt's k_1, or nowhere
        os.ynew = omg[iyn];
        os.store_k = false;
        if (fs) os.kout = omg[0];
        os.te = true;
        os.ct_self = tab.ct[6];
        int covered = 0;
        for (int j = 0; j < 6; ++j) covered += (tab.ct[j] == 0.0 || tab.a[6][j] != 0.0);
        require(covered == 6, "error weights need a stage-7 term the tableau does not carry");
      }
      os.par = (iy0 == 7) ? 0 : 1;
      os.y0xi_resident = true;
      if (fused_used) os.wpart = Wpart[s % 2];
      // the first stage's kernels leave the instances that reuse k_7 alone; the currents
      // of the first stage are J, those of the last (y_new) the next attempt's
      skip_now = (s == 0) ? skip1 : nullptr;
      rhs_stage(s, sin_, os, s, s == 0 ? Jd : ((last && fs) ? Jn : nullptr), true,
                last ? -1 : s + 1);
      skip_now = nullptr;
    }
    // every early product is followed on its stream by a Ψ branch the joins wait for
    if (fused_used) {
      // y_new ξ_c from the last stage's partials: the next attempt's y0 ξ_c if accepted
      const OmegaFusedArgs fa = fused_args(kFusedChunk[tsel[TS_OMEGA]]);
      const int par = (iy0 == 7) ? 0 : 1;
      timed(KC_WSUM, [&] {
        hipLaunchKernelGGL(k_sum_wpart<cs>, dim3(1024), dim3(256), 0, stream,
                           (const cs*)Wpart[0], fa.SA + fa.SP, ff ? fa.SA : fa.SA + fa.SP, ML, NK,
                           Md, nc, PredY[1 - par]);
      });
      HIPCHK(hipGetLastError());
    }
    control();
  }

  void ensure_h0() {
    if (h_dirty) {
      drop_graphs();
      install_h(H0h.data());
      HIPCHK(hipStreamSynchronize(stream));
      h_dirty = false;
    }
  }

  long bytes() const override { return nbytes; }

  // Candidate 2 of the FP64 lists is switched off: a graph-mode attempt of the
  // complex-H FP64 parity case (n = 12, M = 28) with every product on candidate 2
  // never completed once, and no cause has been shown in the kernels' index
  // arithmetic or the stream fork/join of the capture.  Until one is, neither the
  // autotuner nor set_tiles reaches those tiles.
  static constexpr int kCandsOn = std::is_same<R, double>::value ? 2 : kCands;

  void set_tiles(const int* sel) override {
    for (int i = 0; i < TS_N; ++i)
      require(sel[i] >= 0 && sel[i] < kCands, "tile candidate index out of range");
    for (int i = 0; i < TS_N; ++i)
      require(sel[i] < kCandsOn, "tile candidate 2 is switched off in FP64 (an attempt with it "
                                 "once did not complete; no cause has been shown)");
    for (int i = 0; i < TS_N; ++i) tsel[i] = sel[i];
    drop_graphs();
  }

  TuneKey tune_key() const {
    return TuneKey{(int)sizeof(R), pow2_bucket(n), pow2_bucket(Md), pow2_bucket(std::max(nc, 1)),
                   h_sparse ? 2 : (h_real ? 1 : 0), tri_on ? 1 : 0};
  }

  // Time each candidate tile of every product on rejected attempts from the
  // resident state (eps = 0: nothing is committed) and keep the fastest.
  void autotune(int force, int* out) override {
    require(have_consts && have_state, "autotune needs constants and a resident state", ZH_ESTATE);
    ensure_h0();
    const TuneKey key = tune_key();
    bool cached = false;
    if (!force) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) {
        for (int i = 0; i < TS_N; ++i) tsel[i] = it->second[i];
        fused_tuned = it->second[TS_N] != 0;
        xi_tuned = it->second[TS_N + 1] != 0;
        cached = true;
      }
    }
    if (!cached) {
      fused_tuned = false;  // the tile candidates are timed on the unfused path
      xi_tuned = false;     // and on the tiled Ωξ GEMM
      static const int kc_of[TS_N] = {KC_GEMM_PSI, KC_GEMM_HSIG, KC_GEMM_OMEGA, KC_GEMM_OMGXI,
                                      KC_GEMM_PI};
      double best[TS_N];
      int arg[TS_N];
      for (int i = 0; i < TS_N; ++i) { best[i] = 1e300; arg[i] = 0; }
      // profile mode and the tile choice back as they were if an attempt throws
      struct Restore {
        bool& prof;
        int* sel;
        bool p0;
        int s0[TS_N];
        bool done = false;
        ~Restore() {
          if (done) return;
          prof = p0;
          for (int i = 0; i < TS_N; ++i) sel[i] = s0[i];
        }
      } restore{profile, tsel, profile, {tsel[0], tsel[1], tsel[2], tsel[3], tsel[4]}};
      const bool prof0 = profile;
      profile = true;
      zh_attempt_info info1;
      std::memset(&info1, 0, sizeof(info1));
      info1.h = 1e-6;
      info1.eps = 0.0;  // never accepted: the resident state is untouched
      std::vector<zh_attempt_info> infos(NI, info1);
      zh_attempt_info* info = infos.data();
      for (int c = 0; c < kCandsOn; ++c) {
        for (int i = 0; i < TS_N; ++i) tsel[i] = c;
        attempt(info, nullptr);  // warm-up (code load)
        for (int k = 0; k < KC_N; ++k) { kc_ms[k] = 0; kc_count[k] = 0; }
        for (int r = 0; r < 2; ++r) attempt(info, nullptr);
        for (int i = 0; i < TS_N; ++i) {
          const int k = kc_of[i];
          const double t = kc_count[k] ? kc_ms[k] / kc_count[k] : 0.0;
          if (t < best[i]) { best[i] = t; arg[i] = c; }
        }
      }
      for (int i = 0; i < TS_N; ++i) tsel[i] = arg[i];
      // the streaming Ωξ kernel against the chosen Ωξ tile
      if (xi_ok && xi_req == 0) {
        xi_tuned = true;
        attempt(info, nullptr);
        for (int k = 0; k < KC_N; ++k) { kc_ms[k] = 0; kc_count[k] = 0; }
        for (int r = 0; r < 2; ++r) attempt(info, nullptr);
        const double t = kc_count[KC_GEMM_OMGXI] ? kc_ms[KC_GEMM_OMGXI] / kc_count[KC_GEMM_OMGXI]
                                                 : 1e300;
        xi_tuned = t < best[TS_OMGXI];
      }
      // the fused Ω kernel (mixed precision) against the Ω and Ωξ products it replaces,
      // for each chunk width (the Ω candidate index selects it)
      fused_tuned = false;
      if (std::is_same<R, float>::value && fusion_req == 0 && nc > 0 && nc <= 16 * kFusedNG &&
          !overlap && !tri_on) {
        // per-attempt time of the Ω kernels, the Ωξ products and their sums, on the
        // chosen tiles unfused, then fused for each chunk width
        auto omega_side = [&]() {
          attempt(info, nullptr);
          for (int k = 0; k < KC_N; ++k) { kc_ms[k] = 0; kc_count[k] = 0; }
          for (int r = 0; r < 2; ++r) attempt(info, nullptr);
          return 0.5 * (kc_ms[KC_GEMM_OMEGA] + kc_ms[KC_GEMM_OMGXI] + kc_ms[KC_WSUM]);
        };
        const double unfused = omega_side();
        double fbest = 1e300;
        int fc = 0;
        fused_tuned = true;
        for (int c = 0; c < kCands; ++c) {
          tsel[TS_OMEGA] = c;
          const double t = omega_side();
          if (t < fbest) { fbest = t; fc = c; }
        }
        fused_tuned = fbest < unfused;
        tsel[TS_OMEGA] = fused_tuned ? fc : arg[TS_OMEGA];
      }
      for (int k = 0; k < KC_N; ++k) { kc_ms[k] = 0; kc_count[k] = 0; }
      attempt_ms = 0; attempt_count = 0;
      profile = prof0;
      restore.done = true;
      std::lock_guard<std::mutex> lk(g_tune_mu);
      std::array<int, TS_N + 2> a;
      for (int i = 0; i < TS_N; ++i) a[i] = tsel[i];
      a[TS_N] = fused_tuned ? 1 : 0;
      a[TS_N + 1] = xi_tuned ? 1 : 0;
      g_tune_cache[key] = a;
    }
    drop_graphs();
    if (out)
      for (int i = 0; i < TS_N; ++i) out[i] = tsel[i];
  }

  // Enqueue one attempt on the problem's stream and return; collect() waits.
  void launch(const zh_attempt_info* info) override {
    require(have_consts, "constants not set", ZH_ESTATE);
    require(have_state, "no resident state", ZH_ESTATE);
    require(!pending, "an attempt is already in flight on this problem", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    ensure_h0();
    prepare_model();
    if (ff && ff_dirty) { clear_ff(); ff_dirty = false; }
    upload_pulses();
    const int par = (iy0 == 7) ? 0 : 1;
    const bool fz = fusable();
    if (fz != fused_used) drop_graphs();
    fused_used = fz;
    const bool fs = fsal_on();
    for (int g = 0; g < NI; ++g) {  // info[g]: instance g
      ctrl_h[g].t0 = info[g].t0;
      ctrl_h[g].h = info[g].h;
      ctrl_h[g].eps = info[g].eps;
      ctrl_h[g].force = info[g].force_accept ? 1 : 0;
      ctrl_h[g].norm_scale = norm_scale();
      ctrl_h[g].fsal = (fs && fsal_ready[g]) ? 1 : 0;
      ctrl_h[g].hprev = fsal_hprev[g];
    }
    // every instance known to take k_7 (its start at the time of the last stage before,
    // or where the controller found the same parameters): the first stage is left out of
    // the attempt altogether, not just skipped by its kernels
    bool omit1 = fs;
    for (int g = 0; g < NI && omit1; ++g)
      omit1 = fsal_ready[g] && info[g].h > 0.0 &&
              (info[g].t0 == fsal_t7[g] || (fsal_hn[g] && info[g].t0 == fsal_tn[g]));
    HIPCHK(hipMemcpyAsync(ctrl, ctrl_h, sizeof(Ctrl) * NI, hipMemcpyHostToDevice, stream));
    if (nc > 0 && !y0xi_valid) {
      compute_y0xi(par, stream);
      y0xi_valid = true;
    }
    HIPCHK(hipEventRecord(ev[2 * KC_N], stream));
    const int gi = par + (omit1 ? 2 : 0);
    if (use_graph && !profile) {
      if (!graph[gi]) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        attempt_body(omit1);
        HIPCHK(hipStreamEndCapture(stream, &g));
        HIPCHK(hipGraphInstantiate(&graph[gi], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
      }
      HIPCHK(hipGraphLaunch(graph[gi], stream));
    } else {
      attempt_body(omit1);
    }
    omit1_pending = omit1;
    HIPCHK(hipEventRecord(ev[2 * KC_N + 1], stream));
    HIPCHK(hipMemcpyAsync(ctrl_h, ctrl, sizeof(Ctrl) * NI, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(J_h, Jd, sizeof(double) * NL * NK, hipMemcpyDeviceToHost, stream));
    pend_info.assign(info, info + NI);
    pending = true;
  }

  // 1 when the attempt in flight has finished (collect will not wait), else 0.
  int ready() override {
    require(pending, "no attempt in flight on this problem", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    const hipError_t e = hipStreamQuery(stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    HIPCHK(e);
    return 0;
  }

  void collect(zh_attempt_info* info, double* J, bool finite_next = false) override {
    require(pending, "no attempt in flight on this problem", ZH_ESTATE);
    HIPCHK(hipSetDevice(device));
    pending = false;
    HIPCHK(hipStreamSynchronize(stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * KC_N], ev[2 * KC_N + 1]));
    attempt_ms += ms;
    attempt_count += 1;
    int nacc = 0;
    const bool fs = fsal_on();
    for (int g = 0; g < NI; ++g) {
      info[g] = pend_info[g];
      info[g].te = ctrl_h[g].te;
      info[g].h_next = ctrl_h[g].h_next;
      info[g].accepted = ctrl_h[g].accept;
      // several instances: one whose next step size is not finite is not committed
      // (unless forced: a finished instance advanced with h = 0) and reports it
      if (NI > 1 && info[g].accepted && !info[g].force_accept && !std::isfinite(info[g].h_next))
        info[g].accepted = 0;
      nacc += info[g].accepted ? 1 : 0;
      fsal_taken += ctrl_h[g].fsal_used ? 1 : 0;
      if (omit1_pending && !ctrl_h[g].fsal_used) {
        fsal_reset();
        throw Error{ZH_ESTATE, "first same as last: an attempt without its first stage found "
                               "different stage parameters (its results are not valid)"};
      }
      // the last stage left k_7 = h f(t0 + h, y_new) in the k_1 slot: the next attempt
      // may take it as its k_1 when this step is committed
      fsal_ready[g] = (fs && info[g].accepted && info[g].h > 0.0) ? 1 : 0;
      fsal_hprev[g] = info[g].h;
      fsal_t7[g] = info[g].t0 + info[g].h;
      fsal_tn[g] = info[g].t0 + info[g].h_next;
      fsal_hn[g] = ctrl_h[g].fsal_hn ? 1 : 0;
    }
    if (J) std::memcpy(J, J_h, sizeof(double) * NL * NK);
    if (NI == 1 && info->accepted && finite_next && !std::isfinite(info->h_next)) {
      // not committed: the resident state stays at the step's start
      info->accepted = 0;
      fsal_ready[0] = 0;
      throw Error{ZH_ENONFINITE, "non-finite next step size (error estimate " +
                                     std::to_string(info->te) + ") at t=" +
                                     std::to_string(info->t0) + "; the step is not committed"};
    }
    if (nacc == 0) return;
    if (nacc < NI) {
      // instances that stay: their step's start goes into the slot that becomes y0
      for (int g = 0; g < NI; ++g)
        if (!info[g].accepted) keep_instance(g);
      HIPCHK(hipStreamSynchronize(stream));
    }
    std::swap(iy0, iyn);
    y0xi_valid = true;  // the last stage left y_new ξ_c in the new y0's PredY slot
  }

  // Instance g's y0 (σ, Ψ, Ω and its y0 ξ_c) copied over its block of the y_new slots,
  // so that the slot swap of the others leaves g at its step's start.
  void keep_instance(int g) {
    const size_t k0 = (size_t)g * KI;
    const size_t nn = (size_t)n * n, mn = (size_t)Md * n, mm = (size_t)Md * Md,
                 mc = (size_t)Md * nc;
    const int par = (iy0 == 7) ? 0 : 1;
    HIPCHK(hipMemcpyAsync(sig[iyn] + k0 * nn, sig[iy0] + k0 * nn, sizeof(double2) * KI * nn,
                          hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemcpyAsync(psi[iyn] + k0 * mn, psi[iy0] + k0 * mn, sizeof(cs) * KI * mn,
                          hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemcpyAsync(omg[iyn] + k0 * mm, omg[iy0] + k0 * mm, sizeof(cs) * KI * mm,
                          hipMemcpyDeviceToDevice, stream));
    if (nc > 0)
      HIPCHK(hipMemcpyAsync(PredY[1 - par] + k0 * mc, PredY[par] + k0 * mc, sizeof(cs) * KI * mc,
                            hipMemcpyDeviceToDevice, stream));
  }

  void attempt(zh_attempt_info* info, double* J, bool finite_next = false) override {
    launch(info);
    collect(info, J, finite_next);
  }

  // ---------------------------------------------------------------- steady state
  // Ψ₀ and Ω₀ of cmdtools/psinought (:180-449) for the resident σ (taken as σ₀), the
  // bias at t0 and H(t0, σ₀): BiCGSTAB over the real inner product (the operator is
  // real-linear, ψ* enters), right-preconditioned with (χ⁺_i + Δ - H)^{-1} from the
  // eigenpairs of H(t0, σ₀) when given (psinought's block preconditioner without
  // its coupling terms).  Writes (σ₀, Ψ₀, Ω₀) as the resident state.
  void steady_state(double t0, double tol, int maxiter, const zh_precond* pc, int* iters,
                    double* resid) override {
    require(NI == 1, "the steady-state solver takes one instance per problem");
    require(!overlap, "the steady-state solver is defined for an orthogonal basis");
    require(have_consts && have_state, "steady state needs constants and a resident sigma",
            ZH_ESTATE);
    require(tol > 0.0 && maxiter > 0, "tolerance and iteration limit must be positive");
    HIPCHK(hipSetDevice(device));
    ensure_h0();
    drop_graphs();
    struct Restore {  // profile mode back as it was, also on error
      bool& f;
      bool v;
      ~Restore() { f = v; }
    } restore{profile, profile};
    profile = false;
    // stage parameters at t0 with h = 1 (sp[0]: Δ(t0), the dissipator rate unused)
    ctrl_h->t0 = t0; ctrl_h->h = 1.0; ctrl_h->eps = 1.0; ctrl_h->force = 1;
    ctrl_h->norm_scale = 1.0;
    ctrl_h->fsal = 0;
    HIPCHK(hipMemcpyAsync(ctrl, ctrl_h, sizeof(Ctrl), hipMemcpyHostToDevice, stream));
    upload_pulses();
    hipLaunchKernelGGL(k_stage_params, dim3(NI), dim3(64), 0, stream, ctrl,
                       (const PulseDesc*)pulse_d, diss, drive, tab, NL, NI, sp, skip1, NK,
                       (const double*)Jn, Jd);
    HIPCHK(hipGetLastError());
    const StageParams* sp0 = sp;
    apply_model(sp0, stream);  // drive terms at t0 (an overlap is refused above)
    const long N = (long)NK * Md * n;
    const long mc = (long)NK * Md * nc;
    const double2* sig0 = sig[iy0];
    DevBuf buf;
    double2 *x = buf.get<double2>(N), *r = buf.get<double2>(N), *rh = buf.get<double2>(N),
            *pv = buf.get<double2>(N), *v = buf.get<double2>(N), *sv = buf.get<double2>(N),
            *tv = buf.get<double2>(N), *bv = buf.get<double2>(N);
    cs* hpsi = buf.get<cs>(N);
    double* part = buf.get<double>(kDotBlocks);
    const bool prec = pc && pc->evecs && pc->evals;
    const double* evecs = prec ? pc->evecs : nullptr;
    const double* evals = prec ? pc->evals : nullptr;
    const int ncr = prec && pc->ncr > 0 ? pc->ncr : 0;
    if (ncr) {
      require(pc->corb && pc->group && pc->K && pc->ngroups > 0, "incomplete preconditioner");
      for (int c = 0; c < ncr; ++c)
        require(pc->corb[c] >= 0 && pc->corb[c] < n, "preconditioner orbital out of range");
      for (long m = 0; m < (long)NK * M; ++m)
        require(pc->group[m] >= 0 && pc->group[m] < pc->ngroups,
                "preconditioner group out of range");
    }
    double2 *ph = pv, *sh = sv;
    cs *Vc = nullptr, *Vt = nullptr, *y1 = nullptr, *y2 = nullptr;
    double* ev = nullptr;
    if (prec) {
      ph = buf.get<double2>(N);
      sh = buf.get<double2>(N);
      y1 = buf.get<cs>(N);
      y2 = buf.get<cs>(N);
      Vc = buf.get<cs>((size_t)NK * n * n);
      Vt = buf.get<cs>((size_t)NK * n * n);
      ev = buf.get<double>((size_t)NK * n);
      std::vector<cs> hc((size_t)NK * n * n), ht((size_t)NK * n * n);
      for (int b = 0; b < NK; ++b)
        for (int o = 0; o < n; ++o)
          for (int k = 0; k < n; ++k) {
            const size_t at = ((size_t)b * n + o) * n + k;
            hc[at] = from_d2<cs>(make_double2(evecs[2 * at], -evecs[2 * at + 1]));
            ht[((size_t)b * n + k) * n + o] = from_d2<cs>(make_double2(evecs[2 * at], evecs[2 * at + 1]));
          }
      HIPCHK(hipMemcpyAsync(Vc, hc.data(), sizeof(cs) * hc.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(Vt, ht.data(), sizeof(cs) * ht.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(ev, evals, sizeof(double) * NK * n, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    // self-energy correction: V_c (n x ncr, V[corb_c][k] at (k, c)), conj V_c (ncr x n),
    // K per group, group per device mode
    cs *VcT = nullptr, *Vcc = nullptr, *Zc = nullptr, *Wc = nullptr;
    double2* Kd = nullptr;
    int* gdev = nullptr;
    if (ncr) {
      VcT = buf.get<cs>((size_t)NK * n * ncr);
      Vcc = buf.get<cs>((size_t)NK * ncr * n);
      Zc = buf.get<cs>((size_t)NK * Md * ncr);
      Wc = buf.get<cs>((size_t)NK * Md * ncr);
      Kd = buf.get<double2>((size_t)pc->ngroups * ncr * ncr);
      gdev = buf.get<int>((size_t)NK * Md);
      std::vector<cs> a((size_t)NK * n * ncr), c((size_t)NK * ncr * n);
      for (int b = 0; b < NK; ++b)
        for (int q = 0; q < ncr; ++q)
          for (int k = 0; k < n; ++k) {
            const size_t at = ((size_t)b * n + pc->corb[q]) * n + k;
            a[((size_t)b * n + k) * ncr + q] =
                from_d2<cs>(make_double2(evecs[2 * at], evecs[2 * at + 1]));
            c[((size_t)b * ncr + q) * n + k] =
                from_d2<cs>(make_double2(evecs[2 * at], -evecs[2 * at + 1]));
          }
      std::vector<int> gd((size_t)NK * Md, -1);
      for (int b = 0; b < NK; ++b)
        for (int d = 0; d < Md; ++d)
          if (perm[d] >= 0) gd[(size_t)b * Md + d] = pc->group[(size_t)b * M + perm[d]];
      HIPCHK(hipMemcpyAsync(VcT, a.data(), sizeof(cs) * a.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(Vcc, c.data(), sizeof(cs) * c.size(), hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(Kd, pc->K, sizeof(double2) * pc->ngroups * ncr * ncr,
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync(gdev, gd.data(), sizeof(int) * gd.size(), hipMemcpyHostToDevice,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    const dim3 eg(2048), eb(256);
    auto lincomb = [&](double a, const double2* xa, double b, const double2* yb, double c,
                       const double2* zc, double2* out) {
      hipLaunchKernelGGL(k_lincomb, eg, eb, 0, stream, N, a, xa, b, yb, c, zc, out);
      HIPCHK(hipGetLastError());
    };
    std::vector<double> ph_(kDotBlocks);
    auto dot = [&](const double2* a, const double2* b) {
      hipLaunchKernelGGL(k_dot, dim3(kDotBlocks), dim3(256), 0, stream, N, a, b, part);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(ph_.data(), part, sizeof(double) * kDotBlocks, hipMemcpyDeviceToHost,
                            stream));
      HIPCHK(hipStreamSynchronize(stream));
      double acc = 0.0;
      for (double q : ph_) acc += q;
      return acc;
    };
    // Ω₀(ψ) of the ψ block of Ap into om (FF tiles skipped when the block is zero)
    auto omega0 = [&](cs* om) {
      KSegOmega ks{t1, t2, nc, ML, ff ? 1 : 0, (Md + kTile - 1) / kTile};
      ks.tri = tri_on ? 1 : 0;
      GemmArgs<R, R> g{Md, Md, 2 * nc, Aomg, 2 * nc, (long)Md * 2 * nc, Bomg, 2 * nc,
                       (long)Md * 2 * nc};
      g.xcd_order = 0;
      EpiOmegaSteady<cs> e{om, uvec, vvec, Md};
      HIPCHK((launch_sel<R, R, R, false, true, 0>(tsel[TS_OMEGA], g, NK, e, stream, ks,
                                                  typename Cands<R>::omega())));
    };
    auto to_ap = [&](const double2* xin) {
      hipLaunchKernelGGL(k_to_ap<cs>, eg, eb, 0, stream, N, n, xin, Ap, (long)(n + nc));
      HIPCHK(hipGetLastError());
    };
    // y = A(x)
    auto apply_A = [&](const double2* xin, double2* yout) {
      to_ap(xin);
      gather(sp0, sig0, true);
      omega0(omg[0]);
      omgxi_run<0>(omg[0], stream, false);
      WSumArgs<cs> w{};
      w.part = W;
      for (int j = 0; j < 7; ++j) w.pred[j] = Pred[j];
      w.W = Wst;
      w.s = 0; w.parts = kWParts; w.NK = NK; w.M = Md; w.nc = nc;
      hipLaunchKernelGGL(k_wsum<cs>, dim3(1024), dim3(256), 0, stream, w);
      HIPCHK(hipGetLastError());
      if (h_sparse) {
        hipLaunchKernelGGL(k_ell_rows<cs>, eg, eb, 0, stream, NK, Md, (const cs*)Ap,
                           (long)(n + nc), ell(), hpsi);
        HIPCHK(hipGetLastError());
      } else {
        GemmArgs<R, R> gh{Md, n, n, Ap, n + nc, (long)Md * (n + nc), Bp, n, (long)(n + nc) * n};
        EpiStore<cs> eh{hpsi, n, (long)Md * n, 1.0, 0.0};
        HIPCHK((launch_sel<R, R, R, false, false, 0>(tsel[TS_PSI], gh, NK, eh, stream, KFull(),
                                                     typename Cands<R>::psi())));
      }
      SteadyApplyArgs<cs> a{NK, Md, n, nc, xin, hpsi, Wst, chip, dvec, lead, cpos, sp0,
                            1.0 / hbar, yout};
      hipLaunchKernelGGL(k_steady_apply<cs>, eg, eb, 0, stream, a);
      HIPCHK(hipGetLastError());
    };
    // z = M^{-1} r (identity without eigenpairs)
    auto apply_M = [&](const double2* rin, double2* zout) {
      hipLaunchKernelGGL(k_to_ap<cs>, eg, eb, 0, stream, N, n, rin, y1, (long)n);
      HIPCHK(hipGetLastError());
      GemmArgs<R, R> g1{Md, n, n, y1, n, (long)Md * n, Vc, n, (long)n * n};
      EpiPrecScale<cs> e1{y2, ev, chip, lead, sp0, Md, n};
      HIPCHK((launch_sel<R, R, R, false, false, 0>(tsel[TS_PSI], g1, NK, e1, stream, KFull(),
                                                   typename Cands<R>::psi())));
      if (ncr) {
        // Woodbury: y2 -= D V† U K Uᵀ V y2 (U: the corb orbitals)
        GemmArgs<R, R> ga{Md, ncr, n, y2, n, (long)Md * n, VcT, ncr, (long)n * ncr};
        EpiStore<cs> ea{Zc, ncr, (long)Md * ncr, 1.0, 0.0};
        HIPCHK((launch_sel<R, R, R, false, false, 0>(tsel[TS_PSI], ga, NK, ea, stream, KFull(),
                                                     typename Cands<R>::psi())));
        hipLaunchKernelGGL(k_group_matvec<cs>, eg, eb, 0, stream, NK, Md, ncr, (const int*)gdev,
                           (const double2*)Kd, (const cs*)Zc, Wc);
        HIPCHK(hipGetLastError());
        GemmArgs<R, R> gb{Md, n, ncr, Wc, ncr, (long)Md * ncr, Vcc, n, (long)ncr * n};
        EpiPrecCorr<cs> eb2{y2, ev, chip, lead, sp0, Md, n};
        HIPCHK((launch_sel<R, R, R, false, false, 0>(tsel[TS_PSI], gb, NK, eb2, stream, KFull(),
                                                     typename Cands<R>::psi())));
      }
      GemmArgs<R, R> g2{Md, n, n, y2, n, (long)Md * n, Vt, n, (long)n * n};
      EpiStore<double2> e2{zout, n, (long)Md * n, 1.0, 0.0};
      HIPCHK((launch_sel<R, R, R, false, false, 0>(tsel[TS_PSI], g2, NK, e2, stream, KFull(),
                                                   typename Cands<R>::psi())));
    };
    // b: gather with ψ = 0 (σ₀ columns, Hubbard diagonal, shifts), then the ξσ₀ product
    lincomb(0.0, nullptr, 0.0, nullptr, 0.0, nullptr, x);
    to_ap(x);
    gather(sp0, sig0, true);
    if (nc > 0) {
      EpiStore<cs> e0{hpsi, n, (long)Md * n, 1.0, 0.0};
      GemmArgs<R, R> g0{Md, n, nc, Ap + n, n + nc, (long)Md * (n + nc), Bp + (long)n * n, n,
                        (long)(n + nc) * n};
      HIPCHK((launch_sel<R, R, R, false, false, 0>(tsel[TS_PSI], g0, NK, e0, stream, kseg_g0(),
                                                   typename Cands<R>::psi())));
    }
    SteadyRhsArgs<cs> ra{NK, Md, n, nc, nc > 0 ? g0_rows() : 0, hpsi, xic, glp, cidx, hbar, bv};
    hipLaunchKernelGGL(k_steady_rhs<cs>, eg, eb, 0, stream, ra);
    hipLaunchKernelGGL(k_steady_rhs_xi<cs>, eg, eb, 0, stream, ra);
    HIPCHK(hipGetLastError());
    const double bnorm = std::sqrt(dot(bv, bv));
    int it = 0;
    double res = 0.0;
    bool ok = bnorm == 0.0;
    if (!ok) {
      // x0 = M^{-1} b (psinought starts from preconditioner @ b), else 0
      if (prec) apply_M(bv, x);
      apply_A(x, r);
      lincomb(1.0, bv, -1.0, r, 0.0, nullptr, r);
      lincomb(1.0, r, 0.0, nullptr, 0.0, nullptr, rh);
      lincomb(0.0, nullptr, 0.0, nullptr, 0.0, nullptr, pv);
      lincomb(0.0, nullptr, 0.0, nullptr, 0.0, nullptr, v);
      double rho = 1.0, alpha = 1.0, omega = 1.0;
      res = std::sqrt(dot(r, r)) / bnorm;
      ok = res < tol;
      while (!ok && it < maxiter) {
        ++it;
        const double rho1 = dot(rh, r);
        require(rho1 != 0.0 && std::isfinite(rho1), "steady state: Krylov breakdown (rho = 0)",
                ZH_ENONFINITE);
        const double beta = (rho1 / rho) * (alpha / omega);
        lincomb(1.0, r, beta, pv, -beta * omega, v, pv);
        if (prec) apply_M(pv, ph);
        apply_A(ph, v);
        const double rv = dot(rh, v);
        require(rv != 0.0 && std::isfinite(rv), "steady state: Krylov breakdown (r·v = 0)",
                ZH_ENONFINITE);
        alpha = rho1 / rv;
        lincomb(1.0, r, -alpha, v, 0.0, nullptr, sv);
        const double snorm = std::sqrt(dot(sv, sv)) / bnorm;
        if (snorm < tol) {
          lincomb(1.0, x, alpha, ph, 0.0, nullptr, x);
          lincomb(1.0, sv, 0.0, nullptr, 0.0, nullptr, r);
          res = snorm;
          ok = true;
          break;
        }
        if (prec) apply_M(sv, sh);
        apply_A(sh, tv);
        const double tt = dot(tv, tv);
        require(tt != 0.0 && std::isfinite(tt), "steady state: Krylov breakdown (t = 0)",
                ZH_ENONFINITE);
        omega = dot(tv, sv) / tt;
        lincomb(1.0, x, alpha, ph, omega, sh, x);
        lincomb(1.0, sv, -omega, tv, 0.0, nullptr, r);
        rho = rho1;
        res = std::sqrt(dot(r, r)) / bnorm;
        require(std::isfinite(res), "steady state: non-finite residual", ZH_ENONFINITE);
        ok = res < tol;
      }
    }
    // resident state (σ₀, Ψ₀, Ω₀(Ψ₀)); the passive block of Ω₀ is zero
    hipLaunchKernelGGL(k_to_ap<cs>, eg, eb, 0, stream, N, n, (const double2*)x, psi[iy0], (long)n);
    HIPCHK(hipGetLastError());
    to_ap(x);
    gather(sp0, sig0, true);
    omega0(omg[iy0]);
    HIPCHK(hipStreamSynchronize(stream));
    set_ff(true);
    y0xi_valid = false;
    if (iters) *iters = it;
    if (resid) *resid = res;
    require(ok, "steady state: not converged to the tolerance within the iteration limit "
                "(the state holds the last iterate)", ZH_ELIMIT);
  }

  void rhs(const double* Ht, const double* delta, const double* s, const double* p,
           const double* o, double* ds, double* dp, double* dom, double* J) override {
    require(have_consts, "constants not set", ZH_ESTATE);
    require(NI == 1, "one evaluation takes one instance per problem");
    HIPCHK(hipSetDevice(device));
    install_h(Ht, false);
    h_dirty = true;
    fsal_reset();  // the k slots and stage parameters are used here
    StageParams st;
    std::memset(&st, 0, sizeof(st));
    st.ts = 0.0;
    st.h = 1.0;
    st.r1 = 1.0;
    for (int a = 0; a < NL; ++a) st.delta[a] = delta[a];
    HIPCHK(hipMemcpyAsync(sp, &st, sizeof(st), hipMemcpyHostToDevice, stream));
    upload(sig[9], n, s, (long)NK * n, n);
    upload_psi(Ap, n + nc, p, psi[1]);  // k slot 1 is free outside an attempt
    upload_omega(omg[9], o);
    // the input's passive block may be non-zero: evaluate it in full; its block upper
    // triangle stands for all of it only when the input is anti-Hermitian
    const bool ff_keep = ff, tri_keep = tri_on;
    struct Keep {
      bool& f; bool& t; bool fv, tv;
      ~Keep() { f = fv; t = tv; }
    } keep{ff, tri_on, ff_keep, tri_keep};
    ff = false;
    tri_on = anti_ok && anti_req != 1 && omega_antihermitian(o);
    OmegaStage os;
    os.y0 = omg[9];
    rhs_stage(0, sig[9], os, 0, J ? Jd : nullptr, false);
    HIPCHK(hipStreamSynchronize(stream));
    ff_dirty = true;
    if (ds) download(ds, sig[0], (size_t)NK * n * n);
    if (dp) download_psi(dp, psi[0]);
    if (dom) download_omega(dom, omg[0]);
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
  HIPCHK((launch_zgemm<R, R, R, TA, TB, Tile<64, 64, 16, 2>>(g, 1, e, nullptr)));
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
  zh_dims dims;       // dims.NK: k blocks of all instances together
  int instances = 1;  // independent instances stacked on the k axis
};

// ---------------------------------------------------------- equilibrium density
// Device buffers and stream of one window_density call, released on every exit.
struct ScfScratch {
  hipStream_t st = nullptr;
  std::vector<void*> bufs;
  ~ScfScratch() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : bufs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  template <typename T> T* alloc(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
      throw Error{ZH_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    bufs.push_back(p);
    return (T*)p;
  }
};

// Bytes a buffer of `count` elements takes in the scratch array (256-byte aligned).
static inline size_t arena_bytes(size_t count, size_t elem) {
  return ((std::max<size_t>(count, 1) * elem + 255) / 256) * 256;
}
// Counts what a buffer set needs, allocates nothing.
struct CountArena {
  size_t used = 0;
  template <typename T> T* alloc(size_t count) {
    used += arena_bytes(count, sizeof(T));
    return nullptr;
  }
};
// Slices of a caller-provided device region (the scratch array of the allocation-free
// entry point); the region's size is checked before any buffer set is laid out.
struct SliceArena {
  char* base;
  size_t cap, used = 0;
  template <typename T> T* alloc(size_t count) {
    const size_t b = arena_bytes(count, sizeof(T));
    require(used + b <= cap, "scratch layout exceeds the scratch array", ZH_EINVAL);
    T* p = reinterpret_cast<T*>(base + used);
    used += b;
    return p;
  }
};

// Device buffers of one contour_density evaluation with P contour points per chunk.
// X of V Y V† shares Y's buffer and ρ shares Z's (each is dead when the other is
// written).
struct DensityBufs {
  double2 *Ed, *Wd;
  double* ed;
  int* cd;
  double2 *V, *Vh, *Vch, *w, *Y, *Z;
  double2 *Tp, *T, *U, *Sg, *G0, *Mm, *Kk;
};
template <typename Arena>
static DensityBufs density_bufs(Arena& a, int n, int nE, int nc, int P) {
  const size_t nn = (size_t)n * n, cc = (size_t)nc * nc, nL = (size_t)n * P * nc;
  DensityBufs b;
  b.Ed = a.template alloc<double2>(nE);
  b.Wd = a.template alloc<double2>(nE);
  b.ed = a.template alloc<double>(n);
  b.cd = a.template alloc<int>(nc);
  b.V = a.template alloc<double2>(nn);
  b.Vh = a.template alloc<double2>(nn);
  b.Vch = a.template alloc<double2>((size_t)n * nc);
  b.w = a.template alloc<double2>(n);
  b.Y = a.template alloc<double2>(nn);
  b.Z = a.template alloc<double2>(nn);
  b.Tp = a.template alloc<double2>(nL);
  b.T = a.template alloc<double2>(nL);
  b.U = a.template alloc<double2>(nL);
  b.Sg = a.template alloc<double2>((size_t)P * cc);
  b.G0 = a.template alloc<double2>((size_t)P * cc);
  b.Mm = a.template alloc<double2>((size_t)P * cc);
  b.Kk = a.template alloc<double2>((size_t)P * cc);
  return b;
}

static void density_args(int NK, int n, int nE, int nc, const int* corb, const double* sigma_cc) {
  require(NK > 0 && n > 0 && nE > 0, "NK, n and the number of contour points must be positive");
  require(nc >= 0 && nc <= n && nc <= kGJMax, "coupled orbitals: 0 <= nc <= min(n, 1024)");
  require(nc == 0 || (corb && sigma_cc), "the self-energy needs its orbitals and values");
  for (int c = 0; c < nc; ++c) require(corb[c] >= 0 && corb[c] < n, "orbital index out of range");
}

// The contour sum on stream st over the buffers b, P contour points per chunk: the
// chunks' (n x P n_c x n) products accumulate into Y = Σ_i W_i T'_i K_i T_i +
// diag(Σ_i W_i D_i).  With P = nE it is one product.
static void contour_density_run(const DensityBufs& b, hipStream_t st, int P, int NK, int n,
                                int nE, int nc, const double* EW, const double* evals,
                                const double* evecs, const int* corb, const double* sigma_cc,
                                double* rho) {
  const long nn = (long)n * n, cc = (long)nc * nc;
  std::vector<double2> Eh(nE), Wh(nE);
  for (int i = 0; i < nE; ++i) {
    Eh[i] = make_double2(EW[4 * i], EW[4 * i + 1]);
    Wh[i] = make_double2(EW[4 * i + 2], EW[4 * i + 3]);
  }
  HIPCHK(hipMemcpyAsync(b.Ed, Eh.data(), sizeof(double2) * nE, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(b.Wd, Wh.data(), sizeof(double2) * nE, hipMemcpyHostToDevice, st));
  if (nc > 0) HIPCHK(hipMemcpyAsync(b.cd, corb, sizeof(int) * nc, hipMemcpyHostToDevice, st));
  // FP64 products, three real MFMAs per complex tile product
  typedef Tile<64, 64, 16, 2, false, 3> TBig;
  typedef Tile<32, 32, 16, 2, false, 4> TSmall;
  const dim3 g1(1024), b1(256);
  for (int k = 0; k < NK; ++k) {
    HIPCHK(hipMemcpyAsync(b.V, evecs + 2 * (size_t)k * nn, sizeof(double2) * nn,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b.ed, evals + (size_t)k * n, sizeof(double) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scf_adjoint, g1, b1, 0, st, n, nc, (const double2*)b.V,
                       (const int*)b.cd, b.Vh, b.Vch);
    hipLaunchKernelGGL(k_scf_diag, dim3((n + 255) / 256), b1, 0, st, n, nE, (const double2*)b.Ed,
                       (const double2*)b.Wd, (const double*)b.ed, b.w);
    HIPCHK(hipGetLastError());
    for (int i0 = 0; i0 < (nc > 0 ? nE : 1); i0 += P) {
      const int Pc = nc > 0 ? std::min(P, nE - i0) : 0;
      const long L = (long)Pc * nc;
      if (nc > 0) {
        HIPCHK(hipMemcpyAsync(b.Sg, sigma_cc + 2 * ((size_t)k * nE + i0) * cc,
                              sizeof(double2) * Pc * cc, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_scf_factors, g1, b1, 0, st, n, Pc, nc, (const double2*)(b.Ed + i0),
                           (const double*)b.ed, (const double2*)b.V, (const int*)b.cd, b.Tp, b.T);
        HIPCHK(hipGetLastError());
        // G0_i,cc = T_i V_c†   (nc x n)(n x nc), one batch entry per contour point
        GemmArgs<double, double> ga{nc, nc, n, b.T, n, (long)nc * n, b.Vch, nc, 0};
        HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
            ga, Pc, EpiStore<double2>{b.G0, nc, cc, 1.0, 0.0}, st)));
        // I - Σ_i G0_i,cc
        GemmArgs<double, double> gb{nc, nc, nc, b.Sg, nc, cc, b.G0, nc, cc};
        HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
            gb, Pc, EpiIdMinus{b.Mm, nc, cc}, st)));
        // K_i = (I - Σ_i G0_i,cc)⁻¹ Σ_i
        HIPCHK(hipMemcpyAsync(b.Kk, b.Sg, sizeof(double2) * Pc * cc, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_gauss_jordan, dim3(Pc), dim3(256), 0, st, nc, b.Mm, b.Kk);
        HIPCHK(hipGetLastError());
        // U_i = W_i T'_i K_i  (n x nc)(nc x nc), into the column block i of U (n, Pc nc)
        GemmArgs<double, double> gc{n, nc, nc, b.Tp, L, (long)nc, b.Kk, nc, cc};
        HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
            gc, Pc, EpiScaleZ{b.U, L, (long)nc, b.Wd + i0}, st)));
      }
      // Y = U T + diag(Σ_i W_i D_i) for the first chunk, Y += U T for the others
      // (n x Pc nc)(Pc nc x n)
      GemmArgs<double, double> gy{n, n, (int)L, b.U, L, 0, b.T, n, 0};
      if (i0 == 0)
        HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
            gy, 1, EpiAddDiag{b.Y, n, b.w}, st)));
      else
        HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
            gy, 1, EpiAddTo{b.Y, n}, st)));
    }
    // X = V Y V† (X into Y's buffer), ρ = (i/2π)(X - X†) (into Z's)
    GemmArgs<double, double> gz{n, n, n, b.V, n, 0, b.Y, n, 0};
    HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
        gz, 1, EpiStore<double2>{b.Z, n, nn, 1.0, 0.0}, st)));
    GemmArgs<double, double> gx{n, n, n, b.Z, n, 0, b.Vh, n, 0};
    HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
        gx, 1, EpiStore<double2>{b.Y, n, nn, 1.0, 0.0}, st)));
    hipLaunchKernelGGL(k_scf_rho, g1, b1, 0, st, n, (const double2*)b.Y, b.Z);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rho + 2 * (size_t)k * nn, b.Z, sizeof(double2) * nn,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  for (size_t i = 0; i < 2 * (size_t)NK * nn; ++i)
    require(std::isfinite(rho[i]), "non-finite density (singular Woodbury matrix?)", ZH_ENONFINITE);
}

static void contour_density(int NK, int n, int nE, int nc, const double* EW, const double* evals,
                            const double* evecs, const int* corb, const double* sigma_cc,
                            double* rho, int device) {
  density_args(NK, n, nE, nc, corb, sigma_cc);
  HIPCHK(hipSetDevice(device));
  ScfScratch sc;
  HIPCHK(hipStreamCreateWithFlags(&sc.st, hipStreamNonBlocking));
  const DensityBufs b = density_bufs(sc, n, nE, nc, nE);  // one chunk: every point at once
  contour_density_run(b, sc.st, nE, NK, n, nE, nc, EW, evals, evecs, corb, sigma_cc, rho);
}

// ------------------------------------------- allocation-free contour density
// Scratch of zh_contour_density_noalloc: one copy per GPU, held until the process
// ends.  In host code `scratch` names a host-side stand-in; the device address of
// this GPU's copy comes from hipGetSymbolAddress (scratch_ptr), once per device.
static __device__ __align__(256) unsigned char scratch[ZH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {nullptr};

static char* scratch_ptr(int device) {
  require(device >= 0 && device < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, "no device address for the scratch array", ZH_EHIP);
    g_scratch_ptr[device] = ptr;
  }
  return static_cast<char*>(g_scratch_ptr[device]);
}

// Contour points per chunk that fit in `cap` bytes (0: not even one).
static int density_chunk(int n, int nE, int nc, size_t cap) {
  auto need = [&](int P) {
    CountArena a;
    (void)density_bufs(a, n, nE, nc, P);
    return a.used;
  };
  if (need(1) > cap) return 0;
  if (nc == 0 || need(nE) <= cap) return nE;
  int lo = 1, hi = nE;  // need(lo) fits, need(hi) does not
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    (need(mid) <= cap ? lo : hi) = mid;
  }
  return lo;
}

static void contour_density_noalloc(int NK, int n, int nE, int nc, const double* EW,
                                    const double* evals, const double* evecs, const int* corb,
                                    const double* sigma_cc, double* rho, int device) {
  density_args(NK, n, nE, nc, corb, sigma_cc);
  const size_t cap = ZH_NOALLOC_SCRATCH_BYTES;
  const int P = density_chunk(n, nE, nc, cap);
  if (P == 0) {
    CountArena a;
    (void)density_bufs(a, n, nE, nc, 1);
    throw Error{ZH_EINVAL, "zh_contour_density_noalloc: n = " + std::to_string(n) + ", nc = " +
                               std::to_string(nc) + " needs " + std::to_string(a.used) +
                               " bytes of scratch, more than the " + std::to_string(cap) +
                               " it has; use zh_contour_density"};
  }
  char* base = scratch_ptr(device);  // also makes `device` current
  SliceArena arena{base, cap};
  const DensityBufs b = density_bufs(arena, n, nE, nc, P);
  // the null stream: no stream is created (every buffer is written before it is read,
  // so nothing needs clearing between calls)
  contour_density_run(b, 0, P, NK, n, nE, nc, EW, evals, evecs, corb, sigma_cc, rho);
}

// Bias-window (nonequilibrium) density of Zandpack/cmdtools/SCF:236-253, 299-306,
//   ρ_win = ∫ dE (1/2π) G(E)[:, c] M(E) G(E)[:, c]†,  M = Σ_α (f_α - f_min) Γ_α,
// on real nodes E_i with weights w_i.  With H = V diag(e) V† and the Woodbury identity
// on the coupled orbitals c, G[:, c] = V P_i with P_i = T'_i (I - Σ_i G0_i,cc)⁻¹
// (T'_i = D_i V_c†, as for the contour sum), so
//   ρ_win = V (Σ_i (w_i/2π) P_i M_i P_i†) V†,
// the inner sum one (n x nE nc)(nE nc x n) product.
static void window_density(int NK, int n, int nE, int nc, const double* Ew, const double* evals,
                           const double* evecs, const int* corb, const double* sigma_cc,
                           const double* gamma_w, double* rho, int device) {
  require(NK > 0 && n > 0 && nE > 0, "NK, n and the number of nodes must be positive");
  require(nc > 0 && nc <= n && nc <= kGJMax, "coupled orbitals: 0 < nc <= min(n, 1024)");
  require(corb && sigma_cc && gamma_w, "the window needs its orbitals, self-energy and weights");
  for (int c = 0; c < nc; ++c) require(corb[c] >= 0 && corb[c] < n, "orbital index out of range");
  for (int i = 0; i < nE; ++i)
    require(std::isfinite(Ew[2 * i]) && std::isfinite(Ew[2 * i + 1]), "nodes must be finite");
  HIPCHK(hipSetDevice(device));
  ScfScratch sc;
  HIPCHK(hipStreamCreateWithFlags(&sc.st, hipStreamNonBlocking));
  hipStream_t st = sc.st;
  const long nn = (long)n * n, L = (long)nE * nc, cc = (long)nc * nc;
  std::vector<double2> Eh(nE), Wh(nE);
  for (int i = 0; i < nE; ++i) {
    Eh[i] = make_double2(Ew[2 * i], 0.0);
    Wh[i] = make_double2(Ew[2 * i + 1] / (2.0 * M_PI), 0.0);
  }
  double2* Ed = sc.alloc<double2>(nE);
  double2* Wd = sc.alloc<double2>(nE);
  double* ed = sc.alloc<double>(n);
  int* cd = sc.alloc<int>(nc);
  double2 *V = sc.alloc<double2>(nn), *Vh = sc.alloc<double2>(nn), *Vch = sc.alloc<double2>((size_t)n * nc);
  double2 *Y = sc.alloc<double2>(nn), *Z = sc.alloc<double2>(nn), *X = sc.alloc<double2>(nn);
  double2 *R = sc.alloc<double2>(nn);
  double2 *Tp = sc.alloc<double2>((size_t)n * L), *T = sc.alloc<double2>((size_t)n * L);
  double2 *P = sc.alloc<double2>((size_t)n * L), *Q = sc.alloc<double2>((size_t)n * L);
  double2 *Ph = sc.alloc<double2>((size_t)n * L);
  double2 *Sg = sc.alloc<double2>((size_t)nE * cc), *Gm = sc.alloc<double2>((size_t)nE * cc);
  double2 *G0 = sc.alloc<double2>((size_t)nE * cc), *Mm = sc.alloc<double2>((size_t)nE * cc);
  double2 *Ki = sc.alloc<double2>((size_t)nE * cc), *S = sc.alloc<double2>((size_t)nE * cc);
  HIPCHK(hipMemcpyAsync(Ed, Eh.data(), sizeof(double2) * nE, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(Wd, Wh.data(), sizeof(double2) * nE, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cd, corb, sizeof(int) * nc, hipMemcpyHostToDevice, st));
  typedef Tile<64, 64, 16, 2, false, 3> TBig;
  typedef Tile<32, 32, 16, 2, false, 4> TSmall;
  const dim3 g1(1024), b1(256);
  for (int k = 0; k < NK; ++k) {
    HIPCHK(hipMemcpyAsync(V, evecs + 2 * (size_t)k * nn, sizeof(double2) * nn,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ed, evals + (size_t)k * n, sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(Sg, sigma_cc + 2 * (size_t)k * nE * cc, sizeof(double2) * nE * cc,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(Gm, gamma_w + 2 * (size_t)k * nE * cc, sizeof(double2) * nE * cc,
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scf_adjoint, g1, b1, 0, st, n, nc, (const double2*)V, (const int*)cd, Vh,
                       Vch);
    hipLaunchKernelGGL(k_scf_factors, g1, b1, 0, st, n, nE, nc, (const double2*)Ed,
                       (const double*)ed, (const double2*)V, (const int*)cd, Tp, T);
    HIPCHK(hipGetLastError());
    // G0_i,cc = T_i V_c†, then I - Σ_i G0_i,cc, then its inverse K_i (R = I, solved in place)
    GemmArgs<double, double> ga{nc, nc, n, T, n, (long)nc * n, Vch, nc, 0};
    HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
        ga, nE, EpiStore<double2>{G0, nc, cc, 1.0, 0.0}, st)));
    GemmArgs<double, double> gb{nc, nc, nc, Sg, nc, cc, G0, nc, cc};
    HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
        gb, nE, EpiIdMinus{Mm, nc, cc}, st)));
    hipLaunchKernelGGL(k_identity_batch, g1, b1, 0, st, nE, nc, Ki);
    hipLaunchKernelGGL(k_gauss_jordan, dim3(nE), dim3(256), 0, st, nc, Mm, Ki);
    HIPCHK(hipGetLastError());
    // S_i = (w_i/2π) K_i M_i  (nc x nc x nc)
    GemmArgs<double, double> gs{nc, nc, nc, Ki, nc, cc, Gm, nc, cc};
    HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
        gs, nE, EpiScaleZ{S, nc, cc, Wd}, st)));
    // P_i = T'_i K_i and Q_i = T'_i S_i, column block i of (n, nE nc)
    GemmArgs<double, double> gp{n, nc, nc, Tp, L, (long)nc, Ki, nc, cc};
    HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
        gp, nE, EpiStore<double2>{P, L, (long)nc, 1.0, 0.0}, st)));
    GemmArgs<double, double> gq{n, nc, nc, Tp, L, (long)nc, S, nc, cc};
    HIPCHK((launch_zgemm<double, double, double, false, false, TSmall, 3>(
        gq, nE, EpiStore<double2>{Q, L, (long)nc, 1.0, 0.0}, st)));
    hipLaunchKernelGGL(k_conj_transpose, g1, b1, 0, st, n, (int)L, (const double2*)P, Ph);
    HIPCHK(hipGetLastError());
    // Y = Q P†  (n x nE nc)(nE nc x n); X = V Y V†; ρ = Hermitian part of X
    GemmArgs<double, double> gy{n, n, (int)L, Q, L, 0, Ph, n, 0};
    HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
        gy, 1, EpiStore<double2>{Y, n, nn, 1.0, 0.0}, st)));
    GemmArgs<double, double> gz{n, n, n, V, n, 0, Y, n, 0};
    HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
        gz, 1, EpiStore<double2>{Z, n, nn, 1.0, 0.0}, st)));
    GemmArgs<double, double> gx{n, n, n, Z, n, 0, Vh, n, 0};
    HIPCHK((launch_zgemm<double, double, double, false, false, TBig, 3>(
        gx, 1, EpiStore<double2>{X, n, nn, 1.0, 0.0}, st)));
    hipLaunchKernelGGL(k_herm_part, g1, b1, 0, st, n, (const double2*)X, R);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rho + 2 * (size_t)k * nn, R, sizeof(double2) * nn, hipMemcpyDeviceToHost,
                          st));
    HIPCHK(hipStreamSynchronize(st));
  }
  for (size_t i = 0; i < 2 * (size_t)NK * nn; ++i)
    require(std::isfinite(rho[i]), "non-finite density (singular Woodbury matrix?)", ZH_ENONFINITE);
}

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

static void problem_create(const zh_dims* d, int instances, zh_problem** out) {
  require(d && out, "null argument");
  require(d->NK > 0 && d->n > 0 && d->NL > 0 && d->NX > 0 && d->NOT > 0, "dimensions must be positive");
  require(instances >= 1, "at least one instance");
  require((long)d->NK * instances <= (1 << 20), "too many k blocks");
  require(d->NL <= ZH_MAX_LEADS, "too many leads");
  require(d->precision == ZH_FP64 || d->precision == ZH_MIXED, "unknown precision");
  require(d->hbar > 0, "hbar must be positive");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  require(d->device >= 0 && d->device < ndev, "no such HIP device");
  auto* p = new zh_problem();
  p->dims = *d;
  p->dims.NK = d->NK * instances;
  p->instances = instances;
  try {
    if (d->precision == ZH_FP64) p->eng.reset(new Engine<double>(*d, instances));
    else p->eng.reset(new Engine<float>(*d, instances));
  } catch (...) {
    delete p;
    throw;
  }
  *out = p;
}

int zh_problem_create(const zh_dims* d, zh_problem** out) {
  return zh_try([&] { problem_create(d, 1, out); });
}

int zh_problem_create_instances(const zh_dims* d, int instances, zh_problem** out) {
  return zh_try([&] { problem_create(d, instances, out); });
}

int zh_instances(zh_problem* p, int* instances) {
  return zh_try([&] {
    require(p && instances, "null argument");
    *instances = p->instances;
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

int zh_set_drive(zh_problem* p, int nterms, const int* kinds, const double* amps,
                 const double* par, const int* nnz, const int* ks, const int* rows,
                 const int* cols, const double* vals) {
  return zh_try([&] {
    require(p, "null argument");
    require(nterms == 0 || (kinds && amps && par && nnz), "null argument");
    long tot = 0;
    for (int i = 0; i < nterms; ++i) tot += nnz[i] > 0 ? nnz[i] : 0;
    require(tot == 0 || (ks && rows && cols && vals), "null argument");
    p->eng->set_drive(nterms, kinds, amps, par, nnz, ks, rows, cols, vals);
  });
}

static PulseDesc pulse_desc(zh_problem* p, int kind, double amp, const double* par,
                            const double* w) {
  require(p && w && par, "null argument");
  require(kind >= ZH_PULSE_NONE && kind <= ZH_PULSE_SPLINE, "unknown pulse kind");
  require(!((kind == ZH_PULSE_GAUSS || kind == ZH_PULSE_BOX) && par[1] == 0.0), "zero pulse width");
  require(!(kind == ZH_PULSE_RAMP && par[0] == 0.0), "zero ramp time");
  require(!(kind == ZH_PULSE_GENERIC && par[0] == 0.0), "zero envelope width");
  require(!(kind == ZH_PULSE_STAIRS && (par[2] == 0.0 || par[1] < 0.0 || par[1] > 1e6)),
          "stairs: softness must be non-zero and the count in [0, 1e6]");
  require(!(kind == ZH_PULSE_SPLINE && par[2] == 0.0), "zero window softness");
  PulseDesc d;
  std::memset(&d, 0, sizeof(d));
  d.kind = kind; d.amp = amp; d.p1 = par[0]; d.p2 = par[1]; d.p3 = par[2]; d.p4 = par[3];
  for (int a = 0; a < p->dims.NL; ++a) d.w[a] = w[a];
  return d;
}

int zh_set_pulse_ex(zh_problem* p, int kind, double amp, const double* par, const double* w) {
  return zh_try([&] { p->eng->set_pulse(pulse_desc(p, kind, amp, par, w)); });
}

int zh_instance_pulse(zh_problem* p, int instance, int kind, double amp, const double* par,
                      const double* w) {
  return zh_try([&] { p->eng->set_instance_pulse(instance, pulse_desc(p, kind, amp, par, w)); });
}

int zh_set_pulse_spline(zh_problem* p, int nknots, const double* x, const double* c) {
  return zh_try([&] {
    require(p && x && c, "null argument");
    require(nknots >= 2 && nknots <= (1 << 20), "spline: 2 to 2^20 knots");
    for (int i = 1; i < nknots; ++i) require(x[i] > x[i - 1], "spline knots must increase");
    p->eng->set_pulse_spline(nknots, x, c);
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

int zh_steady_state(zh_problem* p, double t0, double tol, int maxiter, const zh_precond* pc,
                    int* iters, double* rel_resid) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->steady_state(t0, tol, maxiter, pc, iters, rel_resid);
  });
}

int zh_set_dissipator(zh_problem* p, int kind, double eta0, double t0, double width,
                      const double* sig_eq) {
  return zh_try([&] {
    require(p, "null argument");
    require(kind >= ZH_DISS_NONE && kind <= ZH_DISS_STEP, "unknown dissipator kind");
    require(kind == ZH_DISS_NONE || sig_eq, "the dissipator needs sigma_eq");
    require(kind != ZH_DISS_STEP || width != 0.0, "zero dissipator step width");
    DissDesc d{kind, eta0, t0, width == 0.0 ? 1.0 : width};
    p->eng->set_dissipator(d, sig_eq);
  });
}

int zh_window_density(int NK, int n, int nE, int nc, const double* nodes, const double* evals,
                      const double* evecs, const int* corb, const double* sigma_cc,
                      const double* gamma_w, double* rho, int device) {
  return zh_try([&] {
    require(nodes && evals && evecs && rho, "null argument");
    window_density(NK, n, nE, nc, nodes, evals, evecs, corb, sigma_cc, gamma_w, rho, device);
  });
}

int zh_contour_density(int NK, int n, int nE, int nc, const double* contour,
                       const double* evals, const double* evecs, const int* corb,
                       const double* sigma_cc, double* rho, int device) {
  return zh_try([&] {
    require(contour && evals && evecs && rho, "null argument");
    contour_density(NK, n, nE, nc, contour, evals, evecs, corb, sigma_cc, rho, device);
  });
}

int zh_contour_density_noalloc(int NK, int n, int nE, int nc, const double* contour,
                               const double* evals, const double* evecs, const int* corb,
                               const double* sigma_cc, double* rho, int device) {
  return zh_try([&] {
    require(contour && evals && evecs && rho, "null argument");
    contour_density_noalloc(NK, n, nE, nc, contour, evals, evecs, corb, sigma_cc, rho, device);
  });
}

int zh_pi_get(zh_problem* p, int hermitian, double* Q) {
  return zh_try([&] {
    require(p && Q, "null argument");
    p->eng->pi_get(hermitian, Q);
  });
}

int zh_set_tolerance(zh_problem* p, double rtol, double atol) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_tolerance(rtol, atol);
  });
}

int zh_set_overlap(zh_problem* p, const double* Sinv) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_overlap(Sinv);
  });
}

int zh_state_put(zh_problem* p, const double* sig, const double* psi, const double* omg) {
  return zh_try([&] {
    require(p && sig, "null argument");
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

static void one_instance(zh_problem* p) {
  require(p->instances == 1, "a problem with several instances attempts all of them together "
                             "(zh_attempt_instances)");
}

int zh_attempt(zh_problem* p, zh_attempt_info* info, double* J) {
  return zh_try([&] {
    require(p && info, "null argument");
    one_instance(p);
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
    one_instance(p);
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
      try {
        p->eng->attempt(&info, Jtmp.data(), true);
      } catch (...) {
        // the attempt was not committed: the counts describe the resident state
        *n_steps = steps; *n_rejected = rej; if (h_final) *h_final = h;
        throw;
      }
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

int zh_attempt_launch(zh_problem* p, const zh_attempt_info* info) {
  return zh_try([&] {
    require(p && info, "null argument");
    one_instance(p);
    require(info->h > 0 && std::isfinite(info->h), "step size must be positive and finite");
    p->eng->launch(info);
  });
}

int zh_attempt_collect(zh_problem* p, zh_attempt_info* info, double* J) {
  return zh_try([&] {
    require(p && info, "null argument");
    one_instance(p);
    p->eng->collect(info, J);
    if (!std::isfinite(info->te)) throw Error{ZH_ENONFINITE, "non-finite error estimate"};
  });
}

static void check_instance_infos(zh_problem* p, const zh_attempt_info* info) {
  for (int g = 0; g < p->instances; ++g) {
    const bool frozen = info[g].h == 0.0 && info[g].force_accept;
    require(frozen || (info[g].h > 0 && std::isfinite(info[g].h)),
            "instance " + std::to_string(g) + ": step size must be positive and finite (or 0 "
            "with force_accept to hold a finished instance)");
  }
}

int zh_launch_instances(zh_problem* p, const zh_attempt_info* info) {
  return zh_try([&] {
    require(p && info, "null argument");
    check_instance_infos(p, info);
    p->eng->launch(info);
  });
}

int zh_collect_instances(zh_problem* p, zh_attempt_info* info, double* J) {
  return zh_try([&] {
    require(p && info, "null argument");
    p->eng->collect(info, J);
  });
}

int zh_attempt_instances(zh_problem* p, zh_attempt_info* info, double* J) {
  return zh_try([&] {
    require(p && info, "null argument");
    check_instance_infos(p, info);
    p->eng->attempt(info, J);
  });
}

int zh_attempt_ready(zh_problem* p, int* ready) {
  return zh_try([&] {
    require(p && ready, "null argument");
    *ready = p->eng->ready();
  });
}

int zh_batch_attempt(zh_problem* const* ps, int count, zh_attempt_info* infos, double* const* J,
                     int* status) {
  if (count < 0 || (count > 0 && (!ps || !infos || !J || !status))) {
    g_err = "null argument";
    return ZH_EINVAL;
  }
  // every launch first (each problem has its own streams), then every collect; one
  // problem's failure is reported in its status and does not stop the others
  int first = ZH_OK;
  std::string msg;
  for (int i = 0; i < count; ++i) {
    status[i] = zh_attempt_launch(ps[i], &infos[i]);
    if (status[i] != ZH_OK && first == ZH_OK) { first = status[i]; msg = g_err; }
  }
  for (int i = 0; i < count; ++i) {
    if (status[i] != ZH_OK) continue;
    status[i] = zh_attempt_collect(ps[i], &infos[i], J[i]);
    if (status[i] != ZH_OK && first == ZH_OK) { first = status[i]; msg = g_err; }
  }
  if (first != ZH_OK) g_err = msg;
  return first;
}

int zh_autotune(zh_problem* p, int force, int* choice) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->autotune(force, choice);
  });
}

int zh_set_tiles(zh_problem* p, const int* choice) {
  return zh_try([&] {
    require(p && choice, "null argument");
    p->eng->set_tiles(choice);
  });
}

int zh_problem_bytes(zh_problem* p, long* bytes) {
  return zh_try([&] {
    require(p && bytes, "null argument");
    *bytes = p->eng->bytes();
  });
}

int zh_set_layout(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_layout(mode);
  });
}

int zh_layout_info(zh_problem* p, int* out) {
  return zh_try([&] {
    require(p && out, "null argument");
    p->eng->layout_info(out);
  });
}

int zh_set_sparse(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_sparse(mode);
  });
}

int zh_hamiltonian_info(zh_problem* p, int* out) {
  return zh_try([&] {
    require(p && out, "null argument");
    p->eng->hamiltonian_info(out);
  });
}

int zh_set_fusion(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_fusion(mode);
  });
}

int zh_set_psi_kernel(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_psi_kernel(mode);
  });
}

int zh_psi_kernel_info(zh_problem* p, int* fused) {
  return zh_try([&] {
    require(p && fused, "null argument");
    *fused = p->eng->psi_kernel_info();
  });
}

int zh_set_omgxi(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_omgxi(mode);
  });
}

int zh_set_fsal(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_fsal(mode);
  });
}

int zh_fsal_info(zh_problem* p, long* taken, int* on) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->fsal_info(taken, on);
  });
}

int zh_omgxi_info(zh_problem* p, int* out) {
  return zh_try([&] {
    require(p && out, "null argument");
    p->eng->omgxi_info(out);
  });
}

int zh_set_antihermitian(zh_problem* p, int mode) {
  return zh_try([&] {
    require(p, "null argument");
    p->eng->set_antihermitian(mode);
  });
}

int zh_antihermitian_info(zh_problem* p, int* out) {
  return zh_try([&] {
    require(p && out, "null argument");
    *out = p->eng->antihermitian_info();
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
