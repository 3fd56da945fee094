// Elementwise / reduction kernels and GEMM epilogues of the AME right-hand side
// and the DOP54 attempt.  All are HBM-streaming: one pass over their operands,
// coalesced along the fastest (orbital / mode) index.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "zh_gemm.hpp"
#include "../../include/zandhip.h"

namespace zh {

struct StageParams {
  double ts, h;
  double delta[ZH_MAX_LEADS];
  double eta;  // dissipator rate at ts
  double fdrive[ZH_MAX_DRIVE];  // time factors of the drive terms at ts
  // first same as last: k_1 of this attempt is the previous attempt's k_7, left in the
  // k_1 slot and scaled by r1 = h / h_prev where it is used; skip (set in the first
  // stage's entry only) tells the first stage's kernels to leave the instance alone
  double r1;
  int skip;
};

// Time factors f_i(t) of the drive terms H(t) = H0 + Σ_i f_i(t) M_i (the waveforms the
// reference's examples write into dH(t, σ); ZH_WAVE_* in zandhip.h).
struct DriveDesc {
  int nt;
  int kind[ZH_MAX_DRIVE];
  double amp[ZH_MAX_DRIVE];
  double par[ZH_MAX_DRIVE][4];
};
__device__ inline double drive_value(const DriveDesc& d, int i, double t) {
  const double a = d.par[i][0], b = d.par[i][1], c = d.par[i][2], e = d.par[i][3];
  switch (d.kind[i]) {
    case ZH_WAVE_CONST: return d.amp[i];
    case ZH_WAVE_COS_WINDOW: return (t > c && t <= e) ? d.amp[i] * cos(a * t + b) : 0.0;
    case ZH_WAVE_SIN_FERMI: return d.amp[i] * sin(a * t + b) / (1.0 + exp((t - c) / e));
    case ZH_WAVE_GAUSS_SIN: {
      const double x = (t - e) / a;
      return d.amp[i] * exp(-x * x) * sin(2.0 * M_PI * b * (t - e) - c);
    }
    default: return 0.0;
  }
}

// dissipator(t, σ) = η(t)(σ - σ_eq) (Zandpack/mpi/zand:343 hook; the form of
// examples/Hubbard/Calculation/Bias.py:29-39).
struct DissDesc {
  int kind;
  double eta0, t0, width;
};
__device__ inline double diss_rate(const DissDesc& d, double t) {
  switch (d.kind) {
    case ZH_DISS_CONST: return d.eta0;
    case ZH_DISS_STEP: return d.eta0 / (1.0 + exp((t - d.t0) / d.width));
    default: return 0.0;
  }
}

struct Ctrl {
  double t0, h, eps;
  int force;
  int accept;
  double te2, te, h_next;
  double norm_scale;  // TE = sqrt(norm_scale * sum): 1 (absolute norm) or 1/count (scaled)
  // first same as last: fsal = 1 when the k_1 slot holds the k_7 of the previous attempt
  // (accepted, step hprev > 0, nothing changed since); fsal_used reports whether this
  // attempt took it (the stage parameters at t0 equal those of that k_7, bit for bit)
  int fsal;
  double hprev;
  int fsal_used;
  // the controller's forecast: 1 when the stage parameters at t0 + h_next equal those of
  // this attempt's last stage (a next attempt started there may reuse its k_7)
  int fsal_hn;
};

// Error weight of one complex entry in the scaled norm: 1/(atol + rtol max(|y0|, |y1|))²
__device__ inline double te_weight(double atol, double rtol, double y0r, double y0i,
                                   double y1r, double y1i) {
  const double sc = atol + rtol * fmax(sqrt(y0r * y0r + y0i * y0i), sqrt(y1r * y1r + y1i * y1i));
  return 1.0 / (sc * sc);
}

struct PulseDesc {
  int kind;
  double amp, p1, p2, hbar;
  double w[ZH_MAX_LEADS];
  double p3, p4;
  const double* sx;  // spline knots (device), ns of them
  const double* sc;  // spline coefficients (4, ns-1) (device)
  int ns;
};

// Dormand-Prince 5(4) (Zandpack/mpi_RK4pars_dev.py:121-138): nodes c, stage
// matrix a (row i = coefficients of k_0..k_{i-1}), error weights ct = b5 - b4.
struct Tableau {
  double c[7];
  double a[7][7];
  double ct[7];
};

__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __host__ __device__ inline double2 to_d2(C v) {
  return make_double2((double)v.x, (double)v.y);
}
template <typename C> __host__ __device__ inline C from_d2(double2 v);
template <> __host__ __device__ inline double2 from_d2<double2>(double2 v) { return v; }
template <> __host__ __device__ inline float2 from_d2<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}

// Load of a slot element that is streamed once per pass (non-temporal: it does not
// displace the reused operands of the concurrent GEMMs from L2).
template <typename C> __device__ inline C ld_stream(const C* p) {
  typedef decltype(p->x) E;
  typedef E v2 __attribute__((ext_vector_type(2)));
  const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  C v;
  v.x = x.x;
  v.y = x.y;
  return v;
}

template <typename C> __device__ inline void st_stream(C* p, C v) {
  typedef decltype(p->x) E;
  typedef E v2 __attribute__((ext_vector_type(2)));
  const v2 x = {v.x, v.y};
  __builtin_nontemporal_store(x, reinterpret_cast<v2*>(p));
}

__device__ inline double pulse_value(const PulseDesc& p, double t) {
  switch (p.kind) {
    case ZH_PULSE_STEP: return t >= p.p1 ? p.amp : 0.0;
    case ZH_PULSE_GAUSS: { const double x = (t - p.p1) / p.p2; return p.amp * exp(-x * x); }
    case ZH_PULSE_RAMP: return p.amp * (1.0 - exp(-t / p.p1));
    case ZH_PULSE_SIN: return p.amp * sin(p.p1 * t / p.hbar);
    case ZH_PULSE_BOX: return p.amp * (tanh(t / p.p2) - tanh((t - p.p1) / p.p2)) * 0.5;
    case ZH_PULSE_GENERIC: {  // p1 = s, p2 = f, p3 = phase, p4 = delay
      const double x = (t - p.p4) / p.p1;
      return p.amp * exp(-x * x) * sin(2.0 * M_PI * p.p2 * (t - p.p4) - p.p3);
    }
    case ZH_PULSE_STAIRS: {  // p1 = width, p2 = n, p3 = softness
      double r = 0.0;
      const int ns = (int)p.p2;
      for (int i = 1; i < ns; ++i) r += 1.0 - 1.0 / (1.0 + exp((t - i * p.p1) / p.p3));
      return p.amp * r;
    }
    case ZH_PULSE_SPLINE: {  // p1 = a, p2 = b, p3 = ts
      if (p.ns < 2 || t < p.sx[0] || t > p.sx[p.ns - 1]) return 0.0;
      int lo = 0, hi = p.ns - 2;  // interval i: the last knot x_i <= t
      while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (p.sx[mid] <= t) lo = mid; else hi = mid - 1;
      }
      const double d = t - p.sx[lo];
      const int m1 = p.ns - 1;
      const double s = ((p.sc[lo] * d + p.sc[m1 + lo]) * d + p.sc[2 * m1 + lo]) * d + p.sc[3 * m1 + lo];
      return p.amp * s * (tanh((t - p.p1) / p.p3) - tanh((t - p.p2) / p.p3)) * 0.5;
    }
    default: return 0.0;
  }
}

// The right-hand side depends on t only through these parameters: equal parameters
// (bit for bit) give the same f(t, y).
__device__ inline bool same_params(const StageParams& a, const StageParams& b, int NL,
                                   int nt) {
  bool eq = a.eta == b.eta;
  for (int i = 0; i < NL; ++i) eq = eq && a.delta[i] == b.delta[i];
  for (int i = 0; i < nt; ++i) eq = eq && a.fdrive[i] == b.fdrive[i];
  return eq;
}

// Stage times and biases of one attempt from the device control blocks, one workgroup
// per instance g of a problem that carries G independent instances (each its own
// control block, pulse and step): sp[s * G + g].  First same as last: when the host
// offers it (ctrl.fsal) and the parameters at t0 equal those of the previous attempt's
// last stage (still in sp[6 G + g]), that stage's k_7 = h_prev f(t0, y0) stands in for
// k_1 (skip1[g] = 1, r1 = h / h_prev); the tableau's last row is its weights and
// c_7 = 1, so the last stage's input is this attempt's y0.
__device__ inline StageParams stage_params_at(const PulseDesc& pulse, const DissDesc& diss,
                                              const DriveDesc& drive, int NL, double ts,
                                              double h) {
  StageParams q;
  q.ts = ts;
  q.h = h;
  const double v = pulse_value(pulse, ts);
  for (int a = 0; a < ZH_MAX_LEADS; ++a) q.delta[a] = a < NL ? pulse.w[a] * v : 0.0;
  q.eta = diss_rate(diss, ts);
  for (int i = 0; i < ZH_MAX_DRIVE; ++i) q.fdrive[i] = i < drive.nt ? drive_value(drive, i, ts) : 0.0;
  q.r1 = 1.0;
  q.skip = 0;
  return q;
}
// One workgroup of 64 threads per instance: thread 0 decides first same as last (before
// the last stage's entry is replaced), then thread s < 7 writes stage s.
__global__ void k_stage_params(Ctrl* ctrl, const PulseDesc* pulses, DissDesc diss,
                               DriveDesc drive, Tableau tab, int NL, int G, StageParams* sp,
                               int* skip1, int NK, const double* Jn, double* J) {
  const int g = blockIdx.x, s = threadIdx.x;
  const PulseDesc pulse = pulses[g];
  const double t0 = ctrl[g].t0, h = ctrl[g].h;
  __shared__ int fsal_s;
  __shared__ double r1_s;
  StageParams q;
  if (s < 7) q = stage_params_at(pulse, diss, drive, NL, t0 + tab.c[s] * h, h);
  if (s == 0) {
    // the first stage's parameters against those of the last attempt's seventh stage
    // (read here, rewritten below after the barrier)
    const Ctrl& c = ctrl[g];
    const bool fsal = c.fsal && c.hprev > 0.0 && h > 0.0 && same_params(q, sp[6 * G + g], NL,
                                                                        drive.nt);
    fsal_s = fsal ? 1 : 0;
    r1_s = fsal ? h / c.hprev : 1.0;
  }
  __syncthreads();
  if (s < 7) {
    q.r1 = r1_s;
    q.skip = (s == 0) ? fsal_s : 0;
    sp[s * G + g] = q;
  }
  if (s == 0) {
    skip1[g] = fsal_s;
    ctrl[g].fsal_used = fsal_s;
  }
  // the currents at y0 of an instance that reuses k_7: those its last stage formed
  const int KI = NK / G;
  if (fsal_s)
    for (int i = s; i < NL * KI; i += blockDim.x) {
      const int a = i / KI, b = g * KI + i % KI;
      J[a * NK + b] = Jn[a * NK + b];
    }
}

// H(t_s) = H0 + Σ_i f_i(t_s) M_i written at the drive's entries into the Hamiltonian
// operands of the right-hand side: the complex copies (A of Hσ, Hᵀ block of the Ψ
// product's B), the real copies when H is real, the row-compressed values when the H
// products are gathers.  Entry e: (k, row, col), H0 value, its terms' coefficients
// term[beg[e] .. beg[e+1]).  The entries are distinct, so writes never collide.
template <typename C, typename RT>
struct DriveArgs {
  int E, n, nc, R;
  const int *k, *r, *c, *slot;  // slot: position in the row-compressed row (sparse)
  const double2* h0;
  const int* beg;
  const int* term;
  const double2* coef;
  double2* Hcur; C* Bp; double* Hre; RT* HreT; double2* ell_val;
  const StageParams* sp;  // the stage's parameters, one per instance of ki k blocks
  int ki = 1 << 30;
};
template <typename C, typename RT>
__global__ void k_drive_update(DriveArgs<C, RT> a) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < a.E; e += gridDim.x * blockDim.x) {
    double2 v = a.h0[e];
    for (int q = a.beg[e]; q < a.beg[e + 1]; ++q) {
      const double f = a.sp[a.k[e] / a.ki].fdrive[a.term[q]];
      v.x += f * a.coef[q].x;
      v.y += f * a.coef[q].y;
    }
    const long b = a.k[e], r = a.r[e], c = a.c[e], n = a.n;
    a.Hcur[(b * n + r) * n + c] = v;
    a.Bp[(b * (n + a.nc) + c) * n + r] = from_d2<C>(v);
    if (a.Hre) a.Hre[(b * n + r) * n + c] = v.x;
    if (a.HreT) a.HreT[(b * n + c) * n + r] = (RT)v.x;
    if (a.ell_val && a.slot[e] >= 0) a.ell_val[(b * a.R + a.slot[e]) * n + r] = v;
  }
}

// Non-orthogonal form with a σ- or t-dependent Hamiltonian (Zandpack/mpi/nozand:427,
// 456-461, 1075-1085): S⁻¹H(t, σ) = S⁻¹H0 + Σ_i f_i(t) S⁻¹M_i + S⁻¹ diag(d), d the Hubbard
// diagonal of the stage, into the complex A operand of Hσ and (transposed) the Hᵀ block
// of the Ψ product's B.
template <typename C>
__global__ void k_overlap_model(int NK, int n, int nc, int nt, const double2* SH0,
                                const double2* SM, const double2* Sinv, const double2* dvec,
                                const StageParams* sp, int ki, double2* Hcur, C* Bp) {
  const long nn = (long)n * n, total = (long)NK * nn;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / nn, rc = idx - b * nn;
    const long r = rc / n, c = rc - r * n;
    double2 v = SH0[idx];
    for (int i = 0; i < nt; ++i) {
      const double f = sp[b / ki].fdrive[i];
      const double2 m = SM[(long)i * total + idx];
      v.x += f * m.x;
      v.y += f * m.y;
    }
    if (dvec) {
      const double2 t = cmul(Sinv[idx], dvec[b * n + c]);
      v.x += t.x;
      v.y += t.y;
    }
    Hcur[idx] = v;
    Bp[(b * (n + nc) + c) * n + r] = from_d2<C>(v);
  }
}

// Row maps of the elementwise slot kernels: visited row r -> offsets of its first
// element in the source slots and in the destination(s), and its visited length.
struct MapDense {  // rows x cols block, source row stride lds, destination strides ld1 / ld2
  int rows, cols; long lds, ld1, ld2;
  __host__ __device__ inline int nrows() const { return rows; }
  // every row's offsets and length even: complex64 elements go in 16-byte pairs
  __host__ __device__ inline bool pairs() const {
    return cols % 2 == 0 && lds % 2 == 0 && ld1 % 2 == 0 && ld2 % 2 == 0;
  }
  __device__ inline int row(int r, long& src, long& d1, long& d2) const {
    src = r * lds; d1 = r * ld1; d2 = r * ld2;
    return cols;
  }
};

// dst = y0 + sum_j a_j k_j   (Zandpack/mpi/zand:306-313, 722-735)
// With sp: rows of instance (r / rows_b) / ki scale term k1 (the k_1 slot) by that
// instance's r1, and skip the instance when its skip flag is set.
template <typename C>
struct CombineArgs {
  const C* y0;
  const C* k[7];
  double a[7];
  int nterms;
  C* d1;
  C* d2;
  const StageParams* sp = nullptr;
  int ki = 1 << 30, k1 = -1;
  int rows_b = 1 << 30;
  // also e_out = Σ_j ect[j] k_j over the same terms (slot layout), or none: the last
  // stage's input reads the k_j the error norm weights, all but k_7
  C* e_out = nullptr;
  double ect[7] = {0};
};
// the coefficients of one row's instance; false: the row is left alone
template <typename C>
__device__ inline bool combine_coefs(const CombineArgs<C>& g, int r, double* a, double* e) {
#pragma unroll
  for (int j = 0; j < 7; ++j) { a[j] = g.a[j]; e[j] = g.ect[j]; }
  if (!g.sp) return true;
  const StageParams& q = g.sp[(r / g.rows_b) / g.ki];
  if (q.skip) return false;
  if (g.k1 >= 0) { a[g.k1] *= q.r1; e[g.k1] *= q.r1; }
  return true;
}

// NT = number of stage terms, a compile-time constant so that all NT+1 slot
// loads of an element are in flight together.
template <typename C, typename Map, int NT>
__global__ __launch_bounds__(256) void k_combine(CombineArgs<C> g, Map mp) {
  const int nr = mp.nrows();
  for (int r = blockIdx.x; r < nr; r += gridDim.x) {
    double a[7], ec[7];
    if (!combine_coefs(g, r, a, ec)) continue;
    long src, o1, o2;
    const int cols = mp.row(r, src, o1, o2);
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
      C kv[NT + 1];
      kv[0] = ld_stream(g.y0 + src + c);
#pragma unroll
      for (int j = 0; j < NT; ++j) kv[j + 1] = ld_stream(g.k[j] + src + c);
      double vr = (double)kv[0].x, vi = (double)kv[0].y;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        vr += a[j] * (double)kv[j + 1].x;
        vi += a[j] * (double)kv[j + 1].y;
      }
      const C out = from_d2<C>(make_double2(vr, vi));
      g.d1[o1 + c] = out;
      if (g.d2) g.d2[o2 + c] = out;
      if (g.e_out) {
        double er = 0.0, ei = 0.0;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          er += ec[j] * (double)kv[j + 1].x;
          ei += ec[j] * (double)kv[j + 1].y;
        }
        st_stream(g.e_out + src + c, from_d2<C>(make_double2(er, ei)));
      }
    }
  }
}

// Complex64 elements two per thread (16-byte loads and stores), for maps whose rows
// have even offsets and lengths (slot bases are allocation-aligned).
template <typename C, typename Map, int NT>
__global__ __launch_bounds__(256) void k_combine_pairs(CombineArgs<C> g, Map mp) {
  static_assert(sizeof(C) == 8, "complex64 pairs");
  const int nr = mp.nrows();
  for (int r = blockIdx.x; r < nr; r += gridDim.x) {
    double a[7], ec[7];
    if (!combine_coefs(g, r, a, ec)) continue;
    long src, o1, o2;
    const int cols = mp.row(r, src, o1, o2);
    for (int c = 2 * threadIdx.x; c < cols; c += 2 * blockDim.x) {
      v4f kv[NT + 1];
      kv[0] = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(g.y0 + src + c));
#pragma unroll
      for (int j = 0; j < NT; ++j)
        kv[j + 1] = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(g.k[j] + src + c));
      double v0r = kv[0].x, v0i = kv[0].y, v1r = kv[0].z, v1i = kv[0].w;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        v0r += a[j] * (double)kv[j + 1].x;
        v0i += a[j] * (double)kv[j + 1].y;
        v1r += a[j] * (double)kv[j + 1].z;
        v1i += a[j] * (double)kv[j + 1].w;
      }
      const v4f out = {(float)v0r, (float)v0i, (float)v1r, (float)v1i};
      *reinterpret_cast<v4f*>(g.d1 + o1 + c) = out;
      if (g.d2) *reinterpret_cast<v4f*>(g.d2 + o2 + c) = out;
      if (g.e_out) {
        double e0r = 0.0, e0i = 0.0, e1r = 0.0, e1i = 0.0;
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          e0r += ec[j] * (double)kv[j + 1].x;
          e0i += ec[j] * (double)kv[j + 1].y;
          e1r += ec[j] * (double)kv[j + 1].z;
          e1i += ec[j] * (double)kv[j + 1].w;
        }
        const v4f e = {(float)e0r, (float)e0i, (float)e1r, (float)e1i};
        __builtin_nontemporal_store(e, reinterpret_cast<v4f*>(g.e_out + src + c));
      }
    }
  }
}

template <typename C, typename Map>
inline void launch_combine(const CombineArgs<C>& g, const Map& mp, int blocks, hipStream_t s) {
  if constexpr (sizeof(C) == 8) {
    if (mp.pairs()) {
      switch (g.nterms) {
#define ZH_COMBINE_CASE(N)                                                                   \
  case N:                                                                                    \
    hipLaunchKernelGGL((k_combine_pairs<C, Map, N>), dim3(blocks), dim3(256), 0, s, g, mp); break;
        ZH_COMBINE_CASE(0) ZH_COMBINE_CASE(1) ZH_COMBINE_CASE(2) ZH_COMBINE_CASE(3)
        ZH_COMBINE_CASE(4) ZH_COMBINE_CASE(5) ZH_COMBINE_CASE(6)
#undef ZH_COMBINE_CASE
        default: break;
      }
      return;
    }
  }
  switch (g.nterms) {
#define ZH_COMBINE_CASE(N)                                                                   \
  case N:                                                                                    \
    hipLaunchKernelGGL((k_combine<C, Map, N>), dim3(blocks), dim3(256), 0, s, g, mp); break;
    ZH_COMBINE_CASE(0) ZH_COMBINE_CASE(1) ZH_COMBINE_CASE(2) ZH_COMBINE_CASE(3)
    ZH_COMBINE_CASE(4) ZH_COMBINE_CASE(5) ZH_COMBINE_CASE(6)
#undef ZH_COMBINE_CASE
    default: break;
  }
}

// Per-stage gathers of the stage input onto the coupled orbitals:
//   Aomg[:, :nc] = h psi[:, cidx], Bomg[:, nc:] = h conj(psi[:, cidx])   (the Ω
//     sources come out of their GEMM already scaled by the step),
//   Bp[n + c, o] = hbar * sigma[o, cidx[c]], dvec = Hubbard diagonal,
//   u_i = (h/ħ)(χ⁺_i + Δ_α(i)), v_j = (h/ħ)(χ⁻_j + Δ_α(j))   (real parts shifted),
// so that the Ω epilogue reads two per-mode values instead of the lead table,
// the bias and the step.
template <typename C>
struct GatherArgs {
  int NK, n, M, nc;
  const C* Ap; long ldAp; long sAp;          // stage psi (left n columns)
  C* Aomg; C* Bomg;                          // (NK, M, 2nc)
  C* Bp;                                     // (NK, n+nc, n)
  const double2* sig;                        // (NK, n, n)
  const int* cidx;
  double2* dvec;                             // (NK, n)
  const double* n0;                          // (NK, n)
  double U, hbar;
  int hubbard;
  const StageParams* sp;                     // per instance of ki k blocks
  int ki = 1 << 30;
  const double2* chip; const double2* chim;  // (NK, M)
  const int* lead;                           // (M)
  double2* u; double2* v;                    // (NK, M)
};

template <typename C>
__global__ void k_gather(GatherArgs<C> g) {
  const long stride = (long)gridDim.x * blockDim.x;
  const long tid0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (blockIdx.y == 0) {
    const long total = (long)g.NK * g.M * g.nc;
    for (long idx = tid0; idx < total; idx += stride) {
      const int c = (int)(idx % g.nc);
      const long bm = idx / g.nc;                   // b*M + m
      const int b = (int)(bm / g.M);
      const long m = bm - (long)b * g.M;
      if (g.sp[b / g.ki].skip) continue;  // k_1 of the instance is already there
      const double h = g.sp[b / g.ki].h;
      const int o = g.cidx[c];  // -1: padding column of a lead group
      const double2 v = o < 0 ? make_double2(0.0, 0.0) : to_d2(g.Ap[b * g.sAp + m * g.ldAp + o]);
      g.Aomg[bm * 2 * g.nc + c] = from_d2<C>(make_double2(h * v.x, h * v.y));
      g.Bomg[bm * 2 * g.nc + g.nc + c] = from_d2<C>(make_double2(h * v.x, -h * v.y));
    }
  } else if (blockIdx.y == 3) {
    const long total = (long)g.NK * g.M;
    for (long idx = tid0; idx < total; idx += stride) {
      const StageParams& q = g.sp[(int)(idx / g.M) / g.ki];
      const double s = q.h / g.hbar;
      const double d = q.delta[g.lead[idx % g.M]];
      const double2 a = g.chip[idx], c = g.chim[idx];
      g.u[idx] = make_double2(s * (a.x + d), s * a.y);
      g.v[idx] = make_double2(s * (c.x + d), s * c.y);
    }
  } else if (blockIdx.y == 1) {
    const long total = (long)g.NK * g.nc * g.n;
    for (long idx = tid0; idx < total; idx += stride) {
      const int o = (int)(idx % g.n);
      const long bc = idx / g.n;
      const int b = (int)(bc / g.nc);
      const int c = (int)(bc - (long)b * g.nc);
      if (g.sp[b / g.ki].skip) continue;
      const int oc = g.cidx[c];
      double2 s = oc < 0 ? make_double2(0.0, 0.0) : g.sig[((long)b * g.n + o) * g.n + oc];
      s.x *= g.hbar; s.y *= g.hbar;
      g.Bp[((long)b * (g.n + g.nc) + g.n + c) * g.n + o] = from_d2<C>(s);
    }
  } else {
    const long total = (long)g.NK * g.n;
    for (long idx = tid0; idx < total; idx += stride) {
      const int b = (int)(idx / g.n);
      const int i = (int)(idx - (long)b * g.n);
      double2 d = make_double2(0.0, 0.0);
      if (g.hubbard) {
        const int i0 = b - b % g.ki;          // the instance's first k block
        const int pb = i0 + g.ki - 1 - (b - i0);  // partner spin block
        const double2 s = g.sig[((long)pb * g.n + i) * g.n + i];
        d = make_double2(g.U * (s.x - g.n0[(long)pb * g.n + i]), g.U * s.y);
      }
      g.dvec[idx] = d;
    }
  }
}

// Bp[b][o'][o] = H[b][o][o'] (top n rows of the Psi GEMM's B operand); when the
// Hamiltonian is real also its real copies Hre = Re H and HreT = Re Hᵀ.
template <typename C, typename R>
__global__ void k_transpose_h(int n, long ldBp_batch, const double2* H, C* Bp, double* Hre,
                              R* HreT) {
  __shared__ double2 tile[32][33];
  const int b = blockIdx.z;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    if (i < n && j < n) {
      const double2 v = H[((long)b * n + i) * n + j];
      tile[r][tx] = v;
      if (Hre) Hre[((long)b * n + i) * n + j] = v.x;
    }
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = bx + r, i = by + tx;  // out[j][i] = H[i][j]
    if (i < n && j < n) {
      Bp[b * ldBp_batch + (long)j * n + i] = from_d2<C>(tile[tx][r]);
      if (HreT) HreT[((long)b * n + j) * n + i] = (R)tile[tx][r].x;
    }
  }
}

// dsig = (h/hbar) (G + G^H)   (Zandpack/mpi/zand:340-342, 384-387 with G = -iHσ + Π/ħ)
// With sig_eq, also the dissipator term -h η(ts)(σ - σ_eq) of the stage input σ.
__global__ void k_sigma_sym(int n, const double2* G, double2* out, const StageParams* sps,
                            int ki, double inv_hbar, const double2* sig_in,
                            const double2* sig_eq) {
  __shared__ double2 tile[32][33];
  const int b = blockIdx.z;
  if (sps[b / ki].skip) return;  // k_1 of the instance is already in its slot
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const double2* Gb = G + (long)b * n * n;
  // load the transposed block G[bx.., by..]
  for (int r = ty; r < 32; r += 8) {
    const int i = bx + r, j = by + tx;
    if (i < n && j < n) tile[r][tx] = Gb[(long)i * n + j];
  }
  __syncthreads();
  const StageParams* sp = sps + b / ki;
  const double f = sp->h * inv_hbar;
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    if (i < n && j < n) {
      const double2 g = Gb[(long)i * n + j];
      const double2 gt = tile[tx][r];  // G[j][i]
      const long off = ((long)b * n + i) * n + j;
      double2 v = make_double2(f * (g.x + gt.x), f * (g.y - gt.y));
      if (sig_eq) {
        const double e = sp->h * sp->eta;
        const double2 s = sig_in[off], q = sig_eq[off];
        v.x -= e * (s.x - q.x);
        v.y -= e * (s.y - q.y);
      }
      out[off] = v;
    }
  }
}

// J_a[b] = 2 Re Tr(Π_a) / ħ² from the gathered stage psi (Zandpack/mpi/zand:368-375):
// kCurChunks partial sums per (lead, k), then a fixed-order final sum.
constexpr int kCurChunks = 64;

// Π_α traces from the stage psi on the coupled orbitals (Ap[:, cidx]).
template <typename C>
__global__ void k_currents(int NK, int M, int nc, const int* lead, const C* Ap, long ldAp,
                           long sAp, const int* cidx, const double2* Ixic, double* part,
                           const StageParams* sp, int ki) {
  const int chunk = blockIdx.x, a = blockIdx.y, b = blockIdx.z;
  if (sp[b / ki].skip) return;  // the currents at y0 came with the reused k_1
  const long total = (long)M * nc;
  const long per = (total + kCurChunks - 1) / kCurChunks;
  const long lo = chunk * per, hi = min(total, lo + per);
  double acc = 0.0;
  for (long idx = lo + threadIdx.x; idx < hi; idx += blockDim.x) {
    const long m = idx / nc;
    if (lead[m] != a) continue;
    const int c = (int)(idx % nc);
    if (cidx[c] < 0) continue;
    const double2 p = to_d2(Ap[b * sAp + m * ldAp + cidx[c]]);
    const double2 x = Ixic[((long)b * M + m) * nc + c];
    acc += p.x * x.x - p.y * x.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[((long)a * NK + b) * kCurChunks + chunk] = red[0];
}

// J[a NK + b]; an instance whose k_1 is the previous attempt's k_7 takes the currents
// that attempt's last stage formed from the same state (Jprev)
__global__ void k_currents_final(int nlk, const double* part, double inv_hbar2, double* J,
                                 int NK, const StageParams* sp, int ki, const double* Jprev) {
  const int i = threadIdx.x;
  if (i >= nlk) return;
  if (sp[(i % NK) / ki].skip) {
    J[i] = Jprev[i];
    return;
  }
  double s = 0.0;
  for (int c = 0; c < kCurChunks; ++c) s += part[(long)i * kCurChunks + c];
  J[i] = 2.0 * s * inv_hbar2;
}

// Partial sums of |sum_j ct_j k_j|^2 over the visited set of one state part
// (Zandpack/mpi_funcs.py:247-252).
template <typename C>
struct TEArgs {
  const C* k[7];
  double ct[7];
  // scaled norm: the step's start and end states and the tolerances
  const C* y0 = nullptr;
  const C* y1 = nullptr;
  double rtol = 0.0, atol = 0.0;
  int scaled = 0;
  // first same as last: weight k1 (the k_1 slot) times the instance's r1 (sp: the first
  // stage's entries, one per instance), and the last weighted slot (k_7) written over
  // the k_1 slot kdst once read, as the next attempt's k_1
  const StageParams* sp = nullptr;
  int k1 = -1;
  C* kdst = nullptr;
};
// the error weights of instance blockIdx.y
template <typename C> __device__ inline void te_weights(const TEArgs<C>& g, double* ct) {
#pragma unroll
  for (int j = 0; j < 7; ++j) ct[j] = g.ct[j];
  if (g.sp && g.k1 >= 0) ct[g.k1] *= g.sp[blockIdx.y].r1;
}

constexpr int kTEBlocks = 1024;

// Instance blockIdx.y of a problem with several instances: its slots start istride
// elements further on.
template <typename C> __device__ inline void te_instance(TEArgs<C>& g, long istride) {
  const long ib = blockIdx.y * istride;
#pragma unroll
  for (int j = 0; j < 7; ++j)
    if (g.k[j]) g.k[j] += ib;
  if (g.kdst) g.kdst += ib;
  if (g.y0) g.y0 += ib;
  if (g.y1) g.y1 += ib;
}

// The slots with a non-zero error weight are packed in front (k[0..NW)).
template <typename C, typename Map, int NW>
__global__ __launch_bounds__(256) void k_te(TEArgs<C> g, Map mp, double* partial, long istride) {
  te_instance(g, istride);
  double ct[7];
  te_weights(g, ct);
  double acc = 0.0;
  const int nr = mp.nrows();
  for (int r = blockIdx.x; r < nr; r += gridDim.x) {
    long src, o1, o2;
    const int cols = mp.row(r, src, o1, o2);
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
      C kv[NW];
#pragma unroll
      for (int j = 0; j < NW; ++j) kv[j] = ld_stream(g.k[j] + src + c);
      if (g.kdst) st_stream(g.kdst + src + c, kv[NW - 1]);
      double vr = 0.0, vi = 0.0;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        vr += ct[j] * (double)kv[j].x;
        vi += ct[j] * (double)kv[j].y;
      }
      if (g.scaled) {
        const C a = g.y0[src + c], b = g.y1[src + c];
        acc += (vr * vr + vi * vi) *
               te_weight(g.atol, g.rtol, (double)a.x, (double)a.y, (double)b.x, (double)b.y);
      } else {
        acc += vr * vr + vi * vi;
      }
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// k_te for complex64 slots two elements per thread (16-byte loads), for maps whose rows
// have even offsets and lengths.
template <typename C, typename Map, int NW>
__global__ __launch_bounds__(256) void k_te_pairs(TEArgs<C> g, Map mp, double* partial,
                                                  long istride) {
  static_assert(sizeof(C) == 8, "complex64 pairs");
  te_instance(g, istride);
  double ct[7];
  te_weights(g, ct);
  double acc = 0.0;
  const int nr = mp.nrows();
  for (int r = blockIdx.x; r < nr; r += gridDim.x) {
    long src, o1, o2;
    const int cols = mp.row(r, src, o1, o2);
    for (int c = 2 * threadIdx.x; c < cols; c += 2 * blockDim.x) {
      v4f kv[NW];
#pragma unroll
      for (int j = 0; j < NW; ++j)
        kv[j] = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(g.k[j] + src + c));
      if (g.kdst)
        __builtin_nontemporal_store(kv[NW - 1], reinterpret_cast<v4f*>(g.kdst + src + c));
      double v0r = 0.0, v0i = 0.0, v1r = 0.0, v1i = 0.0;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        v0r += ct[j] * (double)kv[j].x;
        v0i += ct[j] * (double)kv[j].y;
        v1r += ct[j] * (double)kv[j].z;
        v1i += ct[j] * (double)kv[j].w;
      }
      if (g.scaled) {
        const C a0 = g.y0[src + c], b0 = g.y1[src + c];
        const C a1 = g.y0[src + c + 1], b1 = g.y1[src + c + 1];
        acc += (v0r * v0r + v0i * v0i) *
               te_weight(g.atol, g.rtol, (double)a0.x, (double)a0.y, (double)b0.x, (double)b0.y);
        acc += (v1r * v1r + v1i * v1i) *
               te_weight(g.atol, g.rtol, (double)a1.x, (double)a1.y, (double)b1.x, (double)b1.y);
      } else {
        acc += v0r * v0r + v0i * v0i;
        acc += v1r * v1r + v1i * v1i;
      }
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// TE = sqrt(sum), accept iff TE < eps, h_next = 0.9 h (eps/TE)^0.2
// (Zandpack/mpi/zand:396-407, Zandpack/mpi_RK4pars_dev.py:148-149).
constexpr int kCtrlThreads = 1024;
// One workgroup per instance g: its partials are nb of σ (partial[g nb ..]), nb of Ψ
// (partial[G nb + g nb ..]) and no of Ω (partial[2 G nb + g no ..]), summed as one
// sequence in that order (with one instance, partial[0 .. 2 nb + no) in place).
__global__ __launch_bounds__(kCtrlThreads) void k_controller(const double* partial, int nb,
                                                             int no, int G, Ctrl* ctrls,
                                                             const PulseDesc* pulses,
                                                             DissDesc diss, DriveDesc drive,
                                                             int NL, const StageParams* sp6) {
  __shared__ double red[kCtrlThreads];
  const int g = blockIdx.x;
  Ctrl* ctrl = ctrls + g;
  double acc = 0.0;
  const int np = 2 * nb + no;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    const long at = i < nb ? (long)g * nb + i
                  : i < 2 * nb ? (long)G * nb + (long)g * nb + (i - nb)
                               : 2L * G * nb + (long)g * no + (i - 2 * nb);
    acc += partial[at];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kCtrlThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double te2 = red[0];
    const double te = sqrt(te2 * ctrl->norm_scale);
    ctrl->te2 = te2;
    ctrl->te = te;
    ctrl->accept = (ctrl->force || te < ctrl->eps) ? 1 : 0;
    ctrl->h_next = 0.9 * ctrl->h * pow(ctrl->eps / te, 0.2);
    const double tn = ctrl->t0 + ctrl->h_next;
    const StageParams q = stage_params_at(pulses[g], diss, drive, NL, tn, ctrl->h_next);
    ctrl->fsal_hn = same_params(q, sp6[g], NL, drive.nt) ? 1 : 0;
  }
}

// ---------------------------------------------------------------- epilogues
// C = alpha * acc (complex alpha), batched, dense.
template <typename C>
struct EpiStore {
  C* out; long ldc; long sC;
  double ar, ai;
  __device__ inline void operator()(int b, int row, int col, double re, double im) const {
    out[b * sC + (long)row * ldc + col] =
        from_d2<C>(make_double2(ar * re - ai * im, ar * im + ai * re));
  }
};

// G = -i (H0 σ) - i d_row σ + Π_c[row, cpos[col]] / ħ, or + Q[row, col] / ħ with
// Q = Π_c S⁻¹[c, :] in the non-orthogonal form (cpos then all -1)
struct EpiSigma {
  double2* G; const double2* sig; const double2* dvec; const double2* Pic; const int* cpos;
  int n, nc; double inv_hbar;
  int pi_parts;  // split-K partials of Π_c, layout (NK, pi_parts, n, nc)
  const double2* Q = nullptr;  // (NK, n, n) or null
  const int* skip = nullptr;   // per instance of ki k blocks: leave it alone (k_hsig_ell)
  int ki = 1 << 30;
  __device__ inline void operator()(int b, int row, int col, double re, double im) const {
    apply(b, row, col, re, im, sig[((long)b * n + row) * n + col]);
  }
  // the same with the stage input's element s = σ[row, col] already loaded
  __device__ inline void apply(int b, int row, int col, double re, double im, double2 s) const {
    const long off = ((long)b * n + row) * n + col;
    const double2 d = dvec[(long)b * n + row];
    const double2 ds = cmul(d, s);
    double gr = im + ds.y, gi = -re - ds.x;  // -i (acc + d s)
    const int c = cpos[col];
    if (c >= 0) {
      double pr = 0.0, pi = 0.0;
      for (int s = 0; s < pi_parts; ++s) {
        const double2 p = Pic[(((long)b * pi_parts + s) * n + row) * nc + c];
        pr += p.x;
        pi += p.y;
      }
      gr += pr * inv_hbar;
      gi += pi * inv_hbar;
    }
    if (Q) {
      const double2 q = Q[off];
      gr += q.x * inv_hbar;
      gi += q.y * inv_hbar;
    }
    G[off] = make_double2(gr, gi);
  }
};

// out (NK, per) = Σ_s in (NK, parts, per): split-K partials summed in a fixed order.
__global__ void k_sum_parts(int NK, int parts, long per, const double2* in, double2* out) {
  const long total = (long)NK * per;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long b = i / per, r = i - b * per;
    double x = 0.0, y = 0.0;
    for (int s = 0; s < parts; ++s) {
      const double2 v = in[(b * parts + s) * per + r];
      x += v.x;
      y += v.y;
    }
    out[i] = make_double2(x, y);
  }
}

// Non-orthogonal Ψ term (Zandpack/mpi/nozand:1003, 1208-1210): the coupled-orbital
// vector ħ GL_P ξ_c + Ω_s ξ_c / ħ that S⁻¹[:, c] carries to every orbital.
// glp (NK, M); xic, W, Y (NK, M, nc).
template <typename C>
__global__ void k_overlap_y(long total, int nc, const double2* glp, const C* xic, const C* W,
                            C* Y, double hbar) {
  const double ih = 1.0 / hbar;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const double2 gx = cmul(glp[i / nc], to_d2(xic[i]));
    const double2 w = to_d2(W[i]);
    Y[i] = from_d2<C>(make_double2(hbar * gx.x + w.x * ih, hbar * gx.y + w.y * ih));
  }
}

// kΨ = -i h [ (acc - (χ⁺+Δ)ψ + d_col ψ)/ħ + (GL_P ξ + W/ħ²) on coupled columns ]
// with acc = ψ H0ᵀ + ħ (ΔP ξ)_c σ_cᵀ  (Zandpack/mpi/zand:746-778).  With acc0 the
// GEMM computed only part of acc; the rest is read from acc0 for rows below
// acc0_rows (same layout as out, the element is read and then overwritten by the
// same thread; rows from acc0_rows on have ΔP = 0 and no such part).
template <typename C>
struct EpiPsi {
  C* out;                       // (NK, M, n)
  const C* acc0;
  const C* Ap; long ldAp;       // stage psi
  const C* xic; const C* W;     // (NK, M, nc)
  const double2* glp; const double2* chip;  // (NK, M)
  const double2* dvec;          // (NK, n)
  const int* lead; const int* cpos;
  const StageParams* sp;        // per instance of ki k blocks
  int M, n, nc; double inv_hbar, inv_hbar2;
  int w_parts;  // split-K partials of W, layout (NK, w_parts, M, nc)
  int acc0_rows;
  int ki = 1 << 30;
  __device__ inline void operator()(int b, int row, int col, double re, double im) const {
    const long bm = (long)b * M + row;
    if (acc0 && row < acc0_rows) {
      const double2 a0 = to_d2(acc0[bm * n + col]);
      re += a0.x;
      im += a0.y;
    }
    const double2 p = to_d2(Ap[bm * ldAp + col]);
    const StageParams& q = sp[b / ki];
    double2 chi = chip[bm];
    chi.x += q.delta[lead[row]];
    const double2 d = dvec[(long)b * n + col];
    const double2 cp = cmul(chi, p), dp = cmul(d, p);
    double vr = (re - cp.x + dp.x) * inv_hbar;
    double vi = (im - cp.y + dp.y) * inv_hbar;
    const int c = cpos[col];
    if (c >= 0) {
      const double2 gx = cmul(glp[bm], to_d2(xic[bm * nc + c]));
      double wr = 0.0, wi = 0.0;
      for (int s = 0; s < w_parts; ++s) {
        const double2 w = to_d2(W[(((long)b * w_parts + s) * M + row) * nc + c]);
        wr += w.x;
        wi += w.y;
      }
      vr += gx.x + wr * inv_hbar2;
      vi += gx.y + wi * inv_hbar2;
    }
    const double h = q.h;
    out[bm * n + col] = from_d2<C>(make_double2(h * vi, -h * vr));
  }
};

// The Ω right-hand side of one stage with the stage input formed in place:
//   Ω_s = y0 + Σ_j a_j k_j                     (Zandpack/mpi/zand:306-313, 722-735)
//   k_s = h acc + i (u_row - v_col) Ω_s        (Zandpack/mpi/zand:784-809)
// so no stage-input copy of Ω is written or read back.  out (k_s) may be null when
// nothing reads it; ynew receives Ω_s (the last stage: the new state).  With TE the
// epilogue also accumulates |Σ_j ct_j k_j + ct_s k_s|² (Zandpack/mpi_funcs.py:247-252)
// over the elements it visits, reduced per workgroup into te_part.
template <typename C, bool TE>
struct EpiOmegaRK {
  static constexpr bool kReduce = TE;
  static constexpr bool kBlock4 = true;
  static constexpr int kMaxTerms = 6;
  C* out;                                 // (NK, M, M) k_s, or null
  const C* y0;                            // (NK, M, M)
  const C* k[kMaxTerms]; double a[kMaxTerms]; int nterms;
  C* ynew;                                // Ω_s, or null
  const double2* u; const double2* v;     // (NK, M), (h/ħ)(χ± + Δ)
  int M;
  double ct[kMaxTerms]; double ct_self;   // error weights of k[] and of k_s (TE)
  double* te_part;                        // one partial per workgroup (TE)
  double rtol, atol; int scaled;          // scaled norm (TE): weights from |y0|, |Ω_7|
  int tri = 0;  // only the block upper triangle is computed: entries outside the diagonal
                // 64 x 64 blocks stand for two in the error norm
  // first same as last: term k1 (the k_1 slot) weighted by the instance's r1 (sp: one
  // entry per instance of ki k blocks)
  const StageParams* sp = nullptr;
  int kper = 1 << 30, k1 = -1;

  __device__ inline double tri_weight(int row, int col) const {
    return (tri && (row >> 6) != (col >> 6)) ? 2.0 : 1.0;
  }

  // Four outputs of one column (rows[r] < 0: outside): every slot load first, then
  // the stores.
  __device__ inline void block(int b, const int* rows, int col, const double* re,
                               const double* im, double& acc) const {
    long off[4];
    double sr[4], si[4], er[4], ei[4], yr0[4], yi0[4];
    const double r1 = (sp && k1 >= 0) ? sp[b / kper].r1 : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      off[r] = ((long)b * M + (rows[r] < 0 ? 0 : rows[r])) * M + col;
      const double2 y = rows[r] < 0 ? make_double2(0.0, 0.0) : ld_nt(y0 + off[r]);
      sr[r] = y.x; si[r] = y.y; er[r] = 0.0; ei[r] = 0.0;
      if constexpr (TE) { yr0[r] = y.x; yi0[r] = y.y; }
    }
#pragma unroll
    for (int j = 0; j < kMaxTerms; ++j)
      if (j < nterms) {
        const double aj = j == k1 ? a[j] * r1 : a[j];
        const double cj = j == k1 ? ct[j] * r1 : ct[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double2 x = rows[r] < 0 ? make_double2(0.0, 0.0) : ld_nt(k[j] + off[r]);
          sr[r] += aj * x.x;
          si[r] += aj * x.y;
          if constexpr (TE) {
            er[r] += cj * x.x;
            ei[r] += cj * x.y;
          }
        }
      }
    // col < M whenever any row of the block is inside
    const bool any = rows[0] >= 0 || rows[1] >= 0 || rows[2] >= 0 || rows[3] >= 0;
    const double2 vv = v[any ? (long)b * M + col : 0];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (rows[r] < 0) continue;
      if (ynew) st_nt(ynew + off[r], make_double2(sr[r], si[r]));
      const double2 uu = u[(long)b * M + rows[r]];
      const double dr = uu.x - vv.x, di = uu.y - vv.y;
      // k_s = h acc + i (dr + i di) Ω_s
      const double2 t = cmul(make_double2(-di, dr), make_double2(sr[r], si[r]));
      const C ks = from_d2<C>(make_double2(re[r] + t.x, im[r] + t.y));
      if (out) st_nt(out + off[r], to_d2(ks));
      if constexpr (TE) {
        const double xr = er[r] + ct_self * (double)ks.x, xi = ei[r] + ct_self * (double)ks.y;
        acc += tri_weight(rows[r], col) *
               (scaled ? (xr * xr + xi * xi) * te_weight(atol, rtol, yr0[r], yi0[r], sr[r], si[r])
                       : xr * xr + xi * xi);
      }
    }
  }
  // Slot values are streamed once: non-temporal loads and stores keep them from
  // evicting the GEMM operands from L2.
  __device__ static inline double2 ld_nt(const C* p) {
    typedef typename std::conditional<sizeof(C) == 16, double, float>::type E;
    typedef E v2 __attribute__((ext_vector_type(2)));
    const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return make_double2((double)x.x, (double)x.y);
  }
  __device__ static inline void st_nt(C* p, double2 v) {
    typedef typename std::conditional<sizeof(C) == 16, double, float>::type E;
    typedef E v2 __attribute__((ext_vector_type(2)));
    const v2 x = {(E)v.x, (E)v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<v2*>(p));
  }
  // Four consecutive columns of one row (FP32 storage): 32-byte loads and stores
  // when all four are inside the matrix; every slot load first, then the stores.
  // qr, qi (optional) receive the four values as stored (k_s, or Ω_s when ynew is set;
  // 0 outside the matrix) for a fused product with them.
  static constexpr bool kRow4 = true;
  __device__ inline void row4(int b, int row, int col0, int ncol, const double* re,
                              const double* im, double& acc, double* qr = nullptr,
                              double* qi = nullptr) const {
    if (ncol <= 0) {
      if (qr) {
#pragma unroll
        for (int r = 0; r < 4; ++r) { qr[r] = 0.0; qi[r] = 0.0; }
      }
      return;
    }
    const long off = ((long)b * M + row) * M + col0;
    const bool vec = ncol == 4 && sizeof(C) == 8 && (off & 3) == 0;
    const double r1 = (sp && k1 >= 0) ? sp[b / kper].r1 : 1.0;
    double sr[4], si[4], er[4], ei[4], yr0[4], yi0[4];
    load4(y0, off, ncol, vec, sr, si);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      er[r] = 0.0; ei[r] = 0.0;
      if constexpr (TE) { yr0[r] = sr[r]; yi0[r] = si[r]; }
    }
#pragma unroll
    for (int j = 0; j < kMaxTerms; ++j)
      if (j < nterms) {
        double xr[4], xi[4];
        load4(k[j], off, ncol, vec, xr, xi);
        const double aj = j == k1 ? a[j] * r1 : a[j];
        const double cj = j == k1 ? ct[j] * r1 : ct[j];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          sr[r] += aj * xr[r];
          si[r] += aj * xi[r];
          if constexpr (TE) {
            er[r] += cj * xr[r];
            ei[r] += cj * xi[r];
          }
        }
      }
    const double2 uu = u[(long)b * M + row];
    const double tw = tri_weight(row, col0);  // col0 .. col0 + 3 lie in one 64-block
    double kr[4], ki[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double2 vv = r < ncol ? v[(long)b * M + col0 + r] : make_double2(0.0, 0.0);
      const double dr = uu.x - vv.x, di = uu.y - vv.y;
      // k_s = h acc + i (dr + i di) Ω_s, rounded to the storage type
      const double2 t = cmul(make_double2(-di, dr), make_double2(sr[r], si[r]));
      const C ks = from_d2<C>(make_double2(re[r] + t.x, im[r] + t.y));
      kr[r] = (double)ks.x;
      ki[r] = (double)ks.y;
      if constexpr (TE) {
        if (r < ncol) {
          const double xr = er[r] + ct_self * kr[r], xi = ei[r] + ct_self * ki[r];
          acc += tw * (scaled ? (xr * xr + xi * xi) *
                                    te_weight(atol, rtol, yr0[r], yi0[r], sr[r], si[r])
                              : xr * xr + xi * xi);
        }
      }
    }
    if (ynew) store4(ynew, off, ncol, vec, sr, si);
    if (out) store4(out, off, ncol, vec, kr, ki);
    if (qr) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool in = r < ncol;
        if (ynew) {
          const C v = from_d2<C>(make_double2(sr[r], si[r]));
          qr[r] = in ? (double)v.x : 0.0;
          qi[r] = in ? (double)v.y : 0.0;
        } else {
          qr[r] = in ? kr[r] : 0.0;
          qi[r] = in ? ki[r] : 0.0;
        }
      }
    }
  }
  __device__ static inline void load4(const C* p, long off, int ncol, bool vec, double* xr,
                                      double* xi) {
    if constexpr (sizeof(C) == 8) {
      if (vec) {
        const v4f* q = reinterpret_cast<const v4f*>(p + off);
        const v4f lo = __builtin_nontemporal_load(q), hi = __builtin_nontemporal_load(q + 1);
        xr[0] = lo.x; xi[0] = lo.y; xr[1] = lo.z; xi[1] = lo.w;
        xr[2] = hi.x; xi[2] = hi.y; xr[3] = hi.z; xi[3] = hi.w;
        return;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double2 x = r < ncol ? to_d2(p[off + r]) : make_double2(0.0, 0.0);
      xr[r] = x.x;
      xi[r] = x.y;
    }
  }
  __device__ static inline void store4(C* p, long off, int ncol, bool vec, const double* xr,
                                       const double* xi) {
    if constexpr (sizeof(C) == 8) {
      if (vec) {
        v4f* q = reinterpret_cast<v4f*>(p + off);
        const v4f lo = {(float)xr[0], (float)xi[0], (float)xr[1], (float)xi[1]};
        const v4f hi = {(float)xr[2], (float)xi[2], (float)xr[3], (float)xi[3]};
        __builtin_nontemporal_store(lo, q);
        __builtin_nontemporal_store(hi, q + 1);
        return;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (r < ncol) p[off + r] = from_d2<C>(make_double2(xr[r], xi[r]));
  }
  __device__ inline void finish(int block, double sum) const { te_part[block] = sum; }
};

// Ω_s ξ_c by linearity: the split-K partials of this stage's product X ξ_c (X = y0
// at the first stage, else the previous stage's k) are summed into pred[s], then
//   W = pred[0] + Σ_{j<s} a_j pred[j+1]
// (pred[0] = y0 ξ_c, pred[j] = k_j ξ_c).
// With the fused Ω stage kernel the partials are per column chunk, and rows from ML on
// (passive, with the passive x passive block skipped) have only parts_pass of them.
template <typename C>
struct WSumArgs {
  const C* part;            // (NK, parts, M, nc)
  C* pred[7];               // (NK, M, nc) each
  C* W;                     // (NK, M, nc)
  double a[6];
  int s, parts, NK, M, nc;
  int parts_pass = -1, ML = 0;
  // first same as last: a[0] (the k_1 ξ_c term) times the instance's r1
  const StageParams* sp = nullptr;
  int ki = 1 << 30;
};

template <typename C>
__global__ __launch_bounds__(256) void k_wsum(WSumArgs<C> g) {
  const long per = (long)g.M * g.nc, total = (long)g.NK * per;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / per, rem = idx - b * per;
    const int np = (g.parts_pass >= 0 && rem / g.nc >= g.ML) ? g.parts_pass : g.parts;
    double pr = 0.0, pi = 0.0;
    for (int p = 0; p < np; ++p) {
      const double2 x = to_d2(g.part[(b * g.parts + p) * per + rem]);
      pr += x.x;
      pi += x.y;
    }
    const C ps = from_d2<C>(make_double2(pr, pi));
    g.pred[g.s][idx] = ps;
    double wr, wi;
    if (g.s == 0) {
      wr = (double)ps.x; wi = (double)ps.y;
    } else {
      const double2 y = to_d2(g.pred[0][idx]);
      wr = y.x; wi = y.y;
      const double r1 = g.sp ? g.sp[b / g.ki].r1 : 1.0;
      for (int j = 0; j < g.s; ++j) {
        const double2 x = (j + 1 == g.s) ? to_d2(ps) : to_d2(g.pred[j + 1][idx]);
        const double aj = j == 0 ? g.a[0] * r1 : g.a[j];
        wr += aj * x.x;
        wi += aj * x.y;
      }
    }
    g.W[idx] = from_d2<C>(make_double2(wr, wi));
  }
}

// out = Σ_p part[p] (fixed order; rows from ML on sum parts_pass partials): the y_new ξ_c
// of the last fused Ω stage, the next attempt's y0 ξ_c.
template <typename C>
__global__ __launch_bounds__(256) void k_sum_wpart(const C* part, int parts, int parts_pass,
                                                   int ML, int NK, int M, int nc, C* out) {
  const long per = (long)M * nc, total = (long)NK * per;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / per, rem = idx - b * per;
    const int np = rem / nc >= ML ? parts_pass : parts;
    double pr = 0.0, pi = 0.0;
    for (int p = 0; p < np; ++p) {
      const double2 x = to_d2(part[(b * parts + p) * per + rem]);
      pr += x.x;
      pi += x.y;
    }
    out[idx] = from_d2<C>(make_double2(pr, pi));
  }
}

// Fill the strictly lower block triangle (64 x 64 blocks) of an anti-Hermitian X held
// as its block upper triangle: X_ij = -conj(X_ji) for block(i) > block(j).  Reads only
// the upper part, writes only the lower.
template <typename C>
__global__ __launch_bounds__(256) void k_mirror_lower(int NK, int M, C* X) {
  const long mm = (long)M * M, total = (long)NK * mm;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long b = idx / mm, rc = idx - b * mm;
    const int i = (int)(rc / M), j = (int)(rc - (long)i * M);
    if ((i >> 6) > (j >> 6)) {
      const C v = X[b * mm + (long)j * M + i];
      C w;
      w.x = -v.x;
      w.y = v.y;
      X[idx] = w;
    }
  }
}

// ------------------------------------------------------------ K segments
// Device mode order: active modes (ΔP or ΔM non-zero) in [0, ML), ML a multiple
// of the Ω tiles, passive (Fermi) modes after.  For the Ω sources
//   term1 (K = [0, nc))   vanishes on passive columns (ΔM_j = 0) and needs only
//                          the coupled orbitals of the column's lead(s);
//   term2 (K = [nc, 2nc)) vanishes on passive rows (ΔP_i = 0);
// the passive x passive block of Ω stays exactly zero when it starts at zero
// (ff = 1) and is skipped.
// Hull of the non-empty ranges of 64-wide tiles a..b of a per-tile range table.
__device__ inline int2 tile_hull(const int2* t, int a, int b) {
  int lo = 1 << 30, hi = 0;
  for (int i = a; i <= b; ++i) {
    const int2 q = t[i];
    if (q.y > q.x) { lo = min(lo, q.x); hi = max(hi, q.y); }
  }
  return hi > lo ? make_int2(lo, hi) : make_int2(0, 0);
}

struct KSegOmega {
  const int2* t1;  // per 64 columns: term1 K range [x, y)
  const int2* t2;  // per 64 rows: term2 K range [nc + x, nc + y) (the row leads' orbitals)
  int nc, ML, ff, nt;  // nt: 64-wide tiles in the table
  // tri: Ω is anti-Hermitian and only its block upper triangle (64 x 64 blocks, the
  // diagonal blocks in full) is computed; tiles (at most 64 x 64, aligned) below it
  // are skipped
  int tri = 0;
  __device__ inline KRange get(int m0, int n0, int bm, int bn, int) const {
    const int2 r = tile_hull(t1, n0 / 64, min(nt - 1, (n0 + bn - 1) / 64));
    const int2 q = tile_hull(t2, m0 / 64, min(nt - 1, (m0 + bm - 1) / 64));
    const bool s1 = r.y > r.x, s2 = q.y > q.x;
    // term1 first when present, then term2; term2 alone otherwise
    return KRange{s1 ? r.x : nc + q.x, s1 ? r.y : nc + q.y, nc + q.x, nc + q.y,
                  (int)s1 + (int)s2};
  }
  __device__ inline bool skip(int m0, int n0, int, int) const {
    return (ff && m0 >= ML && n0 >= ML) || (tri && (m0 >> 6) > (n0 >> 6));
  }
};

// Products X ξ_c of an anti-Hermitian X held as its block upper triangle (64 x 64
// blocks): row block I of X ξ_c is -X[J < I, I]^H ξ_c[J < I] + X[I, J >= I] ξ_c[J >= I].
// One K loop per output tile runs both: the first segment (K < m0) reads A mirrored,
// -conj(X[k][m]) from the column strip above the row block, then up to two segments (K
// >= m0) read the row strip.  The K pieces of the inner selection are cut at m0; the
// pieces below m0 (at most two, ordered) become their hull, where the gap's modes do not
// couple the tile's columns and add zero.  Output tiles are 64 rows, aligned, so m0 is
// the first column of the row's diagonal block.
template <typename KS>
struct KSegTri {
  static constexpr bool kMirrorA = true;
  KS in;
  __device__ inline KRange get(int m0, int n0, int bm, int bn, int K) const {
    const KRange r = in.get(m0, n0, bm, bn, K);
    const bool p0 = r.nseg > 0 && r.hi0 > r.lo0, p1 = r.nseg > 1 && r.hi1 > r.lo1;
    // below m0
    const bool b0 = p0 && r.lo0 < m0, b1 = p1 && r.lo1 < m0;
    const int mlo = b0 ? r.lo0 : (b1 ? r.lo1 : 0);
    const int mhi = b1 ? min(r.hi1, m0) : (b0 ? min(r.hi0, m0) : 0);
    // from m0 on
    const int a0 = max(r.lo0, m0), a1 = max(r.lo1, m0);
    const bool s0 = p0 && r.hi0 > a0, s1 = p1 && r.hi1 > a1;
    KRange o{mlo, mhi, s0 ? a0 : 0, s0 ? r.hi0 : 0, 3};
    o.lo2 = s1 ? a1 : 0;
    o.hi2 = s1 ? r.hi1 : 0;
    return o;
  }
  __device__ inline bool skip(int m0, int n0, int bm, int bn) const {
    return in.skip(m0, n0, bm, bn);
  }
};

// Products over the modes against a coupled-orbital operand (X ξ_c, ψᵀ Ixi_c): the
// operand vanishes unless the mode's lead couples the column, so a column tile
// whose 16-column groups all belong to one lead runs over that lead's modes only
// (its rows in the active block, then in the passive block).  With ff the passive
// rows of X (Ω-shaped) see only the active block.
struct KSegLead {
  const int* gmask;    // per 16-column group: mask of the coupling leads
  const int4* lrange;  // per lead: active rows [x, y), passive rows [z, w)
  int ngroup, ML, ff;
  __device__ inline KRange get(int m0, int n0, int, int bn, int K) const {
    const int g0 = n0 / 16, g1 = min(ngroup, (n0 + bn + 15) / 16);
    int mask = 0;
    for (int g = g0; g < g1; ++g) mask |= gmask[g];
    const bool act_only = ff && m0 >= ML;
    if (mask == 0) return KRange{0, 0, 0, 0, 0};
    if (mask & (mask - 1)) return KRange{0, act_only ? ML : K, 0, 0, 1};
    const int4 r = lrange[__ffs(mask) - 1];
    const bool sa = r.y > r.x, sp = !act_only && r.w > r.z;
    return KRange{sa ? r.x : r.z, sa ? r.y : r.w, r.z, r.w, (int)sa + (int)sp};
  }
  __device__ inline bool skip(int, int, int, int) const { return false; }
};

// The same two K ranges for every tile (a lead's active and passive device modes).
struct KSegFixed {
  int lo0, hi0, lo1, hi1;
  __device__ inline KRange get(int, int, int, int, int) const {
    const bool a = hi0 > lo0, b = hi1 > lo1;
    return KRange{a ? lo0 : lo1, a ? hi0 : hi1, lo1, hi1, (int)a + (int)b};
  }
  __device__ inline bool skip(int, int, int, int) const { return false; }
};

// [ΔP ξ_c][ħ σ_cᵀ] (K = coupled orbitals): a row tile's K range is its leads'
// orbitals; tiles past ML (passive rows, ΔP = 0) are skipped when skip_passive
// (their consumer then reads no partial product for those rows).
struct KSegRowLead {
  const int2* t2;  // per 64 rows: K range
  int ML, skip_passive, nt;
  __device__ inline KRange get(int m0, int, int bm, int, int) const {
    const int2 q = tile_hull(t2, m0 / 64, min(nt - 1, (m0 + bm - 1) / 64));
    return q.y > q.x ? KRange{q.x, q.y, 0, 0, 1} : KRange{0, 0, 0, 0, 0};
  }
  __device__ inline bool skip(int m0, int, int, int) const {
    return skip_passive && m0 >= ML;
  }
};

// ------------------------------------------------------------ sparse H
// Row-compressed Hamiltonian with a fixed count R of entries per row:
// col[(b R + r) n + o], val[(b R + r) n + o] for row o of H[b]; rows with fewer
// non-zeros are padded with (col = o, val = 0).  Tight-binding and most DFT
// Hamiltonians hold a few entries per row, so the two H products of the right-hand
// side (Hσ and ψHᵀ, Zandpack/mpi/zand:340-343, 746-778, dense there) become
// gathers bounded by HBM instead of n³-class matrix products.
struct Ell {
  const int* col;
  const double2* val;
  int R, n;
};

// The whole Ψ right-hand side (EpiPsi's formula) with the row-compressed H, for RB
// mode rows per workgroup held in LDS: per-row constants are read once per
// workgroup, and for its two columns a thread loads every operand (the partial
// product acc0, which shares the output buffer, ξ_c and W on coupled columns)
// before its first store, so the loads of a column pair are in flight together.
template <typename C, int RB>
__global__ __launch_bounds__(256) void k_psi_ell(EpiPsi<C> e, long sAp, Ell h) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C* rows = reinterpret_cast<C*>(lds_raw);  // RB x n
  __shared__ double2 rchi[RB], rglp[RB];
  const int b = blockIdx.y;
  if (e.sp[b / e.ki].skip) return;  // k_1 of the instance is already in its slot
  const int i0 = blockIdx.x * RB;
  const int n = h.n, M = e.M;
  for (int q = 0; q < RB; ++q) {
    const int i = i0 + q;
    for (int o = threadIdx.x; o < n; o += blockDim.x)
      rows[q * n + o] =
          (i < M) ? ld_stream(e.Ap + b * sAp + (long)i * e.ldAp + o) : from_d2<C>(make_double2(0, 0));
  }
  if (threadIdx.x < RB) {
    const int i = i0 + threadIdx.x;
    double2 chi = make_double2(0.0, 0.0), g = make_double2(0.0, 0.0);
    if (i < M) {
      const long bm = (long)b * M + i;
      chi = e.chip[bm];
      chi.x += e.sp[b / e.ki].delta[e.lead[i]];
      g = e.glp[bm];
    }
    rchi[threadIdx.x] = chi;
    rglp[threadIdx.x] = g;
  }
  __syncthreads();
  const double hs = e.sp[b / e.ki].h;
  const int* col = h.col + (long)b * h.R * n;
  const double2* val = h.val + (long)b * h.R * n;
  constexpr int CPT = 2;  // columns per thread per pass
  for (int ob = threadIdx.x; ob < n; ob += CPT * blockDim.x) {
    int oc[CPT], cp[CPT];
    double2 dv[CPT];
    C a0[CPT][RB], xv[CPT][RB], wv[CPT][RB];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int o = ob + u * blockDim.x;
      oc[u] = o;
      const bool ok = o < n;
      dv[u] = ok ? e.dvec[(long)b * n + o] : make_double2(0.0, 0.0);
      cp[u] = ok ? e.cpos[o] : -1;
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int i = i0 + q;
        const long bm = (long)b * M + i;
        const bool rok = ok && i < M;
        a0[u][q] = (rok && e.acc0 && i < e.acc0_rows) ? ld_stream(e.acc0 + bm * n + o)
                                                      : from_d2<C>(make_double2(0, 0));
        const bool cc = rok && cp[u] >= 0;
        xv[u][q] = cc ? e.xic[bm * e.nc + cp[u]] : from_d2<C>(make_double2(0, 0));
        double wr = 0.0, wi = 0.0;
        if (cc)
          for (int s = 0; s < e.w_parts; ++s) {
            const double2 w = to_d2(e.W[(((long)b * e.w_parts + s) * M + i) * e.nc + cp[u]]);
            wr += w.x;
            wi += w.y;
          }
        wv[u][q] = from_d2<C>(make_double2(wr, wi));
      }
    }
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int o = oc[u];
      if (o >= n) continue;
      double ar[RB], ai[RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) { ar[q] = 0.0; ai[q] = 0.0; }
      for (int r = 0; r < h.R; ++r) {
        const int c = col[(long)r * n + o];
        const double2 v = val[(long)r * n + o];
#pragma unroll
        for (int q = 0; q < RB; ++q) {
          const double2 p = to_d2(rows[q * n + c]);
          ar[q] += p.x * v.x - p.y * v.y;
          ai[q] += p.x * v.y + p.y * v.x;
        }
      }
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int i = i0 + q;
        if (i >= M) continue;
        const double2 a = to_d2(a0[u][q]);
        const double re = ar[q] + a.x, im = ai[q] + a.y;
        const double2 p = to_d2(rows[q * n + o]);
        const double2 cpx = cmul(rchi[q], p), dp = cmul(dv[u], p);
        double vr = (re - cpx.x + dp.x) * e.inv_hbar;
        double vi = (im - cpx.y + dp.y) * e.inv_hbar;
        if (cp[u] >= 0) {
          const double2 gx = cmul(rglp[q], to_d2(xv[u][q]));
          const double2 w = to_d2(wv[u][q]);
          vr += gx.x + w.x * e.inv_hbar2;
          vi += gx.y + w.y * e.inv_hbar2;
        }
        st_stream(e.out + ((long)b * M + i) * n + o, from_d2<C>(make_double2(hs * vi, -hs * vr)));
      }
    }
  }
}

// acc[o, j] = (H σ)[o, j] = Σ_r H_r(o) σ[col_r(o), j] for RB rows o per workgroup
// and 256 consecutive columns j; acc goes to epi (the σ right-hand side).  With the
// entries per row R known at compile time, a thread issues every gather of its RB
// rows and their epilogue's σ elements before the first product (RB x (R + 1) loads
// in flight); R = 0 takes e.R at run time, one row after another.
template <int RB, int R, typename Epi>
__global__ __launch_bounds__(256) void k_hsig_ell(const double2* __restrict__ sig, Ell e,
                                                  Epi epi) {
  const int b = blockIdx.z;
  if (epi.skip && epi.skip[b / epi.ki]) return;
  const int n = e.n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int o0 = blockIdx.y * RB;
  if (j >= n) return;
  const int* __restrict__ col = e.col + (long)b * e.R * n;
  const double2* __restrict__ val = e.val + (long)b * e.R * n;
  const double2* __restrict__ s = sig + (long)b * n * n + j;
  if constexpr (R > 0) {
    double2 x[RB][R], sv[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      const int o = min(o0 + q, n - 1);  // rows past n: loads kept in bounds, no output
#pragma unroll
      for (int r = 0; r < R; ++r) x[q][r] = s[(long)col[(long)r * n + o] * n];
      sv[q] = s[(long)o * n];
    }
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      const int o = o0 + q;
      if (o >= n) break;
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 v = val[(long)r * n + o];
        ar += v.x * x[q][r].x - v.y * x[q][r].y;
        ai += v.x * x[q][r].y + v.y * x[q][r].x;
      }
      epi.apply(b, o, j, ar, ai, sv[q]);
    }
  } else {
    for (int q = 0; q < RB; ++q) {
      const int o = o0 + q;
      if (o >= n) break;
      double ar = 0.0, ai = 0.0;
      for (int r = 0; r < e.R; ++r) {
        const int c = col[(long)r * n + o];
        const double2 v = val[(long)r * n + o];
        const double2 x = s[(long)c * n];
        ar += v.x * x.x - v.y * x.y;
        ai += v.x * x.y + v.y * x.x;
      }
      epi(b, o, j, ar, ai);
    }
  }
}

// ------------------------------------------------------------ steady state
// (cmdtools/psinought): Ψ₀ solves the real-linear system A(ψ) = b with
//   A(ψ)_i = (χ⁺_i + Δ_α(i) - H) ψ_i - (Ω₀(ψ) ξ)_i / ħ,   b_i = ħ Λ^{<+}_i ξ_i + ħ ΔP_i σ₀ ξ_i,
// Ω₀(ψ)_ij = -(ΔM_j Ixi_j·ψ_i + ΔP_i ψ_j*·ξ_i) iħ / (χ⁻_j - χ⁺_i)  (psinought:199-239, 447-449).

// Ω₀ from the Ω sources acc (gathered with h = 1, so u - v = (χ⁺_i - χ⁻_j)/ħ):
// Ω₀ = i acc / (u_i - v_j); inert padding modes (u = v = 0, acc = 0) give 0.
template <typename C>
struct EpiOmegaSteady {
  C* out;
  const double2* u; const double2* v;
  int M;
  __device__ inline void operator()(int b, int row, int col, double re, double im) const {
    const double2 a = u[(long)b * M + row], c = v[(long)b * M + col];
    const double dr = a.x - c.x, di = a.y - c.y;
    const double den = dr * dr + di * di;
    double2 o = make_double2(0.0, 0.0);
    if (den > 0.0) {
      // i (re + i im) / (dr + i di) = i (re + i im)(dr - i di) / den
      const double qr = (re * dr + im * di) / den, qi = (im * dr - re * di) / den;
      o = make_double2(-qi, qr);
    }
    out[((long)b * M + row) * M + col] = from_d2<C>(o);
  }
};

// Krylov vector (FP64, (NK, M, n)) -> the ψ block of Ap (storage type).
template <typename C>
__global__ void k_to_ap(long total, int n, const double2* x, C* Ap, long ldAp) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long row = idx / n;
    const int o = (int)(idx - row * n);
    Ap[row * ldAp + o] = from_d2<C>(x[idx]);
  }
}

// HPsi[i, o] = (ψ Hᵀ)[i, o] over the row-compressed H (ψ read from Ap).
template <typename C>
__global__ void k_ell_rows(int NK, int M, const C* Ap, long ldAp, Ell h, C* out) {
  const long total = (long)NK * M * h.n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long row = idx / h.n;  // b*M + i
    const int o = (int)(idx - row * h.n);
    const int b = (int)(row / M);
    const int* col = h.col + (long)b * h.R * h.n;
    const double2* val = h.val + (long)b * h.R * h.n;
    double ar = 0.0, ai = 0.0;
    for (int r = 0; r < h.R; ++r) {
      const double2 p = to_d2(Ap[row * ldAp + col[(long)r * h.n + o]]);
      const double2 v = val[(long)r * h.n + o];
      ar += p.x * v.x - p.y * v.y;
      ai += p.x * v.y + p.y * v.x;
    }
    out[idx] = from_d2<C>(make_double2(ar, ai));
  }
}

// y = (χ⁺_i + Δ_α(i) - d_o) x - HPsi - W[:, c(o)]/ħ   (A(ψ) with HPsi = ψ H0ᵀ, d the
// Hubbard diagonal, W = Ω₀ ξ_c).
template <typename C>
struct SteadyApplyArgs {
  int NK, M, n, nc;
  const double2* x; const C* hpsi; const C* W;
  const double2* chip; const double2* dvec;
  const int* lead; const int* cpos;
  const StageParams* sp;
  double inv_hbar;
  double2* y;
};
template <typename C>
__global__ void k_steady_apply(SteadyApplyArgs<C> g) {
  const long total = (long)g.NK * g.M * g.n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long bm = idx / g.n;
    const int o = (int)(idx - bm * g.n);
    const int b = (int)(bm / g.M);
    const int i = (int)(bm - (long)b * g.M);
    double2 chi = g.chip[bm];
    chi.x += g.sp->delta[g.lead[i]];
    const double2 d = g.dvec[(long)b * g.n + o];
    chi.x -= d.x;
    chi.y -= d.y;
    const double2 xv = g.x[idx];
    const double2 hp = to_d2(g.hpsi[idx]);
    double2 r = cmul(chi, xv);
    r.x -= hp.x;
    r.y -= hp.y;
    const int c = g.cpos[o];
    if (c >= 0) {
      const double2 w = to_d2(g.W[bm * g.nc + c]);
      r.x -= w.x * g.inv_hbar;
      r.y -= w.y * g.inv_hbar;
    }
    g.y[idx] = r;
  }
}

// b = ħ ΔP_i σ₀ ξ_i (from its GEMM, or 0 on rows it skipped) + ħ Λ^{<+}_i ξ_i.
template <typename C>
struct SteadyRhsArgs {
  int NK, M, n, nc, g0_rows;
  const C* g0;                 // (NK, M, n): ħ ΔP ξ_c σ_cᵀ
  const C* xic; const double2* glp;
  const int* cidx;
  double hbar;
  double2* bvec;
};
template <typename C>
__global__ void k_steady_rhs(SteadyRhsArgs<C> g) {
  const long total = (long)g.NK * g.M * g.n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long bm = idx / g.n;
    const int i = (int)(bm % g.M);
    g.bvec[idx] = i < g.g0_rows ? to_d2(g.g0[idx]) : make_double2(0.0, 0.0);
  }
}
template <typename C>
__global__ void k_steady_rhs_xi(SteadyRhsArgs<C> g) {
  // one thread per (mode row, coupled column): columns are distinct orbitals
  const long total = (long)g.NK * g.M * g.nc;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long bm = idx / g.nc;
    const int c = (int)(idx - bm * g.nc);
    const int o = g.cidx[c];
    if (o < 0) continue;
    const double2 gx = cmul(g.glp[bm], to_d2(g.xic[bm * g.nc + c]));
    double2& t = g.bvec[bm * g.n + o];
    t.x += g.hbar * gx.x;
    t.y += g.hbar * gx.y;
  }
}

// out = a x + b y + c z over FP64 complex vectors (real coefficients; null skips).
__global__ void k_lincomb(long total, double a, const double2* x, double b, const double2* y,
                          double c, const double2* z, double2* out) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    double2 r = make_double2(0.0, 0.0);
    if (x) { const double2 v = x[idx]; r.x += a * v.x; r.y += a * v.y; }
    if (y) { const double2 v = y[idx]; r.x += b * v.x; r.y += b * v.y; }
    if (z) { const double2 v = z[idx]; r.x += c * v.x; r.y += c * v.y; }
    out[idx] = r;
  }
}

// Real inner product Σ (x.re y.re + x.im y.im) of FP64 complex vectors: one partial
// per block (fixed-order tree), summed by the host in block order.
constexpr int kDotBlocks = 512;
__global__ __launch_bounds__(256) void k_dot(long total, const double2* x, const double2* y,
                                             double* part) {
  double acc = 0.0;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const double2 a = x[idx], b = y[idx];
    acc += a.x * b.x + a.y * b.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Self-energy correction of the preconditioner (Woodbury on the corb orbitals):
// Wc_i = K_{g(i)} Zc_i for every mode row i (g < 0: padding mode, Wc_i = 0).
template <typename C>
__global__ void k_group_matvec(int NK, int M, int ncr, const int* group, const double2* K,
                               const C* Zc, C* Wc) {
  const long total = (long)NK * M * ncr;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long bm = idx / ncr;
    const int c = (int)(idx - bm * ncr);
    const int g = group[bm];
    double wr = 0.0, wi = 0.0;
    if (g >= 0) {
      const double2* Kr = K + ((long)g * ncr + c) * ncr;
      const C* z = Zc + bm * ncr;
      for (int q = 0; q < ncr; ++q) {
        const double2 k = Kr[q], v = to_d2(z[q]);
        wr += k.x * v.x - k.y * v.y;
        wi += k.x * v.y + k.y * v.x;
      }
    }
    Wc[idx] = from_d2<C>(make_double2(wr, wi));
  }
}

// y[i, k] -= T[i, k] / (χ⁺_i + Δ - e_k)   (T from the GEMM; y read and written here)
template <typename C>
struct EpiPrecCorr {
  C* y; const double* eval; const double2* chip; const int* lead; const StageParams* sp;
  int M, n;
  __device__ inline void operator()(int b, int row, int col, double re, double im) const {
    double2 chi = chip[(long)b * M + row];
    chi.x += sp->delta[lead[row]] - eval[(long)b * n + col];
    const double den = chi.x * chi.x + chi.y * chi.y;
    const long off = ((long)b * M + row) * n + col;
    if (den > 0.0) {
      const double2 yv = to_d2(y[off]);
      const double qr = (re * chi.x + im * chi.y) / den, qi = (im * chi.x - re * chi.y) / den;
      y[off] = from_d2<C>(make_double2(yv.x - qr, yv.y - qi));
    }
  }
};

// Preconditioner (χ⁺_i + Δ - H)^{-1} = V diag(1/(χ⁺_i + Δ - e_k)) V†, applied as
// Y = R conj(V) (GEMM), this scaling, Z = Y Vᵀ (GEMM).
template <typename C>
struct EpiPrecScale {
  C* out; const double* eval; const double2* chip; const int* lead; const StageParams* sp;
  int M, n;
  __device__ inline void operator()(int b, int row, int col, double re, double im) const {
    double2 chi = chip[(long)b * M + row];
    chi.x += sp->delta[lead[row]] - eval[(long)b * n + col];
    const double den = chi.x * chi.x + chi.y * chi.y;
    double2 o = make_double2(0.0, 0.0);
    if (den > 0.0) o = make_double2((re * chi.x + im * chi.y) / den, (im * chi.x - re * chi.y) / den);
    out[((long)b * M + row) * n + col] = from_d2<C>(o);
  }
};

}  // namespace zh
