// The Ψ right-hand side of one stage in one pass when H is row-compressed
// (Zandpack/mpi/zand:746-778):
//   k_i[o] = -i h/ħ [ (H ψ_i)[o] + ħ ΔP_i (σ ξ_i)[o] - χ_i ψ_i[o] + d_o ψ_i[o]
//                     + (ħ Λ ξ_i + Ω_s ξ_c / ħ)[o] on coupled orbitals ]
// with χ_i = χ⁺_i + Δ_α(i) and d the Hubbard diagonal.  The split form writes the
// ΔP (σ ξ) part with a matrix-core product into the k slot and reads it back in the
// gather kernel; here a workgroup owns TR mode rows x 64 orbitals o and
//   * loads its stage-Ψ rows over the orbitals the tile's H rows reach (the tile's
//     columns and their neighbours, [lo, lo + W) from a host table: W <= WMAX) into LDS,
//   * forms ħ ΔP_i (σ ξ_i)[o] = Σ_c (ΔP ξ_c)[i, c] (ħ σ_c)[c, o] over the row tile's
//     lead orbitals on the matrix cores while those loads are in flight,
//   * gathers (H ψ_i)[o] from the LDS rows and applies the epilogue,
// so the stage Ψ is read once and k written once.  Thread layout: wave w owns orbitals
// [o0 + 16 w, o0 + 16 w + 16) of all 16 rows; a lane holds one orbital and 4 rows,
// where the matrix-core product leaves them.
#pragma once
#include <hip/hip_runtime.h>

#include "zh_kernels.hpp"

namespace zh {

constexpr int kPsiFusedWMax = 160;  // widest gather range of a 64-orbital tile
// Whether the product's result goes through LDS so that the gather and epilogue run
// on wave-uniform rows (lane = orbital: scalar loads of the per-row values, 64-orbital
// stores).  Measured at n = 1000, M = 4000: complex64 85 µs with it, 100 without;
// complex128 116 µs without it, 128 with (the 16 KB more LDS per workgroup costs a
// workgroup per CU).
template <typename C> constexpr bool kPsiXL = sizeof(C) == 8;

template <typename C>
struct PsiFusedArgs {
  C* out;                          // k slot (NK, Md, n)
  const C* Ap; long sAp; int ldAp;  // [stage ψ | ΔP ξ_c] (NK, Md, n + nc)
  const C* Bc; long sBc;           // ħ σ_cᵀ (nc, n) per batch
  const C* xic; const C* W;        // ξ_c, Ω_s ξ_c (NK, Md, nc)
  const double2* glp; const double2* chip;  // (NK, Md)
  const double2* dvec;             // Hubbard diagonal (NK, n)
  const int* lead; const int* cpos;
  const int2* t2;                  // per 64 rows: coupled-orbital range of the ΔP ξ_c σ_cᵀ term
  const int2* span;                // per 64 orbitals: first orbital and width of the gather
  const int* col; const double2* val; int R;  // row-compressed H (NK, R, n)
  const StageParams* sp; int ki;
  int Md, n, nc, acc0_rows;
  double inv_hbar, inv_hbar2;
  // the next stage's input, formed here (the stage combination's Ψ part): nx_out (Ap
  // layout) and nx_y (y_new, the last stage's input, or null) = y0 + Σ_j nx_a[j] k_j over
  // nx_n slots nx_k[j] (term nx_self is this kernel's k, taken from registers; term nx_k1,
  // the k_1 slot, weighted by the instance's r1); none when nx_out is null
  C* nx_out = nullptr;
  C* nx_y = nullptr;
  const C* nx_y0 = nullptr;
  const C* nx_k[6] = {nullptr};
  double nx_a[6] = {0};
  int nx_n = 0, nx_self = -1, nx_k1 = -1;
  // and nx_e = Σ_j nx_ct[j] k_j over the same terms (the last stage's input: the error
  // norm's weighted sum of all but k_7), or none
  C* nx_e = nullptr;
  double nx_ct[6] = {0};
};

template <typename C, int TR, int RC, bool XL, bool NXT>
__global__ __launch_bounds__(256) void k_psi_fused(PsiFusedArgs<C> a) {
  static_assert(TR == 16, "one 16-row matrix-core block per workgroup");
  typedef decltype(C::x) T;
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr bool F64 = sizeof(T) == 8;
  extern __shared__ __align__(16) unsigned char psi_lds[];
  C* tile = reinterpret_cast<C*>(psi_lds);  // TR x W
  const int b = blockIdx.z;
  if (a.sp[b / a.ki].skip) return;  // k_1 of the instance is already in its slot
  const int o0 = blockIdx.x * 64, m0 = blockIdx.y * TR;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lk = lane >> 4, lc = lane & 15;
  const int n = a.n, Md = a.Md;
  const int2 spn = a.span[blockIdx.x];
  const int lo = spn.x, W = spn.y;
  // the lane's orbital in the product: column lc of the wave's 16-orbital block; in
  // the gather and epilogue the same with XL false, else orbital o0 + lane (wave-
  // uniform rows, the product exchanged through LDS)
  const int o2 = o0 + 16 * w + lc;
  const int o = XL ? o0 + lane : o2;
  const bool on = o < n;
  const C* __restrict__ Ab = a.Ap + (long)b * a.sAp;

  // 1. the stage-Ψ rows of the tile over [lo, lo + W): batches of LB independent loads
  //    per thread; the first batch stays in flight over the matrix-core part
  constexpr int LB = F64 ? 8 : 16;
  const int total = TR * W;
  C v[LB];
  auto load_batch = [&](int e0) {
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int e = e0 + u * 256 + t;
      const int r = e / W, x = e - r * W;
      const int i = min(m0 + r, Md - 1);  // rows past Md: a valid row, never used
      v[u].x = 0; v[u].y = 0;
      if (e < total) v[u] = Ab[(long)i * a.ldAp + lo + x];
    }
  };
  auto store_batch = [&](int e0) {
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int e = e0 + u * 256 + t;
      if (e < total) tile[e] = v[u];
    }
  };
  load_batch(0);

  // 2. ħ ΔP_i (σ ξ_i)[o] = Σ_c (ΔP ξ_c)[i, c] (ħ σ_c)[c, o] for the tile's 16 rows and
  //    the wave's 16 orbitals on the matrix cores, over the row tile's coupled-orbital
  //    range in steps of 16 modes: lane (lc, lk) holds modes k0 + 4 lk + j, j < 4, of
  //    row m0 + lc (A) and of orbital o (B), and product j sums them over lk
  acc_t cr = MF::zero(), ci = MF::zero(), c3 = MF::zero();
  if (m0 < a.acc0_rows) {
    const int2 kr = a.t2[m0 >> 6];
    const C* __restrict__ arow = Ab + (long)min(m0 + lc, Md - 1) * a.ldAp + n;
    const C* __restrict__ bcol = a.Bc + (long)b * a.sBc + min(o2, n - 1);
    for (int k0 = kr.x; k0 < kr.y; k0 += 16) {
      C av[4], bv[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + 4 * lk + j;
        av[j].x = 0; av[j].y = 0; bv[j].x = 0; bv[j].y = 0;
        if (k < kr.y) {
          av[j] = arow[k];
          bv[j] = bcol[(long)k * n];
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (F64) {  // three real products per complex one
          c3 = MF::op(av[j].x + av[j].y, bv[j].x, c3);
          ci = MF::op(av[j].x, bv[j].y - bv[j].x, ci);
          cr = MF::op(-av[j].y, bv[j].x + bv[j].y, cr);
        } else {
          cr = MF::op(av[j].x, bv[j].x, cr);
          ci = MF::op(av[j].x, bv[j].y, ci);
          cr = MF::op(-av[j].y, bv[j].y, cr);
          ci = MF::op(av[j].y, bv[j].x, ci);
        }
      }
    }
  }
  store_batch(0);
  if constexpr (XL) {
    C* xl = tile + TR * W;  // TR x 64
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      C x;
      x.x = F64 ? cr[q] + c3[q] : cr[q];
      x.y = F64 ? ci[q] + c3[q] : ci[q];
      xl[MF::row(lk, q) * 64 + 16 * w + lc] = x;
    }
  }
  for (int e0 = 256 * LB; e0 < total; e0 += 256 * LB) {
    load_batch(e0);
    store_batch(e0);
  }
  __syncthreads();
  if (!on) return;

  // 3. (H ψ_i)[o] from the LDS rows, and the epilogue, for the lane's 4 rows of the
  //    product's output layout
  const int R = RC > 0 ? RC : a.R;
  const int* __restrict__ colb = a.col + (long)b * a.R * n + o;
  const double2* __restrict__ valb = a.val + (long)b * a.R * n + o;
  int cl[RC > 0 ? RC : 1];
  double2 vl[RC > 0 ? RC : 1];
  if constexpr (RC > 0) {
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      cl[r] = colb[(long)r * n] - lo;
      vl[r] = valb[(long)r * n];
    }
  }
  const double2 d = a.dvec[(long)b * n + o];
  const int cp = a.cpos[o];
  const StageParams& sq = a.sp[b / a.ki];
  const double hs = sq.h;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int rl = XL ? 4 * w + q : MF::row(lk, q);
    const int i = m0 + rl;
    if (XL && i >= Md) break;
    if (i >= Md) continue;
    const C* row = tile + rl * W;
    double re, im;
    if constexpr (XL) {
      const C x = tile[TR * W + rl * 64 + lane];
      re = (double)x.x;
      im = (double)x.y;
    } else {
      re = (double)(F64 ? cr[q] + c3[q] : cr[q]);
      im = (double)(F64 ? ci[q] + c3[q] : ci[q]);
    }
    if constexpr (RC > 0) {
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        const double2 x = to_d2(row[cl[r]]);
        re += x.x * vl[r].x - x.y * vl[r].y;
        im += x.x * vl[r].y + x.y * vl[r].x;
      }
    } else {
      for (int r = 0; r < R; ++r) {
        const double2 x = to_d2(row[colb[(long)r * n] - lo]);
        const double2 v = valb[(long)r * n];
        re += x.x * v.x - x.y * v.y;
        im += x.x * v.y + x.y * v.x;
      }
    }
    const long bm = (long)b * Md + i;
    double2 chi = a.chip[bm];
    chi.x += sq.delta[a.lead[i]];
    const double2 p = to_d2(row[o - lo]);
    const double2 cpx = cmul(chi, p), dp = cmul(d, p);
    double vr = (re - cpx.x + dp.x) * a.inv_hbar;
    double vi = (im - cpx.y + dp.y) * a.inv_hbar;
    if (cp >= 0) {
      const double2 gx = cmul(a.glp[bm], to_d2(a.xic[bm * a.nc + cp]));
      const double2 wv = to_d2(a.W[bm * a.nc + cp]);
      vr += gx.x + wv.x * a.inv_hbar2;
      vi += gx.y + wv.y * a.inv_hbar2;
    }
    const C kv = from_d2<C>(make_double2(hs * vi, -hs * vr));
    st_stream(a.out + bm * n + o, kv);
    if constexpr (NXT) {
      // the next stage's input at (i, o), the terms in the stage combination's order
      const long e = bm * n + o;
      C x[6];
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (j < a.nx_n) x[j] = j == a.nx_self ? kv : ld_stream(a.nx_k[j] + e);
      double2 y = to_d2(ld_stream(a.nx_y0 + e));
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (j < a.nx_n) {
          const double aj = j == a.nx_k1 ? a.nx_a[j] * sq.r1 : a.nx_a[j];
          y.x += aj * (double)x[j].x;
          y.y += aj * (double)x[j].y;
        }
      const C yc = from_d2<C>(y);
      st_stream(a.nx_out + (long)b * a.sAp + (long)i * a.ldAp + o, yc);
      if (a.nx_y) st_stream(a.nx_y + e, yc);
      if (a.nx_e) {
        double2 u = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < 6; ++j)
          if (j < a.nx_n) {
            const double cj = j == a.nx_k1 ? a.nx_ct[j] * sq.r1 : a.nx_ct[j];
            u.x += cj * (double)x[j].x;
            u.y += cj * (double)x[j].y;
          }
        st_stream(a.nx_e + e, from_d2<C>(u));
      }
    }
  }
}

}  // namespace zh
