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
//     lead orbitals on the VALU while those loads are in flight: (ΔP ξ_c)[i, c] is the
//     same for every lane of a wave (its rows), (ħ σ_c)[c, o] is one coalesced load,
//   * gathers (H ψ_i)[o] from the LDS rows and applies the epilogue,
// so the stage Ψ is read once and k written once.  Thread layout: lane = orbital o of
// the tile, wave w = rows [w TR/4, (w+1) TR/4) of the tile.
#pragma once
#include <hip/hip_runtime.h>

#include "zh_kernels.hpp"

namespace zh {

constexpr int kPsiFusedWMax = 160;  // widest gather range of a 64-orbital tile

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
};

template <typename C, int TR, int RC>
__global__ __launch_bounds__(256) void k_psi_fused(PsiFusedArgs<C> a) {
  typedef decltype(C::x) T;
  constexpr int RW = TR / 4;  // rows per wave
  extern __shared__ __align__(16) unsigned char psi_lds[];
  C* tile = reinterpret_cast<C*>(psi_lds);  // TR x W
  const int b = blockIdx.z;
  const int o0 = blockIdx.x * 64, m0 = blockIdx.y * TR;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int n = a.n, Md = a.Md;
  const int o = o0 + lane;
  const bool on = o < n;
  const int2 spn = a.span[blockIdx.x];
  const int lo = spn.x, W = spn.y;
  const C* __restrict__ Ab = a.Ap + (long)b * a.sAp;

  // 1. the stage-Ψ rows of the tile over [lo, lo + W): batches of LB independent loads
  //    per thread (all in flight), then their LDS stores
  constexpr int LB = 16;
  const int total = TR * W;
  for (int e0 = 0; e0 < total; e0 += 256 * LB) {
    C v[LB];
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int e = e0 + u * 256 + t;
      const int r = e / W, x = e - r * W;
      const int i = min(m0 + r, Md - 1);  // rows past Md: a valid row, never used
      v[u].x = 0; v[u].y = 0;
      if (e < total) v[u] = Ab[(long)i * a.ldAp + lo + x];
    }
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int e = e0 + u * 256 + t;
      if (e < total) tile[e] = v[u];
    }
  }

  // 2. ħ ΔP_i (σ ξ_i)[o] for the wave's rows, in chunks of KB modes: the orbital's KB
  //    values of ħ σ_c in registers, each row's KB values of ΔP ξ_c the same for the
  //    whole wave (rows past Md read the last row; they are never stored)
  constexpr int KB = 8;
  T ar[RW], ai[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) { ar[q] = 0; ai[q] = 0; }
  if (m0 < a.acc0_rows) {
    const int2 kr = a.t2[m0 >> 6];
    const C* __restrict__ Bb = a.Bc + (long)b * a.sBc;
    const int i0 = m0 + RW * w;
    const int oc = on ? o : n - 1;
    for (int k0 = kr.x; k0 < kr.y; k0 += KB) {
      C bv[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) {
        bv[u].x = 0; bv[u].y = 0;
        if (k0 + u < kr.y) bv[u] = Bb[(long)(k0 + u) * n + oc];
      }
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        // one wave-uniform row of KB values (contiguous: wide scalar loads); modes past
        // the range read on into the next row or the allocation's zeroed slack, times
        // bv = 0
        const C* __restrict__ arow = Ab + (long)min(i0 + q, Md - 1) * a.ldAp + n + k0;
#pragma unroll
        for (int u = 0; u < KB; ++u) {
          const C av = arow[u];
          ar[q] += av.x * bv[u].x - av.y * bv[u].y;
          ai[q] += av.x * bv[u].y + av.y * bv[u].x;
        }
      }
    }
  }
  __syncthreads();
  if (!on) return;

  // 3. (H ψ_i)[o] from the LDS rows, and the epilogue
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
  for (int q = 0; q < RW; ++q) {
    const int rl = RW * w + q;
    const int i = m0 + rl;
    if (i >= Md) break;
    const C* row = tile + rl * W;
    double re = (double)ar[q], im = (double)ai[q];
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
    st_stream(a.out + bm * n + o, from_d2<C>(make_double2(hs * vi, -hs * vr)));
  }
}

}  // namespace zh
