// The Ω stage kernel of the mixed-precision attempt with the Ωξ product fused in.
//
// Per stage s the Ω branch forms k_s = h src + i(u - v)∘Ω_s (sources on the matrix
// cores, stage input Ω_s = y0 + Σ_j a_sj k_j and the outer term in the epilogue,
// Zandpack/mpi/zand:722-735, 784-809), and the Ψ branch needs Ω_s ξ_c
// (Zandpack/mpi/zand:771-778).  By linearity Ω_s ξ_c = y0 ξ_c + Σ_j a_sj k_j ξ_c, so
// each stage's k_s ξ_c is all the Ψ branches of the later stages need from stage s.
// This kernel computes it from the k_s values its own epilogue has just stored,
// instead of a separate product streaming the k_s slot back from HBM.
//
// A workgroup owns a band of 64 rows and a chunk of columns (a multiple of 32) and
// walks the chunk in 64 x 32 tiles; each of its four waves owns 16 rows of the band
// and both 16-column halves of a tile, so a row's ξ products never leave its wave.
// The FP32 epilogue works on the transposed MFMA tile (lane = row, four consecutive
// columns), which is the A-operand layout of four 16x16x4 MFMA steps with the
// contraction index permuted inside the 16 columns (step r takes column 4 lk + r);
// ξ_c is read in the matching order.  Per tile only the ξ_c column groups of the
// tile's modes' leads are multiplied.  At the end of its chunk a wave writes its
// rows' partial products (one per chunk, in a fixed order summed by the consumer):
// the result is deterministic and the same in graph and direct mode.
//
// Chunks: the active column block [0, ML) is cut into SA chunks and the passive
// block [ML, Md) into SP; bands of active rows visit all SA + SP chunks, bands of
// passive rows only the SA active ones when the passive x passive block is skipped.
// The work items are enumerated band by band, so consecutive workgroups (which the
// hardware deals out round-robin over the XCDs) carry equal work.
#pragma once
#include "zh_kernels.hpp"

namespace zh {

struct OmegaFusedArgs {
  const float2* A;       // Aomg (NK, Md, K): [h ψ_c | ΔP ξ_c]
  const float2* B;       // Bomg (NK, Md, K): [ΔM Ixi_c | h ψ_c*]
  int Md, K;             // K = 2 nc
  const float2* xic;     // ξ_c (NK, Md, nc)
  int nc;
  const int2* txi;       // per 64-column tile: ξ_c column range of its modes' leads
  int C;                 // columns per chunk (multiple of 32)
  int SA, SP;            // chunks of the active / passive column block
  int ML;                // active block [0, ML), a multiple of 64
  int nfull, nbands;     // bands [0, nfull) visit every chunk, the rest SA of them
  float2* wpart;         // (NK, SA + SP, Md, nc) partial products, one per chunk
};

constexpr int kFusedBM = 64, kFusedBN = 32, kFusedBK = 8;

// Work items per k block, and item -> (band, chunk).
__host__ __device__ inline int fused_items(const OmegaFusedArgs& a) {
  return a.nfull * (a.SA + a.SP) + (a.nbands - a.nfull) * a.SA;
}
__device__ inline void fused_item(const OmegaFusedArgs& a, int item, int& band, int& chunk) {
  const int nf = a.nfull * (a.SA + a.SP);
  if (item < nf) {
    band = item / (a.SA + a.SP);
    chunk = item - band * (a.SA + a.SP);
  } else {
    const int r = item - nf;
    band = a.nfull + r / a.SA;
    chunk = r - (band - a.nfull) * a.SA;
  }
}

template <int NG, typename Epi>
__global__ __launch_bounds__(kThreads) void k_omega_fused(OmegaFusedArgs a, Epi epi,
                                                          KSegOmega ks) {
  constexpr int BM = kFusedBM, BN = kFusedBN, BK = kFusedBK, PAD = 4;
  constexpr bool RED = HasReduce<Epi>::value;
  typedef v4f acc_t;
  __shared__ float sAr[BK][BM + PAD];
  __shared__ float sAi[BK][BM + PAD];
  __shared__ float sBr[BK][BN + PAD];
  __shared__ float sBi[BK][BN + PAD];
  __shared__ float2 sX[BN][NG * 16 + 1];  // ξ_c rows of the current sub-tile

  const int b = blockIdx.y;
  int band, chunk;
  fused_item(a, blockIdx.x, band, chunk);
  const int m0 = band * BM;
  const int cs = chunk < a.SA ? chunk * a.C : a.ML + (chunk - a.SA) * a.C;
  const int ce = chunk < a.SA ? min(a.ML, cs + a.C) : min(a.Md, cs + a.C);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const float2* __restrict__ A = a.A + (long)b * a.Md * a.K;
  const float2* __restrict__ B = a.B + (long)b * a.Md * a.K;
  const float2* __restrict__ X = a.xic + (long)b * a.Md * a.nc;

  acc_t wr[NG], wi[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    wr[g] = Mfma<float>::zero();
    wi[g] = Mfma<float>::zero();
  }
  double red = 0.0;
  const int row = m0 + wave * 16 + lr;  // this lane's output row

  for (int n0 = cs; n0 < ce; n0 += BN) {
    // ---- sources [ψ_c | ΔPξ_c][ΔM Ixi_c | ψ_c*]ᵀ on the matrix cores (transposed tile)
    acc_t cr[2], ci[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) { cr[j] = Mfma<float>::zero(); ci[j] = Mfma<float>::zero(); }
    const KRange kr = ks.get(m0, n0, BM, BN, a.K);
    const int nt0 = kr.nseg > 0 ? (kr.hi0 - kr.lo0 + BK - 1) / BK : 0;
    const int nt1 = kr.nseg > 1 ? (kr.hi1 - kr.lo1 + BK - 1) / BK : 0;
    const int ktiles = nt0 + nt1;
    float2 ra[2], rb;
    auto load = [&](int t) {
      const bool first = t < nt0;
      const int k0 = first ? kr.lo0 + t * BK : kr.lo1 + (t - nt0) * BK;
      const int khi = first ? kr.hi0 : kr.hi1;
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const int e = tid + it * kThreads;
        const int m = e / BK, k = e % BK;
        const int gm = m0 + m, gk = k0 + k;
        float2 v = make_float2(0.f, 0.f);
        if (gm < a.Md && gk < khi) v = A[(long)gm * a.K + gk];
        ra[it] = v;
      }
      {
        const int n = tid / BK, k = tid % BK;
        const int gn = n0 + n, gk = k0 + k;
        float2 v = make_float2(0.f, 0.f);
        if (gn < a.Md && gk < khi) v = B[(long)gn * a.K + gk];
        rb = v;
      }
    };
    if (ktiles > 0) load(0);
    for (int kt = 0; kt < ktiles; ++kt) {
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const int e = tid + it * kThreads;
        sAr[e % BK][e / BK] = ra[it].x;
        sAi[e % BK][e / BK] = ra[it].y;
      }
      sBr[tid % BK][tid / BK] = rb.x;
      sBi[tid % BK][tid / BK] = rb.y;
      __syncthreads();
      if (kt + 1 < ktiles) load(kt + 1);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const float ar = sAr[kk + lk][wave * 16 + lr], ai = sAi[kk + lk][wave * 16 + lr];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const float br = sBr[kk + lk][j * 16 + lr], bi = sBi[kk + lk][j * 16 + lr];
          // b x a: the accumulator is the transposed tile (lane = row, four columns)
          cr[j] = Mfma<float>::op(br, ar, cr[j]);
          ci[j] = Mfma<float>::op(bi, ar, ci[j]);
          cr[j] = Mfma<float>::op(bi, -ai, cr[j]);
          ci[j] = Mfma<float>::op(br, ai, ci[j]);
        }
      }
      __syncthreads();
    }
    // ---- ξ_c rows of the sub-tile's 32 modes, the column groups of their leads only,
    // staged in LDS once for the four waves
    const int2 rg = a.txi[n0 / 64];
    const int g0 = rg.x / 16, g1 = min(NG, (rg.y + 15) / 16);
    __syncthreads();  // the previous sub-tile's ξ reads are done
    {
      const int w16 = (g1 - g0) * 16;
      for (int e = tid; e < BN * w16; e += kThreads) {
        const int jj = e / w16, c = g0 * 16 + (e - jj * w16);
        const int jm = n0 + jj;
        sX[jj][c] = (jm < a.Md && c < a.nc) ? X[(long)jm * a.nc + c] : make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    // ---- epilogue (stage input, outer term, store, error norm), then Ω ξ_c from the
    // values just stored, one 16-column half at a time
    auto half = [&](int j, const acc_t& ar_, const acc_t& ai_) {
      const int col0 = n0 + j * 16 + 4 * lk;
      double re[4], im[4], qr[4], qi[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        re[r] = (double)ar_[r];
        im[r] = (double)ai_[r];
      }
      const int ncol = row < a.Md ? max(0, min(4, a.Md - col0)) : 0;
      epi.row4(b, row < a.Md ? row : -1, col0, ncol, re, im, red, qr, qi);
      float pr[4], pi[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) { pr[r] = (float)qr[r]; pi[r] = (float)qi[r]; }
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        if (g < g0 || g >= g1) continue;  // uniform over the workgroup
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // step r: this lane's k index stands for column j*16 + 4 lk + r of the sub-tile
          const float2 x = sX[j * 16 + 4 * lk + r][g * 16 + lr];
          wr[g] = Mfma<float>::op(pr[r], x.x, wr[g]);
          wr[g] = Mfma<float>::op(-pi[r], x.y, wr[g]);
          wi[g] = Mfma<float>::op(pr[r], x.y, wi[g]);
          wi[g] = Mfma<float>::op(pi[r], x.x, wi[g]);
        }
      }
    };
    half(0, cr[0], ci[0]);
    half(1, cr[1], ci[1]);
  }
  // ---- this chunk's partial Ω ξ_c of the wave's 16 rows (item q of lane l: row 4 lk + q)
  float2* __restrict__ W = a.wpart + ((long)b * (a.SA + a.SP) + chunk) * a.Md * a.nc;
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int c = g * 16 + lr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rw = m0 + wave * 16 + 4 * lk + q;
      if (rw < a.Md && c < a.nc) W[(long)rw * a.nc + c] = make_float2(wr[g][q], wi[g][q]);
    }
  }
  if constexpr (RED) {
    __shared__ double sred[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) red += __shfl_xor(red, off, 64);
    if (lane == 0) sred[wave] = red;
    __syncthreads();
    if (tid == 0) epi.finish(blockIdx.y * gridDim.x + blockIdx.x, (sred[0] + sred[1]) + (sred[2] + sred[3]));
  }
}

}  // namespace zh
