// Batched complex GEMM on CDNA4 matrix cores with fused epilogues.
//
//   acc = op(A) * op(B)      (complex, K-loop on MFMA 16x16x4, FP64 or FP32)
//   epi(batch, row, col, acc)
//
// Complex operands are stored interleaved (re, im) as in numpy complex128 /
// complex64.  Tiles are staged through LDS split into real and imaginary
// planes, k-major, so that the MFMA operand fetch of a wavefront reads 16
// consecutive elements per k row.  One complex product is four real MFMAs:
//   Cre += Are*Bre ; Cre += (-Aim)*Bim ; Cim += Are*Bim ; Cim += Aim*Bre
//
// v_mfma_{f64_16x16x4f64, f32_16x16x4f32} lane layout (wave64):
//   A fragment: lane l holds A[l%16][l/16]     (16 x 4)
//   B fragment: lane l holds B[l/16][l%16]     (4 x 16)
//   D fragment, item r of lane l:
//     f32: D[4*(l/16) + r][l%16]   (each lane owns 4 consecutive rows)
//     f64: D[(l/16) + 4*r][l%16]   (rows interleaved with stride 4)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace zh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef v4d acc_t;
  __device__ static inline acc_t zero() { acc_t z = {0.0, 0.0, 0.0, 0.0}; return z; }
  __device__ static inline acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct Mfma<float> {
  typedef v4f acc_t;
  __device__ static inline acc_t zero() { acc_t z = {0.f, 0.f, 0.f, 0.f}; return z; }
  __device__ static inline acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int r) { return 4 * lk + r; }
};

// Interleaved complex element types as stored in memory.
template <typename R> struct Cplx;
template <> struct Cplx<double> { typedef double2 type; };
template <> struct Cplx<float> { typedef float2 type; };

// Operand descriptor.  Element (m,k) of op(A):
//   !TRANS_A : A[m*lda + k]        TRANS_A : A[k*lda + m]
// Element (k,n) of op(B):
//   !TRANS_B : B[k*ldb + n]        TRANS_B : B[n*ldb + k]
// splitk > 1 splits the K range over blockIdx.z = batch * splitk + part; the
// epilogue then receives z (not the batch index) and must keep the partial
// products apart (they are summed by the consumer in a fixed order).
// Real-operand modes (template RMODE of the kernel): 0 complex x complex; 1 the
// A operand is the real matrix Areal (element (m,k) at Areal[m*ldar + k], or
// [k*ldar + m] transposed); 2 the B operand is the real matrix Breal ((k,n) at
// Breal[k*ldbr + n], or [n*ldbr + k]).  A real operand moves half the bytes and
// its product needs two real MFMAs per tile pair instead of four.  Mode 3 is
// complex x complex in the three-multiplication form
//   t1 = (ar + ai) br,  t2 = ar (bi - br),  t3 = ai (br + bi);
//   Re = t1 - t3,  Im = t1 + t2,
// three real MFMAs per tile pair (the operand sums are formed from the LDS values).
template <typename RA, typename RB>
struct GemmArgs {
  int M, N, K;
  const typename Cplx<RA>::type* A; long lda; long sA;
  const typename Cplx<RB>::type* B; long ldb; long sB;
  int splitk = 1;
  int xcd_order = 1;  // XCD-aware tile order (off where tiles are skipped unevenly)
  const RA* Areal = nullptr; long ldar = 0; long sAr = 0;
  const RB* Breal = nullptr; long ldbr = 0; long sBr = 0;
  // per instance of skip_ki batches: 1 = leave the batch alone (every workgroup of it
  // returns at once; a first stage whose result is already known)
  const int* skip = nullptr; int skip_ki = 1 << 30;
};

constexpr int kThreads = 256;

// K-range selection per output tile (first row m0, first column n0, bm x bn): up
// to three half-open segments [lo0, hi0), [lo1, hi1), [lo2, hi2) of the K axis (the
// first nseg of them; an empty one is lo = hi); tiles whose operands are known to
// vanish are skipped, and a tile with no segment gets acc = 0 (its epilogue still
// runs).  Returned by value and built with selects, so that the ranges stay in
// registers.
struct KRange {
  int lo0, hi0, lo1, hi1, nseg;
  int lo2, hi2;
};

// Selections with static constexpr bool kMirrorA = true read A in their first segment
// as the mirror of a square row-major A: op(A)(m, k) = -conj(A[k*lda + m]) there, and
// A[m*lda + k] in the others (the two halves of an anti-Hermitian matrix held as one
// triangle).  Complex A, not transposed.
template <typename KS, typename = void> struct MirrorA : std::false_type {};
template <typename KS>
struct MirrorA<KS, std::void_t<decltype(KS::kMirrorA)>>
    : std::integral_constant<bool, KS::kMirrorA> {};
struct KFull {
  __device__ inline KRange get(int, int, int, int, int K) const { return KRange{0, K, 0, 0, 1}; }
  __device__ inline bool skip(int, int, int, int) const { return false; }
};

// Reducing epilogues (static constexpr bool kReduce = true) accumulate one double
// per thread through operator()(z, row, col, re, im, double& acc); the workgroup's
// sum (a fixed-order tree over its threads) goes to epi.finish(block, sum), where
// block is the workgroup's linear index in the grid.  Skipped tiles report 0.
template <typename Epi, typename = void> struct HasReduce : std::false_type {};
template <typename Epi>
struct HasReduce<Epi, std::void_t<decltype(Epi::kReduce)>>
    : std::integral_constant<bool, Epi::kReduce> {};

// Blocked epilogues (static constexpr bool kBlock4 = true) take the four outputs a
// lane holds in one 16x16 MFMA tile (same column, four rows) in one call,
//   block(z, rows[4], col, re[4], im[4], double& acc),
// rows[r] = -1 for rows outside the matrix, so that they can issue every load of
// the four elements before their first store (stores to a buffer that may alias
// the inputs otherwise keep the loads of the next element behind them).
template <typename Epi, typename = void> struct HasBlock4 : std::false_type {};
template <typename Epi>
struct HasBlock4<Epi, std::void_t<decltype(Epi::kBlock4)>>
    : std::integral_constant<bool, Epi::kBlock4> {};

// Row epilogues (static constexpr bool kRow4 = true) take, in FP32, four
// consecutive columns of one output row per call,
//   row4(z, row, col0, ncol, re[4], im[4], double& acc)   (ncol of them valid;
// row = -1 outside the matrix),
// so that a lane's loads and stores of them are 32 contiguous bytes.  The kernel
// then issues the MFMAs with A and B swapped: the accumulator tile is the
// transpose, whose four items per lane run along the output row.
template <typename Epi, typename = void> struct HasRow4 : std::false_type {};
template <typename Epi>
struct HasRow4<Epi, std::void_t<decltype(Epi::kRow4)>>
    : std::integral_constant<bool, Epi::kRow4> {};

// Workgroup: 256 threads = 4 waves laid out WM x (4/WM) over a BM x BN output
// tile; each wave owns a (BM/WM) x (BN*WM/4) block of 16x16 MFMA tiles.  The
// global loads of K tile t+1 are issued before the MFMAs of tile t.  With DB the
// LDS holds two K tiles and one barrier per tile separates the stages; without,
// the single LDS tile costs a second barrier but keeps more workgroups per CU.
template <typename T, typename RA, typename RB, bool TRANS_A, bool TRANS_B,
          int BM, int BN, int BK, int WM, bool DB, int OCC, int RMODE, typename Epi,
          typename KSeg>
__global__ __launch_bounds__(kThreads, OCC) void zgemm_kernel(GemmArgs<RA, RB> g, Epi epi,
                                                              KSeg ks) {
  constexpr bool RA_ = RMODE == 1, RB_ = RMODE == 2, G3 = RMODE == 3;
  constexpr bool MIR = MirrorA<KSeg>::value;
  static_assert(!MIR || (!TRANS_A && !RA_), "mirrored A: complex, not transposed");
  constexpr bool SWP = std::is_same<T, float>::value && HasRow4<Epi>::value;
  constexpr int NBUF = DB ? 2 : 1;
  constexpr int WN = 4 / WM;
  constexpr int WTM = BM / WM, WTN = BN / WN;  // wave tile
  constexpr int TM = WTM / 16, TN = WTN / 16;  // 16x16 MFMA tiles per wave
  static_assert(WM * WN == 4 && WTM % 16 == 0 && WTN % 16 == 0 && BK % 4 == 0, "tile shape");
  constexpr int PAD = (sizeof(T) == 8) ? 2 : 4;
  constexpr int LA = (BM * BK) / kThreads;  // elements each thread stages
  constexpr int LB = (BN * BK) / kThreads;
  static_assert((BM * BK) % kThreads == 0 && (BN * BK) % kThreads == 0, "loader");
  typedef typename Mfma<T>::acc_t acc_t;
  typedef typename Cplx<RA>::type ca_t;
  typedef typename Cplx<RB>::type cb_t;

  __shared__ T sAr[NBUF][BK][BM + PAD];
  __shared__ T sAi[NBUF][BK][BM + PAD];
  __shared__ T sBr[NBUF][BK][BN + PAD];
  __shared__ T sBi[NBUF][BK][BN + PAD];

  // XCD-aware tile order: consecutive workgroups go round-robin to the 8 XCDs,
  // so give XCD x the contiguous tile range x*per.. (whole row blocks with all
  // their column tiles), keeping each A row block in one XCD's L2.
  int ty = blockIdx.y, tx = blockIdx.x;
  {
    const int nx = gridDim.x, total = gridDim.x * gridDim.y;
    if (g.xcd_order && total >= 16) {
      const int lin = blockIdx.y * nx + blockIdx.x;
      const int xcd = lin & 7, k = lin >> 3, q = total >> 3, r = total & 7;
      const int tile = xcd * q + min(xcd, r) + k;
      ty = tile / nx;
      tx = tile - ty * nx;
    }
  }
  constexpr bool RED = HasReduce<Epi>::value;
  const int blin = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  if (g.skip && g.skip[(blockIdx.z / g.splitk) / g.skip_ki]) {
    if constexpr (RED) {
      if (threadIdx.x == 0) epi.finish(blin, 0.0);
    }
    return;
  }
  if (ks.skip(ty * BM, tx * BN, BM, BN)) {  // uniform over the workgroup
    if constexpr (RED) {
      if (threadIdx.x == 0) epi.finish(blin, 0.0);
    }
    return;
  }
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int batch = blockIdx.z / g.splitk;
  const int part = blockIdx.z - batch * g.splitk;
  const int m0 = ty * BM;
  const int n0 = tx * BN;
  const ca_t* __restrict__ A = g.A + (long)batch * g.sA;
  const cb_t* __restrict__ B = g.B + (long)batch * g.sB;
  const RA* __restrict__ Ar = g.Areal ? g.Areal + (long)batch * g.sAr : nullptr;
  const RB* __restrict__ Br = g.Breal ? g.Breal + (long)batch * g.sBr : nullptr;

  acc_t cr[TM][TN], ci[TM][TN];
  acc_t c3[G3 ? TM : 1][G3 ? TN : 1];  // t1 of the three-multiplication form
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) { cr[i][j] = Mfma<T>::zero(); ci[i][j] = Mfma<T>::zero(); }
  if constexpr (G3) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) c3[i][j] = Mfma<T>::zero();
  }

  ca_t ra[LA];
  cb_t rb[LB];
  bool mir = false;  // the K tile in the registers reads A mirrored (MIR, first segment)
  auto load_tiles = [&](int k0, int khi) {
#pragma unroll
    for (int it = 0; it < LA; ++it) {
      const int e = tid + it * kThreads;
      if constexpr (MIR) {
        // one address form for both segment kinds: element (R0 + r, C0 + c) of A, where
        // the stored segments read (m0 + m, k0 + k) and the mirrored one (k0 + k, m0 + m)
        const bool q = mir;
        const int r = q ? e / BM : e / BK, c = q ? e % BM : e % BK;
        const int R0 = q ? k0 : m0, C0 = q ? m0 : k0;
        const int limR = q ? khi : g.M, limC = q ? g.M : khi;
        ca_t v; v.x = 0; v.y = 0;
        if (R0 + r < limR && C0 + c < limC) {
          v = A[(long)(R0 + r) * g.lda + (C0 + c)];
          if (q) v.x = -v.x;  // -conj
        }
        ra[it] = v;
      } else {
        int m, k;
        if (!TRANS_A) { m = e / BK; k = e % BK; } else { k = e / BM; m = e % BM; }
        const int gm = m0 + m, gk = k0 + k;
        ca_t v; v.x = 0; v.y = 0;
        if (gm < g.M && gk < khi) {
          if (RA_) v.x = TRANS_A ? Ar[(long)gk * g.ldar + gm] : Ar[(long)gm * g.ldar + gk];
          else v = TRANS_A ? A[(long)gk * g.lda + gm] : A[(long)gm * g.lda + gk];
        }
        ra[it] = v;
      }
    }
#pragma unroll
    for (int it = 0; it < LB; ++it) {
      const int e = tid + it * kThreads;
      int n, k;
      if (!TRANS_B) { k = e / BN; n = e % BN; } else { n = e / BK; k = e % BK; }
      const int gn = n0 + n, gk = k0 + k;
      cb_t v; v.x = 0; v.y = 0;
      if (gn < g.N && gk < khi) {
        if (RB_) v.x = TRANS_B ? Br[(long)gn * g.ldbr + gk] : Br[(long)gk * g.ldbr + gn];
        else v = TRANS_B ? B[(long)gn * g.ldb + gk] : B[(long)gk * g.ldb + gn];
      }
      rb[it] = v;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int it = 0; it < LA; ++it) {
      const int e = tid + it * kThreads;
      int m, k;
      if (!TRANS_A && !(MIR && mir)) { m = e / BK; k = e % BK; } else { k = e / BM; m = e % BM; }
      sAr[buf][k][m] = (T)ra[it].x;
      if (!RA_) sAi[buf][k][m] = (T)ra[it].y;
    }
#pragma unroll
    for (int it = 0; it < LB; ++it) {
      const int e = tid + it * kThreads;
      int n, k;
      if (!TRANS_B) { k = e / BN; n = e % BN; } else { n = e / BK; k = e % BK; }
      sBr[buf][k][n] = (T)rb[it].x;
      if (!RB_) sBi[buf][k][n] = (T)rb[it].y;
    }
  };

  // K tiles of this tile's segments, split evenly over the splitk parts
  const KRange kr = ks.get(m0, n0, BM, BN, g.K);
  const int lo0 = kr.lo0, hi0 = kr.hi0, lo1 = kr.lo1, hi1 = kr.hi1, nseg = kr.nseg;
  const int lo2 = MIR ? kr.lo2 : 0, hi2 = MIR ? kr.hi2 : 0;  // third segment: MIR only
  const int nt0 = nseg > 0 ? (hi0 - lo0 + BK - 1) / BK : 0;
  const int nt1 = nseg > 1 ? (hi1 - lo1 + BK - 1) / BK : 0;
  const int nt2 = (MIR && nseg > 2) ? (hi2 - lo2 + BK - 1) / BK : 0;
  const int ktot = nt0 + nt1 + nt2;
  const int kper = (ktot + g.splitk - 1) / g.splitk;
  const int kt0 = part * kper;
  const int ktiles = max(0, min(ktot, kt0 + kper) - kt0);
  auto tile_load = [&](int t) {
    if constexpr (MIR) {
      // a branch on the uniform t and a two-way select (a three-way select here made the
      // compiler index a private-memory table)
      mir = t < nt0;
      if (t < nt0) {
        load_tiles(lo0 + t * BK, hi0);
      } else {
        const int u = t - nt0;
        const bool a = u < nt1;
        load_tiles(a ? lo1 + u * BK : lo2 + (u - nt1) * BK, a ? hi1 : hi2);
      }
    } else {
      const bool first = t < nt0;
      load_tiles(first ? lo0 + t * BK : lo1 + (t - nt0) * BK, first ? hi0 : hi1);
    }
  };
  const int lr = lane & 15, lk = lane >> 4;
  if (ktiles > 0) {
    tile_load(kt0);
    if (DB) store_tiles(0);
  }
  if (DB) __syncthreads();
  for (int kt = 0; kt < ktiles; ++kt) {
    // double buffer: tile kt is in buffer kt&1, tile kt+1 goes to the other one
    // after the MFMAs; single buffer: store, barrier, prefetch, MFMAs, barrier
    const int buf = DB ? (kt & 1) : 0;
    if (!DB) {
      store_tiles(0);
      __syncthreads();
    }
    const bool more = kt + 1 < ktiles;
    if (more) tile_load(kt0 + kt + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int m = wm * WTM + i * 16 + lr;
        ar[i] = sAr[buf][kk + lk][m];
        if (!RA_) ai[i] = sAi[buf][kk + lk][m];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int n = wn * WTN + j * 16 + lr;
        br[j] = sBr[buf][kk + lk][n];
        if (!RB_) bi[j] = sBi[buf][kk + lk][n];
      }
      // a x b on the matrix core, or b x a (the transposed tile) for row epilogues
      auto mm = [&](T a, T b, acc_t c) { return SWP ? Mfma<T>::op(b, a, c) : Mfma<T>::op(a, b, c); };
      if (RA_) {  // Cre += Ar Br, Cim += Ar Bi
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            cr[i][j] = mm(ar[i], br[j], cr[i][j]);
            ci[i][j] = mm(ar[i], bi[j], ci[i][j]);
          }
      } else if (RB_) {  // Cre += Ar Br, Cim += Ai Br
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            cr[i][j] = mm(ar[i], br[j], cr[i][j]);
            ci[i][j] = mm(ai[i], br[j], ci[i][j]);
          }
      } else if constexpr (G3) {  // c3 += (ar+ai) br; ci += ar (bi-br); cr += -ai (br+bi)
        T sa[TM], nai[TM], db[TN], sb[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) { sa[i] = ar[i] + ai[i]; nai[i] = -ai[i]; }
#pragma unroll
        for (int j = 0; j < TN; ++j) { db[j] = bi[j] - br[j]; sb[j] = br[j] + bi[j]; }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            c3[i][j] = mm(sa[i], br[j], c3[i][j]);
            ci[i][j] = mm(ar[i], db[j], ci[i][j]);
            cr[i][j] = mm(nai[i], sb[j], cr[i][j]);
          }
      } else {
        T nai[TM];
#pragma unroll
        for (int i = 0; i < TM; ++i) nai[i] = -ai[i];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            cr[i][j] = mm(ar[i], br[j], cr[i][j]);
            ci[i][j] = mm(ar[i], bi[j], ci[i][j]);
            cr[i][j] = mm(nai[i], bi[j], cr[i][j]);
            ci[i][j] = mm(ai[i], br[j], ci[i][j]);
          }
      }
    }
    if (DB && more) store_tiles(buf ^ 1);
    __syncthreads();
  }

  if constexpr (G3) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cr[i][j][r] += c3[i][j][r];
          ci[i][j][r] += c3[i][j][r];
        }
  }
  double red = 0.0;
  if constexpr (SWP) {
    // transposed tile: item r of lane l is row lr, column 4 lk + r of the 16 x 16 tile
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int row = m0 + wm * WTM + i * 16 + lr;
        const int col0 = n0 + wn * WTN + j * 16 + 4 * lk;
        double re[4], im[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          re[r] = (double)cr[i][j][r];
          im[r] = (double)ci[i][j][r];
        }
        const int ncol = row < g.M ? max(0, min(4, g.N - col0)) : 0;
        epi.row4(blockIdx.z, row < g.M ? row : -1, col0, ncol, re, im, red);
      }
  } else if constexpr (HasBlock4<Epi>::value) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = n0 + wn * WTN + j * 16 + lr;
        int rows[4];
        double re[4], im[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WTM + i * 16 + Mfma<T>::row(lk, r);
          rows[r] = (row < g.M && col < g.N) ? row : -1;
          re[r] = (double)cr[i][j][r];
          im[r] = (double)ci[i][j][r];
        }
        epi.block(blockIdx.z, rows, col, re, im, red);
      }
  } else {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WTM + i * 16 + Mfma<T>::row(lk, r);
          const int col = n0 + wn * WTN + j * 16 + lr;
          if (row < g.M && col < g.N) {
            if constexpr (RED)
              epi(blockIdx.z, row, col, (double)cr[i][j][r], (double)ci[i][j][r], red);
            else
              epi(blockIdx.z, row, col, (double)cr[i][j][r], (double)ci[i][j][r]);
          }
        }
  }
  if constexpr (RED) {
    // fixed-order reduction: butterfly within each wave, then the four wave sums
    __shared__ double sred[kThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) red += __shfl_xor(red, off, 64);
    if (lane == 0) sred[wave] = red;
    __syncthreads();
    if (tid == 0) epi.finish(blin, (sred[0] + sred[1]) + (sred[2] + sred[3]));
  }
}

// Tile configuration of one GEMM: BM x BN output tile, BK K tile, WM waves along M,
// double-buffered LDS or not, and the fewest waves per SIMD the register allocation
// must allow (OCC).
template <int BM_, int BN_, int BK_, int WM_, bool DB_ = false, int OCC_ = 1>
struct Tile {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, OCC = OCC_;
  static constexpr bool DB = DB_;
};

template <typename T, typename RA, typename RB, bool TRANS_A, bool TRANS_B, typename TL,
          int RMODE = 0, typename Epi, typename KSeg = KFull>
inline hipError_t launch_zgemm(const GemmArgs<RA, RB>& g, int batch, const Epi& epi,
                               hipStream_t s, const KSeg& ks = KSeg()) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  dim3 grid((g.N + TL::BN - 1) / TL::BN, (g.M + TL::BM - 1) / TL::BM, batch * g.splitk);
  hipLaunchKernelGGL((zgemm_kernel<T, RA, RB, TRANS_A, TRANS_B, TL::BM, TL::BN, TL::BK, TL::WM,
                                   TL::DB, TL::OCC, RMODE, Epi, KSeg>),
                     grid, dim3(kThreads), 0, s, g, epi, ks);
  return hipGetLastError();
}

// Candidate tile configurations of one GEMM; the engine's autotuner picks one by
// index at run time (launch_sel).
template <typename... TLs> struct TileList {};

// Output tile (BM, BN) of candidate idx of a list.
template <typename TL0>
inline void tile_dims(int, TileList<TL0>, int& bm, int& bn) { bm = TL0::BM; bn = TL0::BN; }
template <typename TL0, typename TL1, typename... TLs>
inline void tile_dims(int idx, TileList<TL0, TL1, TLs...>, int& bm, int& bn) {
  if (idx <= 0) { bm = TL0::BM; bn = TL0::BN; return; }
  tile_dims(idx - 1, TileList<TL1, TLs...>(), bm, bn);
}

template <typename T, typename RA, typename RB, bool TRANS_A, bool TRANS_B, int RMODE,
          typename Epi, typename KSeg, typename TL0>
inline hipError_t launch_sel(int, const GemmArgs<RA, RB>& g, int batch, const Epi& epi,
                             hipStream_t s, const KSeg& ks, TileList<TL0>) {
  return launch_zgemm<T, RA, RB, TRANS_A, TRANS_B, TL0, RMODE>(g, batch, epi, s, ks);
}
template <typename T, typename RA, typename RB, bool TRANS_A, bool TRANS_B, int RMODE,
          typename Epi, typename KSeg, typename TL0, typename TL1, typename... TLs>
inline hipError_t launch_sel(int idx, const GemmArgs<RA, RB>& g, int batch, const Epi& epi,
                             hipStream_t s, const KSeg& ks, TileList<TL0, TL1, TLs...>) {
  if (idx <= 0) return launch_zgemm<T, RA, RB, TRANS_A, TRANS_B, TL0, RMODE>(g, batch, epi, s, ks);
  return launch_sel<T, RA, RB, TRANS_A, TRANS_B, RMODE>(idx - 1, g, batch, epi, s, ks,
                                                        TileList<TL1, TLs...>());
}

}  // namespace zh
