// X ξ_c, the Ω ξ term of the Ψ equation (Zandpack/mpi/zand:771-778), as a streaming
// MFMA kernel: X is an Ω-shaped matrix (an Ω slot, Md x Md per k block), ξ_c its
// coupling vectors on the coupled orbitals (Md x nc).  The product has few columns
// (nc, a lead's orbitals per column tile) against many rows and modes, so X is read once
// and every X element meets only the 16-48 columns of its lead: the kernel has to stream
// X from HBM at the rate its matrix-core work consumes it.
//
// A workgroup (4 waves) owns 64 output rows and the columns of one lead (up to NG
// groups of 16) and walks K chunks of the lead's modes (KC = 128 bytes of a row: 16
// complex64 or 8 complex128 modes).  Every chunk of X (64 rows x KC) and of ξ_c
// (KC x 16 NG) goes from global memory straight into LDS with global_load_lds_dwordx4
// (no registers on the way), into a ring of STAGES chunk buffers: the loads of chunk
// q + STAGES - 1 are in flight while chunk q is multiplied, one barrier per chunk.
// The loads are written as inline asm with explicit vmcnt waits: issued through the
// compiler builtin, every LDS read after them waits for all of them.
//
// X may be held as its block upper triangle (anti-Hermitian: X_ij = -conj X_ji): the
// modes below the row tile's first row are then read from the stored column strip above
// the diagonal block, mirrored (op(X)(m, k) = -conj X[k][m]), into a transposed LDS
// layout.  The K range of a column tile is its lead's modes (active block, then passive
// block; with ff the passive rows see only the active block), cut into chunks; the
// chunks of a tile are split over `parts` partial products (blockIdx.z), summed by the
// consumer in a fixed order.  A chunk may reach past its lead's modes to the chunk
// boundary: the modes there belong to no lead that couples the tile's columns, so ξ_c
// is zero on them, and they lie in X's computed part (the active block and the
// diagonal split are 64-aligned), so X is finite there.
//
// LDS layouts (16-byte slots, XOR-swizzled so that the MFMA operand reads of a wave,
// 16 rows at one mode, are free of bank conflicts):
//   A direct   : row r (64) x 128 bytes, slot ^ ((r >> 1) & 7)
//   A mirrored : mode k (KC) x 64 rows, slot ^ ((k & 1) << 3)
//   B          : mode k (KC) x 16 NG columns (slot ^ ((k & 1) << 3) for complex64 with
//                NG even, whose row stride would otherwise map k and k+1 to one bank)
// The address arithmetic below is __host__ __device__, shared with the CPU emulation
// that checks bounds and results before the kernel runs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zh_gemm.hpp"

namespace zh {

constexpr int kXiMaxItems = 16;
constexpr int kXiMaxLeads = 8;
constexpr int kXiRowBytes = 128;  // bytes of one X row per chunk

template <typename C> struct XiArgs {
  const C* X; long sX;    // X[b], Md x Md row-major
  const C* xi; long sXi;  // ξ_c[b], Md x nc row-major
  C* part; long sPart;    // partial product p of batch b: part + (b * parts + p) * sPart
  int Md, nc, parts, tri, ML, ff, nitem;
  int4 item[kXiMaxItems];  // column tile: (first column, columns, lead or -1 for none, 0)
  int4 lr[kXiMaxLeads];    // per lead: device modes [x, y) (active block), [z, w) (passive)
};

template <typename C, int NG> struct XiGeo {
  static constexpr int E = (int)sizeof(C);
  static constexpr int KC = kXiRowBytes / E;     // modes per chunk
  static constexpr int NCOL = 16 * NG;
  static constexpr int A_BYTES = 64 * kXiRowBytes;
  static constexpr int B_BYTES = KC * NCOL * E;  // 2 KB per column group
  static constexpr int NB = (B_BYTES / 1024 + 3) / 4;  // B loads per wave per chunk
  static constexpr int B_PAD = NB * 4 * 1024;
  static constexpr int STAGE = A_BYTES + B_PAD;
  static constexpr int NL = 2 + NB;              // loads per wave per chunk
  static constexpr bool SWB = (E == 8) && (NG % 2 == 0);
};

// K pieces of one output tile as chunk ranges [cb, ce), in a fixed order: the lead's
// active modes below the tile's first row (mirrored, with the triangle), from it on,
// then its passive modes below and from it on; an empty piece has cb = ce.  Fixed
// positions keep the ranges in registers (indexing them at run time spilled them).
struct XiPieces {
  int cb[4], ce[4];
};
__host__ __device__ constexpr int xi_piece_mirrored(int i) { return (i & 1) == 0; }

template <int KC>
__host__ __device__ inline XiPieces xi_pieces(int m0, int4 r, int lead, int tri, int ff, int ML) {
  XiPieces P;
  const bool act_only = ff && m0 >= ML;
  const int lo[2] = {r.x, r.z};
  const int hi[2] = {lead < 0 ? r.x : r.y, (lead < 0 || act_only) ? r.z : r.w};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const bool any = hi[s] > lo[s];
    // below m0 (mirrored; none without the triangle)
    const int bh = tri ? (hi[s] < m0 ? hi[s] : m0) : lo[s];
    const bool b = any && bh > lo[s];
    P.cb[2 * s] = b ? lo[s] / KC : 0;
    P.ce[2 * s] = b ? (bh + KC - 1) / KC : 0;
    // from m0 on (everything without the triangle)
    const int al = tri ? (lo[s] > m0 ? lo[s] : m0) : lo[s];
    const bool d = any && hi[s] > al;
    P.cb[2 * s + 1] = d ? al / KC : 0;
    P.ce[2 * s + 1] = d ? (hi[s] + KC - 1) / KC : 0;
  }
  return P;
}

__host__ __device__ inline int xi_nchunks(const XiPieces& P) {
  return (P.ce[0] - P.cb[0]) + (P.ce[1] - P.cb[1]) + (P.ce[2] - P.cb[2]) + (P.ce[3] - P.cb[3]);
}

// Chunk q of the tile (0 <= q < xi_nchunks): first mode and mirrored flag.
template <int KC>
__host__ __device__ inline void xi_chunk(const XiPieces& P, int q, int& kc0, int& mir) {
  kc0 = 0;
  mir = 0;
  bool found = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int len = P.ce[i] - P.cb[i];
    const bool here = !found && q < len;
    kc0 = here ? (P.cb[i] + q) * KC : kc0;
    mir = here ? xi_piece_mirrored(i) : mir;
    found = found || here;
    q -= found ? 0 : len;
  }
}

// Per-lane constants of one load: load i (0..7 of the A chunk, 0 .. 4 NB - 1 of the ξ_c
// chunk) and lane l write the 16-byte LDS slot 1024 i + 16 l; the slot holds, after the
// swizzle, A row dr at chunk mode dk (direct), mode mk at tile row mm (mirrored), or
// ξ_c chunk mode bk at tile column bc (bc = -1: padding of the LDS region).
struct XiSlot {
  int dr, dk, mk, mm, bk, bc;
};
template <typename C, int NG>
__host__ __device__ inline XiSlot xi_slot(int i, int l) {
  typedef XiGeo<C, NG> G;
  XiSlot s;
  const int P = 1024 * i + 16 * l;
  s.dr = P >> 7;
  s.dk = (((P >> 4) & 7) ^ ((s.dr >> 1) & 7)) * (16 / G::E);
  const int rowa = 64 * G::E;
  s.mk = P / rowa;
  s.mm = (((P % rowa) >> 4) ^ ((s.mk & 1) << 3)) * (16 / G::E);
  const int rowb = G::NCOL * G::E;
  s.bk = P / rowb;
  const int ps = (P % rowb) >> 4;
  s.bc = P < G::B_BYTES ? (G::SWB ? ps ^ ((s.bk & 1) << 3) : ps) * (16 / G::E) : -1;
  return s;
}

// Source element of an A load for the chunk at mode kc0: offset into X[b], or -1
// (outside the matrix: the lane reads zeros).  Offsets fit in int (Md^2 < 2^31).
__host__ __device__ inline int xi_off_a(const XiSlot& s, int kc0, int m0, int Md, int mir) {
  const int r = mir ? kc0 + s.mk : m0 + s.dr;
  const int c = mir ? m0 + s.mm : kc0 + s.dk;
  return (r < Md && c < Md) ? r * Md + c : -1;
}
// Source element of a ξ_c load: offset into ξ_c[b] or -1 (padding, outside the matrix
// or the tile's columns).
__host__ __device__ inline int xi_off_b(const XiSlot& s, int kc0, int c0, int ncol, int nc,
                                        int Md) {
  const int k = kc0 + s.bk;
  return (s.bc >= 0 && s.bc < ncol && k < Md) ? k * nc + c0 + s.bc : -1;
}

template <typename C, int NG>
__host__ __device__ inline long xi_src_a(int i, int l, int kc0, int m0, int Md, int mir) {
  return xi_off_a(xi_slot<C, NG>(i, l), kc0, m0, Md, mir);
}
template <typename C, int NG>
__host__ __device__ inline long xi_src_b(int i, int l, int kc0, int c0, int ncol, int nc, int Md) {
  return xi_off_b(xi_slot<C, NG>(i, l), kc0, c0, ncol, nc, Md);
}

// LDS byte offsets of the MFMA operand reads: A row r (0..63) at chunk mode k, B at
// chunk mode k and tile column col.
template <typename C, int NG>
__host__ __device__ inline int xi_lds_a(int r, int k, int mir) {
  typedef XiGeo<C, NG> G;
  if (!mir) {
    const int b = k * G::E;
    return r * kXiRowBytes + ((((b >> 4) ^ ((r >> 1) & 7))) << 4) + (b & 15);
  }
  const int b = r * G::E;
  return k * 64 * G::E + (((b >> 4) ^ ((k & 1) << 3)) << 4) + (b & 15);
}

template <typename C, int NG>
__host__ __device__ inline int xi_lds_b(int k, int col) {
  typedef XiGeo<C, NG> G;
  const int b = col * G::E;
  const int s = G::SWB ? ((b >> 4) ^ ((k & 1) << 3)) : (b >> 4);
  return k * G::NCOL * G::E + (s << 4) + (b & 15);
}

// Zeros for the lanes whose source lies outside the matrices.
static __device__ __attribute__((aligned(16))) unsigned int g_xi_zero[4];

__device__ inline void xi_dma16(const void* src, uint32_t lds) {
  asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(src), "{m0}"(lds) : "memory");
}
template <int N> __device__ inline void xi_vmwait() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
__device__ inline void xi_barrier() { asm volatile("s_barrier" ::: "memory"); }

template <typename C, int NG, int STAGES, int OCC>
__global__ __launch_bounds__(256, OCC) void k_xi_stream(XiArgs<C> a) {
  typedef XiGeo<C, NG> G;
  typedef decltype(C::x) T;
  typedef typename Mfma<T>::acc_t acc_t;
  static_assert(STAGES >= 2 && (STAGES - 2) * G::NL < 64, "pipeline depth");
  __shared__ __attribute__((aligned(1024))) unsigned char lds[STAGES * G::STAGE];

  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m0 = blockIdx.x * 64;
  const int4 it = a.item[blockIdx.y];
  const int b = blockIdx.z / a.parts, p = blockIdx.z - b * a.parts;
  const int lead = it.z;
  const int c0 = it.x, ncol = min(it.y, a.nc - it.x);
  const int4 lr = lead >= 0 ? a.lr[lead] : make_int4(0, 0, 0, 0);
  const XiPieces P = xi_pieces<G::KC>(m0, lr, lead, a.tri, a.ff, a.ML);
  const int nch = xi_nchunks(P);
  const int q0 = (int)((long)p * nch / a.parts);
  const int nq = (int)((long)(p + 1) * nch / a.parts) - q0;
  const C* __restrict__ Xb = a.X + (long)b * a.sX;
  const C* __restrict__ xb = a.xi + (long)b * a.sXi;
  const uint32_t lds0 =
      (uint32_t)(uintptr_t)(__attribute__((address_space(3))) unsigned char*)lds;

  XiSlot sa[2], sbv[G::NB];
#pragma unroll
  for (int j = 0; j < 2; ++j) sa[j] = xi_slot<C, NG>(w + 4 * j, lane);
#pragma unroll
  for (int j = 0; j < G::NB; ++j) sbv[j] = xi_slot<C, NG>(w + 4 * j, lane);
  const void* zero = (const void*)g_xi_zero;
  auto issue = [&](int q) {
    int kc0, mir;
    xi_chunk<G::KC>(P, q0 + q, kc0, mir);
    const uint32_t sb = lds0 + (uint32_t)((q % STAGES) * G::STAGE);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int off = xi_off_a(sa[j], kc0, m0, a.Md, mir);
      xi_dma16(off >= 0 ? (const void*)(Xb + off) : zero,
               __builtin_amdgcn_readfirstlane(sb + 1024 * (w + 4 * j)));
    }
#pragma unroll
    for (int j = 0; j < G::NB; ++j) {
      const int off = xi_off_b(sbv[j], kc0, c0, ncol, a.nc, a.Md);
      xi_dma16(off >= 0 ? (const void*)(xb + off) : zero,
               __builtin_amdgcn_readfirstlane(sb + G::A_BYTES + 1024 * (w + 4 * j)));
    }
  };

  constexpr bool M3 = sizeof(T) == 8;
  acc_t cr[NG], ci[NG], c3[M3 ? NG : 1];
#pragma unroll
  for (int t = 0; t < NG; ++t) {
    cr[t] = Mfma<T>::zero();
    ci[t] = Mfma<T>::zero();
    if constexpr (M3) c3[t] = Mfma<T>::zero();
  }
  const int rr = 16 * w + (lane & 15);
  // the operands of one chunk in registers: KC/4 mode steps x (1 + NG) values per lane,
  // the mirrored chunks' sign already applied
  constexpr int NS = G::KC / 4;
  struct Ops {
    C av[NS];
    C bv[NS][NG];
  };
  auto read = [&](int q, Ops& o) {
    int kc0, mir;
    xi_chunk<G::KC>(P, q0 + q, kc0, mir);
    const unsigned char* Ab = lds + (q % STAGES) * G::STAGE;
    const unsigned char* Bb = Ab + G::A_BYTES;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kl = 4 * s + (lane >> 4);
      if (mir) {
        o.av[s] = *(const C*)(Ab + xi_lds_a<C, NG>(rr, kl, 1));
        o.av[s].x = -o.av[s].x;
      } else {
        o.av[s] = *(const C*)(Ab + xi_lds_a<C, NG>(rr, kl, 0));
      }
#pragma unroll
      for (int t = 0; t < NG; ++t)
        o.bv[s][t] = *(const C*)(Bb + xi_lds_b<C, NG>(kl, 16 * t + (lane & 15)));
    }
  };
  // the same with the chunk's sign known at compile time
  auto read_sign = [&](int q, Ops& o, auto mirc) {
    constexpr int MIR = decltype(mirc)::value;
    const unsigned char* Ab = lds + (q % STAGES) * G::STAGE;
    const unsigned char* Bb = Ab + G::A_BYTES;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kl = 4 * s + (lane >> 4);
      o.av[s] = *(const C*)(Ab + xi_lds_a<C, NG>(rr, kl, MIR));
      if (MIR) o.av[s].x = -o.av[s].x;
#pragma unroll
      for (int t = 0; t < NG; ++t)
        o.bv[s][t] = *(const C*)(Bb + xi_lds_b<C, NG>(kl, 16 * t + (lane & 15)));
    }
  };
  // complex128: three real products per complex one,
  //   re = (-ai)(br + bi) + (ar + ai) br,  im = ar (bi - br) + (ar + ai) br;
  // complex64 keeps four (three were 12% faster at n = 1000, M = 4000, but their
  // larger FP32 rounding moves the mixed error norm past 1e-4 of the FP64 one)
  auto mma = [&](const Ops& o) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const T ar = o.av[s].x, ai = o.av[s].y;
      const T nai = -ai;
      if constexpr (M3) {
        const T sa = ar + ai;
#pragma unroll
        for (int t = 0; t < NG; ++t) {
          c3[t] = Mfma<T>::op(sa, o.bv[s][t].x, c3[t]);
          ci[t] = Mfma<T>::op(ar, o.bv[s][t].y - o.bv[s][t].x, ci[t]);
          cr[t] = Mfma<T>::op(nai, o.bv[s][t].x + o.bv[s][t].y, cr[t]);
        }
      } else {
#pragma unroll
        for (int t = 0; t < NG; ++t) {
          cr[t] = Mfma<T>::op(ar, o.bv[s][t].x, cr[t]);
          ci[t] = Mfma<T>::op(ar, o.bv[s][t].y, ci[t]);
          cr[t] = Mfma<T>::op(nai, o.bv[s][t].y, cr[t]);
          ci[t] = Mfma<T>::op(ai, o.bv[s][t].x, ci[t]);
        }
      }
    }
  };

  // chunk q: wait for it, barrier, read its operands, products, then the next loads
  // into chunk q - 1's buffer, issued behind this chunk's products
  for (int q = 0; q < STAGES - 1 && q < nq; ++q) issue(q);
  for (int q = 0; q < nq; ++q) {
    if (q + STAGES - 2 < nq) xi_vmwait<(STAGES - 2) * G::NL>();
    else xi_vmwait<0>();
    xi_barrier();  // chunk q is in LDS for every wave; chunk q - 1's buffer is free
    Ops o;
    if constexpr (M3) {
      read(q, o);
      mma(o);
    } else {
      // complex64: one copy of the reads and products per sign (at n = 1000, M = 4000:
      // 116 against 120 us; complex128 runs out of registers that way, 212 against 132)
      int kc0, mir;
      xi_chunk<G::KC>(P, q0 + q, kc0, mir);
      if (mir) {
        read_sign(q, o, std::integral_constant<int, 1>());
        mma(o);
      } else {
        read_sign(q, o, std::integral_constant<int, 0>());
        mma(o);
      }
    }
    if (q + STAGES - 1 < nq) issue(q + STAGES - 1);
  }

  C* __restrict__ out = a.part + (long)(b * a.parts + p) * a.sPart;
#pragma unroll
  for (int t = 0; t < NG; ++t) {
    const int col = 16 * t + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + 16 * w + Mfma<T>::row(lane >> 4, r);
      if (row < a.Md && col < ncol) {
        C v;
        v.x = M3 ? cr[t][r] + c3[M3 ? t : 0][r] : cr[t][r];
        v.y = M3 ? ci[t][r] + c3[M3 ? t : 0][r] : ci[t][r];
        out[(long)row * a.nc + c0 + col] = v;
      }
    }
  }
}

}  // namespace zh
