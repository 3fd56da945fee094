// CPU emulation of k_xi_stream: the kernel's own address arithmetic (zh_omgxi.hpp),
// DMA into a byte LDS image poisoned with NaN before every chunk, the MFMA operand reads
// from it, accumulation in double; bounds of every global and LDS access checked, the
// result compared with a direct product over the lead's modes.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "zh_omgxi.hpp"

using namespace zh;
typedef std::complex<double> cd;

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { if (fails++ < 10) { printf("FAIL %s:%d ", __FILE__, __LINE__); printf(__VA_ARGS__); printf("\n"); } } } while (0)

template <typename C, int NG>
void run(int Md, int ML, std::vector<int> na_act, std::vector<int> na_pas, std::vector<int> nc_lead,
         int tri, int ff, int parts, int NK, unsigned seed) {
  typedef XiGeo<C, NG> G;
  typedef decltype(C::x) T;
  const int NL = (int)na_act.size();
  // device modes: active block lead by lead, padding to ML; passive block lead by lead
  std::vector<int> lead(Md, -1);
  std::vector<int4> lr(NL);
  int d = 0;
  for (int a = 0; a < NL; ++a) { lr[a].x = d; for (int i = 0; i < na_act[a]; ++i) lead[d++] = a; lr[a].y = d; }
  if (d > ML) { printf("bad ML\n"); return; }
  d = ML;
  for (int a = 0; a < NL; ++a) { lr[a].z = d; for (int i = 0; i < na_pas[a]; ++i) lead[d++] = a; lr[a].w = d; }
  if (d != Md) { printf("bad Md %d %d\n", d, Md); return; }
  for (int a = 0; a < NL; ++a) if (lr[a].x == lr[a].y) lr[a].x = lr[a].y = 0;
  for (int a = 0; a < NL; ++a) if (lr[a].z == lr[a].w) lr[a].z = lr[a].w = 0;
  // coupled columns: per lead nc_lead[a] real columns padded to 16
  std::vector<int> clead;
  for (int a = 0; a < NL; ++a) {
    for (int i = 0; i < nc_lead[a]; ++i) clead.push_back(a);
    while (clead.size() % 16) clead.push_back(-1);
  }
  const int nc = (int)clead.size();
  std::mt19937 rng(seed);
  std::normal_distribution<double> nd;
  const double nan = std::nan("");
  std::vector<C> X((size_t)NK * Md * Md), xi((size_t)NK * Md * nc);
  std::vector<cd> Xf((size_t)NK * Md * Md);
  for (int b = 0; b < NK; ++b)
    for (int m = 0; m < Md; ++m)
      for (int k = 0; k < Md; ++k) {
        C v; v.x = (T)nd(rng); v.y = (T)nd(rng);
        X[((size_t)b * Md + m) * Md + k] = v;
      }
  for (int b = 0; b < NK; ++b)
    for (int m = 0; m < Md; ++m)
      for (int k = 0; k < Md; ++k) {
        const C s = X[((size_t)b * Md + m) * Md + k];
        cd f(s.x, s.y);
        if (tri && k / 64 < m / 64) {
          const C t = X[((size_t)b * Md + k) * Md + m];
          f = -std::conj(cd(t.x, t.y));
        }
        if (ff && m >= ML && k >= ML) f = 0.0;
        Xf[((size_t)b * Md + m) * Md + k] = f;
      }
  // poison what the kernel must never read
  for (int b = 0; b < NK; ++b)
    for (int m = 0; m < Md; ++m)
      for (int k = 0; k < Md; ++k) {
        C& s = X[((size_t)b * Md + m) * Md + k];
        if ((tri && k / 64 < m / 64) || (ff && m >= ML && k >= ML)) { s.x = (T)nan; s.y = (T)nan; }
      }
  for (int b = 0; b < NK; ++b)
    for (int k = 0; k < Md; ++k)
      for (int c = 0; c < nc; ++c) {
        C v; v.x = 0; v.y = 0;
        if (lead[k] >= 0 && lead[k] == clead[c]) { v.x = (T)nd(rng); v.y = (T)nd(rng); }
        xi[((size_t)b * Md + k) * nc + c] = v;
      }
  // column items as the host builds them
  XiArgs<C> a{};
  a.X = X.data(); a.sX = (long)Md * Md; a.xi = xi.data(); a.sXi = (long)Md * nc;
  a.sPart = (long)Md * nc; a.Md = Md; a.nc = nc; a.parts = parts; a.tri = tri; a.ML = ML; a.ff = ff;
  int ng = nc / 16, g = 0;
  a.nitem = 0;
  while (g < ng) {
    const int l = clead[16 * g];
    int e = g + 1;
    while (e < ng && e - g < NG && clead[16 * e] == l) ++e;
    a.item[a.nitem++] = make_int4(16 * g, 16 * (e - g), l, 0);
    g = e;
  }
  for (int l = 0; l < NL; ++l) a.lr[l] = lr[l];
  std::vector<cd> part((size_t)NK * parts * Md * nc, cd(nan, nan));
  std::vector<unsigned char> lds(G::STAGE);
  long nchunks = 0;
  const int ntile = (Md + 63) / 64;
  for (int b = 0; b < NK; ++b)
    for (int ii = 0; ii < a.nitem; ++ii)
      for (int tx = 0; tx < ntile; ++tx)
        for (int p = 0; p < parts; ++p) {
          const int m0 = tx * 64;
          const int4 it = a.item[ii];
          const int c0 = it.x, ncol = std::min(it.y, nc - it.x);
          const int4 l4 = it.z >= 0 ? a.lr[it.z] : make_int4(0, 0, 0, 0);
          const XiPieces P = xi_pieces<G::KC>(m0, l4, it.z, tri, ff, ML);
          const int nch = xi_nchunks(P);
          const int q0 = (int)((long)p * nch / parts);
          const int nq = (int)((long)(p + 1) * nch / parts) - q0;
          std::vector<cd> acc((size_t)64 * G::NCOL, 0.0);
          for (int q = 0; q < nq; ++q) {
            int kc0, mir;
            xi_chunk<G::KC>(P, q0 + q, kc0, mir);
            CHECK(kc0 % G::KC == 0 && kc0 >= 0 && kc0 < Md, "kc0 %d", kc0);
            // DMA: poison, then every wave's loads
            std::vector<float> f(G::STAGE / 4, std::nanf(""));
            std::memcpy(lds.data(), f.data(), G::STAGE);
            std::vector<int> written(G::STAGE / 16, 0);
            for (int w = 0; w < 4; ++w) {
              for (int j = 0; j < 2; ++j) {
                const int i = w + 4 * j;
                for (int l = 0; l < 64; ++l) {
                  const long off = xi_src_a<C, NG>(i, l, kc0, m0, Md, mir);
                  const int dst = 1024 * i + 16 * l;
                  CHECK(dst + 16 <= G::A_BYTES, "A dst");
                  written[dst / 16]++;
                  if (off < 0) { std::memset(&lds[dst], 0, 16); continue; }
                  CHECK(off >= 0 && (off + 16 / G::E) <= (long)Md * Md, "A src %ld", off);
                  std::memcpy(&lds[dst], &X[(size_t)b * Md * Md + off], 16);
                }
              }
              for (int j = 0; j < G::NB; ++j) {
                const int i = w + 4 * j;
                for (int l = 0; l < 64; ++l) {
                  const long off = xi_src_b<C, NG>(i, l, kc0, c0, ncol, nc, Md);
                  const int dst = G::A_BYTES + 1024 * i + 16 * l;
                  CHECK(dst + 16 <= G::STAGE, "B dst");
                  written[dst / 16]++;
                  if (off < 0) { std::memset(&lds[dst], 0, 16); continue; }
                  CHECK(off >= 0 && (off + 16 / G::E) <= (long)Md * nc, "B src %ld", off);
                  CHECK((off % nc) + 16 / G::E <= nc, "B row overrun");
                  std::memcpy(&lds[dst], &xi[(size_t)b * Md * nc + off], 16);
                }
              }
            }
            for (size_t s = 0; s < written.size(); ++s) CHECK(written[s] == 1, "slot %zu written %d", s, written[s]);
            // MFMA operand reads
            for (int w = 0; w < 4; ++w)
              for (int s = 0; s < G::KC / 4; ++s)
                for (int l = 0; l < 64; ++l) {
                  const int kl = 4 * s + (l >> 4), rr = 16 * w + (l & 15);
                  const int oa = xi_lds_a<C, NG>(rr, kl, mir);
                  CHECK(oa >= 0 && oa + G::E <= G::A_BYTES, "A read");
                  C av; std::memcpy(&av, &lds[oa], sizeof(C));
                  const cd A = mir ? -std::conj(cd(av.x, av.y)) : cd(av.x, av.y);
                  // the lane's A value is row rr, mode kl; it meets B[kl][col] for the
                  // 16 columns of each tile (MFMA: rows of A x columns of B)
                  for (int t = 0; t < NG; ++t)
                    for (int cc = 0; cc < 16; ++cc) {
                      const int ob = xi_lds_b<C, NG>(kl, 16 * t + cc);
                      CHECK(ob >= 0 && ob + G::E <= G::B_BYTES, "B read");
                      C bv; std::memcpy(&bv, &lds[G::A_BYTES + ob], sizeof(C));
                      const cd B(bv.x, bv.y);
                      if (B == cd(0.0) && !std::isfinite(A.real() + A.imag())) {
                        // a NaN times zero is still NaN on the device
                        acc[(size_t)rr * G::NCOL + 16 * t + cc] += cd(nan, nan);
                      } else {
                        acc[(size_t)rr * G::NCOL + 16 * t + cc] += A * B;
                      }
                    }
                }
            ++nchunks;
          }
          for (int r = 0; r < 64; ++r)
            for (int c = 0; c < ncol; ++c) {
              const int row = m0 + r;
              if (row >= Md) continue;
              part[(((size_t)b * parts + p) * Md + row) * nc + c0 + c] = acc[(size_t)r * G::NCOL + c];
            }
        }
  // compare the summed partials with the direct product
  double maxerr = 0, maxref = 0;
  long bad = 0;
  for (int b = 0; b < NK; ++b)
    for (int m = 0; m < Md; ++m)
      for (int c = 0; c < nc; ++c) {
        cd ref = 0.0;
        for (int k = 0; k < Md; ++k) ref += Xf[((size_t)b * Md + m) * Md + k] * cd(xi[((size_t)b * Md + k) * nc + c].x, xi[((size_t)b * Md + k) * nc + c].y);
        cd got = 0.0;
        for (int p = 0; p < parts; ++p) got += part[(((size_t)b * parts + p) * Md + m) * nc + c];
        if (clead[c] < 0 && !(std::isfinite(got.real()))) { ++bad; continue; }
        const double e = std::abs(got - ref);
        if (!std::isfinite(e)) ++bad;
        maxerr = std::max(maxerr, e);
        maxref = std::max(maxref, std::abs(ref));
      }
  printf("%-7s NG=%d Md=%d ML=%d nc=%d tri=%d ff=%d parts=%d: chunks %ld, max err %.2e of %.2e, non-finite %ld %s\n",
         sizeof(T) == 8 ? "c128" : "c64", NG, Md, ML, nc, tri, ff, parts, nchunks, maxerr, maxref, bad,
         (bad == 0 && maxerr <= (sizeof(T) == 8 ? 1e-12 : 1e-4) * maxref) ? "ok" : "FAIL");
  if (!(bad == 0 && maxerr <= (sizeof(T) == 8 ? 1e-12 : 1e-4) * maxref)) ++fails;
}

int main() {
  // C3-like shape scaled down: 2 leads, active block 64-aligned, passive block, 3 groups per lead
  run<float2, 3>(448, 256, {100, 100}, {96, 96}, {40, 40}, 1, 1, 4, 2, 1);
  run<double2, 3>(448, 256, {100, 100}, {96, 96}, {40, 40}, 1, 1, 4, 2, 2);
  run<float2, 3>(448, 256, {100, 100}, {96, 96}, {40, 40}, 0, 1, 4, 1, 3);
  run<double2, 3>(448, 256, {100, 100}, {96, 96}, {40, 40}, 0, 0, 3, 1, 4);
  // odd lead sizes, Md not a multiple of 64, few columns
  run<float2, 1>(330, 192, {61, 67}, {70, 68}, {5, 9}, 1, 1, 4, 2, 5);
  run<double2, 1>(330, 192, {61, 67}, {70, 68}, {5, 9}, 1, 1, 2, 1, 6);
  run<float2, 2>(330, 192, {61, 67}, {70, 68}, {20, 30}, 1, 0, 4, 1, 7);
  run<double2, 2>(330, 192, {61, 67}, {70, 68}, {20, 30}, 0, 1, 4, 1, 8);
  run<float2, 4>(330, 192, {61, 67}, {70, 68}, {64, 50}, 1, 1, 4, 1, 9);
  run<double2, 4>(330, 192, {61, 67}, {70, 68}, {64, 50}, 1, 1, 5, 1, 10);
  // one lead, no passive block (reference order)
  run<float2, 2>(200, 200, {200}, {0}, {24}, 0, 0, 4, 1, 11);
  run<double2, 2>(200, 200, {200}, {0}, {24}, 1, 0, 4, 1, 12);
  // three leads
  run<float2, 3>(520, 320, {100, 90, 110}, {70, 60, 70}, {33, 17, 40}, 1, 1, 4, 1, 13);
  printf(fails ? "FAILURES: %d\n" : "all ok\n", fails);
  return fails ? 1 : 0;
}
