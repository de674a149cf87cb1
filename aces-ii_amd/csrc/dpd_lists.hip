// Symmetry-packed lists <-> C1 order (see dpd_lists.h).  One thread per list
// element: the element's group, row and column give the four orbitals, whose
// global indices give its C1 position; gathers and scatters touch each
// element once.
#include <string>
#include "common.h"
#include "dpd_lists.h"

namespace aces {

namespace {

constexpr int kIrr = 8;

struct Side {
  int sp, sq;   // spaces of the pair's members (0 O alpha, 1 V alpha, 2 O beta, 3 V beta)
  int pack;     // 0 full, 1 p < q, 2 p <= q, -1 single element (vector lists)
  long sub[kIrr][kIrr];  // [pair irrep][irrep of q]: sub-block offset (-1: no sub-block)
  long c1n;     // C1 pair count (rows or columns of the C1 order)
};

struct Conv {
  int nirrep;
  int irrepx;   // 0-based
  int cnt[4][kIrr], off[4][kIrr], tot[4];
  Side side[2];  // 0 rows (left pairs), 1 columns (right pairs)
  long nrow[kIrr], ncol[kIrr];  // per group (index = right pair irrep)
  long goff[kIrr + 1];          // flat element offset of each group
  double* grp[kIrr];
};

__device__ inline long tri(long n, int pack) { return pack == 1 ? n * (n - 1) / 2 : n * (n + 1) / 2; }

// Pair index t of pair irrep G on a side -> (irrep, local) of p and q.
__device__ inline void decode(const Conv& c, const Side& S, int G, long t, int& gp, int& lp, int& gq, int& lq) {
  gp = lp = gq = lq = 0;
  if (S.pack < 0) return;
  for (int h = 0; h < c.nirrep; ++h) {
    const long b0 = S.sub[G][h];
    if (b0 < 0) continue;
    const int hp = h ^ G;
    const long np = c.cnt[S.sp][hp], nq = c.cnt[S.sq][h];
    const bool square = S.pack == 0 || hp < h;
    const long sz = square ? np * nq : (hp == h ? tri(np, S.pack) : 0);
    if (t < b0 || t >= b0 + sz) continue;
    const long u = t - b0;
    gp = hp;
    gq = h;
    if (square) {
      lq = (int)(u / np);
      lp = (int)(u - (long)lq * np);
    } else {
      // u = q(q-1)/2 + p (p < q) or q(q+1)/2 + p (p <= q)
      const int shift = S.pack == 1 ? 0 : 1;
      long q = (long)((sqrt(8.0 * (double)u + 1.0) + 1.0) * 0.5) - shift;
      while (q > 0 && (q + shift) * (q + shift - 1) / 2 > u) --q;
      while ((q + 1 + shift) * (q + shift) / 2 <= u) ++q;
      lq = (int)q;
      lp = (int)(u - (q + shift) * (q + shift - 1) / 2);
    }
    return;
  }
}

__device__ inline long c1_index(const Conv& c, const Side& S, int gp, int lp, int gq, int lq) {
  if (S.pack < 0) return 0;
  const long p = c.off[S.sp][gp] + lp, q = c.off[S.sq][gq] + lq;
  if (S.pack == 0) return p + (long)c.tot[S.sp] * q;
  return S.pack == 1 ? q * (q - 1) / 2 + p : q * (q + 1) / 2 + p;
}

// to_c1: c1[...] = group element; else group element = c1[...].
__global__ void dpd_c1_kernel(Conv c, double* __restrict__ c1, int to_c1) {
  const long total = c.goff[c.nirrep];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    int g = 0;
    while (g + 1 < c.nirrep && e >= c.goff[g + 1]) ++g;
    const long x = e - c.goff[g];
    const long nr = c.nrow[g];
    const long col = x / nr, row = x - col * nr;
    const int Gr = g, Gl = g ^ c.irrepx;
    int gp, lp, gq, lq, gr, lr, gs, ls;
    decode(c, c.side[0], Gl, row, gp, lp, gq, lq);
    decode(c, c.side[1], Gr, col, gr, lr, gs, ls);
    const long i = c1_index(c, c.side[0], gp, lp, gq, lq) + c.side[0].c1n * c1_index(c, c.side[1], gr, lr, gs, ls);
    if (to_c1)
      c1[i] = c.grp[g][x];
    else
      c.grp[g][x] = c1[i];
  }
}

int side_of(const DpdSym& sym, int type, Side& S) {
  for (int a = 0; a < kIrr; ++a)
    for (int b = 0; b < kIrr; ++b) S.sub[a][b] = -1;
  if (type == 0) {  // single element
    S.sp = S.sq = 0;
    S.pack = -1;
    S.c1n = 1;
    return ACES_OK;
  }
  ACES_REQUIRE(pair_type(type, &S.sp, &S.sq, &S.pack), "DPD lists: unknown pair type " + std::to_string(type));
  for (int G = 0; G < sym.nirrep; ++G)
    for (int h = 0; h < sym.nirrep; ++h) {
      const int hp = h ^ G;
      if (S.pack != 0 && hp > h) continue;
      S.sub[G][h] = sym.isymoff(h, G, type);
    }
  S.c1n = c1_pairs(sym, type);
  return ACES_OK;
}

// vgroup > 0: the single block of group vgroup of a vector list (tl = 9, tr = 0,
// totally symmetric) instead of one group per irrep.
int build_conv(const ListStore& L, int list, int tl, int tr, int irrepx, Conv& c, int vgroup = 0) {
  const DpdSym& sym = L.sym();
  ACES_REQUIRE(irrepx >= 1 && irrepx <= sym.nirrep, "DPD lists: irrepx out of range");
  c.nirrep = sym.nirrep;
  c.irrepx = irrepx - 1;
  for (int sp = 0; sp < 4; ++sp) {
    int acc = 0;
    for (int h = 0; h < kIrr; ++h) {
      const int n = h < sym.nirrep ? ((sp & 1) ? sym.vrt[sp >> 1][h] : sym.pop[sp >> 1][h]) : 0;
      c.cnt[sp][h] = n;
      c.off[sp][h] = acc;
      acc += n;
    }
    c.tot[sp] = acc;
  }
  ACES_CHECK(side_of(sym, tl, c.side[0]));
  ACES_CHECK(side_of(sym, tr, c.side[1]));
  c.goff[0] = 0;
  for (int g = 0; g < kIrr; ++g) {
    c.grp[g] = nullptr;
    c.nrow[g] = c.ncol[g] = 0;
  }
  if (vgroup > 0) {
    const long want = sym.irpdpd(0, tl);
    long nr = 0, nc = 0;
    ACES_REQUIRE(L.rows(vgroup, list, &nr) == ACES_OK && L.cols(vgroup, list, &nc) == ACES_OK && nr * nc == want,
                 "DPD lists: list " + std::to_string(list) + " group " + std::to_string(vgroup) + " must hold " +
                     std::to_string(want) + " elements");
    c.nrow[0] = want;
    c.ncol[0] = 1;
    c.grp[0] = want > 0 ? L.device_ptr(vgroup, list) : nullptr;
    for (int g = 1; g <= kIrr; ++g) c.goff[g] = want;
    return ACES_OK;
  }
  for (int g = 0; g < sym.nirrep; ++g) {
    const int Gl = g ^ c.irrepx;
    const long want_r = sym.irpdpd(Gl, tl), want_c = tr == 0 ? (g == c.irrepx ? 1 : 0) : sym.irpdpd(g, tr);
    long nr = 0, nc = 0;
    const bool have = L.rows(g + 1, list, &nr) == ACES_OK && L.cols(g + 1, list, &nc) == ACES_OK;
    if (!have) {
      ACES_REQUIRE(want_r * want_c == 0, "DPD lists: list " + std::to_string(list) + " group " + std::to_string(g + 1) +
                                             " is missing");
      nr = nc = 0;
    } else {
      ACES_REQUIRE(nr == want_r && nc == want_c,
                   "DPD lists: list " + std::to_string(list) + " group " + std::to_string(g + 1) + " is " +
                       std::to_string(nr) + " x " + std::to_string(nc) + ", the pair types give " +
                       std::to_string(want_r) + " x " + std::to_string(want_c));
    }
    c.nrow[g] = nr;
    c.ncol[g] = nc;
    c.grp[g] = (nr * nc > 0) ? L.device_ptr(g + 1, list) : nullptr;
    c.goff[g + 1] = c.goff[g] + nr * nc;
  }
  for (int g = sym.nirrep + 1; g <= kIrr; ++g) c.goff[g] = c.goff[sym.nirrep];
  return ACES_OK;
}

int run_conv(const Conv& c, double* c1, int to_c1, hipStream_t s) {
  const long total = c.goff[c.nirrep];
  if (total == 0) return ACES_OK;
  long g = (total + 255) / 256;
  if (g > (1L << 20)) g = 1L << 20;
  hipLaunchKernelGGL(dpd_c1_kernel, dim3((unsigned)g), dim3(256), 0, s, c, c1, to_c1);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

}  // namespace

long c1_pairs(const DpdSym& sym, int type) {
  if (type == 0) return 1;
  int sp, sq, pk;
  if (!pair_type(type, &sp, &sq, &pk)) return -1;
  long np = 0, nq = 0;
  for (int h = 0; h < sym.nirrep; ++h) {
    np += (sp & 1) ? sym.vrt[sp >> 1][h] : sym.pop[sp >> 1][h];
    nq += (sq & 1) ? sym.vrt[sq >> 1][h] : sym.pop[sq >> 1][h];
  }
  if (pk == 0) return np * nq;
  return pk == 1 ? np * (np - 1) / 2 : np * (np + 1) / 2;
}

int dpd_to_c1(const ListStore& L, int list, int tl, int tr, int irrepx, double* dst, hipStream_t s) {
  Conv c;
  ACES_CHECK(build_conv(L, list, tl, tr, irrepx, c));
  const long n = c.side[0].c1n * c.side[1].c1n;
  ACES_HIP(hipMemsetAsync(dst, 0, sizeof(double) * n, s));
  return run_conv(c, dst, 1, s);
}

int c1_to_dpd(ListStore& L, int list, int tl, int tr, int irrepx, const double* src, hipStream_t s) {
  Conv c;
  ACES_CHECK(build_conv(L, list, tl, tr, irrepx, c));
  return run_conv(c, const_cast<double*>(src), 0, s);
}

int t1_to_dense(const ListStore& L, int group, int list, double* t1, hipStream_t s) {
  Conv c;
  ACES_CHECK(build_conv(L, list, 9, 0, 1, c, group));
  ACES_HIP(hipMemsetAsync(t1, 0, sizeof(double) * c.side[0].c1n, s));
  return run_conv(c, t1, 1, s);
}

int dense_to_t1(ListStore& L, int group, int list, const double* t1, hipStream_t s) {
  Conv c;
  ACES_CHECK(build_conv(L, list, 9, 0, 1, c, group));
  return run_conv(c, const_cast<double*>(t1), 0, s);
}

}  // namespace aces
