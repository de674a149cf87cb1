// The integral and denominator lists of the CC path from density-fitted (or
// Cholesky) factors on the GPU, in the role of xintprc (intprc/intproc.F:
// the <..|..> lists; list 233 at intprc/intproc.F:240-266; the denominators at
// intprc/dijab.F:178-185, 300-375).
//
// (pq|rs) = sum_P B[P][p][q] B[P][r][s] with the orbitals in list order
// (occupied then virtual, irrep-sorted in each).  Every integral block is one
// fp64 MFMA product over P of two slices of B; list 233 (rows A<=B, columns
// (C,D), C fastest) is built one column block d at a time,
//   R_d[b][(a,c)] = sum_P B[P][d][b] B[P][a][c] = (ac|bd)  (B[P] symmetric),
// and moved into the list's columns (c, d) through LDS tiles (reads along c,
// writes along the pair index a + b(b+1)/2).  With one irrep the lists are
// written in place; with more, each list is formed in C1 order and packed by
// irrep on the device (dpd_lists.h).
#include <string>
#include <vector>
#include "common.h"
#include "dpd_lists.h"
#include "engine_util.h"
#include "gemm.h"
#include "intprc_df.h"
#include "tensor_ops.h"

namespace aces {

namespace {

using eu::P4;

constexpr int kT = 32;

// L233[pair(a, b) + nvp (c + v d)] = R[b][a][c] for a <= b, one block per
// (b, a tile, c tile); the tile goes through LDS so that reads run along c and
// writes along a.
__global__ void __launch_bounds__(256) list233_columns_kernel(int v, int d, const double* __restrict__ R,
                                                              double* __restrict__ L233, long nvp) {
  __shared__ double tile[kT][kT + 1];
  const int nt = (v + kT - 1) / kT;
  int x = blockIdx.x;
  const int tc = x % nt;
  x /= nt;
  const int ta = x % nt;
  const int b = x / nt;
  if (ta * kT > b) return;  // the tile holds no a <= b
  const long vv = (long)v * v;
  for (int e = threadIdx.x; e < kT * kT; e += blockDim.x) {
    const int la = e / kT, lc = e % kT;
    const int a = ta * kT + la, c = tc * kT + lc;
    tile[la][lc] = (a <= b && c < v) ? R[(long)b * vv + (long)a * v + c] : 0.0;
  }
  __syncthreads();
  const long pb = (long)b * (b + 1) / 2;
  for (int e = threadIdx.x; e < kT * kT; e += blockDim.x) {
    const int lc = e / kT, la = e % kT;
    const int a = ta * kT + la, c = tc * kT + lc;
    if (a <= b && c < v) L233[pb + a + nvp * (c + (long)v * d)] = tile[la][lc];
  }
}

// Reciprocal denominators in C1 list order: list 50 (13,14) at
// a + v b + v^2 (i + o j); list 48 (1,3) at pair(a<b) + nvm pair(i<j) with
// pair(p<q) = q(q-1)/2 + p; the T1 vector (type 9) at a + v i.
__global__ void denominators_kernel(int o, int v, const double* __restrict__ e, double* __restrict__ d50,
                                    double* __restrict__ d48, double* __restrict__ d1) {
  const long vv = (long)v * v, n50 = vv * o * o;
  const long nvm = (long)v * (v - 1) / 2, n48 = nvm * ((long)o * (o - 1) / 2);
  const long n1 = (long)o * v;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n50 + n48 + n1; x += (long)gridDim.x * blockDim.x) {
    if (x < n50) {
      const int a = (int)(x % v), b = (int)((x / v) % v);
      const long ij = x / vv;
      const int i = (int)(ij % o), j = (int)(ij / o);
      d50[x] = 1.0 / (e[i] + e[j] - e[o + a] - e[o + b]);
    } else if (x < n50 + n48) {
      const long y = x - n50;
      const long rab = y % nvm, cij = y / nvm;
      int b = (int)((sqrt(8.0 * (double)rab + 1.0) + 1.0) * 0.5);
      while ((long)b * (b - 1) / 2 > rab) --b;
      while ((long)(b + 1) * b / 2 <= rab) ++b;
      const int a = (int)(rab - (long)b * (b - 1) / 2);
      int j = (int)((sqrt(8.0 * (double)cij + 1.0) + 1.0) * 0.5);
      while ((long)j * (j - 1) / 2 > cij) --j;
      while ((long)(j + 1) * j / 2 <= cij) ++j;
      const int i = (int)(cij - (long)j * (j - 1) / 2);
      d48[y] = 1.0 / (e[i] + e[j] - e[o + a] - e[o + b]);
    } else {
      const long y = x - n50 - n48;
      const int a = (int)(y % v), i = (int)(y / v);
      d1[y] = 1.0 / (e[i] - e[o + a]);
    }
  }
}

struct DevBuf {
  double* p = nullptr;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  int alloc(long n) {
    if (hipMalloc(&p, sizeof(double) * (n > 0 ? n : 1)) != hipSuccess) {
      (void)hipGetLastError();
      set_error("intprc: device allocation of " + std::to_string(n) + " doubles failed");
      return ACES_ERR_NOMEM;
    }
    return ACES_OK;
  }
};

}  // namespace

int intprc_df(ListStore& L, RecordStore& R, const double* B, int naux, const double* eps, hipStream_t s) {
  const DpdSym& sy = L.sym();
  int o = 0, v = 0;
  for (int h = 0; h < sy.nirrep; ++h) {
    ACES_REQUIRE(sy.pop[0][h] == sy.pop[1][h] && sy.vrt[0][h] == sy.vrt[1][h], "intprc: closed-shell symmetry only");
    o += sy.pop[0][h];
    v += sy.vrt[0][h];
  }
  ACES_REQUIRE(o >= 1 && v >= 1 && naux >= 1 && B != nullptr && eps != nullptr, "intprc: bad arguments");
  const int n = o + v;
  const long nn = (long)n * n, oo = (long)o * o, vv = (long)v * v, ov = (long)o * v;
  const long nvp = (long)v * (v + 1) / 2;
  const bool c1 = sy.nirrep == 1;
  // factor slices [P][i][a], [P][i][j], [P][a][b]
  DevBuf Bov, Boo, Bvv, ws;
  ACES_CHECK(Bov.alloc((long)naux * ov));
  ACES_CHECK(Boo.alloc((long)naux * oo));
  ACES_CHECK(Bvv.alloc((long)naux * vv));
  ACES_CHECK(P4(naux, o, v, 1, Bov.p).src(B + o, 1.0, nn, n, 1, 0).run(s));
  ACES_CHECK(P4(naux, o, o, 1, Boo.p).src(B, 1.0, nn, n, 1, 0).run(s));
  ACES_CHECK(P4(naux, v, v, 1, Bvv.p).src(B + (long)o * n + o, 1.0, nn, n, 1, 0).run(s));
  const long ws_elems = 1L << 24;
  ACES_CHECK(ws.alloc(ws_elems));
  auto prod = [&](double* C, long M, long N, const double* X, long ldx, const double* Y, long ldy) {
    // C[m][n] = sum_P X[P][m] Y[P][n]
    GemmDesc d = eu::gd((int)M, (int)N, naux, X, ldx, 0, Y, ldy, 0, C, N, 1.0, 0.0);
    d.ws = ws.p;
    d.ws_elems = ws_elems;
    return gemm(d, GP_F64, s);
  };
  // dense blocks in the engines' conventions
  DevBuf ovov, oovv, ovoo, oooo, ovvv;
  ACES_CHECK(ovov.alloc(ov * ov));
  ACES_CHECK(oovv.alloc(oo * vv));
  ACES_CHECK(ovoo.alloc(ov * oo));
  ACES_CHECK(oooo.alloc(oo * oo));
  ACES_CHECK(ovvv.alloc(ov * vv));
  ACES_CHECK(prod(ovov.p, ov, ov, Bov.p, ov, Bov.p, ov));   // (ia|jb)
  ACES_CHECK(prod(oovv.p, oo, vv, Boo.p, oo, Bvv.p, vv));   // (ij|ab)
  ACES_CHECK(prod(ovoo.p, ov, oo, Bov.p, ov, Boo.p, oo));   // (ia|jk)
  ACES_CHECK(prod(oooo.p, oo, oo, Boo.p, oo, Boo.p, oo));   // (ij|kl)
  ACES_CHECK(prod(ovvv.p, ov, vv, Bov.p, ov, Bvv.p, vv));   // (ia|bc)
  // lists in C1 order (in place for one irrep, else a C1 buffer packed afterwards)
  struct Out { int list, tl, tr; };
  const Out outs[6] = {{16, 13, 14}, {13, 14, 14}, {10, 14, 18}, {30, 13, 11}, {23, 9, 9}, {233, 5, 13}};
  for (const Out& x : outs) {
    int tl = 0, tr = 0;
    if (L.isytyp(x.list, &tl, &tr) != ACES_OK || tl != x.tl || tr != x.tr)
      ACES_CHECK(L.inipck(1, x.tl, x.tr, x.list));
  }
  auto target = [&](int list, DevBuf& tmp, long n1) -> double* {
    if (c1) return L.device_ptr(1, list);
    return tmp.alloc(n1) == ACES_OK ? tmp.p : nullptr;
  };
  auto pack = [&](int list, int tl, int tr, DevBuf& tmp) {
    if (c1) return ACES_OK;
    return c1_to_dpd(L, list, tl, tr, 1, tmp.p, s);
  };
  {
    DevBuf t;
    // list 16 <ab|ij> = (ia|jb) at a + v b + v^2 (i + o j): output [j][i][b][a]
    double* l16 = target(16, t, ov * ov);
    ACES_REQUIRE(l16, "intprc: list 16");
    ACES_CHECK(P4(o, o, v, v, l16).src(ovov.p, 1.0, v, v * o * v, 1, o * v).run(s));
    ACES_CHECK(pack(16, 13, 14, t));
  }
  {
    DevBuf t;
    // list 13 <mn|ij> = (mi|nj) at m + o n + o^2 (i + o j): output [j][i][n][m]
    double* l13 = target(13, t, oo * oo);
    ACES_REQUIRE(l13, "intprc: list 13");
    ACES_CHECK(P4(o, o, o, o, l13).src(oooo.p, 1.0, 1, oo, o, oo * o).run(s));
    ACES_CHECK(pack(13, 14, 14, t));
  }
  {
    DevBuf t;
    // list 10 <ij|ka> = (ja|ik) at i + o j + o^2 (k + o a): output [a][k][j][i]
    double* l10 = target(10, t, oo * ov);
    ACES_REQUIRE(l10, "intprc: list 10");
    ACES_CHECK(P4(v, o, o, o, l10).src(ovoo.p, 1.0, oo, 1, v * oo, o).run(s));
    ACES_CHECK(pack(10, 14, 18, t));
  }
  {
    DevBuf t;
    // list 30 <ab|ci> = (ib|ac) at a + v b + v^2 (c + v i): output [i][c][b][a]
    double* l30 = target(30, t, vv * ov);
    ACES_REQUIRE(l30, "intprc: list 30");
    ACES_CHECK(P4(o, v, v, v, l30).src(ovvv.p, 1.0, v * vv, 1, vv, v).run(s));
    ACES_CHECK(pack(30, 13, 11, t));
  }
  {
    DevBuf t;
    // list 23 <AJ||BI>(AI,BJ) = (ij|ab) - (ia|jb) at a + v i + o v (b + v j): output [j][b][i][a]
    double* l23 = target(23, t, ov * ov);
    ACES_REQUIRE(l23, "intprc: list 23");
    ACES_CHECK(P4(o, v, o, v, l23).src(oovv.p, 1.0, vv, 1, o * vv, v).src(ovov.p, -1.0, v, 1, v * o * v, o * v).run(s));
    ACES_CHECK(pack(23, 9, 9, t));
  }
  // the dense blocks are done with
  for (DevBuf* b : {&ovov, &oovv, &ovoo, &oooo, &ovvv}) b->release();
  {
    DevBuf t, Rd;
    double* l233 = target(233, t, nvp * vv);
    ACES_REQUIRE(l233, "intprc: list 233");
    ACES_CHECK(Rd.alloc(vv * v));
    const int nt = (v + kT - 1) / kT;
    for (int d = 0; d < v; ++d) {
      // R_d[b][(a,c)] = sum_P Bvv[P][d][b] Bvv[P][a][c] (B[P] symmetric: [d][b] = [b][d])
      GemmDesc g = eu::gd(v, (int)vv, naux, Bvv.p + (long)d * v, vv, 0, Bvv.p, vv, 0, Rd.p, vv, 1.0, 0.0);
      g.ws = ws.p;
      g.ws_elems = ws_elems;
      ACES_CHECK(gemm(g, GP_F64, s));
      hipLaunchKernelGGL(list233_columns_kernel, dim3((unsigned)((long)v * nt * nt)), dim3(256), 0, s, v, d, Rd.p,
                         l233, nvp);
      ACES_HIP(hipGetLastError());
    }
    ACES_CHECK(pack(233, 5, 13, t));
  }
  // reciprocal denominators: lists 50 (AB), 48 (AA) and 48 group 9 (T1)
  {
    DevBuf e, d50, d48, d1;
    ACES_CHECK(e.alloc(n));
    ACES_HIP(hipMemcpyAsync(e.p, eps, sizeof(double) * n, hipMemcpyHostToDevice, s));
    const long n48 = (long)v * (v - 1) / 2 * ((long)o * (o - 1) / 2);
    int tl = 0, tr = 0;
    if (L.isytyp(50, &tl, &tr) != ACES_OK || tl != 13 || tr != 14) ACES_CHECK(L.inipck(1, 13, 14, 50));
    if (L.isytyp(48, &tl, &tr) != ACES_OK || tl != 1 || tr != 3) ACES_CHECK(L.inipck(1, 1, 3, 48));
    ACES_CHECK(L.updmoi(1, sy.irpdpd(0, 9), 9, 48));
    ACES_CHECK(d50.alloc(ov * ov));
    ACES_CHECK(d48.alloc(n48));
    ACES_CHECK(d1.alloc(ov));
    const long tot = ov * ov + n48 + ov;
    hipLaunchKernelGGL(denominators_kernel, dim3((unsigned)std::min<long>((tot + 255) / 256, 1L << 20)), dim3(256), 0,
                       s, o, v, e.p, d50.p, d48.p, d1.p);
    ACES_HIP(hipGetLastError());
    ACES_CHECK(c1_to_dpd(L, 50, 13, 14, 1, d50.p, s));
    ACES_CHECK(c1_to_dpd(L, 48, 1, 3, 1, d48.p, s));
    ACES_CHECK(dense_to_t1(L, 9, 48, d1.p, s));
    ACES_HIP(hipStreamSynchronize(s));
  }
  // orbital energies (record SCFEVALA, 8-byte words)
  ACES_CHECK(R.putrec("SCFEVALA", n, eps));
  ACES_HIP(hipStreamSynchronize(s));
  return ACES_OK;
}

}  // namespace aces
