// Non-iterative triples (T) of closed-shell CCSD(T) on the engine's resident
// integrals and amplitudes (the reference's TRPS, trp/trps.F:4, called from
// vcc/vcc.F:1199-1221 after the CC iterations).
//
// Closed-shell form (Rendell, Lee, Komornicki, CPL 178, 462 (1991)), for each
// occupied triple i >= j >= k (weights 6, 3, 1 for three, two, one distinct):
//   W_abc = sum over the six simultaneous permutations of (ia), (jb), (kc) of
//           sum_d (ia|bd) t_kj^cd - sum_l (kc|jl) t_il^ab
//   V_abc = W_abc + (jb|kc) t_i^a + (ia|kc) t_j^b + (ia|jb) t_k^c
//   Z_abc = 4 W_abc + W_bca + W_cab - 2 W_acb - 2 W_bac - 2 W_cba
//   E[4]T += 1/3 sum Z W / D,  E[4]T + E[5]ST += 1/3 sum Z V / D,
//   E4TAAA += 1/36 sum (antisymmetrised W)^2 / D,  D = e_i + e_j + e_k - e_a - e_b - e_c.
//
// W (v^3) is built by six MFMA products per triple, one per permutation, each
// with K = v + o: the particle term (sum over d) and the hole term (sum over l)
// share one contraction over the combined operands
//   A_x[(P,Q)][d] = (xP|Qd),  A_x[(P,Q)][v + l] = -t_xl^PQ
//   B_zy[R][d]    = t_zy^Rd,  B_zy[R][v + l]    = (zR|yl)
// (built once per call, K padded with zeros to the kernels' k step), and each
// product writes straight into W in its permuted order: the permutation is
// carried by the output's leading dimension and the batch stride (batched over
// one virtual index where the order needs it), so no transpose pass is needed.
// The products run in fp64 or as bf16x3 split products (fp32 accumulation).
// The energy pass walks canonical triples a >= b >= c in 8x8x8 tiles: a block
// loads the six permuted sub-cubes of W it needs into LDS and sums over the six
// orderings of each (a, b, c), which share D.
#include <algorithm>
#include <vector>
#include "bf16_util.h"
#include "ccsd_engine.h"
#include "common.h"

namespace aces {

namespace {

constexpr int kTile = 8;
constexpr int kPitch = kTile + 1;  // LDS row pitch (odd, fewer bank conflicts on permuted reads)
constexpr int kCube = kTile * kTile * kPitch;
constexpr long kEnergyBlocks = 512;  // two blocks per CU (221 VGPRs, 48 KB of LDS each)

__device__ __forceinline__ int widx(int p0, int p1, int p2) {
  // index of the ordering (p0, p1, p2) of (0, 1, 2) in {012, 021, 102, 120, 201, 210}
  (void)p2;
  return 2 * p0 + (p1 < p2 ? 0 : 1);
}

// Canonical tile triples (TA >= TB >= TC) tiles[0 .. ntab) with
// tiles[t] = TA | TB << 10 | TC << 20, walked by a grid of 256-thread blocks
// (block b takes t = b, b + gridDim.x, ...).  Per tile triple a block stages in
// LDS the six permuted 8^3 sub-cubes of W, the 8 x 8 blocks of (jb|kc),
// (ia|kc), (ia|jb) between the three tiles, and t1 and the orbital energies on
// them; the next tile triple's values are fetched into registers while the
// current one is summed.  part[3 b + {0,1,2}]: block sums of Z W / D, Z V / D
// and A^2 / D over the canonical (a >= b >= c) elements of its tile triples,
// each summed over the distinct orderings.
constexpr int kElems = kTile * kTile * kTile;
constexpr int kCubes = 6 * kElems;              // staged W values
constexpr int kXblk = 9 * kTile * kTile;        // one integral matrix: blocks (p, q) of the 3 tiles
constexpr int kXall = 3 * kXblk;
constexpr int kVec = 3 * 3 * kTile + 3 * kTile;  // t1i, t1j, t1k and e_v on the 3 tiles
constexpr int kStage = kCubes + kXall + kVec;
constexpr int kLoads = (kStage + 255) / 256;
static_assert(3 * 256 <= 6 * kCube, "block reduction fits the cube storage");

__global__ void __launch_bounds__(256) triples_energy_kernel(
    int v, int ntab, const int* __restrict__ tiles, const double* __restrict__ W, double eijk,
    const double* __restrict__ ev, const double* __restrict__ t1i, const double* __restrict__ t1j,
    const double* __restrict__ t1k, const double* __restrict__ Xjk, const double* __restrict__ Xik,
    const double* __restrict__ Xij, long ldx, double* __restrict__ part) {
  __shared__ double cube[6 * kCube];
  __shared__ double xl[kXall];
  __shared__ double vl[kVec];
  const long vv = (long)v * v;
  // sub-cube s holds W at tile coordinates (T[p0], T[p1], T[p2]) for the ordering p of s
  const int ords[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const double* const X[3] = {Xjk, Xik, Xij};
  const double* const t1v[3] = {t1i, t1j, t1k};
  double pre[kLoads];
  auto fetch = [&](int tb) {
    const int code = tiles[tb];
    const int T[3] = {code & 1023, (code >> 10) & 1023, (code >> 20) & 1023};
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      const int g = threadIdx.x + 256 * q;
      double val = 0.0;
      if (g < kCubes) {
        const int sc = g / kElems, e = g % kElems;
        const int u2 = e % kTile, u1 = (e / kTile) % kTile, u0 = e / (kTile * kTile);
        const int x0 = T[ords[sc][0]] * kTile + u0, x1 = T[ords[sc][1]] * kTile + u1;
        const int x2 = T[ords[sc][2]] * kTile + u2;
        if (x0 < v && x1 < v && x2 < v) val = W[x0 * vv + (long)x1 * v + x2];
      } else if (g < kCubes + kXall) {
        const int h = g - kCubes, mat = h / kXblk, r = h % kXblk;
        const int pq = r / (kTile * kTile), u = (r / kTile) % kTile, w = r % kTile;
        const int row = T[pq / 3] * kTile + u, col = T[pq % 3] * kTile + w;
        if (row < v && col < v) val = X[mat][row * ldx + col];
      } else if (g < kStage) {
        const int h = g - kCubes - kXall;
        if (h < 9 * kTile) {
          const int m = h / (3 * kTile), pp = (h / kTile) % 3, u = h % kTile;
          const int id = T[pp] * kTile + u;
          if (id < v) val = t1v[m][id];
        } else {
          const int h2 = h - 9 * kTile, pp = h2 / kTile, u = h2 % kTile;
          const int id = T[pp] * kTile + u;
          if (id < v) val = ev[id];
        }
      }
      pre[q] = val;
    }
  };
  double s4 = 0.0, st = 0.0, sa = 0.0;
  int tb = blockIdx.x;
  if (tb < ntab) fetch(tb);
  for (; tb < ntab; tb += gridDim.x) {
    __syncthreads();  // the previous tile triple's sums are done with the staged values
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      const int g = threadIdx.x + 256 * q;
      if (g < kCubes) {
        const int sc = g / kElems, e = g % kElems;
        cube[sc * kCube + (e / kTile) * kPitch + e % kTile] = pre[q];
      } else if (g < kCubes + kXall) {
        xl[g - kCubes] = pre[q];
      } else if (g < kStage) {
        vl[g - kCubes - kXall] = pre[q];
      }
    }
    __syncthreads();
    const int code = tiles[tb];
    if (tb + (int)gridDim.x < ntab) fetch(tb + gridDim.x);
    const int T[3] = {code & 1023, (code >> 10) & 1023, (code >> 20) & 1023};
    for (int e = threadIdx.x; e < kElems; e += blockDim.x) {
      const int lc = e % kTile, lb = (e / kTile) % kTile, la = e / (kTile * kTile);
      const int l[3] = {la, lb, lc};
      const int a = T[0] * kTile + la, b = T[1] * kTile + lb, c = T[2] * kTile + lc;
      if (a >= v || b >= v || c >= v || a < b || b < c) continue;
      // w[q] = W at the ordering q of (a, b, c)
      double w[6];
#pragma unroll
      for (int q = 0; q < 6; ++q)
        w[q] = cube[q * kCube + (l[ords[q][0]] * kTile + l[ords[q][1]]) * kPitch + l[ords[q][2]]];
      const double D = eijk - vl[9 * kTile + la] - vl[10 * kTile + lb] - vl[11 * kTile + lc];
      const double inv = 1.0 / D;
      const double mult = (a == b && b == c) ? (1.0 / 6.0) : ((a == b || b == c) ? 0.5 : 1.0);
      double e4 = 0.0, et = 0.0;
#pragma unroll
      for (int o = 0; o < 6; ++o) {
        const int x = ords[o][0], y = ords[o][1], z = ords[o][2];  // ordering (x, y, z) of positions of (a, b, c)
        // W at the patterns xyz, xzy, yxz, yzx, zxy, zyx of this ordering
        const double wxyz = w[widx(x, y, z)], wxzy = w[widx(x, z, y)];
        const double wyxz = w[widx(y, x, z)], wyzx = w[widx(y, z, x)];
        const double wzxy = w[widx(z, x, y)], wzyx = w[widx(z, y, x)];
        const double Z = 4.0 * wxyz + wyzx + wzxy - 2.0 * (wxzy + wyxz + wzyx);
        // (jb|kc) t_i^a + (ia|kc) t_j^b + (ia|jb) t_k^c at (a, b, c) = the indices in positions x, y, z
        const double vd = xl[(y * 3 + z) * kTile * kTile + l[y] * kTile + l[z]] * vl[x * kTile + l[x]] +
                          xl[kXblk + (x * 3 + z) * kTile * kTile + l[x] * kTile + l[z]] * vl[3 * kTile + y * kTile + l[y]] +
                          xl[2 * kXblk + (x * 3 + y) * kTile * kTile + l[x] * kTile + l[y]] * vl[6 * kTile + z * kTile + l[z]];
        e4 += Z * wxyz;
        et += Z * (wxyz + vd);
      }
      const double A = w[0] - w[1] - w[2] - w[5] + w[3] + w[4];
      s4 += mult * e4 * inv;
      st += mult * et * inv;
      sa += mult * 6.0 * A * A * inv;  // each of the 6 orderings carries A^2 (A changes sign only)
    }
  }
  // block sums in fixed order, through the cube storage (3 x 256 <= 6 kCube)
  __syncthreads();
  double* red = cube;
  red[threadIdx.x] = s4;
  red[256 + threadIdx.x] = st;
  red[512 + threadIdx.x] = sa;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < 3; ++q) red[q * 256 + threadIdx.x] += red[q * 256 + threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3L * blockIdx.x + threadIdx.x] = red[threadIdx.x * 256];
}

// Sum of the block partials in block order: out[0] = E[4]T piece (1/3 Z W / D),
// out[1] = E[4]T + E[5]ST piece, out[2] = E4TAAA piece (1/36 A^2 / D).
__global__ void triples_reduce_kernel(int nblk, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[3][256];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int q = 0; q < 3; ++q) acc[q] += part[3L * b + q];
  for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0][0] / 3.0;
    out[1] = red[1][0] / 3.0;
    out[2] = red[2][0] / 36.0;
  }
}

// Combined left operand A_x[(P,Q)][k], rows of kp elements (k >= v + o zero),
// for every occupied x: fp64 (out64) or bf16 hi/lo (hi, lo).
__global__ void triples_operand_a_kernel(int o, int v, int kp, const double* __restrict__ ovvv,
                                         const double* __restrict__ t2, double* __restrict__ out64,
                                         unsigned short* __restrict__ hi, unsigned short* __restrict__ lo) {
  const long vv = (long)v * v, total = (long)o * vv * kp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long row = e / kp;  // x * v^2 + (P,Q)
    const int k = (int)(e - row * kp);
    const long x = row / vv, pq = row - x * vv;
    double val = 0.0;
    if (k < v)
      val = ovvv[row * v + k];
    else if (k < v + o)
      val = -t2[(x * o + (k - v)) * vv + pq];
    if (out64) {
      out64[e] = val;
    } else {
      const unsigned short h = bf16_rn((float)val);
      hi[e] = h;
      lo[e] = bf16_rn((float)(val - (double)bf16_to_f(h)));
    }
  }
}

// Combined right operand B_zy[R][k] for every occupied pair (z, y).
__global__ void triples_operand_b_kernel(int o, int v, int kp, const double* __restrict__ t2,
                                         const double* __restrict__ ovoo, double* __restrict__ out64,
                                         unsigned short* __restrict__ hi, unsigned short* __restrict__ lo) {
  const long total = (long)o * o * v * kp;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long row = e / kp;  // (z * o + y) * v + R
    const int k = (int)(e - row * kp);
    const long zy = row / v, r = row - zy * v;
    const long z = zy / o, y = zy - z * o;
    double val = 0.0;
    if (k < v)
      val = t2[row * v + k];
    else if (k < v + o)
      val = ovoo[((z * v + r) * o + y) * o + (k - v)];
    if (out64) {
      out64[e] = val;
    } else {
      const unsigned short h = bf16_rn((float)val);
      hi[e] = h;
      lo[e] = bf16_rn((float)(val - (double)bf16_to_f(h)));
    }
  }
}

// Device buffer released when the call returns (the combined operands live
// for one call only).
struct ScopedBuf {
  void* p = nullptr;
  ~ScopedBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

int CcsdEngine::triples(GemmPrec prec, double* out, double* d_per_triple) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_REQUIRE(prec == GP_F64 || prec == GP_BF16X3, "triples: precision must be fp64 or bf16x3");
  ACES_REQUIRE(v <= 1023 * kTile, "triples: v too large for the tile table");
  const long v3 = (long)v * v * v, vv = (long)v * v;
  const int ntile = (v + kTile - 1) / kTile;
  const long ntrip = (long)o * (o + 1) * (o + 2) / 6;
  // k step of the contraction kernels: 16 (fp64), 32 (bf16x3)
  const int kstep = prec == GP_F64 ? 16 : 32;
  const int kp = (v + o + kstep - 1) / kstep * kstep;
  ACES_REQUIRE((long)o * vv < (1L << 31), "triples: o v^2 too large");
  if (trW_ == nullptr) {
    ACES_CHECK(alloc(&trW_, v3));
    std::vector<int> tab;
    for (int a = 0; a < ntile; ++a)
      for (int b = 0; b <= a; ++b)
        for (int c = 0; c <= b; ++c) tab.push_back(a | (b << 10) | (c << 20));
    trNblk_ = (long)tab.size();
    void* p;
    ACES_CHECK(alloc_raw(&p, (long)sizeof(int) * trNblk_));
    trTiles_ = (int*)p;
    ACES_HIP(hipMemcpyAsync(trTiles_, tab.data(), sizeof(int) * trNblk_, hipMemcpyHostToDevice, s_));
    ACES_CHECK(alloc(&trPart_, 3 * trNblk_));
    ACES_CHECK(alloc(&trOut_, 3 * ntrip));
    ACES_HIP(hipStreamSynchronize(s_));
  }
  // combined operands (fp64, or bf16 hi then lo halves of one buffer)
  const long na = (long)o * vv * kp, nb = (long)o * o * v * kp;
  const long esz = prec == GP_F64 ? 8 : 4;  // bytes per element (hi + lo for bf16x3)
  ScopedBuf bufA, bufB;
  if (hipMalloc(&bufA.p, (size_t)(na * esz)) != hipSuccess || hipMalloc(&bufB.p, (size_t)(nb * esz)) != hipSuccess) {
    (void)hipGetLastError();
    set_error("triples: no device memory for the combined operands");
    return ACES_ERR_NOMEM;
  }
  const void *Aop, *Alo = nullptr, *Bop, *Blo = nullptr;
  {
    double *a64 = nullptr, *b64 = nullptr;
    unsigned short *ah = nullptr, *al = nullptr, *bh = nullptr, *bl = nullptr;
    if (prec == GP_F64) {
      a64 = (double*)bufA.p;
      b64 = (double*)bufB.p;
      Aop = a64;
      Bop = b64;
    } else {
      ah = (unsigned short*)bufA.p;
      al = ah + na;
      bh = (unsigned short*)bufB.p;
      bl = bh + nb;
      Aop = ah; Alo = al; Bop = bh; Blo = bl;
    }
    hipLaunchKernelGGL(triples_operand_a_kernel, dim3((unsigned)std::min<long>((na + 255) / 256, 1L << 20)), dim3(256),
                       0, s_, o, v, kp, ovvv_, t2_, a64, ah, al);
    ACES_HIP(hipGetLastError());
    hipLaunchKernelGGL(triples_operand_b_kernel, dim3((unsigned)std::min<long>((nb + 255) / 256, 1L << 20)), dim3(256),
                       0, s_, o, v, kp, t2_, ovoo_, b64, bh, bl);
    ACES_HIP(hipGetLastError());
  }
  std::vector<double> eo(o), ev(v);
  ACES_HIP(hipMemcpyAsync(eo.data(), eo_, sizeof(double) * o, hipMemcpyDeviceToHost, s_));
  ACES_HIP(hipMemcpyAsync(ev.data(), ev_, sizeof(double) * v, hipMemcpyDeviceToHost, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  hipEvent_t e0, e1;
  ACES_HIP(hipEventCreate(&e0));
  ACES_HIP(hipEventCreate(&e1));
  ACES_HIP(hipEventRecord(e0, s_));
  const int saved_cat = prof_cat_;
  prof_cat_ = 6;
  double flops = 0.0;
  // element offsets of A_x (rows (P,Q)) and B_zy (rows R) in the combined operands
  const long ra = (long)vv * kp, rb = (long)v * kp;
  auto offA = [&](int x) { return (long)x * ra; };
  auto offB = [&](int z, int y) { return ((long)z * o + y) * rb; };
  auto at = [&](const void* base, long off) -> const void* {
    return prec == GP_F64 ? (const void*)((const double*)base + off) : (const void*)((const unsigned short*)base + off);
  };
  // energy pass grid: one block per tile triple up to two per CU,
  // then each block walks several
  const int egrid = (int)std::min<long>(trNblk_, kEnergyBlocks);
  long t = 0;
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k, ++t) {
        double* Wt = trW_;
        const int occ[3] = {i, j, k};
        const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int m = 0; m < 6; ++m) {
          const int x = occ[perms[m][0]], y = occ[perms[m][1]], z = occ[perms[m][2]];
          const long oa = offA(x), ob = offB(z, y);
          // base_xyz[P][Q][R] = sum_k A_x[(P,Q)][k] B_zy[R][k], accumulated into W in order m
          GemmDesc d;
          d.K = kp;
          switch (m) {
            case 0:  // W[P][Q][R]
              d.M = (int)vv; d.N = v;
              d.A = at(Aop, oa); d.A_lo = Alo ? at(Alo, oa) : nullptr; d.lda = kp;
              d.B = at(Bop, ob); d.B_lo = Blo ? at(Blo, ob) : nullptr; d.ldb = kp;
              d.C = Wt; d.ldc = v;
              break;
            case 3:  // W[R][P][Q]
              d.M = v; d.N = (int)vv;
              d.A = at(Bop, ob); d.A_lo = Blo ? at(Blo, ob) : nullptr; d.lda = kp;
              d.B = at(Aop, oa); d.B_lo = Alo ? at(Alo, oa) : nullptr; d.ldb = kp;
              d.C = Wt; d.ldc = vv;
              break;
            case 1:  // W[P][R][Q]: for each P, C[R][Q] = B_zy A_x[P]^T
              d.M = v; d.N = v; d.batch = v;
              d.A = at(Bop, ob); d.A_lo = Blo ? at(Blo, ob) : nullptr; d.lda = kp; d.sA = 0;
              d.B = at(Aop, oa); d.B_lo = Alo ? at(Alo, oa) : nullptr; d.ldb = kp; d.sB = (long)v * kp;
              d.C = Wt; d.ldc = v; d.sC = vv;
              break;
            case 2:  // W[Q][P][R]: for each P, C[Q][R] = A_x[P] B_zy^T
              d.M = v; d.N = v; d.batch = v;
              d.A = at(Aop, oa); d.A_lo = Alo ? at(Alo, oa) : nullptr; d.lda = kp; d.sA = (long)v * kp;
              d.B = at(Bop, ob); d.B_lo = Blo ? at(Blo, ob) : nullptr; d.ldb = kp; d.sB = 0;
              d.C = Wt; d.ldc = vv; d.sC = v;
              break;
            case 4:  // W[Q][R][P]: for each Q, C[R][P] = B_zy A_x[:, Q]^T
            case 5:  // W[R][Q][P]
              d.M = v; d.N = v; d.batch = v;
              d.A = at(Bop, ob); d.A_lo = Blo ? at(Blo, ob) : nullptr; d.lda = kp; d.sA = 0;
              d.B = at(Aop, oa); d.B_lo = Alo ? at(Alo, oa) : nullptr; d.ldb = (long)v * kp; d.sB = kp;
              d.C = Wt;
              if (m == 4) {
                d.ldc = v; d.sC = vv;
              } else {
                d.ldc = vv; d.sC = v;
              }
              break;
          }
          d.alpha = 1.0;
          d.beta = (m == 0) ? 0.0 : 1.0;
          d.tile_hint = 8;
          ACES_CHECK(timed_gemm(d, prec, 2.0 * (double)v3 * kp * (prec == GP_BF16X3 ? 3.0 : 1.0)));
          flops += 2.0 * (double)v3 * (v + o);
        }
        // energy pass over the canonical abc tiles, then the triple's sums
        const long ldx = (long)o * v;
        const double* Xjk = ovov_ + (long)j * v * ldx + (long)k * v;  // (jb|kc) at [b * ldx + c]
        const double* Xik = ovov_ + (long)i * v * ldx + (long)k * v;
        const double* Xij = ovov_ + (long)i * v * ldx + (long)j * v;
        hipLaunchKernelGGL(triples_energy_kernel, dim3((unsigned)egrid), dim3(256), 0, s_, v, (int)trNblk_, trTiles_,
                           Wt, eo[i] + eo[j] + eo[k], ev_, t1_ + (long)i * v, t1_ + (long)j * v, t1_ + (long)k * v,
                           Xjk, Xik, Xij, ldx, trPart_);
        ACES_HIP(hipGetLastError());
        hipLaunchKernelGGL(triples_reduce_kernel, dim3(1), dim3(256), 0, s_, egrid, trPart_, trOut_ + 3 * t);
        ACES_HIP(hipGetLastError());
      }
  prof_cat_ = saved_cat;
  ACES_HIP(hipEventRecord(e1, s_));
  std::vector<double> per((size_t)3 * ntrip);
  ACES_HIP(hipMemcpyAsync(per.data(), trOut_, sizeof(double) * 3 * ntrip, hipMemcpyDeviceToHost, s_));
  if (d_per_triple)
    ACES_HIP(hipMemcpyAsync(d_per_triple, trOut_, sizeof(double) * 3 * ntrip, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipEventSynchronize(e1));
  ACES_HIP(hipStreamSynchronize(s_));
  float ms = 0.f;
  ACES_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // weighted sums in triple order (deterministic)
  double e4 = 0.0, et = 0.0, ea = 0.0;
  t = 0;
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k, ++t) {
        const double w = (i != j && j != k) ? 6.0 : ((i == j && j == k) ? 1.0 : 3.0);
        e4 += w * per[3 * t];
        et += w * per[3 * t + 1];
        ea += w * per[3 * t + 2];
      }
  out[0] = e4;
  out[1] = et - e4;
  out[2] = ea;
  out[3] = 0.5 * e4 - ea;
  out[4] = ms;
  out[5] = flops;
  out[6] = (double)ntrip;
  out[7] = 0.0;
  return ACES_OK;
}

}  // namespace aces
