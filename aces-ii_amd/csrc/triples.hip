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
// W (v^3) is built by twelve fp64 MFMA products per triple, each written
// straight into W in its permuted order (batched over one virtual index where
// the permutation needs it), so no transpose pass is needed.  The energy pass
// walks canonical triples a >= b >= c in 8x8x8 tiles: a block loads the six
// permuted sub-cubes of W it needs into LDS and sums over the six orderings of
// each (a, b, c), which share D.
#include <algorithm>
#include <vector>
#include "ccsd_engine.h"
#include "common.h"

namespace aces {

namespace {

constexpr int kTile = 8;
constexpr int kPitch = kTile + 1;  // LDS row pitch (odd, fewer bank conflicts on permuted reads)
constexpr int kCube = kTile * kTile * kPitch;

__device__ __forceinline__ int widx(int p0, int p1, int p2) {
  // index of the ordering (p0, p1, p2) of (0, 1, 2) in {012, 021, 102, 120, 201, 210}
  (void)p2;
  return 2 * p0 + (p1 < p2 ? 0 : 1);
}

// One block per canonical tile triple (TA >= TB >= TC), 256 threads.
// tiles[blockIdx.x] = TA | TB << 10 | TC << 20.  part[blockIdx.x * 3 + {0,1,2}]:
// block sums of Z W / D, Z V / D and A^2 / D over the canonical (a >= b >= c)
// elements of the tile triple, each summed over the distinct orderings.
__global__ void __launch_bounds__(256) triples_energy_kernel(
    int v, const int* __restrict__ tiles, const double* __restrict__ W, double eijk, const double* __restrict__ ev,
    const double* __restrict__ t1i, const double* __restrict__ t1j, const double* __restrict__ t1k,
    const double* __restrict__ Xjk, const double* __restrict__ Xik, const double* __restrict__ Xij, long ldx,
    double* __restrict__ part) {
  __shared__ double cube[6][kCube];
  __shared__ double red[3][256];
  const int code = tiles[blockIdx.x];
  const int T[3] = {code & 1023, (code >> 10) & 1023, (code >> 20) & 1023};
  const long vv = (long)v * v;
  // sub-cube s holds W at tile coordinates (T[p0], T[p1], T[p2]) for the ordering p of s
  const int ords[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int s = 0; s < 6; ++s) {
    const int b0 = T[ords[s][0]] * kTile, b1 = T[ords[s][1]] * kTile, b2 = T[ords[s][2]] * kTile;
    for (int e = threadIdx.x; e < kTile * kTile * kTile; e += blockDim.x) {
      const int u2 = e % kTile, u1 = (e / kTile) % kTile, u0 = e / (kTile * kTile);
      const int x0 = b0 + u0, x1 = b1 + u1, x2 = b2 + u2;
      double w = 0.0;
      if (x0 < v && x1 < v && x2 < v) w = W[x0 * vv + (long)x1 * v + x2];
      cube[s][(u0 * kTile + u1) * kPitch + u2] = w;
    }
  }
  __syncthreads();
  double s4 = 0.0, st = 0.0, sa = 0.0;
  for (int e = threadIdx.x; e < kTile * kTile * kTile; e += blockDim.x) {
    const int lc = e % kTile, lb = (e / kTile) % kTile, la = e / (kTile * kTile);
    const int l[3] = {la, lb, lc};
    const int idx[3] = {T[0] * kTile + la, T[1] * kTile + lb, T[2] * kTile + lc};
    const int a = idx[0], b = idx[1], c = idx[2];
    if (a >= v || b >= v || c >= v || a < b || b < c) continue;
    // w[q] = W at the ordering q of (a, b, c)
    double w[6];
    for (int q = 0; q < 6; ++q) w[q] = cube[q][(l[ords[q][0]] * kTile + l[ords[q][1]]) * kPitch + l[ords[q][2]]];
    const double D = eijk - ev[a] - ev[b] - ev[c];
    const double inv = 1.0 / D;
    const double mult = (a == b && b == c) ? (1.0 / 6.0) : ((a == b || b == c) ? 0.5 : 1.0);
    double e4 = 0.0, et = 0.0;
    for (int o = 0; o < 6; ++o) {
      const int x = ords[o][0], y = ords[o][1], z = ords[o][2];  // ordering (x, y, z) of positions of (a, b, c)
      const int p[3] = {x, y, z};
      // W at the patterns xyz, xzy, yxz, yzx, zxy, zyx of this ordering
      const double wxyz = w[widx(p[0], p[1], p[2])], wxzy = w[widx(p[0], p[2], p[1])];
      const double wyxz = w[widx(p[1], p[0], p[2])], wyzx = w[widx(p[1], p[2], p[0])];
      const double wzxy = w[widx(p[2], p[0], p[1])], wzyx = w[widx(p[2], p[1], p[0])];
      const double Z = 4.0 * wxyz + wyzx + wzxy - 2.0 * (wxzy + wyxz + wzyx);
      const int vx = idx[x], vy = idx[y], vz = idx[z];
      const double vd = Xjk[vy * ldx + vz] * t1i[vx] + Xik[vx * ldx + vz] * t1j[vy] + Xij[vx * ldx + vy] * t1k[vz];
      e4 += Z * wxyz;
      et += Z * (wxyz + vd);
    }
    const double A = w[0] - w[1] - w[2] - w[5] + w[3] + w[4];
    s4 += mult * e4 * inv;
    st += mult * et * inv;
    sa += mult * 6.0 * A * A * inv;  // each of the 6 orderings carries A^2 (A changes sign only)
  }
  red[0][threadIdx.x] = s4;
  red[1][threadIdx.x] = st;
  red[2][threadIdx.x] = sa;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3L * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
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

}  // namespace

int CcsdEngine::triples(double* out, double* d_per_triple) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_REQUIRE(v <= 1023 * kTile, "triples: v too large for the tile table");
  const long v3 = (long)v * v * v, vv = (long)v * v, oo = (long)o * o;
  const int ntile = (v + kTile - 1) / kTile;
  const long ntrip = (long)o * (o + 1) * (o + 2) / 6;
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
  const long o2 = oo;
  // operands: Ix = (xP|Qd) as [P][Q][d]; Tzy = t_zy^Rd as [R][d]; Jzy = (zR|yl) as [R][l]
  // (row stride o^2); Tx = t_xl^PQ as [l][(P,Q)]; t2 read as t_lx^QP = [l][x][Q][P].
  auto Ix = [&](int x) { return ovvv_ + (long)x * v3; };
  auto Tzy = [&](int z, int y) { return t2_ + ((long)z * o + y) * vv; };
  auto Jzy = [&](int z, int y) { return ovoo_ + (long)z * v * o2 + (long)y * o; };
  auto Tx = [&](int x) { return t2_ + (long)x * o * vv; };
  long t = 0;
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k, ++t) {
        const int occ[3] = {i, j, k};
        const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        for (int m = 0; m < 6; ++m) {
          const int x = occ[perms[m][0]], y = occ[perms[m][1]], z = occ[perms[m][2]];
          const double beta = (m == 0) ? 0.0 : 1.0;
          GemmDesc dv, dl;  // v-term (K = v) and o-term (K = o, alpha = -1)
          switch (m) {
            case 0:  // W[P][Q][R]
              dv.M = (int)vv; dv.N = v; dv.K = v;
              dv.A = Ix(x); dv.lda = v; dv.B = Tzy(z, y); dv.ldb = v; dv.C = trW_; dv.ldc = v;
              dl.M = (int)vv; dl.N = v; dl.K = o;
              dl.A = Tx(x); dl.lda = vv; dl.a_kc = 0; dl.B = Jzy(z, y); dl.ldb = o2; dl.C = trW_; dl.ldc = v;
              break;
            case 1:  // W[P][R][Q], batched over P
              dv.M = v; dv.N = v; dv.K = v; dv.batch = v;
              dv.A = Tzy(z, y); dv.lda = v; dv.sA = 0; dv.B = Ix(x); dv.ldb = v; dv.sB = vv;
              dv.C = trW_; dv.ldc = v; dv.sC = vv;
              dl.M = v; dl.N = v; dl.K = o; dl.batch = v;
              dl.A = Jzy(z, y); dl.lda = o2; dl.sA = 0; dl.B = Tx(x); dl.ldb = vv; dl.b_kc = 0; dl.sB = v;
              dl.C = trW_; dl.ldc = v; dl.sC = vv;
              break;
            case 2:  // W[Q][P][R], batched over P
              dv.M = v; dv.N = v; dv.K = v; dv.batch = v;
              dv.A = Ix(x); dv.lda = v; dv.sA = vv; dv.B = Tzy(z, y); dv.ldb = v; dv.sB = 0;
              dv.C = trW_; dv.ldc = vv; dv.sC = v;
              dl.M = v; dl.N = v; dl.K = o; dl.batch = v;
              dl.A = Tx(x); dl.lda = vv; dl.a_kc = 0; dl.sA = v; dl.B = Jzy(z, y); dl.ldb = o2; dl.sB = 0;
              dl.C = trW_; dl.ldc = vv; dl.sC = v;
              break;
            case 3:  // W[R][P][Q]
              dv.M = v; dv.N = (int)vv; dv.K = v;
              dv.A = Tzy(z, y); dv.lda = v; dv.B = Ix(x); dv.ldb = v; dv.C = trW_; dv.ldc = vv;
              dl.M = v; dl.N = (int)vv; dl.K = o;
              dl.A = Jzy(z, y); dl.lda = o2; dl.B = Tx(x); dl.ldb = vv; dl.b_kc = 0; dl.C = trW_; dl.ldc = vv;
              break;
            case 4:  // W[Q][R][P], batched over Q
            case 5:  // W[R][Q][P], batched over Q
              dv.M = v; dv.N = v; dv.K = v; dv.batch = v;
              dv.A = Tzy(z, y); dv.lda = v; dv.sA = 0; dv.B = Ix(x); dv.ldb = vv; dv.sB = v;
              dl.M = v; dl.N = v; dl.K = o; dl.batch = v;
              dl.A = Jzy(z, y); dl.lda = o2; dl.sA = 0; dl.B = t2_ + (long)x * vv; dl.ldb = (long)o * vv;
              dl.b_kc = 0; dl.sB = v;
              if (m == 4) {
                dv.C = trW_; dv.ldc = v; dv.sC = vv;
                dl.C = trW_; dl.ldc = v; dl.sC = vv;
              } else {
                dv.C = trW_; dv.ldc = vv; dv.sC = v;
                dl.C = trW_; dl.ldc = vv; dl.sC = v;
              }
              break;
          }
          dv.alpha = 1.0; dv.beta = beta;
          dl.alpha = -1.0; dl.beta = 1.0;
          ACES_CHECK(gemm64(dv));
          ACES_CHECK(gemm64(dl));
          flops += 2.0 * (double)v3 * (v + o);
        }
        // energy pass over the canonical abc tiles, then the triple's sums
        const long ldx = (long)o * v;
        const double* Xjk = ovov_ + (long)j * v * ldx + (long)k * v;  // (jb|kc) at [b * ldx + c]
        const double* Xik = ovov_ + (long)i * v * ldx + (long)k * v;
        const double* Xij = ovov_ + (long)i * v * ldx + (long)j * v;
        hipLaunchKernelGGL(triples_energy_kernel, dim3((unsigned)trNblk_), dim3(256), 0, s_, v, trTiles_, trW_,
                           eo[i] + eo[j] + eo[k], ev_, t1_ + (long)i * v, t1_ + (long)j * v, t1_ + (long)k * v, Xjk,
                           Xik, Xij, ldx, trPart_);
        ACES_HIP(hipGetLastError());
        hipLaunchKernelGGL(triples_reduce_kernel, dim3(1), dim3(256), 0, s_, (int)trNblk_, trPart_, trOut_ + 3 * t);
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
