// HBM-bound kernels of the engine's reverse pass.  Every kernel writes each
// output element from one thread (gathers, no atomics), so the results are
// deterministic.
#include "adjoint_kernels.h"
#include "common.h"
#include "pairs.h"

namespace aces {

// ---------------------------------------------------------------------------
// Forward (unpack_ladder):
//   inc2[i][j][a][b] += Zp[(a<=b)][(i<=j)] + s Zm[(a<b)][(i<j)],  s = sign(b-a) sign(j-i)
//   Yb[b][i][j][k]    = Zp[(k,b)][(i<=j)] + sign(j-i) Zm[(k,b)][(i<j)]
// so each gZp/gZm element gathers the orbit of its packed indices.
__global__ void ladder_unpack_adjoint_kernel(int o, int v, const double* __restrict__ g2,
                                             const double* __restrict__ gYb, double* __restrict__ gZp, long ldzp,
                                             double* __restrict__ gZm, long ldzm) {
  const long nvp = (long)v * (v + 1) / 2, nvm = (long)v * (v - 1) / 2;
  const long nop = (long)o * (o + 1) / 2, nom = (long)o * (o - 1) / 2;
  const long ov = (long)o * v, vv = (long)v * v;
  const long totP = (nvp + ov) * nop, totM = (nvm + ov) * nom;
  auto G2 = [&](int i, int j, int a, int b) { return g2[((long)i * o + j) * vv + (long)a * v + b]; };
  auto GY = [&](int b, int i, int j, int k) { return gYb[(((long)b * o + i) * o + j) * o + k]; };
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < totP + totM;
       idx += (long)gridDim.x * blockDim.x) {
    if (idx < totP) {
      const long row = idx / nop, pij = idx - row * nop;
      int i0, j0;
      unpack_le(pij, i0, j0);
      double z;
      if (row < nvp) {
        int a0, b0;
        unpack_le(row, a0, b0);
        z = G2(i0, j0, a0, b0);
        if (a0 != b0) z += G2(i0, j0, b0, a0);
        if (i0 != j0) {
          z += G2(j0, i0, a0, b0);
          if (a0 != b0) z += G2(j0, i0, b0, a0);
        }
      } else {
        const long r = row - nvp;
        const int k = (int)(r / v), b = (int)(r - (r / v) * v);
        z = GY(b, i0, j0, k);
        if (i0 != j0) z += GY(b, j0, i0, k);
      }
      gZp[row * ldzp + pij] = z;
    } else {
      const long q = idx - totP;
      const long row = q / nom, pij = q - row * nom;
      int i0, j0;
      unpack_lt(pij, i0, j0);
      double z;
      if (row < nvm) {
        int a0, b0;
        unpack_lt(row, a0, b0);
        z = G2(i0, j0, a0, b0) - G2(j0, i0, a0, b0) - G2(i0, j0, b0, a0) + G2(j0, i0, b0, a0);
      } else {
        const long r = row - nvm;
        const int k = (int)(r / v), b = (int)(r - (r / v) * v);
        z = GY(b, i0, j0, k) - GY(b, j0, i0, k);
      }
      gZm[row * ldzm + pij] = z;
    }
  }
}

int ladder_unpack_adjoint(int o, int v, const double* g2, const double* gYb, double* gZp, long ldzp, double* gZm,
                          long ldzm, hipStream_t s) {
  const long nvp = (long)v * (v + 1) / 2, nvm = (long)v * (v - 1) / 2;
  const long nop = (long)o * (o + 1) / 2, nom = (long)o * (o - 1) / 2;
  const long ov = (long)o * v;
  const long n = (nvp + ov) * nop + (nvm + ov) * nom;
  ACES_REQUIRE(ldzp >= nop && (nom == 0 || ldzm >= nom), "ladder_unpack_adjoint: leading dimensions too small");
  hipLaunchKernelGGL(ladder_unpack_adjoint_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, g2, gYb, gZp, ldzp, gZm,
                     ldzm);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Forward (pack_tau), for (i<=j), (c<=d):
//   Sp[(i<=j)][(c<=d)] = (tau_ijcd + tau_jicd)/2,  Dp[(i<j)][(c<d)] = (tau_ijcd - tau_jicd)/2.
// Only the c <= d half of tau is read, so the other half gets no gradient.
__global__ void pack_tau_adjoint_kernel(int o, int v, const double* __restrict__ gSp, long lds,
                                        const double* __restrict__ gDp, long ldd, double* __restrict__ gtau) {
  const long vv = (long)v * v, n = (long)o * o * vv;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    const long ij = idx / vv, cd = idx - ij * vv;
    const int i = (int)(ij / o), j = (int)(ij - (ij / o) * o);
    const int c = (int)(cd / v), d = (int)(cd - (cd / v) * v);
    if (c > d) continue;
    const int i0 = min(i, j), j0 = max(i, j);
    double g = gSp[pk_le(i0, j0) * lds + pk_le(c, d)] * (i == j ? 1.0 : 0.5);
    if (c < d && i != j) g += (i < j ? 0.5 : -0.5) * gDp[pk_lt(i0, j0) * ldd + pk_lt(c, d)];
    gtau[idx] += g;
  }
}

int pack_tau_adjoint(int o, int v, const double* gSp, long lds, const double* gDp, long ldd, double* gtau,
                     hipStream_t s) {
  const long n = (long)o * o * v * v;
  hipLaunchKernelGGL(pack_tau_adjoint_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, gSp, lds, gDp, ldd, gtau);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void adjoint_seed_kernel(int o, int v, const double* __restrict__ eo, const double* __restrict__ ev,
                                    const double* __restrict__ fov, long ldf, const double* __restrict__ l1,
                                    const double* __restrict__ l2, double* __restrict__ g1,
                                    double* __restrict__ g2) {
  const long n1 = (long)o * v, n2 = n1 * n1;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1 + n2;
       idx += (long)gridDim.x * blockDim.x) {
    if (idx < n1) {
      const int i = (int)(idx / v), a = (int)(idx - (idx / v) * v);
      g1[idx] = (fov ? 2.0 * fov[(long)i * ldf + a] : 0.0) - (eo[i] - ev[a]) * l1[idx];
    } else {
      const long r = idx - n1;
      const int b = (int)(r % v);
      const int a = (int)((r / v) % v);
      const int j = (int)((r / ((long)v * v)) % o);
      const int i = (int)(r / ((long)v * v * o));
      g2[r] = -(eo[i] + eo[j] - ev[a] - ev[b]) * l2[r];
    }
  }
}

// X[p][j] *= s (rows) or X[j][p] *= s (cols) for the diagonal pairs p = (c, c)
// of a packed a <= b index, c < v, j < n.
__global__ void scale_diag_pairs_kernel(int v, double* __restrict__ X, long ld, long n, int cols, double sc) {
  const long total = (long)v * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int c = (int)(idx / n);
    const long j = idx - (long)c * n;
    const long p = pk_le(c, c);
    X[cols ? j * ld + p : p * ld + j] *= sc;
  }
}

int scale_diag_pairs(int v, double* X, long ld, long n, int cols, double sc, hipStream_t s) {
  const long total = (long)v * n;
  if (total == 0) return ACES_OK;
  hipLaunchKernelGGL(scale_diag_pairs_kernel, dim3(grid1(total)), dim3(256), 0, s, v, X, ld, n, cols, sc);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

int adjoint_seed(int o, int v, const double* eo, const double* ev, const double* fov, long ldf, const double* l1,
                 const double* l2, double* g1, double* g2, hipStream_t s) {
  const long n = (long)o * v * (1 + (long)o * v);
  hipLaunchKernelGGL(adjoint_seed_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, eo, ev, fov, ldf, l1, l2, g1, g2);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// d/de_i and d/de_a of -sum l1 (e_i - e_a) t1 - sum l2 (e_i + e_j - e_a - e_b) t2.
// One block per orbital; the block sums its slice and one thread writes.
__global__ void denominator_adjoint_kernel(int o, int v, const double* __restrict__ l1,
                                          const double* __restrict__ l2, const double* __restrict__ t1,
                                          const double* __restrict__ t2, double* __restrict__ geo,
                                          double* __restrict__ gev) {
  __shared__ double sh[256];
  const long vv = (long)v * v, ov2 = (long)o * vv, oov = (long)o * o * v;
  const int p = blockIdx.x;
  double acc = 0.0;
  if (p < o) {
    const int i = p;
    for (long k = threadIdx.x; k < v; k += blockDim.x) acc -= l1[(long)i * v + k] * t1[(long)i * v + k];
    for (long k = threadIdx.x; k < ov2; k += blockDim.x) {
      const long x = (long)i * ov2 + k;  // first index i
      acc -= l2[x] * t2[x];
      const long j = k / vv, ab = k - (k / vv) * vv;
      const long y = (j * o + i) * vv + ab;  // second index i
      acc -= l2[y] * t2[y];
    }
  } else {
    const int a = p - o;
    for (long k = threadIdx.x; k < o; k += blockDim.x) acc += l1[k * v + a] * t1[k * v + a];
    for (long k = threadIdx.x; k < oov; k += blockDim.x) {
      const long ij = k / v, b = k - (k / v) * v;
      const long x = ij * vv + (long)a * v + b;  // third index a
      const long y = ij * vv + b * v + a;        // fourth index a
      acc += l2[x] * t2[x] + l2[y] * t2[y];
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (p < o)
      geo[p] += sh[0];
    else
      gev[p - o] += sh[0];
  }
}

int denominator_adjoint(int o, int v, const double* l1, const double* l2, const double* t1, const double* t2,
                        double* geo, double* gev, hipStream_t s) {
  hipLaunchKernelGGL(denominator_adjoint_kernel, dim3(o + v), dim3(256), 0, s, o, v, l1, l2, t1, t2, geo, gev);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void lambda_jacobi_kernel(int o, int v, const double* __restrict__ eo, const double* __restrict__ ev,
                                     const double* __restrict__ L, const double* __restrict__ g,
                                     double* __restrict__ Ln) {
  const long n1 = (long)o * v, n2 = n1 * n1, vv = (long)v * v;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n1 + n2;
       idx += (long)gridDim.x * blockDim.x) {
    if (idx < n1) {
      const int i = (int)(idx / v), a = (int)(idx - (idx / v) * v);
      Ln[idx] = L[idx] + g[idx] / (2.0 * (eo[i] - ev[a]));
    } else {
      const long r = idx - n1;
      const int b = (int)(r % v);
      const int a = (int)((r / v) % v);
      const int j = (int)((r / vv) % o);
      const int i = (int)(r / (vv * o));
      const long pr = ((long)j * o + i) * vv + (long)a * v + b;
      const double d = eo[i] + eo[j] - ev[a] - ev[b];
      Ln[idx] = L[idx] + (2.0 * g[idx] + g[n1 + pr]) / (3.0 * d);
    }
  }
}

int lambda_jacobi(int o, int v, const double* eo, const double* ev, const double* L, const double* g, double* Ln,
                  hipStream_t s) {
  const long n = (long)o * v * (1 + (long)o * v);
  hipLaunchKernelGGL(lambda_jacobi_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, eo, ev, L, g, Ln);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void assemble_density_kernel(int o, int v, const double* __restrict__ gfo, const double* __restrict__ gfv,
                                        const double* __restrict__ gfov, const double* __restrict__ geo,
                                        const double* __restrict__ gev, double* __restrict__ dens) {
  const int n = o + v;
  const long tot = (long)n * n;
  auto G = [&](int p, int q) -> double {
    if (p < o && q < o) return gfo[(long)p * o + q] + (p == q ? geo[p] : 0.0);
    if (p >= o && q >= o) return gfv[(long)(p - o) * v + (q - o)] + (p == q ? gev[p - o] : 0.0);
    if (p < o) return gfov[(long)p * v + (q - o)];
    return 0.0;
  };
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < tot; idx += (long)gridDim.x * blockDim.x) {
    const int p = (int)(idx / n), q = (int)(idx - (idx / n) * n);
    dens[idx] = 0.5 * (G(p, q) + G(q, p));
  }
}

int assemble_density(int o, int v, const double* gfo, const double* gfv, const double* gfov, const double* geo,
                     const double* gev, double* dens, hipStream_t s) {
  const long n = (long)(o + v) * (o + v);
  hipLaunchKernelGGL(assemble_density_kernel, dim3(grid1(n)), dim3(256), 0, s, o, v, gfo, gfv, gfov, geo, gev,
                     dens);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

}  // namespace aces
