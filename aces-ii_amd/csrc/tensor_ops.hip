// HBM-bound tensor kernels for the CCSD amplitude path (gfx950).
#include "common.h"
#include "tensor_ops.h"

namespace aces {

static inline unsigned grid_for(long n, int block, long cap = 1L << 20) {
  long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

__device__ inline unsigned short bf16_rn(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

__global__ void convert_f32_kernel(int rows, int cols, const double* __restrict__ x, long ldx,
                                   float* __restrict__ y, long ldy) {
  const long total = (long)rows * ldy;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldy, c = i - r * ldy;
    y[i] = (c < cols) ? (float)x[r * ldx + c] : 0.0f;
  }
}

__global__ void split_bf16_kernel(int rows, int cols, const double* __restrict__ x, long ldx,
                                  unsigned short* __restrict__ hi, unsigned short* __restrict__ lo, long ldy) {
  const long total = (long)rows * ldy;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldy, c = i - r * ldy;
    unsigned short h = 0, l = 0;
    if (c < cols) {
      const double v = x[r * ldx + c];
      h = bf16_rn((float)v);
      const double rem = v - (double)bf16_to_f(h);
      l = bf16_rn((float)rem);
    }
    hi[i] = h;
    lo[i] = l;
  }
}

int convert_f32(int rows, int cols, const double* x, long ldx, float* y, long ldy, hipStream_t s) {
  ACES_REQUIRE(ldy >= cols && ldx >= cols, "convert_f32: bad leading dimension");
  const long n = (long)rows * ldy;
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(convert_f32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, rows, cols, x, ldx, y, ldy);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

int split_bf16(int rows, int cols, const double* x, long ldx, unsigned short* hi, unsigned short* lo, long ldy,
               hipStream_t s) {
  ACES_REQUIRE(ldy >= cols && ldx >= cols, "split_bf16: bad leading dimension");
  const long n = (long)rows * ldy;
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(split_bf16_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, rows, cols, x, ldx, hi, lo,
                     ldy);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
struct Permute4Args {
  long d1, d2, d3;
  long total;
  double* out;
  double beta;
  const double* in0;
  const double* in1;
  const double* in2;
  double a0, a1, a2;
  long s0[4], s1[4], s2[4];
  const double* X;
  const double* Y;
  double gamma;
  long xs[4], ys[4];
};

__global__ void permute4_kernel(Permute4Args p) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < p.total;
       idx += (long)gridDim.x * blockDim.x) {
    long r = idx;
    const long i3 = r % p.d3;
    r /= p.d3;
    const long i2 = r % p.d2;
    r /= p.d2;
    const long i1 = r % p.d1;
    const long i0 = r / p.d1;
    double v = 0.0;
    if (p.in0) v += p.a0 * p.in0[i0 * p.s0[0] + i1 * p.s0[1] + i2 * p.s0[2] + i3 * p.s0[3]];
    if (p.in1) v += p.a1 * p.in1[i0 * p.s1[0] + i1 * p.s1[1] + i2 * p.s1[2] + i3 * p.s1[3]];
    if (p.in2) v += p.a2 * p.in2[i0 * p.s2[0] + i1 * p.s2[1] + i2 * p.s2[2] + i3 * p.s2[3]];
    if (p.X)
      v += p.gamma * p.X[i0 * p.xs[0] + i1 * p.xs[1] + i2 * p.xs[2] + i3 * p.xs[3]] *
           p.Y[i0 * p.ys[0] + i1 * p.ys[1] + i2 * p.ys[2] + i3 * p.ys[3]];
    if (p.beta != 0.0) v += p.beta * p.out[idx];
    p.out[idx] = v;
  }
}

int permute4(const Permute4& q, hipStream_t s) {
  Permute4Args a;
  a.d1 = q.d[1];
  a.d2 = q.d[2];
  a.d3 = q.d[3];
  a.total = q.d[0] * q.d[1] * q.d[2] * q.d[3];
  if (a.total == 0) return ACES_OK;
  ACES_REQUIRE(q.out != nullptr, "permute4: null output");
  ACES_REQUIRE(q.X == nullptr || q.Y != nullptr, "permute4: product term needs X and Y");
  a.out = q.out;
  a.beta = q.beta;
  a.in0 = q.in[0];
  a.in1 = q.in[1];
  a.in2 = q.in[2];
  a.a0 = q.alpha[0];
  a.a1 = q.alpha[1];
  a.a2 = q.alpha[2];
  for (int k = 0; k < 4; ++k) {
    a.s0[k] = q.st[0][k];
    a.s1[k] = q.st[1][k];
    a.s2[k] = q.st[2][k];
    a.xs[k] = q.xs[k];
    a.ys[k] = q.ys[k];
  }
  a.X = q.X;
  a.Y = q.Y;
  a.gamma = q.gamma;
  hipLaunchKernelGGL(permute4_kernel, dim3(grid_for(a.total, 256)), dim3(256), 0, s, a);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// GEMV: one wave per row, lanes stride over k.
__global__ void gemv_kernel(int M, int K, const double* __restrict__ A, long lda, const double* __restrict__ x,
                            double* __restrict__ y, double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  for (long m = (long)blockIdx.x * wpb + (threadIdx.x >> 6); m < M; m += (long)gridDim.x * wpb) {
    const double* a = A + m * lda;
    double acc = 0.0;
    for (int k = lane; k < K; k += 64) acc += a[k] * x[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) y[m] = alpha * acc + (beta == 0.0 ? 0.0 : beta * y[m]);
  }
}

__global__ void gemv_t_kernel(int K, int N, const double* __restrict__ A, long lda, const double* __restrict__ x,
                              double* __restrict__ y, double alpha, double beta) {
  // Block handles 256 columns and a slice of k; slices are summed with a
  // second pass through shared memory (blockDim.y partial rows).
  __shared__ double part[4][256];
  const int n = blockIdx.x * 256 + threadIdx.x;
  double acc = 0.0;
  if (n < N)
    for (long k = threadIdx.y; k < K; k += blockDim.y) acc += A[k * lda + n] * x[k];
  part[threadIdx.y][threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.y == 0 && n < N) {
    double s = 0.0;
    for (int j = 0; j < (int)blockDim.y; ++j) s += part[j][threadIdx.x];
    y[n] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[n]);
  }
}

int gemv(int M, int K, const double* A, long lda, const double* x, double* y, double alpha, double beta,
         hipStream_t s) {
  if (M == 0) return ACES_OK;
  hipLaunchKernelGGL(gemv_kernel, dim3(grid_for(M, 4, 65536)), dim3(256), 0, s, M, K, A, lda, x, y, alpha, beta);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

int gemv_t(int K, int N, const double* A, long lda, const double* x, double* y, double alpha, double beta,
           hipStream_t s) {
  if (N == 0) return ACES_OK;
  hipLaunchKernelGGL(gemv_t_kernel, dim3((N + 255) / 256), dim3(256, 4), 0, s, K, N, A, lda, x, y, alpha, beta);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void contract2_kernel(int P, int Q, int R, int S, const double* __restrict__ X, long xp, long xq,
                                 long xr, long xs, const double* __restrict__ Y, long yr, long ys,
                                 double* __restrict__ out, double alpha, double beta) {
  const long total = (long)P * Q;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long p = idx / Q, q = idx - (idx / Q) * Q;
    const double* x = X + p * xp + q * xq;
    double acc = 0.0;
    for (int r = 0; r < R; ++r)
      for (int s2 = 0; s2 < S; ++s2) acc += x[r * xr + s2 * xs] * Y[r * yr + s2 * ys];
    out[idx] = alpha * acc + (beta == 0.0 ? 0.0 : beta * out[idx]);
  }
}

int contract2(int P, int Q, int R, int S, const double* X, long xp, long xq, long xr, long xs, const double* Y,
              long yr, long ys, double* out, double alpha, double beta, hipStream_t s) {
  const long n = (long)P * Q;
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(contract2_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, P, Q, R, S, X, xp, xq, xr, xs, Y,
                     yr, ys, out, alpha, beta);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// ---------------------------------------------------------------------------
__global__ void dot_kernel(long n, const double* __restrict__ x, const double* __restrict__ y,
                           double* __restrict__ result) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) acc += x[i] * y[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = sh[0];
}

int dot(long n, const double* x, const double* y, double* result, hipStream_t s) {
  hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(256), 0, s, n, x, y, result);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

}  // namespace aces
