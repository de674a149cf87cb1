// HBM-bound tensor kernels for the CCSD amplitude path (gfx950).
#include <algorithm>
#include <cstdint>
#include <type_traits>
#include "common.h"
#include "tensor_ops.h"
#include "bf16_util.h"

namespace aces {

static inline unsigned grid_for(long n, int block, long cap = 1L << 20) {
  long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}


__global__ void convert_f32_kernel(int rows, int cols, const double* __restrict__ x, long ldx,
                                   float* __restrict__ y, long ldy) {
  const long total = (long)rows * ldy;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldy, c = i - r * ldy;
    y[i] = (c < cols) ? (float)x[r * ldx + c] : 0.0f;
  }
}

__global__ void split_bf16_kernel(int rows, int cols, const double* __restrict__ x, long ldx,
                                  unsigned short* __restrict__ hi, unsigned short* __restrict__ lo, long ldy,
                                  const double* __restrict__ x0) {
  const long total = (long)rows * ldy;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long r = i / ldy, c = i - r * ldy;
    unsigned short h = 0, l = 0;
    if (c < cols) {
      const double v = x0 ? x[r * ldx + c] - x0[r * ldx + c] : x[r * ldx + c];
      h = bf16_rn((float)v);
      const double rem = v - (double)bf16_to_f(h);
      l = bf16_rn((float)rem);
    }
    hi[i] = h;
    lo[i] = l;
  }
}

// One block row per tensor block (blockIdx.y); the block's description is
// picked with constant indices so that the argument block stays in the
// kernarg segment (a runtime index into it, or a loop of selects, copies it
// to scratch).
struct SplitBlock {
  int rows, cols;
  long ldy;
  const double* x;
  const double* x0;
  void* hi;
  void* lo;
};

__device__ inline SplitBlock split_block_of(const SplitBlocks& b, int g) {
  switch (g) {
#define ACES_SPLIT_CASE(k) \
  case k:                  \
    return SplitBlock{b.rows[k], b.cols[k], b.ldy[k], b.x[k], b.x0[k], b.hi[k], b.lo[k]};
    ACES_SPLIT_CASE(1)
    ACES_SPLIT_CASE(2)
    ACES_SPLIT_CASE(3)
    ACES_SPLIT_CASE(4)
    ACES_SPLIT_CASE(5)
    ACES_SPLIT_CASE(6)
    ACES_SPLIT_CASE(7)
#undef ACES_SPLIT_CASE
    default:
      return SplitBlock{b.rows[0], b.cols[0], b.ldy[0], b.x[0], b.x0[0], b.hi[0], b.lo[0]};
  }
}

__global__ void split_blocks_kernel(SplitBlocks bs) {
  const SplitBlock b = split_block_of(bs, (int)blockIdx.y);
  const long total = (long)b.rows * b.ldy;
  const long ldy = b.ldy;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
    const long r = k / ldy, c = k - r * ldy;
    double v = 0.0;
    if (c < b.cols) {
      v = b.x[r * b.cols + c];
      if (b.x0) v -= b.x0[r * b.cols + c];
    }
    if (bs.f64) {
      static_cast<double*>(b.hi)[k] = v;
    } else if (bs.f32) {
      static_cast<float*>(b.hi)[k] = (float)v;
    } else {
      const unsigned short h = bf16_rn((float)v);
      static_cast<unsigned short*>(b.hi)[k] = h;
      static_cast<unsigned short*>(b.lo)[k] = bf16_rn((float)(v - (double)bf16_to_f(h)));
    }
  }
}

int split_blocks(const SplitBlocks& b, hipStream_t s) {
  ACES_REQUIRE(b.n >= 0 && b.n <= 8, "split_blocks: at most 8 blocks");
  long most = 0;
  for (int g = 0; g < b.n; ++g) {
    ACES_REQUIRE(b.ldy[g] >= b.cols[g] && b.rows[g] >= 0 && b.cols[g] >= 0, "split_blocks: bad block shape");
    most = std::max(most, (long)b.rows[g] * b.ldy[g]);
  }
  if (most == 0) return ACES_OK;
  // blocks per tensor block: enough for the largest to keep every CU busy
  const unsigned gx = grid_for(most, 256, std::max<long>(1, 4096 / b.n));
  hipLaunchKernelGGL(split_blocks_kernel, dim3(gx, (unsigned)b.n), dim3(256), 0, s, b);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
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
               hipStream_t s, const double* x0) {
  ACES_REQUIRE(ldy >= cols && ldx >= cols, "split_bf16: bad leading dimension");
  const long n = (long)rows * ldy;
  if (n == 0) return ACES_OK;
  hipLaunchKernelGGL(split_bf16_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, rows, cols, x, ldx, hi, lo,
                     ldy, x0);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

// Transposing conversions for operands stored k-major: x[c*ldx + r] holds
// element (r, c); y[r*ldy + c] is written for c < ldy (columns cols..ldy-1
// zero).  64 x 64 tiles pass through LDS so that both sides are coalesced.
// MODE 0: fp32 y0; 1: bf16 split y0 (hi) / y1 (lo); 2: fp64 y0.
template <int MODE>
__global__ __launch_bounds__(256) void convert_t_kernel(int rows, int cols, const double* __restrict__ x, long ldx,
                                                        void* __restrict__ y0, void* __restrict__ y1, long ldy,
                                                        const double* __restrict__ x0) {
  __shared__ double tile[64][65];
  const long ntr = (rows + 63) / 64;
  const long r0 = (blockIdx.x % ntr) * 64;
  const long c0 = (blockIdx.x / ntr) * 64;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  // branch-free loads (outside the matrix: element (r0, min(c0, cols - 1)),
  // dropped), all in flight together; tiles wholly in the pad columns read
  // nothing
  double v[16];
  const bool any = c0 < cols;
  const long fb = (any ? c0 : 0) * ldx + r0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const long c = c0 + w + 4 * k, r = r0 + l;
    const bool ok = c < cols && r < rows;
    const long i = ok ? c * ldx + r : fb;
    v[k] = any ? (x0 ? x[i] - x0[i] : x[i]) : 0.0;
    if (!ok) v[k] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) tile[w + 4 * k][l] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const long r = r0 + w + 4 * k, c = c0 + l;
    if (r < rows && c < ldy) {
      const double v = tile[l][w + 4 * k];
      if (MODE == 1) {
        const unsigned short h = bf16_rn((float)v);
        reinterpret_cast<unsigned short*>(y0)[r * ldy + c] = h;
        reinterpret_cast<unsigned short*>(y1)[r * ldy + c] = bf16_rn((float)(v - (double)bf16_to_f(h)));
      } else if (MODE == 2) {
        reinterpret_cast<double*>(y0)[r * ldy + c] = v;
      } else {
        reinterpret_cast<float*>(y0)[r * ldy + c] = (float)v;
      }
    }
  }
}

static int convert_t_launch(int mode, int rows, int cols, const double* x, long ldx, void* y0, void* y1, long ldy,
                            hipStream_t s, const double* x0 = nullptr) {
  ACES_REQUIRE(ldy >= cols && ldx >= rows, "transposing conversion: bad leading dimension");
  if (rows == 0 || ldy == 0) return ACES_OK;
  const long nblk = ((rows + 63) / 64) * ((ldy + 63) / 64);
  ACES_REQUIRE(nblk < (1L << 31), "transposing conversion: too many tiles");
  if (mode == 1)
    hipLaunchKernelGGL(convert_t_kernel<1>, dim3((unsigned)nblk), dim3(256), 0, s, rows, cols, x, ldx, y0, y1, ldy,
                       x0);
  else if (mode == 2)
    hipLaunchKernelGGL(convert_t_kernel<2>, dim3((unsigned)nblk), dim3(256), 0, s, rows, cols, x, ldx, y0, y1, ldy,
                       x0);
  else
    hipLaunchKernelGGL(convert_t_kernel<0>, dim3((unsigned)nblk), dim3(256), 0, s, rows, cols, x, ldx, y0, y1, ldy,
                       x0);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

int convert_f32_t(int rows, int cols, const double* x, long ldx, float* y, long ldy, hipStream_t s) {
  return convert_t_launch(0, rows, cols, x, ldx, y, nullptr, ldy, s);
}

int copy_f64_t(int rows, int cols, const double* x, long ldx, double* y, long ldy, hipStream_t s) {
  return convert_t_launch(2, rows, cols, x, ldx, y, nullptr, ldy, s);
}

int split_bf16_t(int rows, int cols, const double* x, long ldx, unsigned short* hi, unsigned short* lo, long ldy,
                 hipStream_t s, const double* x0) {
  return convert_t_launch(1, rows, cols, x, ldx, hi, lo, ldy, s, x0);
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

// Tiled form, used when a source is read across the output's fastest index:
// the output is walked in tiles of 64 (index 3) x 32 (index q, one of 0..2).
// A source whose unit stride lies on index q is read along q (coalesced) and
// transposed through LDS; every other source and the product term are read
// in the output's order.
constexpr int kPT3 = 64, kPTQ = 32;

struct Permute4Tile {
  long d[4];
  long os[4];  // output strides (row major)
  int q, ra, rb;
  long n3, nq, ntiles;
  double* out;
  double beta;
  int nd, nt;  // sources 0..nd-1 read directly, nd..nd+nt-1 through LDS
  const double* in[kP4Src];
  double alpha[kP4Src];
  long st[kP4Src][4];
  const double* X;
  const double* Y;
  double gamma;
  long xs[4], ys[4];
};

__global__ __launch_bounds__(256) void permute4_tiled_kernel(Permute4Tile p) {
  __shared__ double tile[kPT3][kPTQ + 1];
  const int tid = threadIdx.x;
  const int c3 = tid & 63, r0 = tid >> 6;  // output walk: rows r0 + 4k, column c3
  const int cq = tid & 31, s0 = tid >> 5;  // transposed read: rows s0 + 8k, column cq
  const long dq = p.d[p.q], d3 = p.d[3];
  for (long t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
    long b = t;
    const long t3 = b % p.n3;
    b /= p.n3;
    const long tq = b % p.nq;
    b /= p.nq;
    const long ib = b % p.d[p.rb];
    const long ia = b / p.d[p.rb];
    const long i3b = t3 * kPT3, iqb = tq * kPTQ;
    const bool ok3 = i3b + c3 < d3;
    // Loads are branch-free: an element outside the array reads the tile's
    // first element (which exists) and is dropped, so that a thread's loads
    // are in flight together.
    bool ok[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) ok[k] = ok3 && iqb + r0 + 4 * k < dq;
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    for (int s = 0; s < p.nd; ++s) {
      const long* st = p.st[s];
      const double* src = p.in[s] + ia * st[p.ra] + ib * st[p.rb] + iqb * st[p.q] + i3b * st[3];
      const long off = r0 * st[p.q] + c3 * st[3], step = 4 * st[p.q];
      const double al = p.alpha[s];
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = src[ok[k] ? off + k * step : 0];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += ok[k] ? al * v[k] : 0.0;
    }
    if (p.X) {
      const double* xp = p.X + ia * p.xs[p.ra] + ib * p.xs[p.rb] + iqb * p.xs[p.q] + i3b * p.xs[3];
      const double* yp = p.Y + ia * p.ys[p.ra] + ib * p.ys[p.rb] + iqb * p.ys[p.q] + i3b * p.ys[3];
      const long xoff = r0 * p.xs[p.q] + c3 * p.xs[3], yoff = r0 * p.ys[p.q] + c3 * p.ys[3];
      const long xstep = 4 * p.xs[p.q], ystep = 4 * p.ys[p.q];
      double xv[8], yv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        xv[k] = xp[ok[k] ? xoff + k * xstep : 0];
        yv[k] = yp[ok[k] ? yoff + k * ystep : 0];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += ok[k] ? p.gamma * xv[k] * yv[k] : 0.0;
    }
    const bool okq = iqb + cq < dq;
    for (int s = p.nd; s < p.nd + p.nt; ++s) {
      const long* st = p.st[s];
      const double* src = p.in[s] + ia * st[p.ra] + ib * st[p.rb] + iqb + i3b * st[3];
      const long off = cq + s0 * st[3], step = 8 * st[3];
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const bool okt = okq && i3b + s0 + 8 * k < d3;
        v[k] = src[okt ? off + k * step : 0];
        if (!okt) v[k] = 0.0;
      }
      __syncthreads();  // the previous reader of the tile is done
#pragma unroll
      for (int k = 0; k < 8; ++k) tile[s0 + 8 * k][cq] = v[k];
      __syncthreads();
      const double al = p.alpha[s];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += al * tile[c3][r0 + 4 * k];
    }
    double* o0 = p.out + ia * p.os[p.ra] + ib * p.os[p.rb] + iqb * p.os[p.q] + i3b;
    const long ooff = r0 * p.os[p.q] + c3, ostep = 4 * p.os[p.q];
    double prev[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) prev[k] = (p.beta != 0.0) ? o0[ok[k] ? ooff + k * ostep : 0] : 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (ok[k]) {
        double v = acc[k];
        if (p.beta != 0.0) v += p.beta * prev[k];
        o0[ooff + k * ostep] = v;
      }
  }
}

static int permute4_simple(const Permute4& q, hipStream_t s) {
  // three sources per launch; later launches accumulate (beta = 1)
  int idx[kP4Src], n = 0;
  for (int k = 0; k < kP4Src; ++k)
    if (q.in[k]) idx[n++] = k;
  for (int first = 0; first == 0 || first < n; first += 3) {
    Permute4Args a;
    a.d1 = q.d[1];
    a.d2 = q.d[2];
    a.d3 = q.d[3];
    a.total = q.d[0] * q.d[1] * q.d[2] * q.d[3];
    a.out = q.out;
    a.beta = (first == 0) ? q.beta : 1.0;
    const double** ins[3] = {&a.in0, &a.in1, &a.in2};
    double* als[3] = {&a.a0, &a.a1, &a.a2};
    long* sts[3] = {a.s0, a.s1, a.s2};
    for (int j = 0; j < 3; ++j) {
      const int k = first + j < n ? idx[first + j] : -1;
      *ins[j] = k >= 0 ? q.in[k] : nullptr;
      *als[j] = k >= 0 ? q.alpha[k] : 0.0;
      for (int m = 0; m < 4; ++m) sts[j][m] = k >= 0 ? q.st[k][m] : 0;
    }
    for (int m = 0; m < 4; ++m) {
      a.xs[m] = q.xs[m];
      a.ys[m] = q.ys[m];
    }
    a.X = (first == 0) ? q.X : nullptr;
    a.Y = (first == 0) ? q.Y : nullptr;
    a.gamma = q.gamma;
    hipLaunchKernelGGL(permute4_kernel, dim3(grid_for(a.total, 256)), dim3(256), 0, s, a);
    ACES_HIP(hipGetLastError());
  }
  return ACES_OK;
}

// Index (0..2) on which source s has unit stride and that a tile can span, or
// 3 when the source is read well enough in the output's order.
static int permute4_fast_index(const Permute4& q, int s) {
  const long s3 = q.st[s][3] < 0 ? -q.st[s][3] : q.st[s][3];
  if (s3 < 8) return 3;
  for (int k = 2; k >= 0; --k)
    if (q.st[s][k] == 1 && q.d[k] >= 8) return k;
  return 3;
}

int permute4(const Permute4& q, hipStream_t s) {
  const long total = q.d[0] * q.d[1] * q.d[2] * q.d[3];
  if (total == 0) return ACES_OK;
  ACES_REQUIRE(q.out != nullptr, "permute4: null output");
  ACES_REQUIRE(q.X == nullptr || q.Y != nullptr, "permute4: product term needs X and Y");
  int fast[kP4Src];
  for (int k = 0; k < kP4Src; ++k) fast[k] = 3;
  bool any = false, alias = (q.X == q.out || q.Y == q.out);
  for (int k = 0; k < kP4Src; ++k)
    if (q.in[k]) {
      fast[k] = permute4_fast_index(q, k);
      any |= fast[k] < 3;
      alias |= q.in[k] == q.out;
    }
  if (!any || q.d[3] < 16 || total < 65536) return permute4_simple(q, s);
  // one launch per distinct tile index; the first also takes the direct
  // sources, the product term and beta, the later ones accumulate
  int qs[3], nqs = 0;
  for (int k = 0; k < kP4Src; ++k)
    if (q.in[k] && fast[k] < 3) {
      bool seen = false;
      for (int j = 0; j < nqs; ++j) seen |= qs[j] == fast[k];
      if (!seen) qs[nqs++] = fast[k];
    }
  if (nqs > 1 && alias) return permute4_simple(q, s);
  for (int L = 0; L < nqs; ++L) {
    Permute4Tile t;
    for (int k = 0; k < 4; ++k) t.d[k] = q.d[k];
    t.os[3] = 1;
    t.os[2] = q.d[3];
    t.os[1] = q.d[2] * q.d[3];
    t.os[0] = q.d[1] * q.d[2] * q.d[3];
    t.q = qs[L];
    t.ra = (t.q == 0) ? 1 : 0;
    t.rb = (t.q == 2) ? 1 : 2;
    t.n3 = (q.d[3] + kPT3 - 1) / kPT3;
    t.nq = (q.d[t.q] + kPTQ - 1) / kPTQ;
    t.ntiles = t.n3 * t.nq * q.d[t.ra] * q.d[t.rb];
    t.out = q.out;
    t.beta = (L == 0) ? q.beta : 1.0;
    t.nd = 0;
    t.nt = 0;
    int slot = 0;
    if (L == 0)
      for (int k = 0; k < kP4Src; ++k)
        if (q.in[k] && fast[k] == 3) {
          t.in[slot] = q.in[k];
          t.alpha[slot] = q.alpha[k];
          for (int j = 0; j < 4; ++j) t.st[slot][j] = q.st[k][j];
          ++slot;
          ++t.nd;
        }
    for (int k = 0; k < kP4Src; ++k)
      if (q.in[k] && fast[k] == t.q) {
        t.in[slot] = q.in[k];
        t.alpha[slot] = q.alpha[k];
        for (int j = 0; j < 4; ++j) t.st[slot][j] = q.st[k][j];
        ++slot;
        ++t.nt;
      }
    for (int k = slot; k < kP4Src; ++k) {
      t.in[k] = nullptr;
      t.alpha[k] = 0.0;
      for (int j = 0; j < 4; ++j) t.st[k][j] = 0;
    }
    t.X = (L == 0) ? q.X : nullptr;
    t.Y = (L == 0) ? q.Y : nullptr;
    t.gamma = q.gamma;
    for (int j = 0; j < 4; ++j) {
      t.xs[j] = q.xs[j];
      t.ys[j] = q.ys[j];
    }
    hipLaunchKernelGGL(permute4_tiled_kernel, dim3(grid_for(t.ntiles, 1, 1L << 20)), dim3(256), 0, s, t);
    ACES_HIP(hipGetLastError());
  }
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

// Matrix element (k, n) of gemv_t: fp64, or a bf16 hi/lo split (hi + lo).
struct GemvF64 {
  const double* a;
  __device__ double operator()(long q) const { return a[q]; }
};
struct GemvSplit {
  const unsigned short* hi;
  const unsigned short* lo;
  __device__ double operator()(long q) const { return (double)bf16_to_f(hi[q]) + (double)bf16_to_f(lo[q]); }
};

template <class Mat>
__global__ void gemv_t_kernel(int K, int N, Mat A, long lda, const double* __restrict__ x, double* __restrict__ y,
                              double alpha, double beta, int kchunk, double* __restrict__ part) {
  // Block (256 x 4) handles 256 columns and the k slice [by*kchunk, (by+1)*kchunk);
  // its four rows of threads stride that slice and are summed through LDS.
  // With part != nullptr the block writes its slice sum to part[by][n].
  __shared__ double red[4][256];
  const int n = blockIdx.x * 256 + threadIdx.x;
  const long k0 = (long)blockIdx.y * kchunk;
  const long k1 = min((long)K, k0 + kchunk);
  double acc = 0.0;
  if (n < N)
    for (long k = k0 + threadIdx.y; k < k1; k += blockDim.y) acc += A(k * lda + n) * x[k];
  red[threadIdx.y][threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.y == 0 && n < N) {
    double s = 0.0;
    for (int j = 0; j < (int)blockDim.y; ++j) s += red[j][threadIdx.x];
    if (part)
      part[(long)blockIdx.y * N + n] = s;
    else
      y[n] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[n]);
  }
}

// gemv_t on the bf16 hi/lo split with four columns per thread (8-byte loads
// of each half; rows 8-byte aligned): block (64 x 4) covers 256 columns, the
// same k slices and the same order of additions as gemv_t_kernel.
__global__ void gemv_t_split4_kernel(int K, int N, const unsigned short* __restrict__ hi,
                                     const unsigned short* __restrict__ lo, long lda, const double* __restrict__ x,
                                     double* __restrict__ y, double alpha, double beta, int kchunk,
                                     double* __restrict__ part) {
  __shared__ double red[4][256];
  const int n0 = blockIdx.x * 256 + 4 * threadIdx.x;
  const long k0 = (long)blockIdx.y * kchunk;
  const long k1 = min((long)K, k0 + kchunk);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (n0 + 3 < N) {
    for (long k = k0 + threadIdx.y; k < k1; k += blockDim.y) {
      const ushort4 h = *reinterpret_cast<const ushort4*>(hi + k * lda + n0);
      const ushort4 l = *reinterpret_cast<const ushort4*>(lo + k * lda + n0);
      const double xk = x[k];
      acc[0] += ((double)bf16_to_f(h.x) + (double)bf16_to_f(l.x)) * xk;
      acc[1] += ((double)bf16_to_f(h.y) + (double)bf16_to_f(l.y)) * xk;
      acc[2] += ((double)bf16_to_f(h.z) + (double)bf16_to_f(l.z)) * xk;
      acc[3] += ((double)bf16_to_f(h.w) + (double)bf16_to_f(l.w)) * xk;
    }
  } else {
    for (int u = 0; u < 4; ++u)
      if (n0 + u < N)
        for (long k = k0 + threadIdx.y; k < k1; k += blockDim.y)
          acc[u] += ((double)bf16_to_f(hi[k * lda + n0 + u]) + (double)bf16_to_f(lo[k * lda + n0 + u])) * x[k];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) red[threadIdx.y][4 * threadIdx.x + u] = acc[u];
  __syncthreads();
  const int t = threadIdx.y * 64 + threadIdx.x;
  const int n = blockIdx.x * 256 + t;
  if (n < N) {
    double sum = 0.0;
    for (int j = 0; j < 4; ++j) sum += red[j][t];
    if (part)
      part[(long)blockIdx.y * N + n] = sum;
    else
      y[n] = alpha * sum + (beta == 0.0 ? 0.0 : beta * y[n]);
  }
}

__global__ void gemv_t_finish_kernel(int N, int nsplit, const double* __restrict__ part, double* __restrict__ y,
                                     double alpha, double beta) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double s = 0.0;
  for (int j = 0; j < nsplit; ++j) s += part[(long)j * N + n];
  y[n] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[n]);
}

int gemv(int M, int K, const double* A, long lda, const double* x, double* y, double alpha, double beta,
         hipStream_t s) {
  if (M == 0) return ACES_OK;
  hipLaunchKernelGGL(gemv_kernel, dim3(grid_for(M, 4, 65536)), dim3(256), 0, s, M, K, A, lda, x, y, alpha, beta);
  ACES_HIP(hipGetLastError());
  return ACES_OK;
}

template <class Mat>
static int gemv_t_launch(int K, int N, Mat A, long lda, const double* x, double* y, double alpha, double beta,
                         hipStream_t s, double* ws, long ws_elems) {
  if (N == 0) return ACES_OK;
  const long nbx = (N + 255) / 256;
  // split k so that the launch holds about 2048 blocks (deterministic two-pass sum)
  long nsplit = 1;
  if (ws != nullptr && K > 0) {
    nsplit = std::max<long>(1, std::min<long>((2048 + nbx - 1) / nbx, (K + 255) / 256));
    nsplit = std::min<long>(nsplit, std::max<long>(1, ws_elems / N));
    nsplit = std::min<long>(nsplit, 65535);
  }
  const long kchunk = (K + nsplit - 1) / std::max<long>(nsplit, 1);
  bool launched = false;
  if constexpr (std::is_same<Mat, GemvSplit>::value) {
    if (lda % 4 == 0 && (uintptr_t)A.hi % 8 == 0 && (uintptr_t)A.lo % 8 == 0) {
      hipLaunchKernelGGL(gemv_t_split4_kernel, dim3((unsigned)nbx, (unsigned)nsplit), dim3(64, 4), 0, s, K, N, A.hi,
                         A.lo, lda, x, y, alpha, beta, (int)std::max<long>(kchunk, 1), nsplit > 1 ? ws : nullptr);
      launched = true;
    }
  }
  if (!launched)
    hipLaunchKernelGGL(gemv_t_kernel<Mat>, dim3((unsigned)nbx, (unsigned)nsplit), dim3(256, 4), 0, s, K, N, A, lda,
                       x, y, alpha, beta, (int)std::max<long>(kchunk, 1), nsplit > 1 ? ws : nullptr);
  ACES_HIP(hipGetLastError());
  if (nsplit > 1) {
    hipLaunchKernelGGL(gemv_t_finish_kernel, dim3((unsigned)nbx), dim3(256), 0, s, N, (int)nsplit, ws, y, alpha,
                       beta);
    ACES_HIP(hipGetLastError());
  }
  return ACES_OK;
}

int gemv_t(int K, int N, const double* A, long lda, const double* x, double* y, double alpha, double beta,
           hipStream_t s, double* ws, long ws_elems) {
  return gemv_t_launch(K, N, GemvF64{A}, lda, x, y, alpha, beta, s, ws, ws_elems);
}

int gemv_t_split(int K, int N, const unsigned short* hi, const unsigned short* lo, long lda, const double* x,
                 double* y, double alpha, double beta, hipStream_t s, double* ws, long ws_elems) {
  return gemv_t_launch(K, N, GemvSplit{hi, lo}, lda, x, y, alpha, beta, s, ws, ws_elems);
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
