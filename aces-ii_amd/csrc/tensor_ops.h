// HBM-bound tensor kernels: conversions, permute/combine, packing, reductions.
#pragma once
#include <hip/hip_runtime.h>

namespace aces {

// y[r*ldy + c] = (float) x[r*ldx + c] for c < cols; columns cols..ldy-1 zeroed.
int convert_f32(int rows, int cols, const double* x, long ldx, float* y, long ldy, hipStream_t s);
// hi = bf16(x), lo = bf16(x - hi) (round to nearest even); padding columns zeroed.
// With x0 the split is of the difference x - x0 (x0 laid out like x).
int split_bf16(int rows, int cols, const double* x, long ldx, unsigned short* hi, unsigned short* lo,
               long ldy, hipStream_t s, const double* x0 = nullptr);
// The same conversions for an operand stored k-major, x[c*ldx + r] = element
// (r, c): the output is the transposed, k-contiguous y[r*ldy + c].
// The blocks of one irrep-blocked tensor converted in one launch: block g is
// rows[g] x cols[g] at x[g] (leading dimension cols[g]), written with
// leading dimension ldy[g] (columns past cols[g] zero) as bf16 hi/lo
// (split, x - x0 when x0[g] is given), as fp32 (hi holds the floats) or as
// fp64 (hi holds the padded copy).
struct SplitBlocks {
  int n = 0;
  bool f32 = false;
  bool f64 = false;
  int rows[8], cols[8];
  long ldy[8];
  const double* x[8];
  const double* x0[8];
  void* hi[8];
  void* lo[8];
};
int split_blocks(const SplitBlocks& b, hipStream_t s);
int convert_f32_t(int rows, int cols, const double* x, long ldx, float* y, long ldy, hipStream_t s);
// y[r*ldy + c] = x[c*ldx + r] (r < rows, c < cols), zero for cols <= c < ldy
int copy_f64_t(int rows, int cols, const double* x, long ldx, double* y, long ldy, hipStream_t s);
int split_bf16_t(int rows, int cols, const double* x, long ldx, unsigned short* hi, unsigned short* lo, long ldy,
                 hipStream_t s, const double* x0 = nullptr);

// Generic 4-index permute/combine on a row-major output of dims d[0..3]
// (up to kP4Src sources):
//   out[i] = beta*out[i] + sum_s alpha[s] * in[s][ sum_k i_k * st[s][k] ]
//            + gamma * X[ sum_k i_k * xs[k] ] * Y[ sum_k i_k * ys[k] ]
// Unused sources have in[s] == nullptr; the product term is skipped when X == nullptr.
constexpr int kP4Src = 10;  // sources per permute
struct Permute4 {
  long d[4] = {1, 1, 1, 1};
  double* out = nullptr;
  double beta = 0.0;
  const double* in[kP4Src] = {};
  double alpha[kP4Src] = {};
  long st[kP4Src][4] = {};
  const double* X = nullptr;
  const double* Y = nullptr;
  double gamma = 0.0;
  long xs[4] = {0, 0, 0, 0};
  long ys[4] = {0, 0, 0, 0};
};
int permute4(const Permute4& p, hipStream_t s);

// y[m] = alpha * sum_k A[m*lda + k] x[k] + beta * y[m]
int gemv(int M, int K, const double* A, long lda, const double* x, double* y, double alpha, double beta,
         hipStream_t s);
// y[n] = alpha * sum_k A[k*lda + n] x[k] + beta * y[n]   (column access, coalesced over n).
// With a workspace (ws, ws_elems doubles) k is split over blocks and the slice
// sums are added in a second, deterministic pass.
int gemv_t(int K, int N, const double* A, long lda, const double* x, double* y, double alpha, double beta,
           hipStream_t s, double* ws = nullptr, long ws_elems = 0);
// gemv_t on a bf16 hi/lo split matrix (element = hi + lo, fp64 accumulation).
int gemv_t_split(int K, int N, const unsigned short* hi, const unsigned short* lo, long lda, const double* x,
                 double* y, double alpha, double beta, hipStream_t s, double* ws, long ws_elems);

// Two-index strided contraction over a 4-index array:
//   out[p*P1 + q] = beta*out + alpha * sum_{r<R, s<S} X[p*xp + q*xq + r*xr + s*xs] * Y[r*yr + s*ys]
int contract2(int P, int Q, int R, int S, const double* X, long xp, long xq, long xr, long xs,
              const double* Y, long yr, long ys, double* out, double alpha, double beta, hipStream_t s);

// Partial-sum reductions into device scalars (results left in `result`).
// dot = sum x[i]*y[i]
int dot(long n, const double* x, const double* y, double* result, hipStream_t s);

}  // namespace aces
