// Internal helpers shared by the engine's forward and adjoint sources.
#pragma once
#include "gemm.h"
#include "tensor_ops.h"

namespace aces {
namespace eu {

// Builder for one Permute4 launch (tensor_ops.h): up to three strided sources
// and one product term, combined into a row-major 4-index output.
struct P4 {
  Permute4 p;
  int ns = 0;
  P4(long a, long b, long c, long d, double* out, double beta = 0.0) {
    p.d[0] = a; p.d[1] = b; p.d[2] = c; p.d[3] = d;
    p.out = out;
    p.beta = beta;
  }
  P4& src(const double* x, double al, long s0, long s1, long s2, long s3) {
    p.in[ns] = x;
    p.alpha[ns] = al;
    p.st[ns][0] = s0; p.st[ns][1] = s1; p.st[ns][2] = s2; p.st[ns][3] = s3;
    ++ns;
    return *this;
  }
  P4& prod(const double* X, long x0, long x1, long x2, long x3, const double* Y, long y0, long y1, long y2, long y3,
           double g) {
    p.X = X; p.Y = Y; p.gamma = g;
    p.xs[0] = x0; p.xs[1] = x1; p.xs[2] = x2; p.xs[3] = x3;
    p.ys[0] = y0; p.ys[1] = y1; p.ys[2] = y2; p.ys[3] = y3;
    return *this;
  }
  int run(hipStream_t s) { return permute4(p, s); }
};

inline GemmDesc gd(int M, int N, int K, const double* A, long lda, int akc, const double* B, long ldb, int bkc, double* C,
            long ldc, double alpha, double beta) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K;
  d.A = A; d.lda = lda; d.a_kc = akc;
  d.B = B; d.ldb = ldb; d.b_kc = bkc;
  d.C = C; d.ldc = ldc; d.alpha = alpha; d.beta = beta;
  return d;
}

}  // namespace eu
}  // namespace aces
