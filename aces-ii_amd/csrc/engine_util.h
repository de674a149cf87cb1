// Internal helpers shared by the engine's forward and adjoint sources.
#pragma once
#include "gemm.h"
#include "tensor_ops.h"
#include <cmath>
#include <utility>
#include <vector>
#include "common.h"

namespace aces {
namespace eu {

// Builder for one Permute4 launch (tensor_ops.h): up to kP4Src strided sources
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
    if (ns >= kP4Src) {  // reported by run()
      ++ns;
      return *this;
    }
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
  int run(hipStream_t s) {
    ACES_REQUIRE(ns <= kP4Src, "permute4: too many sources");
    return permute4(p, s);
  }
};

// low-precision operand rows are padded (with zeros) to whole 32-wide k tiles
inline long pad32(long x) { return (x + 31) / 32 * 32; }

inline GemmDesc gd(int M, int N, int K, const double* A, long lda, int akc, const double* B, long ldb, int bkc, double* C,
            long ldc, double alpha, double beta) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K;
  d.A = A; d.lda = lda; d.a_kc = akc;
  d.B = B; d.ldb = ldb; d.b_kc = bkc;
  d.C = C; d.ldc = ldc; d.alpha = alpha; d.beta = beta;
  return d;
}

// LU with partial pivoting (dgesv semantics) for the small DIIS system.
inline int solve_small_lu(int n, std::vector<double> A, std::vector<double>& b) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > best) {
        best = std::fabs(A[i * n + k]);
        p = i;
      }
    if (best == 0.0) return ACES_ERR_SOLVE;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      A[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
  return ACES_OK;
}
}  // namespace eu
}  // namespace aces
