// Small dense nonsymmetric eigenproblem (host code), see small_eig.h.
#include <cfloat>
#include <cmath>
#include <vector>
#include "common.h"
#include "engine_util.h"
#include "small_eig.h"

namespace aces {

namespace {

// Householder vector of a 2- or 3-vector x: (I - beta v v^T) x = (alpha, 0, 0).
void house(int m, const double* x, double* v, double& beta) {
  double nrm = 0.0;
  for (int i = 0; i < m; ++i) nrm += x[i] * x[i];
  nrm = std::sqrt(nrm);
  for (int i = 0; i < m; ++i) v[i] = x[i];
  if (nrm == 0.0) {
    beta = 0.0;
    return;
  }
  const double alpha = (x[0] >= 0.0) ? -nrm : nrm;
  v[0] = x[0] - alpha;
  double vv = 0.0;
  for (int i = 0; i < m; ++i) vv += v[i] * v[i];
  beta = (vv == 0.0) ? 0.0 : 2.0 / vv;
}

// Eigenvalues of the real 2 x 2 block [[a, b], [c, d]].
void eig2(double a, double b, double c, double d, double& r1, double& i1, double& r2, double& i2) {
  const double p = 0.5 * (a - d), disc = p * p + b * c, mid = 0.5 * (a + d);
  if (disc >= 0.0) {
    const double z = std::sqrt(disc);
    r1 = mid + (p >= 0.0 ? z : -z);
    const double det = a * d - b * c;
    r2 = (r1 != 0.0) ? det / r1 : mid - (p >= 0.0 ? z : -z);
    i1 = i2 = 0.0;
  } else {
    const double z = std::sqrt(-disc);
    r1 = r2 = mid;
    i1 = z;
    i2 = -z;
  }
}

}  // namespace

int eig_nonsym(int n, const double* A, double* wr, double* wi) {
  ACES_REQUIRE(n >= 1, "eig_nonsym: n >= 1");
  std::vector<double> H(A, A + (size_t)n * n);
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * n + j]; };
  double norm = 0.0;
  for (double x : H) norm += std::fabs(x);
  if (norm == 0.0) {
    for (int i = 0; i < n; ++i) wr[i] = wi[i] = 0.0;
    return ACES_OK;
  }
  // Householder reduction to upper Hessenberg form
  std::vector<double> v(n);
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += h(i, k) * h(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (h(k + 1, k) > 0.0) alpha = -alpha;
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    v[k + 1] = h(k + 1, k) - alpha;
    for (int i = k + 2; i < n; ++i) v[i] = h(i, k);
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double f = 2.0 / vn;
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * h(i, j);
      s *= f;
      for (int i = k + 1; i < n; ++i) h(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += h(i, j) * v[j];
      s *= f;
      for (int j = k + 1; j < n; ++j) h(i, j) -= s * v[j];
    }
    for (int i = k + 2; i < n; ++i) h(i, k) = 0.0;
  }
  // Francis double-shift QR on the active block lo..hi
  int hi = n - 1, iter = 0, total = 0;
  while (hi >= 0) {
    int lo = hi;
    while (lo > 0) {
      double s = std::fabs(h(lo - 1, lo - 1)) + std::fabs(h(lo, lo));
      if (s == 0.0) s = norm;
      if (std::fabs(h(lo, lo - 1)) <= DBL_EPSILON * s) {
        h(lo, lo - 1) = 0.0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      wr[hi] = h(hi, hi);
      wi[hi] = 0.0;
      --hi;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {
      eig2(h(hi - 1, hi - 1), h(hi - 1, hi), h(hi, hi - 1), h(hi, hi), wr[hi - 1], wi[hi - 1], wr[hi], wi[hi]);
      hi -= 2;
      iter = 0;
      continue;
    }
    ACES_REQUIRE(++total <= 60 * n, "eig_nonsym: QR iteration did not converge");
    ++iter;
    double s, t;
    if (iter % 10 == 0) {  // exceptional shift
      const double e = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
      s = 1.5 * e;
      t = e * e;
    } else {
      s = h(hi - 1, hi - 1) + h(hi, hi);
      t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
    }
    double x = h(lo, lo) * h(lo, lo) + h(lo, lo + 1) * h(lo + 1, lo) - s * h(lo, lo) + t;
    double y = h(lo + 1, lo) * (h(lo, lo) + h(lo + 1, lo + 1) - s);
    double z = h(lo + 1, lo) * h(lo + 2, lo + 1);
    const int nb = hi - lo + 1;
    for (int kk = 0; kk + 2 < nb; ++kk) {
      const int r = lo + kk;  // the reflector acts on rows/columns r .. r + 2
      const double xv[3] = {x, y, z};
      double vv[3], beta;
      house(3, xv, vv, beta);
      if (beta != 0.0) {
        const int c0 = std::max(lo, r - 1);
        for (int j = c0; j <= hi; ++j) {
          const double sum = beta * (vv[0] * h(r, j) + vv[1] * h(r + 1, j) + vv[2] * h(r + 2, j));
          h(r, j) -= sum * vv[0];
          h(r + 1, j) -= sum * vv[1];
          h(r + 2, j) -= sum * vv[2];
        }
        const int r1 = std::min(r + 3, hi);
        for (int i = lo; i <= r1; ++i) {
          const double sum = beta * (h(i, r) * vv[0] + h(i, r + 1) * vv[1] + h(i, r + 2) * vv[2]);
          h(i, r) -= sum * vv[0];
          h(i, r + 1) -= sum * vv[1];
          h(i, r + 2) -= sum * vv[2];
        }
      }
      x = h(r + 1, r);
      y = h(r + 2, r);
      z = (kk + 3 < nb) ? h(r + 3, r) : 0.0;
    }
    {
      const double xv[2] = {x, y};
      double vv[2], beta;
      house(2, xv, vv, beta);
      if (beta != 0.0) {
        for (int j = hi - 2; j <= hi; ++j) {
          const double sum = beta * (vv[0] * h(hi - 1, j) + vv[1] * h(hi, j));
          h(hi - 1, j) -= sum * vv[0];
          h(hi, j) -= sum * vv[1];
        }
        for (int i = lo; i <= hi; ++i) {
          const double sum = beta * (h(i, hi - 1) * vv[0] + h(i, hi) * vv[1]);
          h(i, hi - 1) -= sum * vv[0];
          h(i, hi) -= sum * vv[1];
        }
      }
    }
  }
  return ACES_OK;
}

int eigvec_real(int n, const double* A, double w, double* x) {
  const double shift = w + 1e-10 * std::max(1.0, std::fabs(w));
  std::vector<double> M(A, A + (size_t)n * n);
  for (int i = 0; i < n; ++i) M[(size_t)i * n + i] -= shift;
  std::vector<double> y(n, 1.0);
  for (int i = 0; i < n; ++i) y[i] = 1.0 + 0.01 * i;  // generic start
  for (int it = 0; it < 4; ++it) {
    std::vector<double> b = y;
    ACES_CHECK(eu::solve_small_lu(n, M, b));
    double nrm = 0.0;
    for (double q : b) nrm += q * q;
    nrm = std::sqrt(nrm);
    ACES_REQUIRE(nrm > 0.0 && std::isfinite(nrm), "eigvec_real: inverse iteration failed");
    for (int i = 0; i < n; ++i) y[i] = b[i] / nrm;
  }
  // sign: largest component positive
  int im = 0;
  for (int i = 1; i < n; ++i)
    if (std::fabs(y[i]) > std::fabs(y[im])) im = i;
  const double sg = y[im] < 0.0 ? -1.0 : 1.0;
  for (int i = 0; i < n; ++i) x[i] = sg * y[i];
  return ACES_OK;
}

}  // namespace aces
