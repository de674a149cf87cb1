// EOM-CCSD excitation energies of a closed-shell reference (singlets), matrix
// free, on the engine's resident intermediates.
//
// For converged amplitudes the CC Jacobian J = dOmega/dT (Omega the residual
// whose zero gives T) equals the similarity-transformed Hamiltonian minus the
// CC energy on the singles and doubles space, Hbar_{mu nu} - E_CC delta_{mu nu}
// (<mu|[Hbar, tau_nu]|0> with <kappa|Hbar|0> = 0 for every excitation kappa),
// so its eigenvalues are the EOM-CCSD excitation energies.  In the closed-shell
// spatial parametrisation (t1[i][a], pair-symmetric t2[i][j][a][b]) they are
// the singlet ones.  The reverse pass of the increments without the energy
// seeds (compute_adjoint(lin)) is the product J^T l, so the solver works on the
// left eigenvectors (same eigenvalues as the right ones; the reference's xvee
// solves both, vee/doeomee_david.F): a nonsymmetric Davidson iteration
// (vee/nextdav.F plays this part there) with
//  - an orthonormal basis V (rows) and images S = J^T V, the projected matrix
//    G = V S^T kept on the host and extended by one row and column per vector;
//  - Ritz pairs of G by eig_nonsym and inverse iteration, ordered by real part
//    (a complex pair enters once, through its real part, until it splits);
//  - corrections r / (w - d), r = J^T x - w x, with the diagonal estimate d:
//    singles e_a - e_i + 2 (ia|ia) - (ii|aa) (the CIS diagonal), doubles
//    e_a + e_b - e_i - e_j; restricted to the target irrep and to
//    pair-symmetric doubles;
//  - guesses: the unit excitations (singles, and pair-symmetric doubles built
//    from the lowest singles) of the target irrep with the smallest d;
//  - a thick restart when the basis is full: the Ritz vectors of the lowest
//    max(nroots, maxdim / 2) values, orthonormalised in coefficient space
//    (V is orthonormal, so the overlaps of X = Y V are Y Y^T) and formed in
//    place (each column of V and S is read whole before it is written).
#include <algorithm>
#include <cmath>
#include <vector>
#include "ccsd_engine.h"
#include "ccsd_kernels.h"
#include "common.h"
#include "small_eig.h"
#include "tensor_ops.h"

namespace aces {

namespace {

constexpr int kDotBlk = 256;
constexpr int kMaxDim = 64;  // largest Davidson basis (rows of the in-place restart transform)

// part[m * kDotBlk + b] = partial sums of M[m] . x over block b's chunk
__global__ void rowdots_kernel(long n, const double* __restrict__ M, const double* __restrict__ x,
                               double* __restrict__ part) {
  __shared__ double sh[256];
  const int m = blockIdx.y;
  const double* row = M + (long)m * n;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc += row[i] * x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)m * kDotBlk + blockIdx.x] = sh[0];
}

__global__ void rowdots_finish_kernel(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[kDotBlk];
  sh[threadIdx.x] = part[(long)blockIdx.x * kDotBlk + threadIdx.x];
  __syncthreads();
  for (int w = kDotBlk / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// mask[x] = 1 where excitation x (singles [i][a], then doubles [i][j][a][b])
// has irrep `target` (irr: orbital irreps, occupied then virtual; nullptr: all
// ones); diag[x] = the diagonal estimate (W0t[ia][jb] = 2 (ia|jb) - (ij|ab))
__global__ void eom_setup_kernel(int o, int v, const int* __restrict__ irr, int target, const double* __restrict__ eo,
                                 const double* __restrict__ ev, const double* __restrict__ W0t,
                                 double* __restrict__ mask, double* __restrict__ diag) {
  const long ov = (long)o * v, n2 = ov * ov;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < ov + n2; x += (long)gridDim.x * blockDim.x) {
    int i, j, a, b;
    if (x < ov) {
      i = (int)(x / v);
      a = (int)(x % v);
      j = b = -1;
    } else {
      const long y = x - ov;
      b = (int)(y % v);
      a = (int)((y / v) % v);
      const long ij = y / ((long)v * v);
      j = (int)(ij % o);
      i = (int)(ij / o);
    }
    double m = 1.0;
    if (irr) {
      const int g = (x < ov) ? (irr[i] ^ irr[o + a]) : (irr[i] ^ irr[j] ^ irr[o + a] ^ irr[o + b]);
      m = (g == target) ? 1.0 : 0.0;
    }
    mask[x] = m;
    if (x < ov)
      diag[x] = ev[a] - eo[i] + W0t[x * ov + x];
    else
      diag[x] = ev[a] + ev[b] - eo[i] - eo[j];
  }
}

// out = mask * r / (w - diag) (denominators clamped away from zero)
__global__ void eom_precond_kernel(long n, const double* __restrict__ diag, const double* __restrict__ mask,
                                   const double* __restrict__ r, double w, double* __restrict__ out) {
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    double den = w - diag[x];
    if (fabs(den) < 1e-3) den = (den < 0.0) ? -1e-3 : 1e-3;
    out[x] = mask[x] * r[x] / den;
  }
}

__global__ void axpby_kernel(long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

// M[r][c] = sum_j C[r][j] M[j][c] for r < nk, in place: each column c is read
// whole (k <= kMaxDim rows) before its first write.
__global__ void rows_transform_kernel(long n, int k, int nk, const double* __restrict__ C, double* __restrict__ M) {
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
    double col[kMaxDim];
#pragma unroll 8
    for (int j = 0; j < k; ++j) col[j] = M[(long)j * n + c];
    for (int r = 0; r < nk; ++r) {
      double acc = 0.0;
      for (int j = 0; j < k; ++j) acc += C[r * k + j] * col[j];
      M[(long)r * n + c] = acc;
    }
  }
}

unsigned grid_of(long n) { return (unsigned)std::min<long>((n + 255) / 256, 1L << 16); }

}  // namespace

int CcsdEngine::eom(const int* orb_irrep, int target, int nroots, double tol, int maxit, double* omega,
                    double* resid, int* iters, double* d_vecs) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_REQUIRE(nroots >= 1 && nroots <= 16 && maxit >= 1 && tol > 0.0, "eom: nroots 1..16, maxit >= 1, tol > 0");
  ACES_REQUIRE(orb_irrep == nullptr || (target >= 0 && target < 8), "eom: target irrep 0..7");
  ACES_CHECK(adjoint_alloc());
  if (!fwd_valid_) ACES_CHECK(compute_increments(GP_F64));
  const long ov = (long)o * v, nt = ov + ov * ov;
  hipStream_t s = s_;
  const int maxdim = std::min(kMaxDim, std::max(4 * nroots + 8, 24));
  const int nkeep = std::max(nroots, maxdim / 2);
  // device work space for the call: V and S (maxdim x nt), mask, diagonal,
  // residual and spare vectors, dot partials and coefficients
  struct Work {
    double* p = nullptr;
    ~Work() {
      if (p) (void)hipFree(p);
    }
  } work;
  const long wsz = 2L * maxdim * nt + 4 * nt + (long)maxdim * kDotBlk + (long)maxdim * maxdim;
  if (hipMalloc(&work.p, sizeof(double) * wsz) != hipSuccess) {
    (void)hipGetLastError();
    set_error("eom: no device memory for the Davidson subspace");
    return ACES_ERR_NOMEM;
  }
  double* V = work.p;
  double* S = V + (long)maxdim * nt;
  double* mask = S + (long)maxdim * nt;
  double* diag = mask + nt;
  double* rvec = diag + nt;
  double* wvec = rvec + nt;
  double* part = wvec + nt;
  double* small = part + (long)maxdim * kDotBlk;  // maxdim^2 coefficients
  // mask and diagonal
  int* d_irr = nullptr;
  if (orb_irrep) {
    ACES_HIP(hipMalloc(&d_irr, sizeof(int) * n));
    ACES_HIP(hipMemcpyAsync(d_irr, orb_irrep, sizeof(int) * n, hipMemcpyHostToDevice, s));
  }
  hipLaunchKernelGGL(eom_setup_kernel, dim3(grid_of(nt)), dim3(256), 0, s, o, v, d_irr, target, eo_, ev_, W0t_, mask,
                     diag);
  ACES_HIP(hipGetLastError());
  ACES_HIP(hipStreamSynchronize(s));
  if (d_irr) (void)hipFree(d_irr);

  // J^T x (pair-symmetric part)
  auto apply = [&](const double* x, double* y) -> int {
    ACES_HIP(hipMemcpyAsync(lam_, x, sizeof(double) * nt, hipMemcpyDeviceToDevice, s));
    ACES_CHECK(compute_adjoint(lam_, false, GP_F64, true));
    ACES_CHECK(symmetrize_t2(o, v, gT_ + ov, s));
    ACES_HIP(hipMemcpyAsync(y, gT_, sizeof(double) * nt, hipMemcpyDeviceToDevice, s));
    return ACES_OK;
  };
  // dots of the first k rows of M with x, on the host
  auto rowdots = [&](const double* M, int k, const double* x, std::vector<double>& out) -> int {
    out.assign((size_t)k, 0.0);
    if (k == 0) return ACES_OK;
    hipLaunchKernelGGL(rowdots_kernel, dim3(kDotBlk, k), dim3(256), 0, s, nt, M, x, part);
    ACES_HIP(hipGetLastError());
    hipLaunchKernelGGL(rowdots_finish_kernel, dim3(k), dim3(kDotBlk), 0, s, part, small);
    ACES_HIP(hipGetLastError());
    ACES_HIP(hipMemcpyAsync(out.data(), small, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    ACES_HIP(hipStreamSynchronize(s));
    return ACES_OK;
  };
  // out = sum_i c_i M[i] (+ beta out)
  auto combine = [&](const double* M, int k, const std::vector<double>& c, double* out, double beta) -> int {
    ACES_HIP(hipMemcpyAsync(small, c.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    ACES_HIP(hipStreamSynchronize(s));  // c is pageable and may not outlive the call
    return gemv_t(k, (int)nt, M, nt, small, out, 1.0, beta, s, nullptr, 0);
  };
  auto scale = [&](double* x, double a) -> int {
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_of(nt)), dim3(256), 0, s, nt, 0.0, x, a, x);
    ACES_HIP(hipGetLastError());
    return ACES_OK;
  };

  // projected matrix G[i][j] = V_i . S_j (host, row-major with leading dimension maxdim)
  std::vector<double> G((size_t)maxdim * maxdim, 0.0);
  int k = 0;
  // append row k of V (orthonormal to the rows before it): image, then row and column k of G
  auto push = [&]() -> int {
    ACES_CHECK(apply(V + (long)k * nt, S + (long)k * nt));
    std::vector<double> col, row;
    ACES_CHECK(rowdots(V, k + 1, S + (long)k * nt, col));  // V_i . S_k
    ACES_CHECK(rowdots(S, k, V + (long)k * nt, row));      // S_j . V_k
    for (int i = 0; i <= k; ++i) G[(size_t)i * maxdim + k] = col[(size_t)i];
    for (int j = 0; j < k; ++j) G[(size_t)k * maxdim + j] = row[(size_t)j];
    ++k;
    return ACES_OK;
  };

  // guesses: unit singles and pair-symmetric doubles of the target irrep with the smallest diagonal
  std::vector<double> hmask1(ov), hdiag1(ov), eo(o), ev(v);
  ACES_HIP(hipMemcpyAsync(hmask1.data(), mask, sizeof(double) * ov, hipMemcpyDeviceToHost, s));
  ACES_HIP(hipMemcpyAsync(hdiag1.data(), diag, sizeof(double) * ov, hipMemcpyDeviceToHost, s));
  ACES_HIP(hipMemcpyAsync(eo.data(), eo_, sizeof(double) * o, hipMemcpyDeviceToHost, s));
  ACES_HIP(hipMemcpyAsync(ev.data(), ev_, sizeof(double) * v, hipMemcpyDeviceToHost, s));
  ACES_HIP(hipStreamSynchronize(s));
  struct Guess {
    double d;
    long x, y;  // element(s) set: y < 0 for one element
  };
  std::vector<Guess> cand;
  for (long x = 0; x < ov; ++x)
    if (hmask1[(size_t)x] != 0.0) cand.push_back({hdiag1[(size_t)x], x, -1});
  {
    // doubles from pairs of the lowest singles (any irrep) by orbital-energy difference
    std::vector<std::pair<double, long>> low;
    for (long x = 0; x < ov; ++x) low.push_back({ev[(size_t)(x % v)] - eo[(size_t)(x / v)], x});
    std::sort(low.begin(), low.end());
    const size_t m = std::min<size_t>(low.size(), 48);
    for (size_t p = 0; p < m; ++p)
      for (size_t q = p; q < m; ++q) {
        const long ia = low[p].second, jb = low[q].second;
        const int i = (int)(ia / v), a = (int)(ia % v), j = (int)(jb / v), b = (int)(jb % v);
        const long x = ov + (((long)i * o + j) * v + a) * v + b;
        const long y = ov + (((long)j * o + i) * v + b) * v + a;
        if (orb_irrep && ((orb_irrep[i] ^ orb_irrep[j] ^ orb_irrep[o + a] ^ orb_irrep[o + b]) != target)) continue;
        cand.push_back({low[p].first + low[q].first, x, x == y ? -1 : y});
      }
  }
  ACES_REQUIRE(!cand.empty(), "eom: the target irrep has no excitation to start from");
  std::sort(cand.begin(), cand.end(), [](const Guess& p, const Guess& q) { return p.d < q.d; });
  const int nguess = (int)std::min<size_t>(cand.size(), (size_t)std::min(maxdim / 2, nroots + 3));
  ACES_REQUIRE(nguess >= nroots, "eom: fewer start vectors in the target irrep than roots requested");
  for (int g = 0; g < nguess; ++g) {
    double* row = V + (long)k * nt;
    ACES_HIP(hipMemsetAsync(row, 0, sizeof(double) * nt, s));
    const double val[1] = {cand[(size_t)g].y < 0 ? 1.0 : std::sqrt(0.5)};
    ACES_HIP(hipMemcpyAsync(row + cand[(size_t)g].x, val, sizeof(double), hipMemcpyHostToDevice, s));
    if (cand[(size_t)g].y >= 0)
      ACES_HIP(hipMemcpyAsync(row + cand[(size_t)g].y, val, sizeof(double), hipMemcpyHostToDevice, s));
    ACES_HIP(hipStreamSynchronize(s));
    ACES_CHECK(push());
  }

  std::vector<double> w(nroots, 0.0), rn(nroots, 1e300);
  std::vector<std::vector<double>> Y(nroots);
  int it = 0;
  bool done = false;
  for (; it < maxit && !done; ++it) {
    // Ritz values of G (order k0) by real part; a complex pair counts once
    const int k0 = k;
    std::vector<double> Gk((size_t)k0 * k0);
    for (int i = 0; i < k0; ++i)
      for (int j = 0; j < k0; ++j) Gk[(size_t)i * k0 + j] = G[(size_t)i * maxdim + j];
    std::vector<double> wr(k0), wi(k0);
    ACES_CHECK(eig_nonsym(k0, Gk.data(), wr.data(), wi.data()));
    std::vector<double> ritz;
    for (int q = 0; q < k0; ++q)
      if (wi[q] >= 0.0 || std::fabs(wi[q]) <= 1e-10 * std::max(1.0, std::fabs(wr[q]))) ritz.push_back(wr[q]);
    std::sort(ritz.begin(), ritz.end());
    ACES_REQUIRE((int)ritz.size() >= nroots, "eom: fewer Ritz values than roots requested");
    done = true;
    int added = 0;
    for (int r = 0; r < nroots; ++r) {
      w[r] = ritz[(size_t)r];
      Y[r].assign((size_t)k0, 0.0);
      ACES_CHECK(eigvec_real(k0, Gk.data(), w[r], Y[r].data()));
      // residual J^T x - w x with x = V y
      ACES_CHECK(combine(S, k0, Y[r], rvec, 0.0));
      ACES_CHECK(combine(V, k0, Y[r], wvec, 0.0));
      hipLaunchKernelGGL(axpby_kernel, dim3(grid_of(nt)), dim3(256), 0, s, nt, -w[r], wvec, 1.0, rvec);
      ACES_HIP(hipGetLastError());
      std::vector<double> nn;
      ACES_CHECK(rowdots(rvec, 1, rvec, nn));
      rn[r] = std::sqrt(std::max(nn[0], 0.0));
      if (rn[r] < tol) continue;
      done = false;
      if (k >= maxdim) continue;
      double* x = V + (long)k * nt;
      hipLaunchKernelGGL(eom_precond_kernel, dim3(grid_of(nt)), dim3(256), 0, s, nt, diag, mask, rvec, w[r], x);
      ACES_HIP(hipGetLastError());
      ACES_CHECK(symmetrize_t2(o, v, x + ov, s));
      // orthonormalise against the basis (two Gram-Schmidt passes)
      std::vector<double> c;
      for (int pass = 0; pass < 2; ++pass) {
        ACES_CHECK(rowdots(V, k, x, c));
        for (double& q : c) q = -q;
        ACES_CHECK(combine(V, k, c, x, 1.0));
      }
      std::vector<double> xx;
      ACES_CHECK(rowdots(x, 1, x, xx));
      const double nrm = std::sqrt(std::max(xx[0], 0.0));
      if (nrm < 1e-10) continue;
      ACES_CHECK(scale(x, 1.0 / nrm));
      ACES_CHECK(push());
      ++added;
    }
    if (done) break;
    if (added > 0) continue;
    // thick restart: Ritz vectors of the lowest nk values, orthonormalised in
    // coefficient space, formed in place in V and S; G becomes Q G Q^T
    const int nk = std::min<int>((int)ritz.size(), nkeep);
    std::vector<std::vector<double>> Q;
    for (int r = 0; r < nk; ++r) {
      std::vector<double> y(k0);
      ACES_CHECK(eigvec_real(k0, Gk.data(), ritz[(size_t)r], y.data()));
      for (int pass = 0; pass < 2; ++pass)
        for (const auto& q : Q) {
          double d = 0.0;
          for (int j = 0; j < k0; ++j) d += q[(size_t)j] * y[(size_t)j];
          for (int j = 0; j < k0; ++j) y[(size_t)j] -= d * q[(size_t)j];
        }
      double nrm = 0.0;
      for (double q : y) nrm += q * q;
      nrm = std::sqrt(nrm);
      if (nrm < 1e-6) continue;
      for (double& q : y) q /= nrm;
      Q.push_back(std::move(y));
    }
    const int kq = (int)Q.size();
    ACES_REQUIRE(kq >= 1, "eom: restart lost the basis");
    std::vector<double> C((size_t)kq * k0);
    for (int r = 0; r < kq; ++r)
      for (int j = 0; j < k0; ++j) C[(size_t)r * k0 + j] = Q[(size_t)r][(size_t)j];
    ACES_HIP(hipMemcpyAsync(small, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rows_transform_kernel, dim3(grid_of(nt)), dim3(256), 0, s, nt, k0, kq, small, V);
    ACES_HIP(hipGetLastError());
    hipLaunchKernelGGL(rows_transform_kernel, dim3(grid_of(nt)), dim3(256), 0, s, nt, k0, kq, small, S);
    ACES_HIP(hipGetLastError());
    ACES_HIP(hipStreamSynchronize(s));
    std::vector<double> GQ((size_t)k0 * kq, 0.0);  // G Q^T
    for (int i = 0; i < k0; ++i)
      for (int r = 0; r < kq; ++r) {
        double acc = 0.0;
        for (int j = 0; j < k0; ++j) acc += Gk[(size_t)i * k0 + j] * C[(size_t)r * k0 + j];
        GQ[(size_t)i * kq + r] = acc;
      }
    std::fill(G.begin(), G.end(), 0.0);
    for (int p = 0; p < kq; ++p)
      for (int r = 0; r < kq; ++r) {
        double acc = 0.0;
        for (int i = 0; i < k0; ++i) acc += C[(size_t)p * k0 + i] * GQ[(size_t)i * kq + r];
        G[(size_t)p * maxdim + r] = acc;
      }
    k = kq;
  }
  for (int r = 0; r < nroots; ++r) {
    omega[r] = w[r];
    if (resid) resid[r] = rn[r];
  }
  if (d_vecs) {
    // the Ritz vectors of the last step; a restart after it changed the basis, so recompute them
    const int k0 = k;
    std::vector<double> Gk((size_t)k0 * k0);
    for (int i = 0; i < k0; ++i)
      for (int j = 0; j < k0; ++j) Gk[(size_t)i * k0 + j] = G[(size_t)i * maxdim + j];
    for (int r = 0; r < nroots; ++r) {
      std::vector<double> y((size_t)k0);
      ACES_CHECK(eigvec_real(k0, Gk.data(), w[r], y.data()));
      ACES_CHECK(combine(V, k0, y, d_vecs + (long)r * nt, 0.0));
    }
  }
  if (iters) *iters = it;
  ACES_HIP(hipStreamSynchronize(s));
  fwd_valid_ = false;  // the reverse passes share scratch with the forward pass
  return ACES_OK;
}

}  // namespace aces
