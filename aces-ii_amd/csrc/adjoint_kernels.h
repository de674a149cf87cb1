// HBM-bound kernels of the engine's reverse pass (see adjoint_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace aces {

// Adjoint of unpack_ladder: gZp/gZm (rows [(a<=b) | (k,b)] resp. [(a<b) | (k,b)],
// columns (i<=j) resp. (i<j)) from the increment gradient g2[i][j][a][b] and the
// gradient of Yb[b][i][j][k].
int ladder_unpack_adjoint(int o, int v, const double* g2, const double* gYb, double* gZp, long ldzp, double* gZm,
                          long ldzm, hipStream_t s);
// Adjoint of pack_tau: gtau[i][j][c][d] += contributions of gSp, gDp.
int pack_tau_adjoint(int o, int v, const double* gSp, long lds, const double* gDp, long ldd, double* gtau,
                     hipStream_t s);
// X[p][j] *= sc (cols = 0) or X[j][p] *= sc (cols = 1) for the diagonal pairs
// p = (c, c), c < v, of a packed a <= b index; j < n.
int scale_diag_pairs(int v, double* X, long ld, long n, int cols, double sc, hipStream_t s);
// Seeds of the reverse pass for L = E + l.(inc - D*T):
//   g1 = 2 f_ia - D_ia l1,  g2 = -D_ijab l2  (fov == nullptr: without the energy's 2 f_ia).
int adjoint_seed(int o, int v, const double* eo, const double* ev, const double* fov, long ldf, const double* l1,
                 const double* l2, double* g1, double* g2, hipStream_t s);
// Orbital-energy derivatives of -l.(D*T):  geo[i] += dL/de_i, gev[a] += dL/de_a.
int denominator_adjoint(int o, int v, const double* l1, const double* l2, const double* t1, const double* t2,
                        double* geo, double* gev, hipStream_t s);
// Jacobi step of the multipliers in the reference's normalisation:
//   L1n = L1 + g1/(2 D1),  L2n = L2 + (2 g2 + g2(i<->j)) / (3 D2).
int lambda_jacobi(int o, int v, const double* eo, const double* ev, const double* L, const double* g, double* Ln,
                  hipStream_t s);
// dens (n x n) = symmetrised [gfo + diag(geo), gfov ; 0, gfv + diag(gev)].
int assemble_density(int o, int v, const double* gfo, const double* gfv, const double* gfov, const double* geo,
                     const double* gev, double* dens, hipStream_t s);

}  // namespace aces
