// Eigenvalues and eigenvectors of a small real nonsymmetric matrix (the
// projected matrix of the EOM-CC Davidson solver), host code.
#pragma once

namespace aces {

// All eigenvalues of the n x n row-major matrix A: Householder reduction to
// upper Hessenberg form, then the implicit double-shift QR iteration of
// Francis (Golub & Van Loan, Matrix Computations, 4th ed., algorithms 7.4.2
// and 7.5.2) with deflation on negligible subdiagonal elements.  wr/wi: real
// and imaginary parts; complex eigenvalues come in conjugate pairs.
int eig_nonsym(int n, const double* A, double* wr, double* wi);

// Unit eigenvector x of A for its real eigenvalue w, by inverse iteration
// with the shift w (1 + 1e-10) on the LU factors of A - shift I.
int eigvec_real(int n, const double* A, double w, double* x);

}  // namespace aces
