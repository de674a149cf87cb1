// C-ABI entry points for the raw kernels (GEMM engine and operand conversion).
#include "common.h"
#include "gemm.h"
#include "small_eig.h"
#include "tensor_ops.h"

using namespace aces;

extern "C" {

const char* aces_last_error(void) { return aces::last_error(); }

int aces_gemm(int prec, int M, int N, int K, const void* A, const void* A_lo, long lda, int a_kc, const void* B,
              const void* B_lo, long ldb, int b_kc, double* C, long ldc, double alpha, double beta, int batch,
              long sA, long sB, long sC, void* stream) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K;
  d.A = A; d.A_lo = A_lo; d.lda = lda; d.a_kc = a_kc;
  d.B = B; d.B_lo = B_lo; d.ldb = ldb; d.b_kc = b_kc;
  d.C = C; d.ldc = ldc; d.alpha = alpha; d.beta = beta;
  d.batch = batch; d.sA = sA; d.sB = sB; d.sC = sC;
  return gemm(d, (GemmPrec)prec, (hipStream_t)stream);
}

int aces_gemm_grouped(int prec, int ngroups, const int* M, const int* N, const int* K, const void* const* A,
                      const void* const* A_lo, const long* lda, int a_kc, const void* const* B,
                      const void* const* B_lo, const long* ldb, int b_kc, double* const* C, const long* ldc,
                      double alpha, double beta, double* ws, long ws_elems, void* stream) {
  if (ngroups < 1 || ngroups > 8) {
    set_error("aces_gemm_grouped: 1..8 groups");
    return ACES_ERR_ARG;
  }
  GemmDesc d[8];
  for (int g = 0; g < ngroups; ++g) {
    d[g].M = M[g]; d[g].N = N[g]; d[g].K = K[g];
    d[g].A = A[g]; d[g].A_lo = A_lo ? A_lo[g] : nullptr; d[g].lda = lda[g]; d[g].a_kc = a_kc;
    d[g].B = B[g]; d[g].B_lo = B_lo ? B_lo[g] : nullptr; d[g].ldb = ldb[g]; d[g].b_kc = b_kc;
    d[g].C = C[g]; d[g].ldc = ldc[g]; d[g].alpha = alpha; d[g].beta = beta;
    d[g].ws = ws; d[g].ws_elems = ws_elems;
  }
  return gemm_grouped(d, ngroups, (GemmPrec)prec, (hipStream_t)stream);
}

int aces_convert_f32(int rows, int cols, const double* x, long ldx, float* y, long ldy, void* stream) {
  return convert_f32(rows, cols, x, ldx, y, ldy, (hipStream_t)stream);
}

int aces_split_bf16(int rows, int cols, const double* x, long ldx, unsigned short* hi, unsigned short* lo,
                    long ldy, void* stream) {
  return split_bf16(rows, cols, x, ldx, hi, lo, ldy, (hipStream_t)stream);
}

int aces_small_eig(int n, const double* A, double* wr, double* wi) {
  if (n < 1 || !A || !wr || !wi) {
    set_error("aces_small_eig: bad arguments");
    return ACES_ERR_ARG;
  }
  return eig_nonsym(n, A, wr, wi);
}

int aces_small_eigvec(int n, const double* A, double w, double* x) {
  if (n < 1 || !A || !x) {
    set_error("aces_small_eigvec: bad arguments");
    return ACES_ERR_ARG;
  }
  return eigvec_real(n, A, w, x);
}

}  // extern "C"
