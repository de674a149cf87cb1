// CCSD-specific HBM-bound kernels (see ccsd_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace aces {

// Ladder operands Sp / Dp from tau (ccsd_kernels.hip): fp64 rows and/or the
// bf16 hi/lo split rows, wherever the pointers are set.
struct PackTauOut {
  double* S = nullptr;
  long lds = 0;
  double* D = nullptr;
  long ldd = 0;
  unsigned short* Shi = nullptr;
  unsigned short* Slo = nullptr;
  long ldsl = 0;
  unsigned short* Dhi = nullptr;
  unsigned short* Dlo = nullptr;
  long lddl = 0;
};
int pack_tau(int o, int v, const double* tau, const PackTauOut& out, hipStream_t s);
int build_ladder_operands(int o, int v, const double* L233, const double* ovvv, double* Vp, long ldp, double* Vm,
                          long ldm, hipStream_t s);
int unpack_ladder(int o, int v, const double* Zp, long ldzp, const double* Zm, long ldzm, double* inc2, double* Yb,
                  hipStream_t s);
// Ring operands U, X, Ut, Xt ([(i,a)][(j,b)]) from t1 and t2 in one pass
// (ccsd_kernels.hip); index q = 0 U, 1 X, 2 Ut, 3 Xt; null pointers skip.
// The same pass writes tau = t2 + t1 t1 ([i][j][a][b]) and, when t2hi is set,
// the bf16 split of t2 itself (rows (i,j,a), ld = ldt).
struct RingOperands {
  double* f64[4] = {nullptr, nullptr, nullptr, nullptr};  // ld = o v
  unsigned short* hi[4] = {nullptr, nullptr, nullptr, nullptr};
  unsigned short* lo[4] = {nullptr, nullptr, nullptr, nullptr};
  long ldl = 0;  // bf16 row length (>= o v, padding untouched)
  double* tau = nullptr;
  unsigned short* t2hi = nullptr;
  unsigned short* t2lo = nullptr;
  long ldt = 0;  // >= v
  // up to three bf16 splits of tau, element (i,j,a,b) at i*tsi + j*tsj + a*tsa + b
  unsigned short* tauhi[3] = {nullptr, nullptr, nullptr};
  unsigned short* taulo[3] = {nullptr, nullptr, nullptr};
  long tsi[3] = {0, 0, 0}, tsj[3] = {0, 0, 0}, tsa[3] = {0, 0, 0};
};
int ring_operands(int o, int v, const double* t1, const double* t2, const RingOperands& r, hipStream_t s);
int jacobi(int o, int v, const double* eo, const double* ev, const double* inc1, const double* inc2, double* t1n,
           double* t2n, hipStream_t s);
int symmetrize_t2(int o, int v, double* t2, hipStream_t s);
// Jacobi update, energy of the update (before symmetrization) and
// SYMETRIZE_T2 of t2n in one pass (ccsd_kernels.hip); part holds
// jacobi_energy_scratch(o, v) doubles, *out receives the energy.
long jacobi_energy_scratch(int o, int v);
int jacobi_energy_sym(int o, int v, const double* eo, const double* ev, const double* inc1, const double* inc2,
                      const double* fov, long ldf, const double* Lijab, double* t1n, double* t2n, double* part,
                      double* out, hipStream_t s);
int energy(int o, int v, const double* fov, long ldf, const double* Lijab, const double* t1, const double* t2,
           double* part, double* out, hipStream_t s);
int diis_dots(int o, int v, int m, const double* const* d_xs, const double* y, double* part, double* out,
              hipStream_t s);
int diis_absmax(int o, int v, const double* e, void* part, double* out, hipStream_t s);
// sub + diis_absmax + the solution copy + diis_dots of one DODIIS0 step in
// one pass (ccsd_kernels.hip): e_out = xn - x, sol_out = x, out_dots[k] =
// <x_k, e> (d_xs[self] unused: e itself), *out_absmax = ISAMAX value of e.
int diis_error(int o, int v, const double* x, const double* xn, int m, const double* const* d_xs, int self,
               double* e_out, double* sol_out, void* part, double* out_dots, double* out_absmax, hipStream_t s);
long reduction_scratch_bytes();
int lincomb(long n, int m, const double* const* d_xs, const double* d_c, double* out, hipStream_t s);
int sub(long n, const double* x, const double* z, double* y, hipStream_t s);
int sub_diag(int n, double* X, long ld, const double* d, hipStream_t s);

}  // namespace aces
