// C-ABI entry points of the irrep-blocked CCSD engine (sym_engine.h).
#include <new>
#include "common.h"
#include "sym_engine.h"

using namespace aces;

static SymEngine* sym(void* h) { return reinterpret_cast<SymEngine*>(h); }

#define REQUIRE_SYM(h)                          \
  do {                                          \
    if ((h) == nullptr) {                       \
      aces::set_error("null engine handle");    \
      return ACES_ERR_ARG;                      \
    }                                           \
  } while (0)

static void fill_stats(const CcsdStats& st, double* stats) {
  if (!stats) return;
  stats[0] = st.energy_new;
  stats[1] = st.resmiter;
  stats[2] = st.resmdiis;
  stats[3] = st.converged;
  stats[4] = st.energy_final;
  stats[5] = st.ms_increments;
  stats[6] = st.ms_cycle;
}

extern "C" {

int aces_sym_create(int nirr, const int* occ, const int* vir, void* stream, void** handle) {
  if (handle == nullptr || occ == nullptr || vir == nullptr || !(nirr == 1 || nirr == 2 || nirr == 4 || nirr == 8)) {
    set_error("aces_sym_create: nirr must be 1, 2, 4 or 8, with occupied/virtual counts per irrep");
    return ACES_ERR_ARG;
  }
  SymEngine* e = new (std::nothrow) SymEngine(nirr, occ, vir, (hipStream_t)stream);
  if (!e) {
    set_error("aces_sym_create: out of host memory");
    return ACES_ERR_NOMEM;
  }
  *handle = e;
  return ACES_OK;
}

int aces_sym_destroy(void* h) {
  delete sym(h);
  return ACES_OK;
}

int aces_sym_setup(void* h, const double* fock, const double* ovov, const double* oovv, const double* ovoo,
                   const double* oooo, const double* ovvv, const double* list233) {
  REQUIRE_SYM(h);
  return sym(h)->setup(fock, ovov, oovv, ovoo, oooo, ovvv, list233);
}

int aces_sym_set_amplitudes(void* h, const double* t1, const double* t2) {
  REQUIRE_SYM(h);
  return sym(h)->set_amplitudes(t1, t2);
}

int aces_sym_get_amplitudes(void* h, double* t1, double* t2) {
  REQUIRE_SYM(h);
  return sym(h)->get_amplitudes(t1, t2);
}

int aces_sym_start_mp1(void* h) {
  REQUIRE_SYM(h);
  return sym(h)->start_mp1();
}

int aces_sym_increments(void* h, int prec, double* inc1, double* inc2) {
  REQUIRE_SYM(h);
  return sym(h)->increments((GemmPrec)prec, inc1, inc2);
}

int aces_sym_cycle(void* h, int icycle, int prec, double tol, double* stats) {
  REQUIRE_SYM(h);
  CcsdStats st;
  const int rc = sym(h)->cycle(icycle, (GemmPrec)prec, tol, st);
  if (rc == ACES_OK) fill_stats(st, stats);
  return rc;
}

int aces_sym_set_diis_order(void* h, int order) {
  REQUIRE_SYM(h);
  return sym(h)->set_diis_order(order);
}

int aces_sym_reset_diis(void* h) {
  REQUIRE_SYM(h);
  return sym(h)->reset_diis();
}

int aces_sym_reset_anchor(void* h) {
  REQUIRE_SYM(h);
  sym(h)->reset_anchor();
  return ACES_OK;
}

int aces_sym_energy(void* h, double* e) {
  REQUIRE_SYM(h);
  return sym(h)->energy_current(*e);
}

int aces_sym_adjoint(void* h, const double* l1, const double* l2, double* r1, double* r2, double* dens) {
  REQUIRE_SYM(h);
  return sym(h)->adjoint(l1, l2, r1, r2, dens);
}

int aces_sym_lambda_start(void* h) {
  REQUIRE_SYM(h);
  return sym(h)->lambda_start();
}

int aces_sym_set_lambda_precision(void* h, int prec) {
  REQUIRE_SYM(h);
  return sym(h)->set_lambda_precision(prec);
}

int aces_sym_lambda_cycle(void* h, int icycle, double tol, double* stats) {
  REQUIRE_SYM(h);
  CcsdStats st;
  const int rc = sym(h)->lambda_cycle(icycle, tol, st);
  if (rc == ACES_OK) fill_stats(st, stats);
  return rc;
}

int aces_sym_set_lambda_diis_order(void* h, int order) {
  REQUIRE_SYM(h);
  return sym(h)->set_lambda_diis_order(order);
}

int aces_sym_get_lambda(void* h, double* l1, double* l2, int aces_norm) {
  REQUIRE_SYM(h);
  return sym(h)->get_lambda(l1, l2, aces_norm);
}

int aces_sym_density(void* h, double* dens) {
  REQUIRE_SYM(h);
  return sym(h)->density(dens);
}

int aces_sym_set_profiling(void* h, int on) {
  REQUIRE_SYM(h);
  sym(h)->set_profiling(on != 0);
  return ACES_OK;
}

int aces_sym_profile(void* h, double* out, int reset) {
  REQUIRE_SYM(h);
  sym(h)->profile(out, reset);
  return ACES_OK;
}

long aces_sym_device_bytes(void* h) { return h ? sym(h)->device_bytes() : 0; }

long aces_sym_packed_t2(void* h) { return h ? sym(h)->packed_t2() : 0; }

}  // extern "C"
