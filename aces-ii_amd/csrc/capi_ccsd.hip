// C-ABI entry points of the closed-shell CCSD engine.
#include <new>
#include "ccsd_engine.h"
#include "common.h"

using namespace aces;

static CcsdEngine* eng(void* h) { return reinterpret_cast<CcsdEngine*>(h); }

#define REQUIRE_HANDLE(h)                       \
  do {                                          \
    if ((h) == nullptr) {                       \
      aces::set_error("null engine handle");   \
      return ACES_ERR_ARG;                \
    }                                           \
  } while (0)

extern "C" {

int aces_ccsd_create(int o, int v, void* stream, void** handle) {
  if (handle == nullptr || o < 1 || v < 1) {
    set_error("aces_ccsd_create: need o >= 1, v >= 1 and a handle pointer");
    return ACES_ERR_ARG;
  }
  CcsdEngine* e = new (std::nothrow) CcsdEngine(o, v, (hipStream_t)stream);
  if (!e) {
    set_error("aces_ccsd_create: out of host memory");
    return ACES_ERR_NOMEM;
  }
  *handle = e;
  return ACES_OK;
}

int aces_ccsd_destroy(void* handle) {
  delete eng(handle);
  return ACES_OK;
}

int aces_ccsd_setup(void* h, const double* fock, const double* ovov, const double* oovv, const double* ovoo,
                    const double* oooo, const double* ovvv, const double* list233) {
  REQUIRE_HANDLE(h);
  return eng(h)->setup(fock, ovov, oovv, ovoo, oooo, ovvv, list233);
}

int aces_ccsd_set_amplitudes(void* h, const double* t1, const double* t2) {
  REQUIRE_HANDLE(h);
  return eng(h)->set_amplitudes(t1, t2);
}

int aces_ccsd_get_amplitudes(void* h, double* t1, double* t2) {
  REQUIRE_HANDLE(h);
  return eng(h)->get_amplitudes(t1, t2);
}

int aces_ccsd_start_mp1(void* h) {
  REQUIRE_HANDLE(h);
  return eng(h)->start_mp1();
}

int aces_ccsd_reset_anchor(void* h) {
  REQUIRE_HANDLE(h);
  eng(h)->reset_anchor();
  return ACES_OK;
}

int aces_ccsd_increments(void* h, int prec, double* inc1, double* inc2) {
  REQUIRE_HANDLE(h);
  return eng(h)->increments((GemmPrec)prec, inc1, inc2);
}

int aces_ccsd_cycle(void* h, int icycle, int prec, double tol, double* stats) {
  REQUIRE_HANDLE(h);
  CcsdStats st;
  const int rc = eng(h)->cycle(icycle, (GemmPrec)prec, tol, st);
  if (rc == ACES_OK && stats) {
    stats[0] = st.energy_new;
    stats[1] = st.resmiter;
    stats[2] = st.resmdiis;
    stats[3] = st.converged;
    stats[4] = st.energy_final;
    stats[5] = st.ms_increments;
    stats[6] = st.ms_cycle;
  }
  return rc;
}

int aces_ccsd_set_diis_order(void* h, int order) {
  REQUIRE_HANDLE(h);
  return eng(h)->set_diis_order(order);
}

int aces_ccsd_reset_diis(void* h) {
  REQUIRE_HANDLE(h);
  return eng(h)->reset_diis();
}

int aces_ccsd_energy(void* h, const double* t1, const double* t2, double* e) {
  REQUIRE_HANDLE(h);
  return eng(h)->energy_of(t1, t2, *e);
}

long aces_ccsd_device_bytes(void* h) { return h ? eng(h)->device_bytes() : 0; }

int aces_ccsd_set_profiling(void* h, int on) {
  REQUIRE_HANDLE(h);
  eng(h)->set_profiling(on != 0);
  return ACES_OK;
}

int aces_ccsd_profile(void* h, double* out, int reset) {
  REQUIRE_HANDLE(h);
  eng(h)->profile(out, reset);
  return ACES_OK;
}

int aces_ccsd_set_fock(void* h, const double* fock) {
  REQUIRE_HANDLE(h);
  return eng(h)->set_fock(fock);
}

int aces_ccsd_adjoint(void* h, const double* l1, const double* l2, double* r1, double* r2, double* dens) {
  REQUIRE_HANDLE(h);
  return eng(h)->adjoint(l1, l2, r1, r2, dens);
}

int aces_ccsd_lambda_start(void* h) {
  REQUIRE_HANDLE(h);
  return eng(h)->lambda_start();
}

int aces_ccsd_lambda_cycle(void* h, int icycle, double tol, double* stats) {
  REQUIRE_HANDLE(h);
  CcsdStats st;
  const int rc = eng(h)->lambda_cycle(icycle, tol, st);
  if (rc == ACES_OK && stats) {
    stats[0] = 0.0;
    stats[1] = st.resmiter;
    stats[2] = st.resmdiis;
    stats[3] = st.converged;
    stats[4] = 0.0;
    stats[5] = st.ms_increments;
    stats[6] = st.ms_cycle;
  }
  return rc;
}

int aces_ccsd_set_lambda_precision(void* h, int prec) {
  REQUIRE_HANDLE(h);
  return eng(h)->set_lambda_precision(prec);
}

int aces_ccsd_set_lambda_diis_order(void* h, int order) {
  REQUIRE_HANDLE(h);
  return eng(h)->set_lambda_diis_order(order);
}

int aces_ccsd_get_lambda(void* h, double* l1, double* l2, int aces_norm) {
  REQUIRE_HANDLE(h);
  return eng(h)->get_lambda(l1, l2, aces_norm);
}

int aces_ccsd_set_lambda(void* h, const double* l1, const double* l2, int aces_norm) {
  REQUIRE_HANDLE(h);
  return eng(h)->set_lambda(l1, l2, aces_norm);
}

int aces_ccsd_density(void* h, double* dens) {
  REQUIRE_HANDLE(h);
  return eng(h)->density(dens);
}

int aces_ccsd_eom(void* h, const int* orb_irrep, int irrep, int nroots, double tol, int maxit, double* omega,
                  double* resid, int* iters, double* d_vecs) {
  REQUIRE_HANDLE(h);
  if (omega == nullptr) {
    set_error("aces_ccsd_eom: null output");
    return ACES_ERR_ARG;
  }
  return eng(h)->eom(orb_irrep, irrep, nroots, tol, maxit, omega, resid, iters, d_vecs);
}

int aces_ccsd_triples(void* h, int prec, double* out, double* d_per_triple) {
  REQUIRE_HANDLE(h);
  if (out == nullptr) {
    set_error("aces_ccsd_triples: null output");
    return ACES_ERR_ARG;
  }
  if (prec != GP_F64 && prec != GP_BF16X3) {
    set_error("aces_ccsd_triples: prec must be 0 (fp64) or 2 (bf16x3)");
    return ACES_ERR_ARG;
  }
  return eng(h)->triples((GemmPrec)prec, out, d_per_triple);
}

}  // extern "C"
