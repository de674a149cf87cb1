// Closed-shell CCSD amplitude engine (see ccsd_engine.h).
//
// Spatial-orbital restatement of the increments ACES II accumulates on its
// alpha-beta lists (vcc/vcc.F:811-1085).  Chemists' notation (pq|rs);
// t1[i][a]; t2[i][j][a][b] = t_{Ij}^{Ab}; u = 2 t2 - t2(a<->b).
//
// Dominant contractions (each a single MFMA GEMM launch):
//   ladder   Z(+/-) = [V(+/-) ; G(+/-)] . tau(+/-)       symmetric/antisymmetric
//            pair packing of <ab|cd>, plus the (kb) rows of (kc|bd) that carry
//            the T1 dressing of W(abef) (DRLAD/LADAB1 vcc/ladab1.F, T1INW1 vcc/t1inw1.f)
//   W build  Wt = 1/2 (u - 2 t1 t1) . L,  Wv = -1/2 (t^x + 2 t1 t1) . K
//            (DWMBEJ vcc/dwmbej.F; list 56 = 2 W(MbEj) - W(MbeJ))
//   ring     R1 = Wt . u,  R2 = Wv . t^x     (DRRNG vcc/drrng.f, RNGPRD vcc/rngprd.f)
#include <algorithm>
#include <cmath>
#include <cstring>
#include "ccsd_engine.h"
#include "ccsd_kernels.h"
#include "common.h"
#include "tensor_ops.h"
#include "engine_util.h"

namespace aces {

namespace {
using eu::P4;
using eu::gd;

using eu::pad32;

}  // namespace

CcsdEngine::CcsdEngine(int o_, int v_, hipStream_t s) : o(o_), v(v_), n(o_ + v_), s_(s) {
  nvp_ = (long)v * (v + 1) / 2;
  nvm_ = (long)v * (v - 1) / 2;
  nop_ = (long)o * (o + 1) / 2;
  nom_ = (long)o * (o - 1) / 2;
}

CcsdEngine::~CcsdEngine() {
  for (void* p : allocs_) (void)hipFree(p);
}

int CcsdEngine::alloc_raw(void** p, long bytes) {
  if (bytes <= 0) bytes = 8;
  hipError_t e = hipMalloc(p, (size_t)bytes);
  if (e != hipSuccess) {
    char b[256];
    snprintf(b, sizeof(b), "hipMalloc of %ld bytes failed (%s); engine holds %ld bytes", bytes,
             hipGetErrorString(e), bytes_);
    set_error(b);
    return ACES_ERR_NOMEM;
  }
  allocs_.push_back(*p);
  bytes_ += bytes;
  return ACES_OK;
}

int CcsdEngine::alloc(double** p, long count) { return alloc_raw((void**)p, count * (long)sizeof(double)); }

int CcsdEngine::setup(const double* fock, const double* ovov, const double* oovv, const double* ovoo,
                      const double* oooo, const double* ovvv, const double* L233) {
  ACES_REQUIRE(o >= 1 && v >= 1, "ccsd setup: need o >= 1 and v >= 1");
  ACES_REQUIRE(!ready_, "ccsd setup: already set up");
  const long ov = (long)o * v, oo = (long)o * o, vv = (long)v * v;
  const long ov2 = ov * v, o2v = oo * v, o3 = oo * o, ovvo = ov * ov;
  hipStream_t s = s_;

  ACES_CHECK(alloc(&fock_, (long)n * n));
  ACES_HIP(hipMemcpyAsync(fock_, fock, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  ACES_CHECK(alloc(&eo_, o));
  ACES_CHECK(alloc(&ev_, v));
  ACES_CHECK(P4(1, 1, 1, o, eo_).src(fock_, 1, 0, 0, 0, (long)n + 1).run(s));
  ACES_CHECK(P4(1, 1, 1, v, ev_).src(fock_ + (long)o * n + o, 1, 0, 0, 0, (long)n + 1).run(s));

  ACES_CHECK(alloc(&ovov_, ovvo));
  ACES_CHECK(alloc(&oovv_, ovvo));
  ACES_CHECK(alloc(&ovoo_, o3 * v));
  ACES_CHECK(alloc(&ovvv_, ov * vv));
  ACES_HIP(hipMemcpyAsync(ovov_, ovov, sizeof(double) * ovvo, hipMemcpyDeviceToDevice, s));
  ACES_HIP(hipMemcpyAsync(oovv_, oovv, sizeof(double) * ovvo, hipMemcpyDeviceToDevice, s));
  ACES_HIP(hipMemcpyAsync(ovoo_, ovoo, sizeof(double) * o3 * v, hipMemcpyDeviceToDevice, s));
  ACES_HIP(hipMemcpyAsync(ovvv_, ovvv, sizeof(double) * ov * vv, hipMemcpyDeviceToDevice, s));

  // strides of the native blocks
  const long OV[4] = {ov2, ov, v, 1};          // ovov[i][a][j][b]
  const long OO4[4] = {o3, oo, o, 1};          // oooo
  (void)OV;

  ACES_CHECK(alloc(&Lmat_, ovvo));  // [(k,c)][(l,d)] = 2(kc|ld) - (kd|lc)
  ACES_CHECK(P4(o, v, o, v, Lmat_).src(ovov_, 2, ov2, ov, v, 1).src(ovov_, -1, ov2, 1, v, ov).run(s));
  ACES_CHECK(alloc(&Kx_, ovvo));  // (kd|lc)
  ACES_CHECK(P4(o, v, o, v, Kx_).src(ovov_, 1, ov2, 1, v, ov).run(s));
  ACES_CHECK(alloc(&W0t_, ovvo));  // [(i,a)][(k,c)] = 2(ia|kc) - (ik|ac)
  ACES_CHECK(P4(o, v, o, v, W0t_).src(ovov_, 2, ov2, ov, v, 1).src(oovv_, -1, ov2, v, vv, 1).run(s));
  ACES_CHECK(alloc(&Wv0_, ovvo));  // (ki|ac)
  ACES_CHECK(P4(o, v, o, v, Wv0_).src(oovv_, 1, vv, v, ov2, 1).run(s));
  ACES_CHECK(alloc(&Lo_, ovvo));  // [k][l][c][d] = L_kcld
  ACES_CHECK(P4(o, o, v, v, Lo_).src(ovov_, 2, ov2, v, ov, 1).src(ovov_, -1, ov2, v, 1, ov).run(s));
  ACES_CHECK(alloc(&Lv_, ovvo));  // [c][k][l][d] = L_kcld
  ACES_CHECK(P4(v, o, o, v, Lv_).src(ovov_, 2, ov, ov2, v, 1).src(ovov_, -1, 1, ov2, v, ov).run(s));
  ACES_CHECK(alloc(&OVkl_, ovvo));  // [k][l][c][d] = (kc|ld)
  ACES_CHECK(P4(o, o, v, v, OVkl_).src(ovov_, 1, ov2, v, ov, 1).run(s));
  ACES_CHECK(alloc(&Lijab_, ovvo));  // [i][j][a][b] = 2(ia|jb) - (ib|ja)
  ACES_CHECK(P4(o, o, v, v, Lijab_).src(ovov_, 2, ov2, v, ov, 1).src(ovov_, -1, ov2, v, 1, ov).run(s));
  ACES_CHECK(alloc(&ooooP_, oo * oo));  // [k][l][i][j] = (ki|lj)
  {
    double* tmp;
    ACES_CHECK(alloc(&tmp, oo * oo));
    ACES_HIP(hipMemcpyAsync(tmp, oooo, sizeof(double) * oo * oo, hipMemcpyDeviceToDevice, s));
    ACES_CHECK(P4(o, o, o, o, ooooP_).src(tmp, 1, OO4[0], OO4[2], OO4[1], OO4[3]).run(s));
  }
  const long vo2 = v * oo;
  ACES_CHECK(alloc(&ovooQ_, o3 * v));  // [k][l][i][c] = (lc|ki) = ovoo[l][c][k][i]
  ACES_CHECK(P4(o, o, o, v, ovooQ_).src(ovoo_, 1, o, vo2, 1, oo).run(s));
  ACES_CHECK(alloc(&M1_, o3 * v));  // [(k,i)][(l,c)] = 2(lc|ki) - (kc|li)
  ACES_CHECK(P4(o, o, o, v, M1_).src(ovoo_, 2, o, 1, vo2, oo).src(ovoo_, -1, vo2, 1, o, oo).run(s));
  ACES_CHECK(alloc(&M5c_, o3 * v));  // [(k,c,i)][l] = 2(kc|li) - (lc|ki)
  ACES_CHECK(P4(o, v, o, o, M5c_).src(ovoo_, 2, vo2, oo, 1, o).src(ovoo_, -1, o, oo, 1, vo2).run(s));
  ACES_CHECK(alloc(&M5b_, o3 * v));  // [(k,c,i)][l] = (lc|ki)
  ACES_CHECK(P4(o, v, o, o, M5b_).src(ovoo_, 1, o, oo, 1, vo2).run(s));
  ACES_CHECK(alloc(&M7_, o3 * v));  // [i][k][l][c] = 2(lc|ki) - (kc|li)
  ACES_CHECK(P4(o, o, o, v, M7_).src(ovoo_, 2, 1, o, vo2, oo).src(ovoo_, -1, 1, vo2, o, oo).run(s));

  const long v3 = vv * v;
  ACES_CHECK(alloc(&ovvv_kacd_, ov * vv));  // [k][a][c][d] = (kd|ac) = ovvv[k][d][a][c]
  ACES_CHECK(P4(o, v, v, v, ovvv_kacd_).src(ovvv_, 1, v3, v, 1, vv).run(s));
  ACES_CHECK(alloc(&M3_, ov * vv));  // [a][k][c][d] = 2(kd|ac) - (kc|ad)
  ACES_CHECK(P4(v, o, v, v, M3_).src(ovvv_, 2, v, v3, 1, vv).src(ovvv_, -1, v, v3, vv, 1).run(s));
  ACES_CHECK(alloc(&ovvvL_, ov * vv));  // [k][c][a][d] = 2(kc|ad) - (kd|ac)
  ACES_CHECK(P4(o, v, v, v, ovvvL_).src(ovvv_, 2, v3, vv, v, 1).src(ovvv_, -1, v3, 1, v, vv).run(s));

  // ladder operands
  ACES_CHECK(alloc(&Vp_, (nvp_ + ov) * nvp_));
  ACES_CHECK(alloc(&Vm_, (nvm_ + ov) * std::max<long>(nvm_, 1)));
  ACES_CHECK(build_ladder_operands(o, v, L233, ovvv_, Vp_, nvp_, Vm_, std::max<long>(nvm_, 1), s));

  // amplitudes: T = [t1 | t2] contiguous (DIIS vector layout)
  const long nt = ov + ovvo;
  ACES_CHECK(alloc(&t1_, nt));
  t2_ = t1_ + ov;
  ACES_CHECK(alloc(&t1n_, nt));
  t2n_ = t1n_ + ov;
  ACES_CHECK(alloc(&inc1_, nt));
  inc2_ = inc1_ + ov;
  ACES_HIP(hipMemsetAsync(t1_, 0, sizeof(double) * nt, s));
  ACES_CHECK(alloc(&tau_, ovvo));
  ACES_CHECK(alloc(&U_, ovvo));
  ACES_CHECK(alloc(&X_, ovvo));
  ACES_CHECK(alloc(&Ut_, ovvo));
  ACES_CHECK(alloc(&Xt_, ovvo));
  ACES_CHECK(alloc(&Wt_, ovvo));
  ACES_CHECK(alloc(&Wv_, ovvo));
  ACES_CHECK(alloc(&R1_, ovvo));
  ACES_CHECK(alloc(&R2_, ovvo));
  ACES_CHECK(alloc(&Sp_, nop_ * nvp_));
  ACES_CHECK(alloc(&Dp_, std::max<long>(nom_ * nvm_, 1)));
  ACES_CHECK(alloc(&Zp_, (nvp_ + ov) * nop_));
  ACES_CHECK(alloc(&Zm_, std::max<long>((nvm_ + ov) * nom_, 1)));
  ACES_CHECK(alloc(&Yb_, v * o3));
  ACES_CHECK(alloc(&L1_, v * v * oo));
  ACES_CHECK(alloc(&tauK_, ovvo));
  ACES_CHECK(alloc(&P1_, ovvo));
  ACES_CHECK(alloc(&P2_, ovvo));
  ACES_CHECK(alloc(&Q12_, ovvo));
  ACES_CHECK(alloc(&Q2_, ovvo));
  ACES_CHECK(alloc(&Qb_, o2v * o));
  ACES_CHECK(alloc(&T1Q_, o2v * v));
  ACES_CHECK(alloc(&S3_, o2v * o));
  ACES_CHECK(alloc(&G4A_, o2v * o));
  ACES_CHECK(alloc(&G4_, o2v * v));
  ACES_CHECK(alloc(&Wo_, oo * oo));
  ACES_CHECK(alloc(&Qo_, oo * oo));
  ACES_CHECK(alloc(&tmpL_, ovvo));
  ACES_CHECK(alloc(&Fov_, ov));
  ACES_CHECK(alloc(&Foo_, oo));
  ACES_CHECK(alloc(&Fvv_, vv));
  ACES_CHECK(alloc(&Loo_, oo));
  ACES_CHECK(alloc(&Lvv_, vv));
  ACES_CHECK(alloc(&FT_, oo));
  ACES_CHECK(alloc(&FovT_, oo));
  ACES_CHECK(alloc(&Gk_, oo));
  ACES_CHECK(alloc_raw((void**)&red_, reduction_scratch_bytes()));
  ACES_CHECK(alloc_raw((void**)&jpart_, sizeof(double) * jacobi_energy_scratch(o, v)));
  ws_elems_ = 1L << 25;  // split-K workspace: 32 Mi doubles
  ACES_CHECK(alloc(&ws_, ws_elems_));
  ACES_CHECK(alloc(&scal_, 64));
  ACES_CHECK(alloc(&newvec_, nt));
  ACES_CHECK(alloc_raw((void**)&d_ptrs_, 64 * sizeof(double*)));
  ACES_CHECK(alloc(&d_coef_, 64));
  ACES_CHECK(set_diis_order(tdiis_.order));
  ACES_HIP(hipStreamSynchronize(s));
  anc_valid_ = false;  // anchored products hold the previous integrals
  ready_ = true;
  return ACES_OK;
}

int CcsdEngine::diis_resize(DiisState& D, int order) {
  ACES_REQUIRE(order >= 1 && order <= 25, "DIIS order must be 1..25 (ld_diis = 25, vcc/dodiis0.F)");
  const long nt = (long)o * v + (long)o * o * v * v;
  while ((int)D.sol.size() < order) {
    double *a, *b;
    ACES_CHECK(alloc(&a, nt));
    ACES_CHECK(alloc(&b, nt));
    D.sol.push_back(a);
    D.err.push_back(b);
  }
  D.order = order;
  D.R.assign((size_t)order * order, 0.0);
  return ACES_OK;
}

int CcsdEngine::set_diis_order(int order) { return diis_resize(tdiis_, order); }

int CcsdEngine::reset_diis() {
  tdiis_.R.assign((size_t)tdiis_.order * tdiis_.order, 0.0);
  return ACES_OK;
}

int CcsdEngine::set_amplitudes(const double* t1, const double* t2) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  const long ov = (long)o * v;
  t_sym_ = false;
  tn_lists_ = false;
  ACES_HIP(hipMemcpyAsync(t1_, t1, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipMemcpyAsync(t2_, t2, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  fwd_valid_ = false;
  return ACES_OK;
}

int CcsdEngine::get_amplitudes(double* t1, double* t2) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  const long ov = (long)o * v;
  if (t1) ACES_HIP(hipMemcpyAsync(t1, t1_, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  if (t2) ACES_HIP(hipMemcpyAsync(t2, t2_, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

// MP1 guess written by xintprc (intprc/dfrmt2.F): T2 = <ij|ab>/D, T1 = 0.
int CcsdEngine::start_mp1() {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  const long ov = (long)o * v, ov2 = ov * v;
  ACES_HIP(hipMemsetAsync(inc1_, 0, sizeof(double) * ov, s_));
  ACES_CHECK(P4(o, o, v, v, inc2_).src(ovov_, 1, ov2, v, ov, 1).run(s_));
  t_sym_ = false;
  tn_lists_ = false;
  ACES_CHECK(jacobi(o, v, eo_, ev_, inc1_, inc2_, t1_, t2_, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  fwd_valid_ = false;
  return ACES_OK;
}

// Energy into a device scalar; d_out == nullptr targets the slot that
// diis_step() reports as E(T_new).
int CcsdEngine::energy_dev(const double* t1, const double* t2, double* d_out) {
  return energy(o, v, fock_ + o, n, Lijab_, t1, t2, red_, d_out ? d_out : scal_, s_);
}

int CcsdEngine::energy_of(const double* t1, const double* t2, double& e) {
  ACES_CHECK(energy(o, v, fock_ + o, n, Lijab_, t1, t2, red_, scal_, s_));
  ACES_HIP(hipMemcpyAsync(&e, scal_, sizeof(double), hipMemcpyDeviceToHost, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

// ---------------------------------------------------------------------------
// Low-precision operands.  Static slots (integral matrices) are converted
// once per precision; the other slots per call.
int CcsdEngine::lowp_static(GemmPrec prec) {
  if (lp_static_prec_ == prec) return ACES_OK;
  const long ov = (long)o * v;
  struct S { int slot; const double* src; long rows; long K; long ld; };
  const long oo = (long)o * o, vv = (long)v * v;
  const S tab[] = {{0, Vp_, nvp_ + ov, nvp_, nvp_},
                   {1, Vm_, nvm_ + ov, nvm_, std::max<long>(nvm_, 1)},
                   {2, Lmat_, ov, ov, ov},
                   {3, Kx_, ov, ov, ov},
                   {4, OVkl_, oo, vv, vv},
                   {16, ovvv_, vv * o, v, v},             // [(k,c,a)][d]
                   {17, ovvv_kacd_, vv * o, v, v},        // [(k,a,c)][d]
                   {18, M3_, v, ov * v, ov * v},          // [a][(k,c,d)]
                   {19, Lv_, v, oo * v, oo * v},          // [c][(k,l,d)]
                   {26, oovv_, oo * v, v, v},             // [(k,i,b)][c]
                   {27, ovov_, oo * v, v, v},             // [(i,a,k)][c]
                   {28, Lo_, o, ov * v, ov * v},          // [k][(l,c,d)]
                   {29, M7_, o, oo * v, oo * v},          // [i][(k,l,c)]
                   {30, ovvvL_, ov, vv, vv}};             // [(k,d)][(c,a)] (the Lvv gemv)
  for (const S& t : tab) {
    LowP& L = lp_[t.slot];
    const long ld = pad32(std::max<long>(t.K, 1));
    if (L.hi == nullptr || L.ld != ld || L.rows != t.rows) {
      void* blk;
      ACES_CHECK(alloc_raw(&blk, t.rows * ld * 4));
      L.hi = blk;
      L.lo = (char*)blk + t.rows * ld * 2;
      L.ld = ld;
      L.rows = t.rows;
    }
    if (t.K == 0) continue;
    if (prec == GP_F32)
      ACES_CHECK(convert_f32((int)t.rows, (int)t.K, t.src, t.ld, (float*)L.hi, ld, s_));
    else
      ACES_CHECK(split_bf16((int)t.rows, (int)t.K, t.src, t.ld, (unsigned short*)L.hi, (unsigned short*)L.lo, ld,
                            s_));
  }
  lp_static_prec_ = prec;
  return ACES_OK;
}

// Make operand slot `slot` hold rows x pad32(K) low-precision elements
// (allocated zero-filled, so padding columns stay zero).
int CcsdEngine::lowp_slot(int slot, long rows, long K) {
  LowP& L = lp_[slot];
  const long Kp = pad32(std::max<long>(K, 1));
  if (L.hi == nullptr || L.ld != Kp || L.rows < rows) {
    void* blk;
    ACES_CHECK(alloc_raw(&blk, rows * Kp * 4));
    ACES_HIP(hipMemsetAsync(blk, 0, rows * Kp * 4, s_));
    L.hi = blk;
    L.lo = (char*)blk + rows * Kp * 2;
    L.ld = Kp;
    L.rows = rows;
  }
  return ACES_OK;
}

int CcsdEngine::gemm64(GemmDesc d) {
  d.ws = ws_;
  d.ws_elems = ws_elems_;
  return timed_gemm(d, GP_F64, 2.0 * d.M * (double)d.N * d.K * d.batch);
}

int CcsdEngine::set_fock(const double* fock) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_HIP(hipMemcpyAsync(fock_, fock, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s_));
  ACES_CHECK(P4(1, 1, 1, o, eo_).src(fock_, 1, 0, 0, 0, (long)n + 1).run(s_));
  ACES_CHECK(P4(1, 1, 1, v, ev_).src(fock_ + (long)o * n + o, 1, 0, 0, 0, (long)n + 1).run(s_));
  ACES_HIP(hipStreamSynchronize(s_));
  fwd_valid_ = anc_valid_ = false;
  return ACES_OK;
}

void CcsdEngine::profile(double* out, int reset) {
  for (int i = 0; i < 3 * kProfCats; ++i) out[i] = prof_[i];
  if (reset)
    for (int i = 0; i < 3 * kProfCats; ++i) prof_[i] = 0.0;
}

int CcsdEngine::gemm_lp(GemmPrec prec, int M, int N, int K, const double* A64, int aslot, long lda,
                        const double* B64, int bslot, long ldb, double* C, long ldc, double alpha, double beta,
                        int a_kc, int b_kc, int cslot) {
  // the eight large products (profile categories ladder, W build, ring) of a
  // refinement cycle; everything else of such a cycle is plain fp64
  if (anc_mode_ != 0 && prec == GP_F64 && prof_cat_ <= 2)
    return gemm_anchored(M, N, K, A64, aslot, lda, B64, bslot, ldb, C, ldc, alpha, beta, a_kc, b_kc);
  return gemm_lp_impl(prec, M, N, K, A64, aslot, lda, B64, bslot, ldb, C, ldc, alpha, beta, a_kc, b_kc, cslot);
}

// Keep (or refresh) the anchor copy of an amplitude-dependent operand buffer.
// An operand buffer holds the same values at every anchored product that
// reads it within one evaluation, so one copy per buffer serves them all.
int CcsdEngine::anchor_copy(const double* src, long elems, const double** copy) {
  AncSrc* e = nullptr;
  for (AncSrc& a : anc_srcs_)
    if (a.src == src) e = &a;
  if (e == nullptr) {
    ACES_REQUIRE(anc_mode_ == 1, "refinement: operand without an anchor copy");
    anc_srcs_.push_back(AncSrc());
    e = &anc_srcs_.back();
    e->src = src;
  }
  if (e->copy == nullptr || e->elems < elems) {
    ACES_REQUIRE(anc_mode_ == 1, "refinement: anchor copy smaller than the operand");
    ACES_CHECK(alloc(&e->copy, elems));
    e->elems = elems;
  }
  if (anc_mode_ == 1)
    ACES_HIP(hipMemcpyAsync(e->copy, src, sizeof(double) * elems, hipMemcpyDeviceToDevice, s_));
  *copy = e->copy;
  return ACES_OK;
}

// Split src (minus sub, when given) into operand slot `slot`: rows x K,
// k-contiguous (kc) or stored k-major.
int CcsdEngine::split_operand(LowPOp& op, int slot, const double* src, const double* sub, long rows, int K, long ld,
                              int kc) {
  ACES_CHECK(lowp_slot(slot, rows, K));
  LowP& L = lp_[slot];
  if (kc)
    ACES_CHECK(split_bf16((int)rows, K, src, ld, (unsigned short*)L.hi, (unsigned short*)L.lo, L.ld, s_, sub));
  else
    ACES_CHECK(split_bf16_t((int)rows, K, src, ld, (unsigned short*)L.hi, (unsigned short*)L.lo, L.ld, s_, sub));
  lp_fresh_[slot] = false;  // holds a difference, not the operand
  op.hi = L.hi;
  op.lo = L.lo;
  op.ld = L.ld;
  return ACES_OK;
}

int CcsdEngine::lp_product(int M, int N, int K, const LowPOp& A, const LowPOp& B, double* C, long ldc, double alpha,
                           double beta) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = (int)pad32(std::max(K, 1));
  d.A = A.hi; d.A_lo = A.lo; d.lda = A.ld; d.a_kc = 1;
  d.B = B.hi; d.B_lo = B.lo; d.ldb = B.ld; d.b_kc = 1;
  d.C = C; d.ldc = ldc; d.alpha = alpha; d.beta = beta;
  d.ws = ws_;
  d.ws_elems = ws_elems_;
  return timed_gemm(d, GP_BF16X3, 2.0 * M * (double)N * K);
}

int CcsdEngine::gemm_anchored(int M, int N, int K, const double* A64, int aslot, long lda, const double* B64,
                              int bslot, long ldb, double* C, long ldc, double alpha, double beta, int a_kc,
                              int b_kc) {
  const bool sa = lp_static_slot(aslot), sb = lp_static_slot(bslot);
  ACES_REQUIRE(aslot >= 0 && aslot < kLpSlots && bslot >= 0 && bslot < kLpSlots, "gemm_anchored: bad operand slot");
  if ((sa && sb) || ldc != N || M == 0 || N == 0 || K == 0)
    return gemm_lp_impl(GP_F64, M, N, K, A64, aslot, lda, B64, bslot, ldb, C, ldc, alpha, beta, a_kc, b_kc, -1);
  const int site = anc_site_++;
  if (anc_mode_ == 2)
    ACES_REQUIRE(site < (int)anc_sites_.size() && anc_sites_[site].elems == (long)M * N,
                 "refinement: products differ from the anchor's");
  if (site >= (int)anc_sites_.size()) anc_sites_.resize(site + 1);
  AncSite& S = anc_sites_[site];
  // extents of the operands as stored: rows x K (kc) or K x rows (k-major)
  const long ea = a_kc ? (M - 1) * lda + K : (K - 1) * lda + M;
  const long eb = b_kc ? (N - 1) * ldb + K : (K - 1) * ldb + N;
  const double *A0 = nullptr, *B0 = nullptr;
  if (anc_mode_ == 1) {
    if (S.C0 == nullptr || S.elems != (long)M * N) {
      ACES_CHECK(alloc(&S.C0, (long)M * N));
      S.elems = (long)M * N;
    }
    ACES_CHECK(gemm_lp_impl(GP_F64, M, N, K, A64, aslot, lda, B64, bslot, ldb, S.C0, N, 1.0, 0.0, a_kc, b_kc, -1));
  }
  if (!sa) ACES_CHECK(anchor_copy(A64, ea, &A0));
  if (!sb) ACES_CHECK(anchor_copy(B64, eb, &B0));
  if (!sa && !sb && anc_mode_ == 1) {  // the split of B0 stays fixed while the anchor lives
    const long Kp = pad32(K);
    if (S.B0.hi == nullptr || S.B0.ld != Kp || S.B0.rows != N) {
      void* blk;
      ACES_CHECK(alloc_raw(&blk, (long)N * Kp * 4));
      S.B0.hi = blk;
      S.B0.lo = (char*)blk + (long)N * Kp * 2;
      S.B0.ld = Kp;
      S.B0.rows = N;
    }
    if (b_kc)
      ACES_CHECK(split_bf16(N, K, B64, ldb, (unsigned short*)S.B0.hi, (unsigned short*)S.B0.lo, Kp, s_));
    else
      ACES_CHECK(split_bf16_t(N, K, B64, ldb, (unsigned short*)S.B0.hi, (unsigned short*)S.B0.lo, Kp, s_));
  }
  // C = beta C + alpha C0
  ACES_CHECK(P4(1, 1, M, N, C, beta).src(S.C0, alpha, 0, 0, N, 1).run(s_));
  if (anc_mode_ == 1) return ACES_OK;
  ACES_CHECK(lowp_static(GP_BF16X3));
  LowPOp a, b;
  if (sa || sb) {  // static integral operand: one product with the other's difference
    if (sa) {
      a = LowPOp{lp_[aslot].hi, lp_[aslot].lo, lp_[aslot].ld};
      ACES_CHECK(split_operand(b, bslot, B64, B0, N, K, ldb, b_kc));
    } else {
      ACES_CHECK(split_operand(a, aslot, A64, A0, M, K, lda, a_kc));
      b = LowPOp{lp_[bslot].hi, lp_[bslot].lo, lp_[bslot].ld};
    }
    return lp_product(M, N, K, a, b, C, ldc, alpha, 1.0);
  }
  // A B - A0 B0 = A (B - B0) + (A - A0) B0
  ACES_CHECK(split_operand(a, aslot, A64, nullptr, M, K, lda, a_kc));
  ACES_CHECK(split_operand(b, bslot, B64, B0, N, K, ldb, b_kc));
  ACES_CHECK(lp_product(M, N, K, a, b, C, ldc, alpha, 1.0));
  ACES_CHECK(split_operand(a, aslot, A64, A0, M, K, lda, a_kc));
  b = LowPOp{S.B0.hi, S.B0.lo, S.B0.ld};
  return lp_product(M, N, K, a, b, C, ldc, alpha, 1.0);
}

// Launch one MFMA contraction, timed with hipEvents when profiling is on.
int CcsdEngine::timed_gemm(const GemmDesc& d, GemmPrec prec, double flops) {
  if (!profile_) return gemm(d, prec, s_);
  hipEvent_t ev[2];
  ACES_HIP(hipEventCreate(&ev[0]));
  ACES_HIP(hipEventCreate(&ev[1]));
  ACES_HIP(hipEventRecord(ev[0], s_));
  ACES_CHECK(gemm(d, prec, s_));
  ACES_HIP(hipEventRecord(ev[1], s_));
  ACES_HIP(hipEventSynchronize(ev[1]));
  float ms = 0.f;
  ACES_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  prof_[3 * prof_cat_ + 0] += ms;
  prof_[3 * prof_cat_ + 1] += flops;
  prof_[3 * prof_cat_ + 2] += 1.0;
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return ACES_OK;
}

int CcsdEngine::gemm_lp_impl(GemmPrec prec, int M, int N, int K, const double* A64, int aslot, long lda,
                             const double* B64, int bslot, long ldb, double* C, long ldc, double alpha,
                             double beta, int a_kc, int b_kc, int cslot) {
  const double flops = 2.0 * M * (double)N * K;
  if (prec == GP_F64) {
    GemmDesc d = gd(M, N, K, A64, lda, a_kc, B64, ldb, b_kc, C, ldc, alpha, beta);
    d.ws = ws_;
    d.ws_elems = ws_elems_;
    return timed_gemm(d, GP_F64, flops);
  }
  ACES_REQUIRE(aslot >= 0 && aslot < kLpSlots && bslot >= 0 && bslot < kLpSlots, "gemm_lp: bad operand slot");
  ACES_REQUIRE((!lp_static_slot(aslot) || a_kc) && (!lp_static_slot(bslot) || b_kc),
               "gemm_lp: static operands are k-contiguous");
  ACES_CHECK(lowp_static(prec));
  const long Kp = pad32(std::max(K, 1));
  const void* ops[2][2];
  const int slots[2] = {aslot, bslot};
  const double* src[2] = {A64, B64};
  const long rows[2] = {M, N};
  const long lds[2] = {lda, ldb};
  const int kcs[2] = {a_kc, b_kc};
  for (int q = 0; q < 2; ++q) {
    LowP& L = lp_[slots[q]];
    if (!lp_static_slot(slots[q]) && lp_fresh_[slots[q]] && prec == GP_BF16X3) {
      // written in this precision by a fused kernel or an earlier product of this cycle
      if (!lp_keep_[slots[q]]) lp_fresh_[slots[q]] = false;
    } else if (!lp_static_slot(slots[q])) {
      if (prec == GP_BF16X3 && lp_keep_[slots[q]]) lp_fresh_[slots[q]] = true;
      ACES_CHECK(lowp_slot(slots[q], rows[q], K));
      if (prec == GP_F32 && kcs[q])
        ACES_CHECK(convert_f32((int)rows[q], K, src[q], lds[q], (float*)L.hi, Kp, s_));
      else if (prec == GP_F32)
        ACES_CHECK(convert_f32_t((int)rows[q], K, src[q], lds[q], (float*)L.hi, Kp, s_));
      else if (kcs[q])
        ACES_CHECK(split_bf16((int)rows[q], K, src[q], lds[q], (unsigned short*)L.hi, (unsigned short*)L.lo, Kp,
                              s_));
      else
        ACES_CHECK(split_bf16_t((int)rows[q], K, src[q], lds[q], (unsigned short*)L.hi, (unsigned short*)L.lo, Kp,
                                s_));
    }
    ops[q][0] = L.hi;
    ops[q][1] = L.lo;
  }
  GemmDesc d;
  d.M = M; d.N = N; d.K = (prec == GP_BF16X3) ? (int)Kp : K;
  d.A = ops[0][0]; d.A_lo = ops[0][1]; d.lda = lp_[aslot].ld; d.a_kc = 1;
  d.B = ops[1][0]; d.B_lo = ops[1][1]; d.ldb = lp_[bslot].ld; d.b_kc = 1;
  d.C = C; d.ldc = ldc; d.alpha = alpha; d.beta = beta;
  d.ws = ws_;
  d.ws_elems = ws_elems_;
  if (cslot >= 0 && prec == GP_BF16X3) {
    ACES_REQUIRE(!lp_static_slot(cslot) && cslot != aslot && cslot != bslot, "gemm_lp: bad output slot");
    ACES_CHECK(lowp_slot(cslot, M, N));
    d.C_hi = (unsigned short*)lp_[cslot].hi;
    d.C_lo = (unsigned short*)lp_[cslot].lo;
    d.ldc_lp = lp_[cslot].ld;
    ACES_CHECK(timed_gemm(d, prec, flops));
    lp_fresh_[cslot] = true;
    return ACES_OK;
  }
  return timed_gemm(d, prec, flops);
}

// ---------------------------------------------------------------------------
int CcsdEngine::compute_increments(GemmPrec prec) {
  if ((int)prec == kPrecRefine) {
    // the fp64 evaluation with the large products anchored (recorded on the
    // first such evaluation for these amplitudes, refined afterwards)
    anc_mode_ = anc_valid_ ? 2 : 1;
    anc_site_ = 0;
    const int rc = compute_increments(GP_F64);
    const int mode = anc_mode_;
    anc_mode_ = 0;
    ACES_CHECK(rc);
    if (mode == 1) anc_valid_ = true;
    anc_last_ = mode;
    fwd_valid_ = (mode == 1);
    return ACES_OK;
  }
  anc_last_ = 0;
  ACES_REQUIRE(prec == GP_F64 || prec == GP_F32 || prec == GP_BF16X3, "ccsd: unknown precision");
  // the forward intermediates change: the Lambda splits and anchor with them
  lam_lp_valid_ = lam_anc_valid_ = false;
  hipStream_t s = s_;
  const long ov = (long)o * v, oo = (long)o * o, vv = (long)v * v;
  const long ov2 = ov * v, o2v = oo * v, vo2 = v * oo, v2o = vv * o;
  const double* t1 = t1_;
  const double* t2 = t2_;
  const int OV = (int)ov;

  // One pass over t2: tau = t2 + t1 t1 ; U[(i,a)][(j,b)] = 2 t_ijab - t_ijba ;
  // X = t_ijba ; Ut = U - 2 t_ja t_ib ; Xt = X + 2 t_ja t_ib.  The bf16 form
  // writes the split operands of the W-build and ring products directly
  // (slots 11 U, 12 X, 7 Ut, 8 Xt) and the split of t2 (slot 24), and keeps
  // only U in fp64
  for (bool& f : lp_fresh_) f = false;
  // operands that stay fixed through the cycle once split: t1 (slot 20), tau
  // as [i][(k,c,d)] (22), tauK (23)
  for (bool& f : lp_keep_) f = false;
  lp_keep_[20] = lp_keep_[22] = lp_keep_[23] = true;
  {
    RingOperands r;
    r.f64[0] = U_;
    r.tau = tau_;
    if (prec == GP_BF16X3) {
      const int slot[4] = {11, 12, 7, 8};
      for (int q = 0; q < 4; ++q) {
        ACES_CHECK(lowp_slot(slot[q], ov, ov));
        r.hi[q] = (unsigned short*)lp_[slot[q]].hi;
        r.lo[q] = (unsigned short*)lp_[slot[q]].lo;
      }
      r.ldl = lp_[slot[0]].ld;
      ACES_CHECK(lowp_slot(24, o2v, v));
      r.t2hi = (unsigned short*)lp_[24].hi;
      r.t2lo = (unsigned short*)lp_[24].lo;
      r.ldt = lp_[24].ld;
      // tau's splits: slot 13 [(i,j)][(a,b)] (hole ladder), 22 [i][(j,a,b)]
      // (Foo, tau M3), 23 [a][(i,j,b)] (tauK: Fvv, M7 tauK)
      ACES_CHECK(lowp_slot(13, oo, vv));
      ACES_CHECK(lowp_slot(22, o, ov2));
      ACES_CHECK(lowp_slot(23, v, o2v));
      const long ts[3][3] = {{o * lp_[13].ld, lp_[13].ld, v}, {lp_[22].ld, vv, v}, {(long)o * v, v, lp_[23].ld}};
      const int tslot[3] = {13, 22, 23};
      for (int q = 0; q < 3; ++q) {
        r.tauhi[q] = (unsigned short*)lp_[tslot[q]].hi;
        r.taulo[q] = (unsigned short*)lp_[tslot[q]].lo;
        r.tsi[q] = ts[q][0];
        r.tsj[q] = ts[q][1];
        r.tsa[q] = ts[q][2];
      }
      ACES_CHECK(ring_operands(o, v, t1, t2, r, s));
      for (int q = 0; q < 4; ++q) lp_fresh_[slot[q]] = true;
      lp_fresh_[24] = lp_fresh_[13] = lp_fresh_[22] = lp_fresh_[23] = true;
    } else {
      r.f64[1] = X_;
      r.f64[2] = Ut_;
      r.f64[3] = Xt_;
      ACES_CHECK(ring_operands(o, v, t1, t2, r, s));
    }
  }

  // T1 parts of the W(mbej) intermediates (the two o v^3 reads in the ladder precision)
  prof_cat_ = 3;
  ACES_CHECK(gemm_lp(prec, o, (int)v2o, v, t1, 20, v, ovvv_, 16, v, P1_, v2o, 1.0, 0.0));     // [i][(k,c,a)]
  ACES_CHECK(gemm_lp(prec, o, (int)v2o, v, t1, 20, v, ovvv_kacd_, 17, v, P2_, v2o, 1.0, 0.0));  // [i][(k,a,c)]
  ACES_CHECK(gemm64(gd((int)o2v, v, o, M5c_, o, 1, t1, v, 0, Q12_, v, 1.0, 0.0)));        // [(k,c,i)][a]
  ACES_CHECK(gemm64(gd((int)o2v, v, o, M5b_, o, 1, t1, v, 0, Q2_, v, 1.0, 0.0)));         // [(k,c,i)][a]
  // Wt[(i,a)][(k,c)] = W0t + 2 P1[k,c,a,i] - P2[k,a,c,i] - Q12[k,c,i,a] and
  // Wv[(i,a)][(k,c)] = Wv0 + P2[k,a,c,i] - Q2[k,c,i,a]: each source reads
  // along c (direct) or along a, one pass each
  ACES_CHECK(P4(o, v, o, v, Wt_)
                 .src(W0t_, 1, ov2, ov, v, 1)
                 .src(P1_, 2, v2o, 1, vv, v)
                 .src(P2_, -1, v2o, v, vv, 1)
                 .src(Q12_, -1, v, 1, ov2, ov)
                 .run(s));
  ACES_CHECK(P4(o, v, o, v, Wv_).src(Wv0_, 1, ov2, ov, v, 1).src(P2_, 1, v2o, v, vv, 1)
                 .src(Q2_, -1, v, 1, ov2, ov).run(s));
  // T2 parts (MFMA): Wt += 1/2 Ut L^T ; Wv += -1/2 Xt K^T
  prof_cat_ = 1;
  // (in the bf16 form the results go straight to the ring products' operand
  // slots 9 and 10)
  ACES_CHECK(gemm_lp(prec, OV, OV, OV, Ut_, 7, ov, Lmat_, 2, ov, Wt_, ov, 0.5, 1.0, 1, 1, 9));
  ACES_CHECK(gemm_lp(prec, OV, OV, OV, Xt_, 8, ov, Kx_, 3, ov, Wv_, ov, -0.5, 1.0, 1, 1, 10));
  // ring (MFMA): R1 = Wt U^T ; R2 = Wv X^T
  prof_cat_ = 2;
  ACES_CHECK(gemm_lp(prec, OV, OV, OV, Wt_, 9, ov, U_, 11, ov, R1_, ov, 1.0, 0.0));
  ACES_CHECK(gemm_lp(prec, OV, OV, OV, Wv_, 10, ov, X_, 12, ov, R2_, ov, 1.0, 0.0));

  // inc2 = <ij|ab> + ring assembly, one pass:
  //   inc2 = (ia|jb) + RS[iajb] + RS[jbia] - R2[ibja] - R2[jaib],  RS = (R1 - R2)/2
  ACES_CHECK(P4(o, o, v, v, inc2_)
                 .src(ovov_, 1, ov2, v, ov, 1)
                 .src(R1_, 0.5, ov2, v, ov, 1)
                 .src(R2_, -0.5, ov2, v, ov, 1)
                 .src(R1_, 0.5, v, ov2, 1, ov)
                 .src(R2_, -0.5, v, ov2, 1, ov)
                 .src(R2_, -1, ov2, v, 1, ov)
                 .src(R2_, -1, v, ov2, ov, 1)
                 .run(s));

  // particle-particle ladder with the T1-dressing rows (MFMA)
  {
    PackTauOut po;
    if (prec == GP_BF16X3) {  // the split operands of the two ladder products (slots 5, 6)
      ACES_CHECK(lowp_slot(5, nop_, nvp_));
      po.Shi = (unsigned short*)lp_[5].hi;
      po.Slo = (unsigned short*)lp_[5].lo;
      po.ldsl = lp_[5].ld;
      lp_fresh_[5] = true;
      if (nom_ > 0 && nvm_ > 0) {
        ACES_CHECK(lowp_slot(6, nom_, nvm_));
        po.Dhi = (unsigned short*)lp_[6].hi;
        po.Dlo = (unsigned short*)lp_[6].lo;
        po.lddl = lp_[6].ld;
        lp_fresh_[6] = true;
      }
    } else {
      po.S = Sp_;
      po.lds = nvp_;
      po.D = Dp_;
      po.ldd = std::max<long>(nvm_, 1);
    }
    ACES_CHECK(pack_tau(o, v, tau_, po, s));
  }
  prof_cat_ = 0;
  ACES_CHECK(gemm_lp(prec, (int)(nvp_ + ov), (int)nop_, (int)nvp_, Vp_, 0, nvp_, Sp_, 5, nvp_, Zp_, nop_, 1.0, 0.0));
  if (nom_ > 0 && nvm_ > 0)
    ACES_CHECK(gemm_lp(prec, (int)(nvm_ + ov), (int)nom_, (int)nvm_, Vm_, 1, nvm_, Dp_, 6, nvm_, Zm_, nom_, 1.0,
                       0.0));
  else
    ACES_HIP(hipMemsetAsync(Zm_, 0, sizeof(double), s));
  ACES_CHECK(unpack_ladder(o, v, Zp_, nop_, Zm_, std::max<long>(nom_, 1), inc2_, Yb_, s));
  prof_cat_ = 3;
  // L1 = sum_k Yb[(b,i,j)][k] t_ka stored [(b,i,j)][a] ; inc2 -= L1[a,b,i,j] + L1[b,a,j,i]
  ACES_CHECK(gemm64(gd((int)(v * oo), v, o, Yb_, o, 1, t1, v, 0, L1_, v, 1.0, 0.0)));

  // hole-hole ladder: Wo = (ki|lj) + Qo[klij] + Qo[lkji] + (kc|ld) tau_ijcd ;
  // inc2[ij][ab] += sum_kl Wo[kl][ij] tau[kl][ab]  (both products in the ladder precision)
  ACES_CHECK(gemm64(gd((int)(oo * o), o, v, ovooQ_, v, 1, t1, v, 1, Qo_, o, 1.0, 0.0)));
  if (prec == GP_BF16X3) {
    // transposed: Wo^T[(i,j)][(k,l)] += tau[ij][cd] OVkl[kl][cd], written as the
    // split left operand (slot 14) of inc2 += Wo^T tau
    ACES_CHECK(P4(o, o, o, o, Wo_).src(ooooP_, 1, o, 1, oo * o, oo).src(Qo_, 1, o, 1, oo * o, oo)
                   .src(Qo_, 1, 1, o, oo, oo * o).run(s));
    prof_cat_ = 0;
    ACES_CHECK(gemm_lp(prec, (int)oo, (int)oo, (int)vv, tau_, 13, vv, OVkl_, 4, vv, Wo_, oo, 1.0, 1.0, 1, 1, 14));
    ACES_CHECK(gemm_lp(prec, (int)oo, (int)vv, (int)oo, Wo_, 14, oo, tau_, 15, vv, inc2_, vv, 1.0, 1.0, 1, 0));
  } else {
    ACES_CHECK(P4(o, o, o, o, Wo_).src(ooooP_, 1, oo * o, oo, o, 1).src(Qo_, 1, oo * o, oo, o, 1)
                   .src(Qo_, 1, oo, oo * o, 1, o).run(s));
    prof_cat_ = 0;
    ACES_CHECK(gemm_lp(prec, (int)oo, (int)oo, (int)vv, OVkl_, 4, vv, tau_, 13, vv, Wo_, oo, 1.0, 1.0));
    ACES_CHECK(gemm_lp(prec, (int)oo, (int)vv, (int)oo, Wo_, 14, oo, tau_, 15, vv, inc2_, vv, 1.0, 1.0, 0, 0));
  }
  prof_cat_ = 3;

  // F intermediates (QUAD2/QUAD3/MAKFME + T1RING pieces)
  const double* fov = fock_ + o;
  ACES_CHECK(P4(1, 1, o, v, Fov_).src(fov, 1, 0, 0, n, 1).run(s));
  ACES_CHECK(gemv(OV, OV, Lmat_, ov, t1, Fov_, 1.0, 1.0, s));
  ACES_CHECK(P4(1, 1, o, o, Foo_).src(fock_, 1, 0, 0, n, 1).run(s));
  ACES_CHECK(gemm_lp(prec, o, o, (int)ov2, Lo_, 28, ov2, tau_, 22, ov2, Foo_, o, 1.0, 1.0));
  if (prec != GP_BF16X3)  // (the bf16 form has tauK's split from the ring-operand pass)
    ACES_CHECK(P4(v, o, o, v, tauK_).src(tau_, 1, v, ov2, vv, 1).run(s));  // tauK[a][k][l][d] = tau_klad
  ACES_CHECK(P4(1, 1, v, v, Fvv_).src(fock_ + (long)o * n + o, 1, 0, 0, n, 1).run(s));
  ACES_CHECK(gemm_lp(prec, v, v, (int)(oo * v), tauK_, 23, oo * v, Lv_, 19, oo * v, Fvv_, v, -1.0, 1.0));
  ACES_CHECK(gemm64(gd(o, o, v, fov, n, 1, t1, v, 1, FT_, o, 1.0, 0.0)));       // FT[k][i] = f_kc t_ic
  ACES_CHECK(gemm64(gd(o, o, v, Fov_, v, 1, t1, v, 1, FovT_, o, 1.0, 0.0)));    // FovT[k][i]
  ACES_CHECK(P4(1, 1, o, o, Loo_).src(Foo_, 1, 0, 0, o, 1).src(FT_, 1, 0, 0, o, 1).run(s));
  ACES_CHECK(gemv((int)oo, OV, M1_, ov, t1, Loo_, 1.0, 1.0, s));
  ACES_CHECK(P4(1, 1, v, v, Lvv_).src(Fvv_, 1, 0, 0, v, 1).run(s));
  ACES_CHECK(gemm64(gd(v, v, o, t1, v, 0, fov, n, 0, Lvv_, v, -1.0, 1.0)));
  if (prec == GP_BF16X3) {  // the 12.6 GB integral matrix read as its bf16 split (half the bytes)
    ACES_CHECK(lowp_static(prec));
    ACES_CHECK(gemv_t_split(OV, (int)vv, (const unsigned short*)lp_[30].hi, (const unsigned short*)lp_[30].lo,
                            lp_[30].ld, t1, Lvv_, 1.0, 1.0, s, ws_, ws_elems_));
  } else {
    ACES_CHECK(gemv_t(OV, (int)vv, ovvvL_, vv, t1, Lvv_, 1.0, 1.0, s, ws_, ws_elems_));
  }
  // diagonal (orbital energies) moved into the denominators
  ACES_CHECK(sub_diag(o, Foo_, o, eo_, s));
  ACES_CHECK(sub_diag(v, Fvv_, v, ev_, s));
  ACES_CHECK(sub_diag(o, Loo_, o, eo_, s));
  ACES_CHECK(sub_diag(v, Lvv_, v, ev_, s));

  // FEACONT/FMICONT: tmpL = Lvv_ac t_ijcb - Loo_ki t_kjab ; inc2 += tmpL + tmpL^T
  if (prec == GP_F64) {
    GemmDesc d = gd(v, v, v, Lvv_, v, 1, t2, v, 0, tmpL_, v, 1.0, 0.0);
    d.batch = (int)oo;
    d.sA = 0;
    d.sB = vv;
    d.sC = vv;
    ACES_CHECK(gemm64(d));
    ACES_CHECK(gemm64(gd(o, (int)ov2, o, Loo_, o, 0, t2, ov2, 0, tmpL_, ov2, -1.0, 1.0)));
  } else {
    // With t_ijcb = t_jibc (closed-shell amplitudes; the engine's iterates are
    // symmetrized) the transpose F[j][i][b][a] = tmpL[i][j][a][b] is one product
    // with t2's own rows (j,i,b) as the left operand:
    //   F[(j,i,b)][a] = sum_c t_jibc Lvv_ac,   F_j[i][(b,a)] -= sum_k Loo_ki t_jkba,
    // and tmpL + tmpL^T = F + F^T.
    ACES_CHECK(gemm_lp(prec, (int)o2v, v, v, t2, 24, v, Lvv_, 25, v, tmpL_, v, 1.0, 0.0));
    GemmDesc d = gd(o, (int)vv, o, Loo_, o, 0, t2, vv, 0, tmpL_, vv, -1.0, 1.0);
    d.batch = o;
    d.sA = 0;
    d.sB = ov2;
    d.sC = ov2;
    ACES_CHECK(gemm64(d));
  }

  // T1INT2A/T1INT2B/T12INT2 pieces:
  //   inc2 += P1[iabj] - T1Q[aibj] + P1[jbai] - T1Q[bjai]
  ACES_CHECK(gemm_lp(prec, (int)o2v, o, v, oovv_, 26, v, t1, 20, v, Qb_, o, 1.0, 0.0));     // [(k,i,b)][j]
  ACES_CHECK(gemm64(gd((int)o2v, v, o, Qb_, o2v, 0, t1, v, 0, T1Q_, v, 1.0, 0.0)));       // [(i,b,j)][a]
  //   inc2 -= G4[iajb] + G4[jbia],  G4 = (ovoo[iajk] + S3[iakj]) t_kb
  ACES_CHECK(gemm_lp(prec, (int)o2v, o, v, ovov_, 27, v, t1, 20, v, S3_, o, 1.0, 0.0));      // [(i,a,k)][j]
  ACES_CHECK(P4(o, v, o, o, G4A_).src(ovoo_, 1, vo2, oo, o, 1).src(S3_, 1, vo2, oo, 1, o).run(s));
  ACES_CHECK(gemm64(gd((int)o2v, v, o, G4A_, o, 1, t1, v, 0, G4_, v, 1.0, 0.0)));         // [(i,a,j)][b]

  // the pair-symmetric accumulations into inc2, in one permute: P1, L1 and
  // T1Q are stored so that each term reads either along b (direct) or along
  // a (one transposed index), so the permute is a single pass over inc2
  //   inc2 -= L1[a,b,i,j] + L1[b,a,j,i]              (ladder T1 rows)
  //   inc2 += tmpL + tmpL^T                           (FEACONT / FMICONT)
  //   inc2 += P1[iabj] - T1Q[aibj] + P1[jbai] - T1Q[bjai]
  //   inc2 -= G4[iajb] + G4[jbia]
  ACES_CHECK(P4(o, o, v, v, inc2_, 1.0)
                 .src(L1_, -1, ov, v, 1, o2v)
                 .src(L1_, -1, v, ov, o2v, 1)
                 .src(tmpL_, 1, ov2, vv, v, 1)
                 .src(tmpL_, 1, vv, ov2, 1, v)
                 .src(P1_, 1, vv, v2o, v, 1)
                 .src(T1Q_, -1, ov2, v, 1, ov)
                 .src(P1_, 1, v2o, vv, 1, v)
                 .src(T1Q_, -1, v, ov2, ov, 1)
                 .src(G4_, -1, ov2, v, ov, 1)
                 .src(G4_, -1, v, ov2, 1, ov)
                 .run(s));

  // ---- T1 increment ----
  ACES_CHECK(P4(1, 1, o, v, inc1_).src(fov, 1, 0, 0, n, 1).run(s));
  ACES_CHECK(P4(1, 1, o, o, Gk_).src(Foo_, 1, 0, 0, o, 1).src(FT_, 2, 0, 0, o, 1).src(FovT_, -1, 0, 0, o, 1).run(s));
  ACES_CHECK(gemm64(gd(o, v, o, Gk_, o, 0, t1, v, 0, inc1_, v, -1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, v, t1, v, 1, Fvv_, v, 1, inc1_, v, 1.0, 1.0)));
  ACES_CHECK(gemv(OV, OV, U_, ov, Fov_, inc1_, 1.0, 1.0, s));
  ACES_CHECK(gemv(OV, OV, W0t_, ov, t1, inc1_, 1.0, 1.0, s));
  ACES_CHECK(gemm_lp(prec, o, v, (int)ov2, tau_, 22, ov2, M3_, 18, ov2, inc1_, v, 1.0, 1.0));
  ACES_CHECK(gemm_lp(prec, o, v, (int)o2v, M7_, 29, o2v, tauK_, 23, o2v, inc1_, v, -1.0, 1.0));
  // the reverse pass needs fp64 intermediates of these amplitudes
  fwd_valid_ = (prec == GP_F64);
  return ACES_OK;
}

int CcsdEngine::increments(GemmPrec prec, double* inc1_out, double* inc2_out) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_CHECK(compute_increments(prec));
  if (inc1_out == nullptr && inc2_out == nullptr) return ACES_OK;
  const long ov = (long)o * v;
  if (inc1_out) ACES_HIP(hipMemcpyAsync(inc1_out, inc1_, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  if (inc2_out) ACES_HIP(hipMemcpyAsync(inc2_out, inc2_, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

// One cycle of the reference loop (vcc/vcc.F:811-1085) with DODIIS0 semantics.
int CcsdEngine::cycle(int icycle, GemmPrec prec, double tol, CcsdStats& st) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_REQUIRE(icycle >= 1, "ccsd cycle: icycle starts at 1");
  hipEvent_t e0, e1, e2;
  ACES_HIP(hipEventCreate(&e0));
  ACES_HIP(hipEventCreate(&e1));
  ACES_HIP(hipEventCreate(&e2));
  ACES_HIP(hipEventRecord(e0, s_));
  tn_lists_ = false;
  ACES_CHECK(compute_increments(prec));
  ACES_HIP(hipEventRecord(e1, s_));
  // Jacobi update, its energy (CMPENG before SYMETRIZE) and SYMETRIZE_T2 of
  // the update in one pass
  ACES_CHECK(jacobi_energy_sym(o, v, eo_, ev_, inc1_, inc2_, fock_ + o, n, Lijab_, t1n_, t2n_, jpart_, scal_, s_));
  ACES_CHECK(diis_step(icycle, tol, st, true));
  ACES_HIP(hipEventRecord(e2, s_));
  ACES_HIP(hipEventSynchronize(e2));
  float ms1 = 0, ms2 = 0;
  ACES_HIP(hipEventElapsedTime(&ms1, e0, e1));
  ACES_HIP(hipEventElapsedTime(&ms2, e0, e2));
  st.ms_increments = ms1;
  st.ms_cycle = ms2;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  return ACES_OK;
}

// NEWT2 / E4S: T_new = increment / D.
int CcsdEngine::jacobi_step() {
  tn_lists_ = false;
  return jacobi(o, v, eo_, ev_, inc1_, inc2_, t1n_, t2n_, s_);
}

// DODIIS0 (vcc/dodiis0.F) on T (t1_/t2_) and T_new (t1n_/t2n_); scal_[0] holds
// the energy of T_new computed by the caller (CMPENG(60,2)).  On return t1_/t2_
// hold the next amplitudes unless converged; converged, t1n_/t2n_ hold the
// increment lists' content, the Lambda start.
int CcsdEngine::diis_step(int icycle, double tol, CcsdStats& st, bool xn_sym) {
  ACES_CHECK(diis_core(tdiis_, t1_, t1n_, icycle, tol, false, st, t_sym_, xn_sym));
  tn_lists_ = st.converged != 0;
  // refinement anchor: remember the anchor cycle's residual; move the anchor
  // once the residual has fallen 1e3-fold below it
  if (anc_last_ == 1)
    anc_res_ = std::fabs(st.resmiter);
  else if (anc_last_ == 2 && std::fabs(st.resmiter) < 1e-3 * anc_res_)
    anc_valid_ = false;
  anc_last_ = 0;
  // a DIIS combination of symmetrized vectors (or the symmetrized update
  // itself) is exactly pair-symmetric
  if (!st.converged) t_sym_ = true;
  if (st.converged) {
    // CMPENG(43,0) on the amplitudes of this cycle; amplitudes are not moved.
    double e;
    ACES_CHECK(energy_of(t1_, t2_, e));
    st.energy_final = e;
  }
  fwd_valid_ = false;
  return ACES_OK;
}

int CcsdEngine::diis_core(DiisState& D, double* x, double* xn, int icycle, double tol, bool lambda_rule,
                          CcsdStats& st, bool x_sym, bool xn_sym) {
  const long ov = (long)o * v;
  const long nt = ov + ov * ov;
  const int order = D.order;
  if (!x_sym) ACES_CHECK(symmetrize_t2(o, v, x + ov, s_));
  if (!xn_sym) ACES_CHECK(symmetrize_t2(o, v, xn + ov, s_));
  const int ndimr = std::min(icycle, order);
  const int idimr = 1 + (icycle - 1) % order;  // 1-based slot
  // one pass: e = xn - x into the error slot idimr, x into the solution slot,
  // RESMAX = ISAMAX(e) and the dots <err_i, e> for i = 1..ndimr (slot idimr
  // is e itself)
  std::vector<const double*> xs(ndimr);
  for (int i = 1; i <= ndimr; ++i) xs[i - 1] = (i == idimr) ? nullptr : D.err[i - 1];
  ACES_HIP(hipMemcpyAsync(d_ptrs_, xs.data(), sizeof(double*) * ndimr, hipMemcpyHostToDevice, s_));
  ACES_CHECK(diis_error(o, v, x, xn, ndimr, d_ptrs_, idimr - 1, D.err[idimr - 1], D.sol[idimr - 1], red_, scal_ + 8,
                        scal_ + 1, s_));
  double h[40];
  ACES_HIP(hipMemcpyAsync(h, scal_, sizeof(double) * 40, hipMemcpyDeviceToHost, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  st.energy_new = h[0];
  st.resmiter = h[1];
  for (int i = 1; i <= ndimr; ++i) {
    D.R[(size_t)(i - 1) * order + (idimr - 1)] = h[8 + i - 1];
    D.R[(size_t)(idimr - 1) * order + (i - 1)] = h[8 + i - 1];
  }
  // DODIIS: augmented Pulay system
  const int na = ndimr + 1;
  std::vector<double> A((size_t)na * na, 0.0), b(na, 0.0);
  for (int j = 1; j < na; ++j) {
    A[j] = -1.0;
    A[(size_t)j * na] = -1.0;
    for (int i = 1; i < na; ++i) A[(size_t)i * na + j] = D.R[(size_t)(i - 1) * order + (j - 1)];
  }
  b[0] = -1.0;
  if (eu::solve_small_lu(na, A, b) != ACES_OK) {
    set_error("DODIIS: linear solver failed (singular DIIS matrix)");
    return ACES_ERR_SOLVE;
  }
  std::vector<double> c(b.begin() + 1, b.end());
  // DIIS residual: the amplitude DODIIS0 sums slot idimr first, then the
  // others downwards; the Lambda DODIIS0 sums slots 1..ndimr in order.
  std::vector<const double*> rv;
  std::vector<double> rc;
  if (lambda_rule) {
    for (int i = 1; i <= ndimr; ++i) {
      rv.push_back(D.err[i - 1]);
      rc.push_back(c[i - 1]);
    }
  } else {
    rv.push_back(D.err[idimr - 1]);
    rc.push_back(c[idimr - 1]);
    for (int i = ndimr; i > idimr; --i) {
      rv.push_back(D.err[i - 1]);
      rc.push_back(c[i - 1]);
    }
    for (int i = idimr - 1; i >= 1; --i) {
      rv.push_back(D.err[i - 1]);
      rc.push_back(c[i - 1]);
    }
  }
  ACES_HIP(hipMemcpyAsync(d_ptrs_, rv.data(), sizeof(double*) * rv.size(), hipMemcpyHostToDevice, s_));
  ACES_HIP(hipMemcpyAsync(d_coef_, rc.data(), sizeof(double) * rc.size(), hipMemcpyHostToDevice, s_));
  ACES_CHECK(lincomb(nt, (int)rv.size(), d_ptrs_, d_coef_, newvec_, s_));
  ACES_CHECK(diis_absmax(o, v, newvec_, red_, scal_ + 2, s_));
  ACES_HIP(hipMemcpyAsync(h, scal_ + 2, sizeof(double), hipMemcpyDeviceToHost, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  st.resmdiis = h[0];
  if (lambda_rule)
    st.converged = (std::fabs(st.resmiter) < tol || std::fabs(st.resmdiis) < tol) ? 1 : 0;
  else
    st.converged = std::fabs(st.resmiter) < tol ? 1 : 0;
  // Both DODIIS0s extrapolate in the converged cycle too and write the result
  // over the increment lists (vcc/dodiis0.F:181-233, lambda/dodiis0.F:131-170).
  // Converged amplitudes are not moved (vcc/vcc.F:1077-1087): the extrapolant
  // goes to xn, from where xvcc hands it to xlambda as the Lambda start
  // (DRMOVE(...,100), vcc/vcc.F:1130-1143).  xlambda moves the increment lists
  // to the multipliers unconditionally (lambda/vlamcc.F:811-815): x.
  double* dst = (st.converged && !lambda_rule) ? xn : x;
  if (ndimr >= 2) {
    std::vector<const double*> sv;
    std::vector<double> sc;
    sv.push_back(newvec_);
    sc.push_back(1.0);
    for (int i = 1; i <= ndimr; ++i) {
      sv.push_back(D.sol[i - 1]);
      sc.push_back(c[i - 1]);
    }
    ACES_HIP(hipMemcpyAsync(d_ptrs_, sv.data(), sizeof(double*) * sv.size(), hipMemcpyHostToDevice, s_));
    ACES_HIP(hipMemcpyAsync(d_coef_, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice, s_));
    ACES_CHECK(lincomb(nt, (int)sv.size(), d_ptrs_, d_coef_, dst, s_));
  } else if (dst != xn) {
    ACES_HIP(hipMemcpyAsync(x, xn, sizeof(double) * nt, hipMemcpyDeviceToDevice, s_));
  }
  return ACES_OK;
}

}  // namespace aces
