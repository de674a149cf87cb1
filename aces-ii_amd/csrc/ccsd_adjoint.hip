// Reverse pass of the closed-shell CCSD increments: Lambda residual and
// response one-particle density (differentiation form).
//
// The Lagrangian of the amplitude equations,
//   L(T, l; f, g) = E(T) + sum_ia l1_ia Omega1_ia + sum_ijab l2_ijab Omega2_ijab,
//   Omega = inc(T) - D*T,
// is stationary in T at the Lambda solution (lambda/vlamcc.F; the equations of
// Gauss, Stanton and Bartlett, J. Chem. Phys. 95, 2623 (1991), listed in
// lambda/README:1-40), and dL/df_pq is the response density of dens/vdens.F
// (DENSOO/DENSVV/DENSVO, :362-388) for perturbations of the MO Fock matrix.
// compute_adjoint() walks the operations of compute_increments() backwards:
// a GEMM C = A.B^T contributes gA += gC.B and gB += gC^T.A (only for operands
// that depend on T or f), a permutation contributes the inverse gather, a
// product term t1 x t1 contributes two matrix-vector products.  All
// contractions run on the fp64 MFMA engine.
//
// Spatial multipliers l relate to the reference's lists (L1 list 190,
// L2(Ij,Ab) list 146) as l1 = 2 L1, l2 = 2 L2 - L2(i<->j); in that
// normalisation the reference's Jacobi step L <- L + R/D is
// l <- l + (dL/dT)/D here, with R1 = r1/2 and R2 = (2 + P(ij)) r2 / 3.
#include <algorithm>
#include <cmath>
#include "adjoint_kernels.h"
#include "ccsd_engine.h"
#include "ccsd_kernels.h"
#include "common.h"
#include "engine_util.h"
#include "tensor_ops.h"

namespace aces {

using eu::gd;
using eu::P4;
using eu::pad32;

int CcsdEngine::adjoint_alloc() {
  if (adj_ready_) return ACES_OK;
  const long ov = (long)o * v, oo = (long)o * o, vv = (long)v * v;
  const long ovvo = ov * ov, o2v = oo * v, o3 = oo * o;
  const long nt = ov + ovvo;
  ACES_CHECK(alloc(&lam_, nt));
  ACES_CHECK(alloc(&gT_, nt));
  ACES_CHECK(alloc(&Lam_, nt));
  ACES_CHECK(alloc(&Lamn_, nt));
  double** big[] = {&gtau_, &gU_, &gX_, &gUt_, &gXt_, &gWt_, &gWv_, &gRS_, &gR2_, &gP1_, &gP2_, &gQ12_, &gQ2_,
                    &gtmpL_, &gtauK_, &gL1_, &gT1Q_, &gG4_};
  for (double** p : big) ACES_CHECK(alloc(p, ovvo));
  ACES_CHECK(alloc(&gZp_, (nvp_ + ov) * nop_));
  ACES_CHECK(alloc(&gZm_, std::max<long>((nvm_ + ov) * nom_, 1)));
  ACES_CHECK(alloc(&gSp_, nop_ * nvp_));
  ACES_CHECK(alloc(&gDp_, std::max<long>(nom_ * nvm_, 1)));
  ACES_CHECK(alloc(&gYb_, v * o3));
  ACES_CHECK(alloc(&gWo_, oo * oo));
  ACES_CHECK(alloc(&gQo_, oo * oo));
  ACES_CHECK(alloc(&gQb_, o2v * o));
  ACES_CHECK(alloc(&gG4A_, o2v * o));
  ACES_CHECK(alloc(&gS3_, o2v * o));
  ACES_CHECK(alloc(&gFovI_, ov));
  ACES_CHECK(alloc(&gFooI_, oo));
  ACES_CHECK(alloc(&gFvvI_, vv));
  ACES_CHECK(alloc(&gLoo_, oo));
  ACES_CHECK(alloc(&gLvv_, vv));
  ACES_CHECK(alloc(&gFT_, oo));
  ACES_CHECK(alloc(&gFovT_, oo));
  ACES_CHECK(alloc(&gGk_, oo));
  ACES_CHECK(alloc(&gfo_, oo));
  ACES_CHECK(alloc(&gfv_, vv));
  ACES_CHECK(alloc(&gfov_, ov));
  ACES_CHECK(alloc(&geo_, o));
  ACES_CHECK(alloc(&gev_, v));
  ACES_CHECK(diis_resize(ldiis_, tdiis_.order));
  adj_ready_ = true;
  return ACES_OK;
}

// Reverse pass.  lam = [l1 | l2]; on return gT_ = [dL/dt1 | dL/dt2] (t2 part not
// yet symmetrised) and, when density is set, gfo_/gfv_/gfov_/geo_/gev_ hold
// dL/d(f_oo, f_vv, f_ov, e_o, e_v).  Operation numbers refer to the forward
// order in compute_increments().
// The fixed right operands of the mixed reverse products, split once per
// forward state: U^T, Wt^T, X^T, Wv^T (rows = the product's output columns).
int CcsdEngine::lambda_lowp_operands() {
  if (lam_lp_valid_) return ACES_OK;
  ACES_CHECK(lowp_static(GP_BF16X3));
  const long ov = (long)o * v, oo = (long)o * o, vv = (long)v * v;
  const long ldm = std::max<long>(nvm_, 1);
  // {source, rows, K, source leading dimension, k-contiguous source}
  struct S { const double* src; long rows, K, ld; bool kc; };
  const S tab[kLamLp] = {{U_, ov, ov, ov, false},    {Wt_, ov, ov, ov, false},
                         {X_, ov, ov, ov, false},    {Wv_, ov, ov, ov, false},
                         {Vp_ + nvp_ * nvp_, nvp_, ov, nvp_, false},
                         {Vm_ + nvm_ * ldm, nvm_, ov, ldm, false},
                         {Wo_, oo, oo, oo, true},    {tau_, oo, vv, vv, true},
                         {OVkl_, vv, oo, vv, false}};
  for (int q = 0; q < kLamLp; ++q) {
    LowP& L = lam_lp_[q];
    if (tab[q].rows == 0) continue;
    const long Kp = pad32(tab[q].K);
    if (L.hi == nullptr) {
      void* blk;
      ACES_CHECK(alloc_raw(&blk, tab[q].rows * Kp * 4));
      L.hi = blk;
      L.lo = (char*)blk + tab[q].rows * Kp * 2;
      L.ld = Kp;
      L.rows = tab[q].rows;
    }
    if (tab[q].kc)
      ACES_CHECK(split_bf16((int)tab[q].rows, (int)tab[q].K, tab[q].src, tab[q].ld, (unsigned short*)L.hi,
                            (unsigned short*)L.lo, Kp, s_));
    else
      ACES_CHECK(split_bf16_t((int)tab[q].rows, (int)tab[q].K, tab[q].src, tab[q].ld, (unsigned short*)L.hi,
                              (unsigned short*)L.lo, Kp, s_));
  }
  lam_lp_valid_ = true;
  return ACES_OK;
}

int CcsdEngine::adj_gemm_lp(int M, int N, int K, const double* A, long lda, int akc, int aslot, const LowPOp& B,
                            double* C, long ldc, double alpha, double beta) {
  LowPOp a;
  ACES_CHECK(split_operand(a, aslot, A, nullptr, M, K, lda, akc));
  const int cat = prof_cat_;
  prof_cat_ = 5;
  const int rc = lp_product(M, N, K, a, B, C, ldc, alpha, beta);
  prof_cat_ = cat;
  return rc;
}

int CcsdEngine::compute_adjoint(const double* lam, bool density, GemmPrec prec, bool lin) {
  ACES_REQUIRE(prec == GP_F64 || (prec == GP_BF16X3 && !density), "adjoint: bf16x3 only without the density");
  const bool lp = (prec == GP_BF16X3);
  if (lp) ACES_CHECK(lambda_lowp_operands());
  hipStream_t s = s_;
  const long ov = (long)o * v, oo = (long)o * o, vv = (long)v * v;
  const long ov2 = ov * v, o2v = oo * v, vo2 = v * oo, v2o = vv * o, o3 = oo * o;
  const long ovvo = ov * ov;
  const int OV = (int)ov;
  const double* t1 = t1_;
  const double* t2 = t2_;
  const double* l1 = lam;
  const double* l2 = lam + ov;
  double* gt1 = gT_;
  double* gt2 = gT_ + ov;
  const double* fov = fock_ + o;

  prof_cat_ = 4;
  // seeds: energy (2 f_ia t_ia + L_ijab tau_ijab) and -D*T (lin: -D*T only)
  ACES_CHECK(adjoint_seed(o, v, eo_, ev_, lin ? nullptr : fov, n, l1, l2, gt1, gt2, s));
  if (lin)
    ACES_HIP(hipMemsetAsync(gtau_, 0, sizeof(double) * ovvo, s));
  else
    ACES_HIP(hipMemcpyAsync(gtau_, Lijab_, sizeof(double) * ovvo, hipMemcpyDeviceToDevice, s));
  if (density) {
    ACES_HIP(hipMemsetAsync(gfo_, 0, sizeof(double) * oo, s));
    ACES_HIP(hipMemsetAsync(gfv_, 0, sizeof(double) * vv, s));
    ACES_HIP(hipMemsetAsync(geo_, 0, sizeof(double) * o, s));
    ACES_HIP(hipMemsetAsync(gev_, 0, sizeof(double) * v, s));
    // gfov = 2 t1 (energy) + l1 (inc1 = f_ia + ...)
    ACES_CHECK(P4(1, 1, 1, ov, gfov_).src(t1, 2.0, 0, 0, 0, 1).src(l1, 1.0, 0, 0, 0, 1).run(s));
    ACES_CHECK(denominator_adjoint(o, v, l1, l2, t1, t2, geo_, gev_, s));
  }

  // ---- T1 increment ----
  // inc1 -= Gk^T t1 ; inc1 += t1 Fvv^T ; inc1 += U Fov ; inc1 += W0t t1 ; inc1 += tau M3^T ; inc1 -= M7 tauK^T
  ACES_CHECK(gemm64(gd(o, o, v, t1, v, 1, l1, v, 1, gGk_, o, -1.0, 0.0)));
  ACES_CHECK(gemm64(gd(o, v, o, Gk_, o, 1, l1, v, 0, gt1, v, -1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, v, l1, v, 1, Fvv_, v, 0, gt1, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(v, v, o, l1, v, 0, t1, v, 0, gFvvI_, v, 1.0, 0.0)));
  ACES_CHECK(P4(o, v, o, v, gU_).prod(l1, v, 1, 0, 0, Fov_, 0, 0, v, 1, 1.0).run(s));
  ACES_CHECK(gemv_t(OV, OV, U_, ov, l1, gFovI_, 1.0, 0.0, s, ws_, ws_elems_));
  ACES_CHECK(gemv_t(OV, OV, W0t_, ov, l1, gt1, 1.0, 1.0, s, ws_, ws_elems_));
  ACES_CHECK(gemm64(gd(o, (int)ov2, v, l1, v, 1, M3_, ov2, 0, gtau_, ov2, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(v, (int)o2v, o, l1, v, 0, M7_, o2v, 0, gtauK_, o2v, -1.0, 0.0)));
  // Gk = Foo + 2 FT - FovT
  ACES_CHECK(P4(1, 1, 1, oo, gFooI_).src(gGk_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(1, 1, 1, oo, gFT_).src(gGk_, 2.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(1, 1, 1, oo, gFovT_).src(gGk_, -1.0, 0, 0, 0, 1).run(s));

  // ---- (16) inc2 -= G4[iajb] + G4[jbia] ; G4 = G4A t1 ; G4A = ovoo + S3 ; S3 = ovov t1 ----
  ACES_CHECK(P4(o, v, o, v, gG4_).src(l2, -1.0, ov2, v, vv, 1).src(l2, -1.0, vv, 1, ov2, v).run(s));
  ACES_CHECK(gemm64(gd((int)o2v, o, v, gG4_, v, 1, t1, v, 1, gG4A_, o, 1.0, 0.0)));
  ACES_CHECK(gemm64(gd(o, v, (int)o2v, G4A_, o, 0, gG4_, v, 0, gt1, v, 1.0, 1.0)));
  ACES_CHECK(P4(o, v, o, o, gS3_).src(gG4A_, 1.0, vo2, oo, 1, o).run(s));
  ACES_CHECK(gemm64(gd(o, v, (int)o2v, gS3_, o, 0, ovov_, v, 0, gt1, v, 1.0, 1.0)));

  // ---- (15) inc2 += P1[iabj] - T1Q[aibj] + P1[jbai] - T1Q[bjai] ; T1Q = t1^T Qb ; Qb = oovv t1 ----
  ACES_CHECK(P4(v, o, v, o, gT1Q_).src(l2, -1.0, v, ov2, 1, vv).src(l2, -1.0, 1, vv, v, ov2).run(s));
  ACES_CHECK(P4(o, v, v, o, gP1_).src(l2, 1.0, ov2, v, 1, vv).src(l2, 1.0, vv, 1, v, ov2).run(s));
  ACES_CHECK(gemm64(gd(o, v, (int)o2v, Qb_, o2v, 1, gT1Q_, o2v, 1, gt1, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, (int)o2v, v, t1, v, 1, gT1Q_, o2v, 0, gQb_, o2v, 1.0, 0.0)));
  ACES_CHECK(gemm64(gd(o, v, (int)o2v, gQb_, o, 0, oovv_, v, 0, gt1, v, 1.0, 1.0)));

  // ---- (14) tmpL = Lvv t2 (per ij) - Loo^T t2 ; inc2 += tmpL + tmpL^T ----
  ACES_CHECK(P4(o, o, v, v, gtmpL_).src(l2, 1.0, ov2, vv, v, 1).src(l2, 1.0, vv, ov2, 1, v).run(s));
  // gLvv[a][c] = sum_{ij,b} gtmpL[ij][a][b] t2[ij][c][b]  via [a][(ij,b)] and [c][(ij,b)] copies
  ACES_CHECK(P4(v, o, o, v, gRS_).src(gtmpL_, 1.0, v, ov2, vv, 1).run(s));
  ACES_CHECK(P4(v, o, o, v, gR2_).src(t2, 1.0, v, ov2, vv, 1).run(s));
  ACES_CHECK(gemm64(gd(v, v, (int)(oo * v), gRS_, oo * v, 1, gR2_, oo * v, 1, gLvv_, v, 1.0, 0.0)));
  {
    GemmDesc d = gd(v, v, v, Lvv_, v, 0, gtmpL_, v, 0, gt2, v, 1.0, 1.0);
    d.batch = (int)oo;
    d.sA = 0;
    d.sB = vv;
    d.sC = vv;
    ACES_CHECK(gemm64(d));
  }
  ACES_CHECK(gemm64(gd(o, o, (int)ov2, t2, ov2, 1, gtmpL_, ov2, 1, gLoo_, o, -1.0, 0.0)));
  ACES_CHECK(gemm64(gd(o, (int)ov2, o, Loo_, o, 1, gtmpL_, ov2, 0, gt2, ov2, -1.0, 1.0)));

  // ---- (13) F intermediates, reverse order ----
  // Foo, Fvv, Loo, Lvv lost their diagonal (orbital energies) after use in Loo/Lvv.
  if (density) {
    ACES_CHECK(P4(1, 1, 1, o, geo_, 1.0).src(gFooI_, -1.0, 0, 0, 0, o + 1).src(gLoo_, -1.0, 0, 0, 0, o + 1).run(s));
    ACES_CHECK(P4(1, 1, 1, v, gev_, 1.0).src(gFvvI_, -1.0, 0, 0, 0, v + 1).src(gLvv_, -1.0, 0, 0, 0, v + 1).run(s));
  }
  // Lvv = Fvv - t1^T fov + 2 ovvv.t1 - (kc|ad) t_kd
  ACES_CHECK(P4(1, 1, 1, vv, gFvvI_, 1.0).src(gLvv_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(gemm64(gd(o, v, v, fov, n, 1, gLvv_, v, 1, gt1, v, -1.0, 1.0)));
  if (density) ACES_CHECK(gemm64(gd(o, v, v, t1, v, 1, gLvv_, v, 0, gfov_, v, -1.0, 1.0)));
  ACES_CHECK(gemv(OV, (int)vv, ovvvL_, vv, gLvv_, gt1, 1.0, 1.0, s));
  // Loo = Foo + FT + M1 t1
  ACES_CHECK(P4(1, 1, 1, oo, gFooI_, 1.0).src(gLoo_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(1, 1, 1, oo, gFT_, 1.0).src(gLoo_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(gemv_t((int)oo, OV, M1_, ov, gLoo_, gt1, 1.0, 1.0, s, ws_, ws_elems_));
  // FovT = Fov t1^T ; FT = fov t1^T
  ACES_CHECK(gemm64(gd(o, v, o, gFovT_, o, 1, t1, v, 0, gFovI_, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, o, gFovT_, o, 0, Fov_, v, 0, gt1, v, 1.0, 1.0)));
  if (density) ACES_CHECK(gemm64(gd(o, v, o, gFT_, o, 1, t1, v, 0, gfov_, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, o, gFT_, o, 0, fov, n, 0, gt1, v, 1.0, 1.0)));
  // Fvv = fvv - tauK Lv^T ; tauK[a][k][l][d] = tau[k][l][a][d]
  if (density) ACES_CHECK(P4(1, 1, 1, vv, gfv_, 1.0).src(gFvvI_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(gemm64(gd(v, (int)o2v, v, gFvvI_, v, 1, Lv_, o2v, 0, gtauK_, o2v, -1.0, 1.0)));
  ACES_CHECK(P4(o, o, v, v, gtau_, 1.0).src(gtauK_, 1.0, ov, v, o2v, 1).run(s));
  // Foo = foo + Lo tau^T
  if (density) ACES_CHECK(P4(1, 1, 1, oo, gfo_, 1.0).src(gFooI_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(gemm64(gd(o, (int)ov2, o, gFooI_, o, 0, Lo_, ov2, 0, gtau_, ov2, 1.0, 1.0)));
  // Fov = fov + Lmat t1   (Lmat symmetric)
  if (density) ACES_CHECK(P4(1, 1, 1, ov, gfov_, 1.0).src(gFovI_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(gemv(OV, OV, Lmat_, ov, gFovI_, gt1, 1.0, 1.0, s));

  // ---- (12) hole-hole ladder ----
  if (lp) {
    auto op = [](const LowP& L) { return LowPOp{L.hi, L.lo, L.ld}; };
    LowPOp b;
    const int cat = prof_cat_;
    prof_cat_ = 5;
    ACES_CHECK(split_operand(b, 15, l2, nullptr, vv, (int)oo, vv, 0));
    ACES_CHECK(lp_product((int)oo, (int)vv, (int)oo, op(lam_lp_[6]), b, gtau_, vv, 1.0, 1.0));
    ACES_CHECK(split_operand(b, 13, l2, nullptr, oo, (int)vv, vv, 1));
    ACES_CHECK(lp_product((int)oo, (int)oo, (int)vv, op(lam_lp_[7]), b, gWo_, oo, 1.0, 0.0));
    ACES_CHECK(split_operand(b, 14, gWo_, nullptr, oo, (int)oo, oo, 0));
    ACES_CHECK(lp_product((int)oo, (int)vv, (int)oo, b, op(lam_lp_[8]), gtau_, vv, 1.0, 1.0));
    prof_cat_ = cat;
  } else {
    ACES_CHECK(gemm64(gd((int)oo, (int)vv, (int)oo, Wo_, oo, 1, l2, vv, 0, gtau_, vv, 1.0, 1.0)));
    ACES_CHECK(gemm64(gd((int)oo, (int)oo, (int)vv, tau_, vv, 1, l2, vv, 1, gWo_, oo, 1.0, 0.0)));
    ACES_CHECK(gemm64(gd((int)oo, (int)vv, (int)oo, gWo_, oo, 0, OVkl_, vv, 0, gtau_, vv, 1.0, 1.0)));
  }
  ACES_CHECK(P4(o, o, o, o, gQo_).src(gWo_, 1.0, o3, oo, o, 1).src(gWo_, 1.0, oo, o3, 1, o).run(s));
  ACES_CHECK(gemm64(gd(o, v, (int)o3, gQo_, o, 0, ovooQ_, v, 0, gt1, v, 1.0, 1.0)));

  // ---- (11) inc2 -= L1[a,b,i,j] + L1[b,a,j,i] ; L1 = t1^T Yb ----
  ACES_CHECK(P4(v, v, o, o, gL1_).src(l2, -1.0, v, 1, ov2, vv).src(l2, -1.0, 1, v, vv, ov2).run(s));
  ACES_CHECK(gemm64(gd(o, v, (int)vo2, Yb_, o, 0, gL1_, vo2, 1, gt1, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd((int)vo2, o, v, gL1_, vo2, 0, t1, v, 1, gYb_, o, 1.0, 0.0)));

  // ---- (10)-(9) particle-particle ladder ----
  ACES_CHECK(ladder_unpack_adjoint(o, v, l2, gYb_, gZp_, nop_, gZm_, std::max<long>(nom_, 1), s));
  if (lp) {
    // gS = gZ^T V in bf16x3: over V's pair rows with V's k-contiguous split
    // read as V^T, over its T1-dressing rows with their transposed split.
    // V-'s square block is symmetric; V+'s is G W with G symmetric and
    // W = diag(1/2 on the diagonal pairs (c,c), else 1), so
    // V+[k][n] = V+[n][k] w_n / w_k: gZ's diagonal-pair rows are doubled
    // (nothing else reads them) and the product's diagonal-pair columns halved.
    ACES_CHECK(scale_diag_pairs(v, gZp_, nop_, nop_, 0, 2.0, s));
    ACES_CHECK(adj_gemm_lp((int)nop_, (int)nvp_, (int)nvp_, gZp_, nop_, 0, 5, LowPOp{lp_[0].hi, lp_[0].lo, lp_[0].ld},
                           gSp_, nvp_, 1.0, 0.0));
    ACES_CHECK(scale_diag_pairs(v, gSp_, nvp_, nop_, 1, 0.5, s));
    ACES_CHECK(adj_gemm_lp((int)nop_, (int)nvp_, (int)ov, gZp_ + nvp_ * nop_, nop_, 0, 31,
                           LowPOp{lam_lp_[4].hi, lam_lp_[4].lo, lam_lp_[4].ld}, gSp_, nvp_, 1.0, 1.0));
    if (nom_ > 0 && nvm_ > 0) {
      ACES_CHECK(adj_gemm_lp((int)nom_, (int)nvm_, (int)nvm_, gZm_, nom_, 0, 6,
                             LowPOp{lp_[1].hi, lp_[1].lo, lp_[1].ld}, gDp_, nvm_, 1.0, 0.0));
      ACES_CHECK(adj_gemm_lp((int)nom_, (int)nvm_, (int)ov, gZm_ + nvm_ * nom_, nom_, 0, 32,
                             LowPOp{lam_lp_[5].hi, lam_lp_[5].lo, lam_lp_[5].ld}, gDp_, nvm_, 1.0, 1.0));
    }
  } else {
    ACES_CHECK(
        gemm64(gd((int)nop_, (int)nvp_, (int)(nvp_ + ov), gZp_, nop_, 0, Vp_, nvp_, 0, gSp_, nvp_, 1.0, 0.0)));
    if (nom_ > 0 && nvm_ > 0)
      ACES_CHECK(
          gemm64(gd((int)nom_, (int)nvm_, (int)(nvm_ + ov), gZm_, nom_, 0, Vm_, nvm_, 0, gDp_, nvm_, 1.0, 0.0)));
  }
  ACES_CHECK(pack_tau_adjoint(o, v, gSp_, nvp_, gDp_, std::max<long>(nvm_, 1), gtau_, s));

  // ---- (8) ring assembly: RS = (R1 - R2)/2 ; inc2 += RS[iajb] + RS[jbia] - R2[ibja] - R2[jaib] ----
  ACES_CHECK(P4(o, v, o, v, gRS_).src(l2, 1.0, ov2, v, vv, 1).src(l2, 1.0, vv, 1, ov2, v).run(s));
  ACES_CHECK(P4(o, v, o, v, gR2_)
                 .src(gRS_, -0.5, ov2, ov, v, 1)
                 .src(l2, -1.0, ov2, 1, vv, v)
                 .src(l2, -1.0, vv, v, ov2, 1)
                 .run(s));
  // ---- (7) rings: R1 = Wt U^T (gR1 = gRS/2), R2 = Wv X^T ----
  if (lp) {
    auto op = [](const LowP& L) { return LowPOp{L.hi, L.lo, L.ld}; };
    ACES_CHECK(adj_gemm_lp(OV, OV, OV, gRS_, ov, 1, 9, op(lam_lp_[0]), gWt_, ov, 0.5, 0.0));
    ACES_CHECK(adj_gemm_lp(OV, OV, OV, gRS_, ov, 0, 10, op(lam_lp_[1]), gU_, ov, 0.5, 1.0));
    ACES_CHECK(adj_gemm_lp(OV, OV, OV, gR2_, ov, 1, 11, op(lam_lp_[2]), gWv_, ov, 1.0, 0.0));
    ACES_CHECK(adj_gemm_lp(OV, OV, OV, gR2_, ov, 0, 12, op(lam_lp_[3]), gX_, ov, 1.0, 0.0));
    ACES_CHECK(adj_gemm_lp(OV, OV, OV, gWt_, ov, 1, 7, op(lp_[2]), gUt_, ov, 0.5, 0.0));
    ACES_CHECK(adj_gemm_lp(OV, OV, OV, gWv_, ov, 1, 8, op(lp_[3]), gXt_, ov, -0.5, 0.0));
  } else {
    ACES_CHECK(gemm64(gd(OV, OV, OV, gRS_, ov, 1, U_, ov, 0, gWt_, ov, 0.5, 0.0)));
    ACES_CHECK(gemm64(gd(OV, OV, OV, gRS_, ov, 0, Wt_, ov, 0, gU_, ov, 0.5, 1.0)));
    ACES_CHECK(gemm64(gd(OV, OV, OV, gR2_, ov, 1, X_, ov, 0, gWv_, ov, 1.0, 0.0)));
    ACES_CHECK(gemm64(gd(OV, OV, OV, gR2_, ov, 0, Wv_, ov, 0, gX_, ov, 1.0, 0.0)));
    // ---- (6) W build: Wt += Ut L^T / 2, Wv -= Xt K^T / 2  (L, K symmetric) ----
    ACES_CHECK(gemm64(gd(OV, OV, OV, gWt_, ov, 1, Lmat_, ov, 1, gUt_, ov, 0.5, 0.0)));
    ACES_CHECK(gemm64(gd(OV, OV, OV, gWv_, ov, 1, Kx_, ov, 1, gXt_, ov, -0.5, 0.0)));
  }
  // ---- (5) Wt = W0t + 2 P1[kcai] - P2[kaci] - Q12[kcia] ; Wv = Wv0 + P2[kaci] - Q2[kcia] ----
  ACES_CHECK(P4(o, v, v, o, gP1_, 1.0).src(gWt_, 2.0, v, 1, ov, ov2).run(s));
  ACES_CHECK(P4(o, v, v, o, gP2_).src(gWt_, -1.0, v, ov, 1, ov2).src(gWv_, 1.0, v, ov, 1, ov2).run(s));
  ACES_CHECK(P4(o, v, o, v, gQ12_).src(gWt_, -1.0, v, 1, ov2, ov).run(s));
  ACES_CHECK(P4(o, v, o, v, gQ2_).src(gWv_, -1.0, v, 1, ov2, ov).run(s));
  // ---- (4) P1 = ovvv t1, P2 = (kd|ac) t1, Q12 = M5c t1, Q2 = M5b t1 ----
  ACES_CHECK(gemm64(gd(o, v, (int)v2o, gP1_, o, 0, ovvv_, v, 0, gt1, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, (int)v2o, gP2_, o, 0, ovvv_kacd_, v, 0, gt1, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, (int)o2v, M5c_, o, 0, gQ12_, v, 0, gt1, v, 1.0, 1.0)));
  ACES_CHECK(gemm64(gd(o, v, (int)o2v, M5b_, o, 0, gQ2_, v, 0, gt1, v, 1.0, 1.0)));
  // ---- (3) Ut = U - 2 t_la t_id, Xt = X + 2 t_id t_la ----
  // H[(l,a)][(i,d)] = -2 gUt[i,a,l,d] + 2 gXt[i,a,l,d]
  ACES_CHECK(P4(o, v, o, v, gRS_).src(gUt_, -2.0, v, ov, ov2, 1).src(gXt_, 2.0, v, ov, ov2, 1).run(s));
  ACES_CHECK(gemv(OV, OV, gRS_, ov, t1, gt1, 1.0, 1.0, s));
  ACES_CHECK(gemv_t(OV, OV, gRS_, ov, t1, gt1, 1.0, 1.0, s, ws_, ws_elems_));
  // ---- (2) U = 2 t_ijab - t_ijba, X = t_ijba (with gU += gUt, gX += gXt) ----
  ACES_CHECK(P4(1, 1, 1, ovvo, gU_, 1.0).src(gUt_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(1, 1, 1, ovvo, gX_, 1.0).src(gXt_, 1.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(o, o, v, v, gt2, 1.0)
                 .src(gU_, 2.0, ov2, v, ov, 1)
                 .src(gU_, -1.0, ov2, v, 1, ov)
                 .src(gX_, 1.0, ov2, v, 1, ov)
                 .run(s));
  // ---- (1) tau = t2 + t1 t1 ----
  ACES_CHECK(P4(o, v, o, v, gR2_).src(gtau_, 1.0, ov2, v, vv, 1).run(s));
  ACES_CHECK(gemv(OV, OV, gR2_, ov, t1, gt1, 1.0, 1.0, s));
  ACES_CHECK(gemv_t(OV, OV, gR2_, ov, t1, gt1, 1.0, 1.0, s, ws_, ws_elems_));
  ACES_CHECK(P4(1, 1, 1, ovvo, gt2, 1.0).src(gtau_, 1.0, 0, 0, 0, 1).run(s));
  return ACES_OK;
}

int CcsdEngine::adjoint(const double* l1, const double* l2, double* r1, double* r2, double* dens) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_CHECK(adjoint_alloc());
  if (!fwd_valid_) ACES_CHECK(compute_increments(GP_F64));
  const long ov = (long)o * v;
  ACES_HIP(hipMemcpyAsync(lam_, l1, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipMemcpyAsync(lam_ + ov, l2, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  ACES_CHECK(compute_adjoint(lam_, dens != nullptr));
  ACES_CHECK(symmetrize_t2(o, v, gT_ + ov, s_));
  if (r1) ACES_HIP(hipMemcpyAsync(r1, gT_, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  if (r2) ACES_HIP(hipMemcpyAsync(r2, gT_ + ov, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  if (dens) ACES_CHECK(assemble_density(o, v, gfo_, gfv_, gfov_, geo_, gev_, dens, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

// l = from_aces(L): l1 = 2 L1, l2 = 2 L2 - L2(i<->j)
static int lambda_from_aces(int o, int v, const double* L, double* lam, hipStream_t s) {
  const long ov = (long)o * v, vv = (long)v * v, ov2 = ov * v;
  ACES_CHECK(P4(1, 1, 1, ov, lam).src(L, 2.0, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(o, o, v, v, lam + ov).src(L + ov, 2.0, ov2, vv, v, 1).src(L + ov, -1.0, vv, ov2, v, 1).run(s));
  return ACES_OK;
}

// L = to_aces(l): L1 = l1/2, L2 = (2 l2 + l2(i<->j))/3
static int lambda_to_aces(int o, int v, const double* lam, double* L, hipStream_t s) {
  const long ov = (long)o * v, vv = (long)v * v, ov2 = ov * v;
  ACES_CHECK(P4(1, 1, 1, ov, L).src(lam, 0.5, 0, 0, 0, 1).run(s));
  ACES_CHECK(P4(o, o, v, v, L + ov)
                 .src(lam + ov, 2.0 / 3.0, ov2, vv, v, 1)
                 .src(lam + ov, 1.0 / 3.0, vv, ov2, v, 1)
                 .run(s));
  return ACES_OK;
}

int CcsdEngine::lambda_start() {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_CHECK(adjoint_alloc());
  const bool lists = tn_lists_;
  ACES_CHECK(compute_increments(GP_F64));
  const long ov = (long)o * v;
  // Lambda start: what xvcc leaves on the increment lists (61/63, 90 group 3)
  // after the converged cycle and moves to the Lambda lists 144/146/190 with
  // INIT2 + DRMOVE(...,100,SING1) (vcc/vcc.F:1130-1143): the converged cycle's
  // DIIS extrapolant (vcc/dodiis0.F:181-233), kept in t1n_/t2n_ by diis_step.
  // Without a converged cycle at these amplitudes: their Jacobi update,
  // Lambda1 = T1_new and Lambda2 = SYMETRIZE_T2(T2_new) (the lists' content
  // when the run converges before the subspace holds two vectors).
  if (lists) {
    ACES_HIP(hipMemcpyAsync(Lam_, t1n_, sizeof(double) * (ov + ov * ov), hipMemcpyDeviceToDevice, s_));
  } else {
    ACES_CHECK(jacobi(o, v, eo_, ev_, inc1_, inc2_, Lam_, Lam_ + ov, s_));
    ACES_CHECK(symmetrize_t2(o, v, Lam_ + ov, s_));
  }
  ldiis_.R.assign((size_t)ldiis_.order * ldiis_.order, 0.0);
  lam_anc_valid_ = false;
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

int CcsdEngine::set_lambda_precision(int prec) {
  ACES_REQUIRE(prec == GP_F64 || prec == kPrecRefine, "lambda precision must be 0 (fp64) or 3 (refinement)");
  ACES_CHECK(adjoint_alloc());
  if (prec == kPrecRefine && lamA_ == nullptr) {
    const long nt = (long)o * v + (long)o * v * o * v;
    ACES_CHECK(alloc(&lamA_, nt));
    ACES_CHECK(alloc(&gTA_, nt));
    ACES_CHECK(alloc(&dlam_, nt));
  }
  lam_prec_ = prec;
  lam_anc_valid_ = false;
  return ACES_OK;
}

int CcsdEngine::set_lambda_diis_order(int order) {
  ACES_CHECK(adjoint_alloc());
  return diis_resize(ldiis_, order);
}

int CcsdEngine::lambda_cycle(int icycle, double tol, CcsdStats& st) {
  ACES_REQUIRE(ready_ && adj_ready_, "lambda: call lambda_start first");
  ACES_REQUIRE(fwd_valid_, "lambda: amplitudes changed since lambda_start");
  ACES_REQUIRE(icycle >= 1, "lambda cycle: icycle starts at 1");
  const long ov = (long)o * v;
  hipEvent_t e0, e1;
  ACES_HIP(hipEventCreate(&e0));
  ACES_HIP(hipEventCreate(&e1));
  ACES_HIP(hipEventRecord(e0, s_));
  ACES_CHECK(lambda_from_aces(o, v, Lam_, lam_, s_));
  const long nt = ov + ov * ov;
  const bool anchor = (lam_prec_ == kPrecRefine && !lam_anc_valid_);
  if (lam_prec_ == kPrecRefine && lam_anc_valid_) {
    // r(l) = r(l_a) + J^T (l - l_a)
    ACES_CHECK(sub(nt, lam_, lamA_, dlam_, s_));
    ACES_CHECK(compute_adjoint(dlam_, false, GP_BF16X3, true));
    ACES_CHECK(P4(1, 1, 1, nt, gT_, 1.0).src(gTA_, 1.0, 0, 0, 0, 1).run(s_));
  } else {
    ACES_CHECK(compute_adjoint(lam_, false));
    if (anchor) {
      ACES_HIP(hipMemcpyAsync(lamA_, lam_, sizeof(double) * nt, hipMemcpyDeviceToDevice, s_));
      ACES_HIP(hipMemcpyAsync(gTA_, gT_, sizeof(double) * nt, hipMemcpyDeviceToDevice, s_));
    }
  }
  ACES_CHECK(symmetrize_t2(o, v, gT_ + ov, s_));
  ACES_CHECK(lambda_jacobi(o, v, eo_, ev_, Lam_, gT_, Lamn_, s_));
  ACES_HIP(hipEventRecord(e1, s_));
  ACES_CHECK(diis_core(ldiis_, Lam_, Lamn_, icycle, tol, true, st));
  if (anchor) {
    lam_anc_valid_ = true;
    lam_anc_res_ = std::fabs(st.resmiter);
  } else if (lam_prec_ == kPrecRefine && std::fabs(st.resmiter) < 1e-3 * lam_anc_res_) {
    lam_anc_valid_ = false;  // the next cycle anchors at the current multipliers
  }
  float ms = 0.f;
  ACES_HIP(hipEventSynchronize(e1));
  ACES_HIP(hipEventElapsedTime(&ms, e0, e1));
  st.ms_increments = ms;
  st.ms_cycle = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ACES_OK;
}

int CcsdEngine::get_lambda(double* l1, double* l2, int aces_normalisation) {
  ACES_REQUIRE(ready_ && adj_ready_, "lambda: no multipliers (call lambda_start)");
  const long ov = (long)o * v;
  const double* src = Lam_;
  if (!aces_normalisation) {
    ACES_CHECK(lambda_from_aces(o, v, Lam_, lam_, s_));
    src = lam_;
  }
  if (l1) ACES_HIP(hipMemcpyAsync(l1, src, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  if (l2) ACES_HIP(hipMemcpyAsync(l2, src + ov, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

int CcsdEngine::set_lambda(const double* l1, const double* l2, int aces_normalisation) {
  ACES_REQUIRE(ready_, "ccsd: not set up");
  ACES_CHECK(adjoint_alloc());
  const long ov = (long)o * v;
  double* dst = aces_normalisation ? Lam_ : lam_;
  ACES_HIP(hipMemcpyAsync(dst, l1, sizeof(double) * ov, hipMemcpyDeviceToDevice, s_));
  ACES_HIP(hipMemcpyAsync(dst + ov, l2, sizeof(double) * ov * ov, hipMemcpyDeviceToDevice, s_));
  if (!aces_normalisation) ACES_CHECK(lambda_to_aces(o, v, lam_, Lam_, s_));
  lam_anc_valid_ = false;
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

int CcsdEngine::density(double* dens) {
  ACES_REQUIRE(ready_ && adj_ready_, "density: no multipliers (call lambda_start)");
  if (!fwd_valid_) ACES_CHECK(compute_increments(GP_F64));
  ACES_CHECK(lambda_from_aces(o, v, Lam_, lam_, s_));
  ACES_CHECK(compute_adjoint(lam_, true));
  ACES_CHECK(assemble_density(o, v, gfo_, gfv_, gfov_, geo_, gev_, dens, s_));
  ACES_HIP(hipStreamSynchronize(s_));
  return ACES_OK;
}

}  // namespace aces
