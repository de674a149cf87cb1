// Closed-shell CCSD amplitude engine (RHF, C1) on one gfx950 GPU.
//
// One Jacobi/DIIS cycle of the reference's xvcc loop (vcc/vcc.F:798-1437):
// increments (GENINT + FEACONT + FMICONT + FMECONT + T1INT* + DRE3EN + E4S),
// NEWT2, CMPENG, DODIIS0, DRMOVE.  Integrals and amplitudes stay resident in
// HBM; the contractions run on the MFMA GEMM engine (gemm.h), the
// permutations/packings on HBM-bound kernels, the DIIS solve on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "gemm.h"

namespace aces {

// Cycle precision of the engine beyond the GEMM library's three: the eight
// large products (ladders, W build, ring) as bf16x3 products of the operands'
// differences to an fp64 anchor evaluation (CcsdEngine::gemm_anchored).  The
// first such cycle after set-up or reset_anchor() records the anchor; a new
// one is recorded each time |RESMITER| has fallen 1e3-fold below the anchor
// cycle's (the differences, and with them the bf16x3 error, stay small).
constexpr int kPrecRefine = 3;

struct CcsdStats {
  double energy_new = 0.0;   // CMPENG(60,2): E(T_new) printed each cycle
  double resmiter = 0.0;     // largest element of the residual (signed)
  double resmdiis = 0.0;     // largest element of the DIIS residual
  int converged = 0;         // |resmiter| < tol
  double energy_final = 0.0; // CMPENG(43,0) on the T of the converged cycle
  double ms_increments = 0.0;
  double ms_cycle = 0.0;
};

class CcsdEngine {
 public:
  CcsdEngine(int o, int v, hipStream_t s);
  ~CcsdEngine();

  // Integral input (device pointers, chemists' notation, row major):
  //   fock[n][n]; ovov[i][a][j][b]=(ia|jb); oovv[i][j][a][b]=(ij|ab);
  //   ovoo[i][a][j][k]=(ia|jk); oooo[i][j][k][l]=(ij|kl); ovvv[i][a][b][c]=(ia|bc);
  //   L233 = reference list 233 (rows A<=B, columns Cd) holding <ab|cd>=(ac|bd).
  int setup(const double* fock, const double* ovov, const double* oovv, const double* ovoo, const double* oooo,
            const double* ovvv, const double* L233);
  // Amplitudes in/out: t1[i][a], t2[i][j][a][b] (device).
  int set_amplitudes(const double* t1, const double* t2);
  int get_amplitudes(double* t1, double* t2);
  int start_mp1();
  // The residual increments for the current amplitudes (device outputs).
  int increments(GemmPrec prec, double* inc1_out, double* inc2_out);
  // One full cycle (increments, Jacobi, energy, DIIS, move).
  int cycle(int icycle, GemmPrec prec, double tol, CcsdStats& st);
  // Pieces of one cycle, used by the list-level driver entry points.
  int jacobi_step();
  int diis_step(int icycle, double tol, CcsdStats& st, bool xn_sym = false);
  // Writable views; writing through T() invalidates the resident forward
  // intermediates, writing through either the Lambda start kept in Tnew.
  double* T() { fwd_valid_ = false; tn_lists_ = false; return t1_; }  // [t1 | t2] current amplitudes
  double* Tnew() { tn_lists_ = false; return t1n_; }  // [t1 | t2] Jacobi update
  const double* T_ro() const { return t1_; }
  const double* Tnew_ro() const { return t1n_; }
  double* Inc() { return inc1_; }    // [inc1 | inc2]
  int energy_dev(const double* t1, const double* t2, double* d_out);
  int reset_diis();
  // Forget the refinement anchor (the next kPrecRefine cycle records a new one).
  void reset_anchor() { anc_valid_ = false; }
  int set_diis_order(int order);
  int energy_of(const double* t1, const double* t2, double& e);
  // Replace the MO Fock matrix (n x n, device) keeping the integrals.
  int set_fock(const double* fock);
  // Per-launch timing of the MFMA contractions (hipEvents on the engine's stream).
  // out[3*c + 0..2] = {ms, algorithmic flops, launches} for c = 0 ladder, 1 W build,
  // 2 ring, 3 the other fp64 contractions of the increments, 4 the reverse pass,
  // 5 the reverse pass's bf16x3 products (mixed Lambda cycles), 6 the (T) products.
  void set_profiling(bool on) { profile_ = on; }
  void profile(double* out, int reset);
  long device_bytes() const { return bytes_; }

  // ---- differentiation form (ccsd_adjoint.hip) ----
  // Lagrangian L = E(T) + sum l1*Omega1 + sum l2*Omega2, Omega = inc - D*T, with
  // spatial multipliers l1[i][a], l2[i][j][a][b].  adjoint() is the reverse pass
  // of compute_increments() plus the energy: r = dL/dT at the amplitudes whose
  // forward intermediates are resident (r2 symmetrised); with dens != nullptr
  // it also returns dL/dfock, the response one-particle density (n x n, symmetric).
  int adjoint(const double* l1, const double* l2, double* r1, double* r2, double* dens);
  // xlambda loop (lambda/vlamcc.F:599-830): lambda_start() forms the forward
  // intermediates at the current amplitudes and starts from their Jacobi update
  // (Lambda1 = T1_new, Lambda2 = T2_new, the lists xvcc hands to xlambda);
  // each lambda_cycle() is one Jacobi step plus the Lambda DODIIS0.
  int lambda_start();
  int lambda_cycle(int icycle, double tol, CcsdStats& st);
  int set_lambda_diis_order(int order);
  // Lambda cycle precision: GP_F64 (0), or kPrecRefine (3): the residual as
  // r(l) = r(l_a) + J^T (l - l_a) with r(l_a) an fp64 reverse pass at an anchor
  // l_a and the linear part (the reverse pass without the energy seeds) on the
  // difference with its eight large products in bf16x3; the anchor moves to
  // the current multipliers each time |RESMITER| has fallen 1e3-fold below
  // the anchor cycle's.
  int set_lambda_precision(int prec);
  // Multipliers in/out: spatial l1/l2 (device) or the reference's lists
  // (L1 = Lambda1 of list 190, L2 = Lambda(Ij,Ab) of list 146).
  int get_lambda(double* l1, double* l2, int aces_normalisation);
  int set_lambda(const double* l1, const double* l2, int aces_normalisation);
  // Response density at the current multipliers (n x n, device).
  int density(double* dens);

  // ---- EOM-CCSD excitation energies (ccsd_eom.hip) ----
  // The nroots lowest singlet excitation energies at the current (converged)
  // amplitudes: eigenvalues of the CC Jacobian, from its left eigenvectors
  // (products J^T l by the reverse pass).  orb_irrep (host, n entries,
  // occupied then virtual; nullptr: no symmetry) selects the excitations of
  // irrep `target` (XOR of the orbital irreps).  omega/resid (host, nroots):
  // eigenvalues and residual norms; iters: Davidson iterations; d_vecs (device,
  // optional, nroots x (o v + o^2 v^2)): the left eigenvectors [l1 | l2].
  int eom(const int* orb_irrep, int target, int nroots, double tol, int maxit, double* omega, double* resid,
          int* iters, double* d_vecs);

  // ---- non-iterative triples (triples.hip) ----
  // (T) of CCSD(T) at the current amplitudes (pair-symmetric t2, canonical HF
  // orbital energies from the Fock diagonal): out[0] E[4]T, out[1] E[5]ST,
  // out[2] E4TAAA, out[3] E4TBAA (E[4]T = 2 (E4TAAA + E4TBAA)), out[4] ms,
  // out[5] MFMA flops, out[6] number of occupied triples i >= j >= k.
  // d_per_triple (device, 3 per triple, optional): each triple's unweighted
  // sums (E[4]T, E[4]T + E[5]ST, E4TAAA pieces) in the order i, j <= i, k <= j.
  // prec: GP_F64 or GP_BF16X3 (the W products; everything else fp64).
  int triples(GemmPrec prec, double* out, double* d_per_triple);
  double* Lam() { return Lam_; }     // [Lambda1 | Lambda2(AB)] current
  double* Lamnew() { return Lamn_; }

  int o, v, n;

 private:
  struct LowP {
    void* hi = nullptr;
    void* lo = nullptr;
    long ld = 0;
    long rows = 0;
  };
  struct LowPOp {  // a bf16 split operand as the GEMM reads it
    const void* hi = nullptr;
    const void* lo = nullptr;
    long ld = 0;
  };
  int alloc(double** p, long count);
  int alloc_raw(void** p, long bytes);
  int compute_increments(GemmPrec prec);
  int lowp_static(GemmPrec prec);
  // Contraction with operands of the given precision; A64/B64 are the fp64
  // sources (a_kc/b_kc as in GemmDesc), converted into their operand slot
  // unless the slot is static.
  // cslot >= 0 (bf16x3 only): the result goes to that operand slot as its
  // bf16 split for a later product, and C is not written.
  int gemm_lp(GemmPrec prec, int M, int N, int K, const double* A64, int aslot, long lda, const double* B64,
              int bslot, long ldb, double* C, long ldc, double alpha, double beta, int a_kc = 1, int b_kc = 1,
              int cslot = -1);
  int timed_gemm(const GemmDesc& d, GemmPrec prec, double flops);
  // Anchored refinement of one product C = beta C + alpha op(A) op(B):
  //   recording:  C0 = op(A0) op(B0) in fp64, copies of the amplitude-dependent
  //               operands A0/B0 kept; C = beta C + alpha C0
  //   refining:   C = beta C + alpha (C0 + op(A) op(B - B0) + op(A - A0) op(B0))
  //               with the last two products in bf16x3 (one of them when an
  //               operand is a static integral block, whose difference is zero)
  // The bf16x3 error of a difference product scales with the difference, so
  // near convergence the result carries fp64 accuracy at bf16x3 speed.
  int gemm_anchored(int M, int N, int K, const double* A64, int aslot, long lda, const double* B64, int bslot,
                    long ldb, double* C, long ldc, double alpha, double beta, int a_kc, int b_kc);
  int anchor_copy(const double* src, long elems, const double** copy);
  int split_operand(LowPOp& op, int slot, const double* src, const double* sub, long rows, int K, long ld, int kc);
  int lp_product(int M, int N, int K, const LowPOp& A, const LowPOp& B, double* C, long ldc, double alpha,
                 double beta);
  int gemm64(GemmDesc d);  // fp64 GEMM with the engine's split-K workspace (timed as prof_cat_)
  int gemm_lp_impl(GemmPrec prec, int M, int N, int K, const double* A64, int aslot, long lda, const double* B64,
                   int bslot, long ldb, double* C, long ldc, double alpha, double beta, int a_kc, int b_kc,
                   int cslot);
  struct DiisState {
    int order = 6;
    std::vector<double*> sol, err;
    std::vector<double> R;  // order x order (row major)
  };
  // DODIIS0 core on (x, xn): error, B matrix, dgesv, DIIS residual and
  // extrapolation into x.  lambda_rule: converged when |RESMITER| or
  // |RESMDIIS| < tol and the DIIS residual summed in slot order
  // (lambda/dodiis0.F); otherwise the amplitude rule of vcc/dodiis0.F.
  // x_sym / xn_sym: the T2 part of x / xn is already exactly pair-symmetric
  // (SYMETRIZE_T2 would leave it unchanged and is skipped)
  int diis_core(DiisState& D, double* x, double* xn, int icycle, double tol, bool lambda_rule, CcsdStats& st,
                bool x_sym = false, bool xn_sym = false);
  int diis_resize(DiisState& D, int order);
  int adjoint_alloc();
  // lin: the linear part only (no energy seeds); prec GP_BF16X3 runs the ring,
  // W-build and ladder products in bf16x3
  int compute_adjoint(const double* lam, bool density, GemmPrec prec = GP_F64, bool lin = false);
  // reverse-pass product C = alpha op(A) B^T + beta C in bf16x3: A (fp64,
  // k-contiguous when akc) split into operand slot aslot, B a fixed split
  int adj_gemm_lp(int M, int N, int K, const double* A, long lda, int akc, int aslot, const LowPOp& B, double* C,
                  long ldc, double alpha, double beta);
  int lambda_lowp_operands();

  hipStream_t s_;
  long bytes_ = 0;
  bool profile_ = false;
  static constexpr int kProfCats = 7;
  double prof_[3 * kProfCats] = {};
  int prof_cat_ = 0;
  std::vector<void*> allocs_;
  bool ready_ = false;

  // static
  double *fock_ = nullptr, *eo_ = nullptr, *ev_ = nullptr;
  double *ovov_ = nullptr, *oovv_ = nullptr, *ovoo_ = nullptr, *ovvv_ = nullptr, *ovvv_kacd_ = nullptr;
  double *M3_ = nullptr, *Lmat_ = nullptr, *Kx_ = nullptr, *W0t_ = nullptr, *Wv0_ = nullptr;
  double *Lo_ = nullptr, *Lv_ = nullptr, *OVkl_ = nullptr, *ooooP_ = nullptr, *ovooQ_ = nullptr;
  double *M1_ = nullptr, *M5c_ = nullptr, *M5b_ = nullptr, *M7_ = nullptr, *Lijab_ = nullptr;
  double *Vp_ = nullptr, *Vm_ = nullptr;
  double *ovvvL_ = nullptr;  // [k][c][a][d] = 2(kc|ad) - (kd|ac): the t1 term of Lvv
  long nvp_, nvm_, nop_, nom_;
  // amplitudes and work
  double *t1_, *t2_, *t1n_, *t2n_, *inc1_, *inc2_, *tau_;
  double *U_, *X_, *Ut_, *Xt_, *Wt_, *Wv_, *R1_, *R2_;
  double *Sp_, *Dp_, *Zp_, *Zm_, *Yb_, *L1_, *tauK_, *P1_, *P2_, *Q12_, *Q2_, *Qb_, *T1Q_, *S3_, *G4A_, *G4_;
  double *Wo_, *Qo_, *tmpL_;
  double *Fov_, *Foo_, *Fvv_, *Loo_, *Lvv_, *FT_, *FovT_, *Gk_;
  double *red_ = nullptr;  // reduction scratch
  double* jpart_ = nullptr;  // energy partials of the fused Jacobi pass
  bool t_sym_ = false;       // t2_ is known to be exactly pair-symmetric (set by DIIS)
  // t1n_/t2n_ hold what the reference leaves on the increment lists after a
  // converged cycle at the current amplitudes (the DIIS extrapolant, or the
  // symmetrized Jacobi update before the subspace has two vectors): the
  // Lambda start (set by diis_step on convergence, cleared by any change)
  bool tn_lists_ = false;
  double *ws_ = nullptr;   // split-K workspace
  long ws_elems_ = 0;
  double *scal_ = nullptr; // device scalars
  // low precision operand slots: static (converted once per precision)
  // 0 Vp 1 Vm 2 Lmat 3 Kx 4 OVkl; per call 5 Sp 6 Dp 7 Ut 8 Xt 9 Wt 10 Wv
  // 11 U 12 X 13 tau 14 Wo^T 15 tau^T; 31, 32 the ladder adjoints' T1-dressing rows
  // slots 0..4, 16..19 and 26..30 hold static operands (converted once per
  // precision), the others this cycle's operands
  static constexpr int kLpStatic = 5, kLpStatic2 = 16, kLpSlots = 33;
  static constexpr bool lp_static_slot(int slot) {
    return slot < kLpStatic || (slot >= kLpStatic2 && slot < 20) || (slot >= 26 && slot < 31);
  }
  LowP lp_[kLpSlots];
  bool lp_fresh_[kLpSlots] = {};  // slot already holds this cycle's operand (bf16x3)
  bool lp_keep_[kLpSlots] = {};   // operand fixed for the whole cycle: convert once, use many times
  int lowp_slot(int slot, long rows, long K);
  GemmPrec lp_static_prec_ = GP_F64;
  // refinement anchor (gemm_anchored): 0 off, 1 recording, 2 refining
  int anc_mode_ = 0;
  bool anc_valid_ = false;
  int anc_last_ = 0;       // mode of the last evaluation of the increments
  double anc_res_ = 0.0;   // |RESMITER| of the anchor cycle
  int anc_site_ = 0;  // product index within one evaluation of the increments
  struct AncSite {
    double* C0 = nullptr;
    long elems = 0;
    LowP B0;  // split of the anchor's right operand (both operands amplitude-dependent)
  };
  std::vector<AncSite> anc_sites_;
  struct AncSrc {
    const double* src = nullptr;
    double* copy = nullptr;
    long elems = 0;
  };
  std::vector<AncSrc> anc_srcs_;
  // DIIS (amplitudes and multipliers)
  DiisState tdiis_, ldiis_;
  double* newvec_ = nullptr;
  const double** d_ptrs_ = nullptr;  // device pointer table
  double* d_coef_ = nullptr;

  // adjoint buffers (allocated on first use)
  bool adj_ready_ = false;
  bool fwd_valid_ = false;  // forward intermediates belong to the current t1_/t2_
  double *lam_ = nullptr, *gT_ = nullptr, *Lam_ = nullptr, *Lamn_ = nullptr;
  double *gtau_, *gU_, *gX_, *gUt_, *gXt_, *gWt_, *gWv_, *gRS_, *gR2_, *gP1_, *gP2_, *gQ12_, *gQ2_;
  double *gZp_, *gZm_, *gSp_, *gDp_, *gYb_, *gL1_, *gWo_, *gQo_, *gtmpL_, *gT1Q_, *gQb_, *gG4_, *gG4A_;
  double *gS3_, *gtauK_;
  double *gFovI_, *gFooI_, *gFvvI_, *gLoo_, *gLvv_, *gFT_, *gFovT_, *gGk_;
  double *gfo_, *gfv_, *gfov_, *geo_, *gev_;
  // mixed Lambda cycles (set_lambda_precision)
  int lam_prec_ = 0;
  bool lam_anc_valid_ = false;
  double lam_anc_res_ = 0.0;
  double *lamA_ = nullptr, *gTA_ = nullptr, *dlam_ = nullptr;
  // splits fixed while T is: U^T, Wt^T, X^T, Wv^T, the transposed T1-dressing
  // rows of V+ and V-, Wo, tau, OVkl^T
  static constexpr int kLamLp = 9;
  LowP lam_lp_[kLamLp];
  bool lam_lp_valid_ = false;
  // (T) buffers (allocated on first use): W (v^3), canonical abc tile table,
  // energy-pass partials, per-triple sums
  double *trW_ = nullptr, *trPart_ = nullptr, *trOut_ = nullptr;
  int* trTiles_ = nullptr;
  long trNblk_ = 0;
};

}  // namespace aces
