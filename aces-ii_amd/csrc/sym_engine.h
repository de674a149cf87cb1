// Closed-shell CCSD on irrep-blocked lists (abelian point groups up to D2h),
// one gfx950 GPU: amplitude equations, Lambda equations and response density.
//
// Same equations as CcsdEngine (C1), restated on irrep-blocked storage
// (dpd.h): every contraction term is one grouped MFMA launch over the irrep
// blocks (gemm_grouped), every index permutation one box-plan launch.  The
// increments are recorded once as a tape of operations; running the tape
// forwards gives the CC increments, running it backwards with the generic
// adjoint of each operation gives the Lambda residual and dL/dfock.
//
// Orbitals are irrep-sorted inside the occupied and the virtual space; the
// dense interfaces (setup, amplitudes in/out, densities) use C1 arrays in
// that order, with symmetry-forbidden elements zero.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "ccsd_engine.h"
#include "dpd.h"
#include "gemm.h"
#include "sym_ladder.h"

namespace aces {

class SymEngine {
 public:
  SymEngine(int nirr, const int* occ, const int* vir, hipStream_t s);
  ~SymEngine();

  // Dense C1 integral blocks (device, chemists' notation, row major, the
  // CcsdEngine::setup conventions); only their symmetry-allowed elements are read.
  int setup(const double* fock, const double* ovov, const double* oovv, const double* ovoo, const double* oooo,
            const double* ovvv, const double* L233);
  int set_amplitudes(const double* t1, const double* t2);  // dense
  int get_amplitudes(double* t1, double* t2);              // dense
  int start_mp1();
  int increments(GemmPrec prec, double* inc1, double* inc2);  // dense outputs (may be null)
  int cycle(int icycle, GemmPrec prec, double tol, CcsdStats& st);
  // DODIIS0 of the list driver (vcc/dodiis0.F) on T and T_new given as dense
  // arrays: st.energy_new = E(T_new); unless converged the next amplitudes
  // (the extrapolated vector) are the engine's current ones (get_amplitudes),
  // converged: the current ones are T, pair-symmetrized, st.energy_final E(T),
  // and get_increment_lists returns the extrapolant DODIIS0 writes over the
  // increment lists in that cycle too (the Lambda start).
  int diis_dense(int icycle, double tol, const double* t1, const double* t2, const double* t1n, const double* t2n,
                 CcsdStats& st);
  int energy_current(double& e);
  // The content of the increment lists after the last DODIIS0 (dense).
  int get_increment_lists(double* t1, double* t2);
  int set_diis_order(int order);
  int reset_diis();
  // prec kPrecRefine (3) as in CcsdEngine: the large products (cat <= 2) as
  // bf16x3 products of the operands' differences to an fp64 anchor evaluation
  void reset_anchor() { anc_valid_ = false; }
  // differentiation form (dense multipliers/densities, spatial normalisation
  // of CcsdEngine::adjoint)
  int adjoint(const double* l1, const double* l2, double* r1, double* r2, double* dens);
  int lambda_start();
  int lambda_cycle(int icycle, double tol, CcsdStats& st);
  // prec 0 (fp64) or kPrecRefine (3): residual = fp64 reverse pass at an anchor
  // plus the linear reverse pass on the difference to it, whose large products
  // (ladder, W build, ring) run as bf16x3 products (CcsdEngine's mixed cycles)
  int set_lambda_precision(int prec);
  int set_lambda_diis_order(int order);
  int get_lambda(double* l1, double* l2, int aces_normalisation);
  int density(double* dens);
  void set_profiling(bool on) { profile_ = on; }
  void profile(double* out, int reset);
  long device_bytes() const { return bytes_; }
  long packed_t2() const;

  int o, v, n, nirr;

 private:
  struct Tensor {
    dpd::Layout L;
    bool has_layout = false;
    long nrow[dpd::kMaxIrr] = {0}, ncol[dpd::kMaxIrr] = {0}, boff[dpd::kMaxIrr + 1] = {0};
    double* d = nullptr;
    double* g = nullptr;
    bool grad = false;
    long ver = 0;  // bumped whenever the forward pass (re)writes the tensor
  };
  enum Kind { OP_PERM, OP_GEMM, OP_GEMV, OP_SUBDIAG, OP_PACK, OP_UNPACK };
  struct Op {
    int kind = OP_PERM;
    int out = -1;
    double beta = 0.0;
    // permutation
    int nsrc = 0;
    int src[3] = {-1, -1, -1};
    double alpha[3] = {0, 0, 0};
    int map[3][4] = {{0, 1, 2, 3}, {0, 1, 2, 3}, {0, 1, 2, 3}};
    int X = -1, Y = -1;
    int xmap[2] = {0, 1}, ymap[2] = {0, 1};
    double gamma = 0.0;
    dpd::Plan fwd, bwd[3], hplan;
    int H = -1;
    // gemm / gemv: out = alpha op(A) op(B)^T + beta out ; gemv: out = alpha A0 x (trans: A0^T x)
    int A = -1, B = -1, a_kc = 1, b_kc = 1, trans = 0;
    double al = 1.0;
    int cat = 3;
    bool lowp = false;
    // subdiag: out += al * diag(e), e = 0: eo, 1: ev
    int evec = 0;
  };
  struct LowBuf {
    void* hi = nullptr;
    void* lo = nullptr;
    long off[dpd::kMaxIrr + 1] = {0};
    long ld[dpd::kMaxIrr] = {0};
    GemmPrec prec = GP_F64;
    bool valid = false;
    long ver = -1;  // tensor version the split was made from
  };

  int alloc(double** p, long count);
  int tensor(const dpd::Layout& L, bool grad, double* data = nullptr, double* gdata = nullptr);
  int block_tensor(const long* nrow, const long* ncol, bool grad);
  int perm(int out, double beta, std::vector<std::pair<int, std::vector<int>>> srcs, std::vector<double> alpha,
           int X = -1, std::vector<int> xmap = {}, int Y = -1, std::vector<int> ymap = {}, double gamma = 0.0,
           bool record = true);
  int gemm_op(int C, int A, int a_kc, int B, int b_kc, double alpha, double beta, int cat, bool lowp);
  int gemv_op(int y, int A, int x, int trans, double alpha, double beta);
  int ggemm(int C, const double* Cd, int A, const double* Ad, int a_kc, int B, const double* Bd, int b_kc,
            double alpha, double beta, GemmPrec prec, int cat, const LowBuf* la, const LowBuf* lb);
  int gemv_run(int y, double* yd, int A, const double* Ad, const double* xd, int trans, double alpha, double beta);
  int outer0(int A, double* Ad, const double* u, const double* w, double alpha);  // A0 += alpha u w^T
  // sub: split the difference T - sub (sub laid out like the tensor)
  int lowp_convert(int t, GemmPrec prec, LowBuf& lb, const double* sub = nullptr);
  int gemm_anchored(size_t k, int mode);
  int run_perm(Op& op, double* out, const double* const* src, const double* X, const double* Y,
               const dpd::Denom* den = nullptr);
  int forward(GemmPrec prec);
  // seeds: add dE/dT (false: the linear part J^T lam alone); lowp: the adjoint
  // products of the low-precision large forward products as bf16x3 products
  int backward(const double* lam1, const double* lam2, bool density, bool seeds = true, bool lowp = false);
  // bf16 hi/lo split of tensor t's blocks from data (laid out like t), as
  // stored ([row][col], trans false) or turned round ([col][row])
  int lowp_split(int t, const double* data, bool trans, LowBuf& lb, bool f64 = false);
  // fp64 copy of tensor t's blocks with rows padded to a multiple of 32
  // (zeros): operands of the direct-to-LDS fp64 kernel whatever the block widths
  int copy64(int t, LowBuf& lb);
  int padded_f64(int t);
  bool f64_wide(const Op& op) const;
  bool skinny(const Op& op) const;
  int low_alloc(int t, LowBuf& lb);
  void plan_split_fusion();
  std::vector<char> fuse_split_;  // per op: the permute also writes its output's bf16x3 split
  int pack_dense(int t, const double* dense, const long* strides, long base_off);
  int unpack_dense(int t, const double* packed, double* dense, const long* strides, long base_off);
  int denom_apply(int t, double* out, double beta, const double* in, double alpha, const double* in2,
                  double alpha2, int pmap, int pow);
  int pair_symmetrize(double* t2);
  int energy_dev(const double* t1, const double* t2, double& e);
  int diis_core(std::vector<double*>& sol, std::vector<double*>& err, std::vector<double>& R, int order, double* x,
                double* xn, int icycle, double tol, bool lambda_rule, CcsdStats& st);
  int diis_grow(std::vector<double*>& sol, std::vector<double*>& err, int order);
  int absmax_vec(const double* e, double& val);
  int assemble_dense_density(double* dens);
  int lam_from_aces(const double* L, double* lam);
  int timed(int cat, double flops, double ms);

  hipStream_t s_;
  long bytes_ = 0;
  std::vector<void*> allocs_;
  bool ready_ = false, profile_ = false;
  double prof_[15] = {0};
  dpd::Space O_, V_;
  std::vector<Tensor> T_;
  std::vector<Op> ops_;
  std::vector<dpd::Plan> plans_;  // auxiliary plans owned by the engine
  std::vector<LowBuf> low_;
  dpd::LadderDesc lad_;
  dpd::LadderDesc* d_lad_ = nullptr;
  dpd::UnpackTile* d_unpack_tiles_ = nullptr;  // tiles of the ladder unpack into inc2
  long n_unpack_tiles_ = 0;
  dpd::Layout* d_layouts_ = nullptr;  // device copies: 0 tau/t2 (O,O|V,V), 1 Yb
  double *eo_ = nullptr, *ev_ = nullptr, *geo_ = nullptr, *gev_ = nullptr;
  double *scal_ = nullptr, *ws_ = nullptr, *scr_ = nullptr, *scr2_ = nullptr;
  long ws_elems_ = 0, scr_elems_ = 0;
  void* red_ = nullptr;
  const double** dptr_ = nullptr;
  double* coef_ = nullptr;
  double* dscr_ = nullptr;
  double* absp_ = nullptr;
  // tensor ids
  int t1_, t2_, fov_, foo_, fvv_, tau_, inc1_, inc2_, Lijab_, Yb_, Sp_, Dp_, Zp_, Zm_, Vp_, Vm_;
  int pack_op_ = -1, unpack_op_ = -1, tau_op_ = -1, last_op_ = -1;
  dpd::Plan jac1_, jac2_, sym2_, perm2_, absplan_, comb_;
  bool t_sym_ = false;  // T_cur_'s t2 is known to be exactly pair-symmetric (set by DIIS)
  void* d_pair_tiles_ = nullptr;  // tiles of the in-place pair symmetrization (sym2_)
  long n_pair_tiles_ = 0;
  dpd::Layout aux_o_, aux_v_, aux_vo_;
  // amplitude state (packed [t1 | t2]) and DIIS
  double *T_cur_ = nullptr, *T_new_ = nullptr, *newvec_ = nullptr, *eperm_ = nullptr;
  long nt_ = 0, n1_ = 0;
  int order_ = 6, lorder_ = 6;
  std::vector<double*> sol_, errv_, lsol_, lerr_;
  std::vector<double> R_, LR_;
  // Lambda state
  double *Lam_ = nullptr, *Lamn_ = nullptr, *lam_ = nullptr, *gT_ = nullptr;
  // mixed Lambda cycles: anchor multipliers and residual, difference; splits of
  // the forward operands turned round (per forward state) and of the output
  // gradients of the large products (per pass)
  int lam_prec_ = 0;
  double *lamA_ = nullptr, *gTA_ = nullptr, *dlam_ = nullptr;
  bool lam_anc_valid_ = false;
  double lam_anc_res_ = 0.0;
  std::vector<LowBuf> lowT_, lowg_, lowgT_;
  std::vector<LowBuf> low64_;  // padded fp64 copies (per tensor version)
  std::vector<LowBuf> low64g_, low64gT_, low64T_;  // the same for the reverse pass's large products
  bool fwd_valid_ = false, lam_ready_ = false;
  // T_new_ holds the increment lists' content after a converged cycle at the
  // current amplitudes (the Lambda start; set on convergence, cleared by any change)
  bool tn_lists_ = false;
  // refinement anchor (forward(kPrecRefine)); per op: C0 = A0 B0 and, when both
  // operands are amplitude-dependent, the split of B0; per tensor: the anchor
  // copy and the split of the difference to it
  bool anc_valid_ = false;
  int anc_last_ = 0;
  double anc_res_ = 0.0;
  long anc_pass_ = 0;
  std::vector<double*> anc_C0_;
  std::vector<LowBuf> anc_B0_;
  std::vector<double*> anc_T_;
  std::vector<long> anc_Tpass_, anc_Tver_;
  std::vector<LowBuf> low_diff_;
};

}  // namespace aces
