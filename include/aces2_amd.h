/*
 * aces2_amd.h -- C ABI of the MI355X (gfx950) CCSD amplitude library.
 *
 * Drop-in boundary for the closed-shell CCSD amplitude path of ACES II
 * ``xvcc``.  All functions return 0 on success and a nonzero status code
 * otherwise; the message of the last failure is available from
 * aces_last_error().  Pointers named d_* are device pointers; ``stream`` is a
 * hipStream_t passed as void* (NULL = default stream).  No torch types cross
 * this boundary.
 */
#ifndef ACES2_AMD_H
#define ACES2_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define ACES_OK 0
#define ACES_ERR_HIP 1
#define ACES_ERR_ARG 2
#define ACES_ERR_STATE 3
#define ACES_ERR_NOMEM 4
#define ACES_ERR_RANGE 5
#define ACES_ERR_NOCONV 6
#define ACES_ERR_SOLVE 7

/* Precision of the dense contractions. */
#define ACES_PREC_F64 0
#define ACES_PREC_F32 1
#define ACES_PREC_BF16X3 2

const char* aces_last_error(void);

/* ------------------------------------------------------------------------
 * Contraction engine.  Replaces the XGEMM funnel every xvcc contraction
 * goes through (tools/xgemm.F:101 -> dgemm; crust/xgemm.F:30-48).
 *   C[m*ldc+n] = alpha * sum_k A(m,k) B(n,k) + beta * C[m*ldc+n]
 *   A(m,k) = a_kc ? A[m*lda+k] : A[k*lda+m];  B(n,k) = b_kc ? B[n*ldb+k] : B[k*ldb+n]
 * prec ACES_PREC_BF16X3 takes bf16 hi/lo operand pairs (A, A_lo, B, B_lo).
 * ---------------------------------------------------------------------- */
int aces_gemm(int prec, int M, int N, int K, const void* d_A, const void* d_A_lo, long lda, int a_kc,
              const void* d_B, const void* d_B_lo, long ldb, int b_kc, double* d_C, long ldc, double alpha,
              double beta, int batch, long sA, long sB, long sC, void* stream);
/* Grouped launch: ngroups (<= 8) independent problems C_g = alpha A_g B_g^T +
 * beta C_g (per-group sizes, leading dimensions and device pointers; shared
 * operand orders) in one kernel launch, blockIdx.y selecting the problem -- the
 * per-irrep blocks of one contraction term.  ws (optional, ws_elems doubles)
 * holds split-K partials.  No reference counterpart (the reference loops over
 * irreps with one XGEMM per block, e.g. vcc/ladab1.F, vcc/rngprd.f). */
int aces_gemm_grouped(int prec, int ngroups, const int* M, const int* N, const int* K, const void* const* d_A,
                      const void* const* d_A_lo, const long* lda, int a_kc, const void* const* d_B,
                      const void* const* d_B_lo, const long* ldb, int b_kc, double* const* d_C, const long* ldc,
                      double alpha, double beta, double* d_ws, long ws_elems, void* stream);
/* aces_gemm_grouped without a workspace pointer: the split-K partials (up to
 * ws_elems doubles, at most 4194304 = 32 MiB) go to a static device array of
 * the library, one per GPU, held until the process ends.  ws_elems above that
 * maximum is rejected (ACES_ERR_RANGE).  Allocates nothing (no hipMalloc, no
 * allocating library routine).  Calls that overlap in time on one GPU share
 * that array and race; calls on one stream are ordered.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (HIP graph
 * capture, serving loops). */
int aces_gemm_grouped_noalloc(int prec, int ngroups, const int* M, const int* N, const int* K,
                              const void* const* d_A, const void* const* d_A_lo, const long* lda, int a_kc,
                              const void* const* d_B, const void* const* d_B_lo, const long* ldb, int b_kc,
                              double* const* d_C, const long* ldc, double alpha, double beta, long ws_elems,
                              void* stream);
int aces_convert_f32(int rows, int cols, const double* d_x, long ldx, float* d_y, long ldy, void* stream);
int aces_split_bf16(int rows, int cols, const double* d_x, long ldx, unsigned short* d_hi, unsigned short* d_lo,
                    long ldy, void* stream);

/* ------------------------------------------------------------------------
 * Closed-shell CCSD amplitude engine (RHF, one GPU).  Replaces the body of the
 * xvcc CC loop (vcc/vcc.F:798-1437) for CCSD: GENINT (vcc/genint.F:1),
 * FEACONT/FMICONT/FMECONT/T1INT1/T1INT2A/T1INT2B (vcc/vcc.F:885-920), DRE3EN
 * (vcc/dre3en.F: DRLAD vcc/drlad.f, DRRNG vcc/drrng.f, T12INT2), E4S
 * (vcc/e4s.F), NEWT2 (libr2/newt2.F:233-244), CMPENG (libr/cmpeng.F),
 * DODIIS0 (vcc/dodiis0.F, libr/dodiis.F) and DRMOVE (libr/drmove.f).
 *
 * Integrals (device, row major, chemists' notation, n = o + v):
 *   fock[n][n], ovov[i][a][j][b]=(ia|jb), oovv[i][j][a][b]=(ij|ab),
 *   ovoo[i][a][j][k]=(ia|jk), oooo[i][j][k][l]=(ij|kl), ovvv[i][a][b][c]=(ia|bc),
 *   list233 = the reference's compressed RHF <Ab|Cd> list (rows A<=B, columns
 *   Cd, column major; intprc/intproc.F:250-257).
 * Amplitudes: t1[i][a], t2[i][j][a][b] = T(Ab,Ij) (list 46 storage for C1).
 * prec: 0 fp64, 1 fp32 and 2 bf16x3 products (fp32 accumulation) for the
 * ladders, W(mbej) build, ring and read-bound T1 products; 3 refinement: an
 * fp64 cycle whose ladder, W build and ring products run as bf16x3 products of
 * the operands' differences to an fp64 anchor evaluation, so that their error
 * scales with the distance to the anchor (the first prec-3 cycle after set-up
 * or aces_ccsd_reset_anchor records the anchor, and a new one is recorded each
 * time |RESMITER| has fallen 1e3-fold below the anchor cycle's); the reference
 * has fp64 only.
 * stats (host double[7]) from aces_ccsd_cycle: E(T_new) printed by CMPENG,
 * RESMITER, RESMDIIS, converged flag, final energy (valid when converged),
 * ms in the increments, ms in the whole cycle.
 * ---------------------------------------------------------------------- */
int aces_ccsd_create(int o, int v, void* stream, void** handle);
int aces_ccsd_destroy(void* handle);
int aces_ccsd_setup(void* handle, const double* d_fock, const double* d_ovov, const double* d_oovv,
                    const double* d_ovoo, const double* d_oooo, const double* d_ovvv, const double* d_list233);
int aces_ccsd_set_amplitudes(void* handle, const double* d_t1, const double* d_t2);
int aces_ccsd_get_amplitudes(void* handle, double* d_t1, double* d_t2);
int aces_ccsd_start_mp1(void* handle);
int aces_ccsd_increments(void* handle, int prec, double* d_inc1, double* d_inc2);
int aces_ccsd_cycle(void* handle, int icycle, int prec, double tol, double* stats);
/* No counterpart in the reference (fp64 only): drop the prec-3 anchor, so that
 * the next prec-3 cycle records a new one at the current amplitudes. */
int aces_ccsd_reset_anchor(void* handle);
int aces_ccsd_set_diis_order(void* handle, int order);
int aces_ccsd_reset_diis(void* handle);
int aces_ccsd_energy(void* handle, const double* d_t1, const double* d_t2, double* e_out);
long aces_ccsd_device_bytes(void* handle);
/* Replace the MO Fock matrix (device n x n) of a set-up engine, keeping the
 * integrals (a one-electron perturbation of the same system). */
int aces_ccsd_set_fock(void* handle, const double* d_fock);
/* Per-launch hipEvent timing of the MFMA contractions (on the engine's stream):
 * out[3c+0] ms, out[3c+1] algorithmic flops (2mnk), out[3c+2] launches, for
 * c = 0 particle ladder, 1 W(mbej) build, 2 ring, 3 the other fp64
 * contractions of the increments, 4 the reverse pass (fp64), 5 the reverse
 * pass's bf16x3 products (mixed Lambda cycles), 6 the (T) products (out holds
 * 21 values; the irrep-blocked engine's aces_sym_profile fills the first 15).
 * No reference counterpart
 * (the reference has only whole-program timers, acescore/aces_fin.F:29-55). */
int aces_ccsd_set_profiling(void* handle, int on);
int aces_ccsd_profile(void* handle, double* out, int reset);
/* Non-iterative triples of CCSD(T) at the engine's current amplitudes (TRPS,
 * trp/trps.F:4, called from vcc/vcc.F:1199-1221; closed-shell HF reference,
 * pair-symmetric T2 as every CC iterate is).  out (host double[8]): E[4]T,
 * E[5]ST, E4TAAA, E4TBAA (E[4]T = 2 (E4TAAA + E4TBAA)), ms, MFMA flops,
 * number of occupied triples i >= j >= k; CCSD(T) = E_CCSD + E[4]T + E[5]ST.
 * d_per_triple (device, optional, 3 doubles per triple in the order i,
 * j <= i, k <= j): each triple's unweighted E[4]T, E[4]T + E[5]ST and E4TAAA
 * sums.  prec: 0 = fp64 W products (the reference's arithmetic), 2 = bf16x3
 * split products with fp32 accumulation (relative error of the triples energy
 * ~1e-6); the energy pass is fp64 either way.  Uses o v^2 (v + o) + o^2 v (v + o)
 * elements (8 bytes fp64, 4 bytes bf16x3) of device memory during the call. */
int aces_ccsd_triples(void* handle, int prec, double* out, double* d_per_triple);
/* EOM-CCSD excitation energies (the xvee EOMEE singlet roots of a closed-shell
 * CCSD, vee/; Hbar enters matrix free): the nroots lowest eigenvalues of the CC
 * Jacobian at the converged amplitudes, Hbar - E_CC on the singles and doubles
 * space, by a Davidson iteration on its left eigenvectors whose products J^T l
 * are the engine's reverse pass.  orb_irrep (host, n ints: occupied then
 * virtual; NULL = no symmetry) with irrep (0-based, XOR product of orbital
 * irreps) selects one symmetry block, as ESTATE_SYM does.  omega (host): the
 * excitation energies (hartree); resid (host, optional): residual norms; iters
 * (optional): iterations; d_vecs (device, optional, nroots x (o v + o^2 v^2)):
 * the left eigenvectors [l1 | l2].  Converges to |residual| < tol (maxit
 * iterations at most; check resid).  Overwrites the engine's reverse-pass scratch. */
int aces_ccsd_eom(void* handle, const int* orb_irrep, int irrep, int nroots, double tol, int maxit,
                  double* omega, double* resid, int* iters, double* d_vecs);

/* ------------------------------------------------------------------------
 * Differentiation form: Lambda equations and response density (reverse pass
 * of the engine's increments).  Replaces the body of the xlambda loop
 * (lambda/vlamcc.F:599-830: INITIN/INITSN, LGENINT lambda/lgenint.F,
 * F1INL2/F2INL2, G1INL2/G2INL2, L1INL2A/B, L1INL1, GT1INL1, GWINL1,
 * L2INL2 lambda/l2inl2.F:26-54, L1W1, L2INL1, NEWT2, lambda/dodiis0.F) and
 * the one-particle density of xdens (DENSOO/DENSVV/DENSVO, dens/vdens.F:362-388).
 *
 * Lagrangian L = E(T) + sum l1*Omega1 + sum l2*Omega2 with Omega = inc - D*T and
 * spatial multipliers l1[i][a], l2[i][j][a][b]; the reference's lists hold
 * L1 = l1/2 (list 190) and L2(Ij,Ab) = (2 l2 + l2(i<->j))/3 (list 146).
 *   aces_ccsd_adjoint        r = dL/dT at the engine's amplitudes for given l
 *                            (r2 symmetrised) and, if d_dens != NULL, the
 *                            response density dL/dfock (n x n, symmetric).
 *   aces_ccsd_lambda_start   forward intermediates at the current amplitudes;
 *                            start (L1, L2) = what xvcc leaves on the increment
 *                            lists and moves to lists 190/146 with
 *                            DRMOVE(...,100,SING1) after convergence
 *                            (vcc/vcc.F:1130-1143): the converged cycle's DIIS
 *                            extrapolant (vcc/dodiis0.F:181-233) when the last
 *                            cycle at these amplitudes converged, else their
 *                            Jacobi update T1_new, SYMETRIZE_T2(T2_new)
 *   aces_ccsd_lambda_cycle   one Jacobi step + Lambda DODIIS0 (converged when
 *                            |RESMITER| or |RESMDIIS| < tol); stats as for
 *                            aces_ccsd_cycle (stats[0], stats[4] unused)
 *   aces_ccsd_get_lambda / aces_ccsd_set_lambda   aces_norm = 1: lists 190/146
 *                            normalisation, 0: spatial l1/l2
 *   aces_ccsd_density        response density of the current multipliers:
 *                            dE_corr/dfock_pq for a perturbation of the MO Fock
 *                            matrix at fixed integrals (n x n, symmetric; the
 *                            reference part 2 delta_ij is not included)
 * ---------------------------------------------------------------------- */
int aces_ccsd_adjoint(void* handle, const double* d_l1, const double* d_l2, double* d_r1, double* d_r2,
                      double* d_dens);
int aces_ccsd_lambda_start(void* handle);
int aces_ccsd_lambda_cycle(void* handle, int icycle, double tol, double* stats);
int aces_ccsd_set_lambda_diis_order(void* handle, int order);
/* No counterpart in the reference (fp64 only): 0 fp64 Lambda cycles, 3 mixed
 * cycles r(l) = r(l_a) + J^T (l - l_a) with r(l_a) an fp64 reverse pass at an
 * anchor l_a and the linear part on the difference with its ring, W-build and
 * ladder products in bf16x3; the anchor moves to the current multipliers each
 * time |RESMITER| has fallen 1e3-fold below the anchor cycle's. */
int aces_ccsd_set_lambda_precision(void* handle, int prec);
int aces_ccsd_get_lambda(void* handle, double* d_l1, double* d_l2, int aces_norm);
int aces_ccsd_set_lambda(void* handle, const double* d_l1, const double* d_l2, int aces_norm);
int aces_ccsd_density(void* handle, double* d_dens);

/* ------------------------------------------------------------------------
 * Irrep-blocked CCSD (abelian point groups up to D2h, nirr = 1, 2, 4, 8).
 * Replaces the same xvcc/xlambda/xdens loop bodies as the engine above on
 * symmetry-blocked lists (the reference's DPD lists, acescore/aces_com_symloc.F,
 * libr/inipck.F, contracted per irrep pair, e.g. vcc/ladab1.F, vcc/rngprd.f):
 * every contraction term is one grouped MFMA launch over the irrep blocks.
 * occ[g], vir[g]: orbitals per irrep; orbitals are irrep-sorted inside the
 * occupied and the virtual space (irrep product = XOR of 0-based labels).
 * Inputs and outputs are dense C1 arrays in that order, with the conventions
 * of aces_ccsd_* (symmetry-forbidden elements ignored on input, zero on output).
 * ---------------------------------------------------------------------- */
int aces_sym_create(int nirr, const int* occ, const int* vir, void* stream, void** handle);
int aces_sym_destroy(void* handle);
int aces_sym_setup(void* handle, const double* d_fock, const double* d_ovov, const double* d_oovv,
                   const double* d_ovoo, const double* d_oooo, const double* d_ovvv, const double* d_list233);
int aces_sym_set_amplitudes(void* handle, const double* d_t1, const double* d_t2);
int aces_sym_get_amplitudes(void* handle, double* d_t1, double* d_t2);
int aces_sym_start_mp1(void* handle);
int aces_sym_increments(void* handle, int prec, double* d_inc1, double* d_inc2);
int aces_sym_cycle(void* handle, int icycle, int prec, double tol, double* stats);
int aces_sym_set_diis_order(void* handle, int order);
int aces_sym_reset_diis(void* handle);
/* As aces_ccsd_reset_anchor for the irrep-blocked engine (prec 3 there runs its
 * ring, W-build and ladder products anchored). No reference counterpart. */
int aces_sym_reset_anchor(void* handle);
int aces_sym_energy(void* handle, double* e_out);
int aces_sym_adjoint(void* handle, const double* d_l1, const double* d_l2, double* d_r1, double* d_r2,
                     double* d_dens);
int aces_sym_lambda_start(void* handle);
/* Lambda cycle precision of the irrep-blocked engine, as aces_ccsd_set_lambda_precision:
 * 0 fp64, 3 mixed (fp64 residual at an anchor plus the linear reverse pass on the
 * difference, its ladder / W-build / ring adjoint products in bf16x3; the anchor moves
 * after every 1e3-fold residual drop).  No counterpart in the reference (xlambda is fp64). */
int aces_sym_set_lambda_precision(void* handle, int prec);
int aces_sym_lambda_cycle(void* handle, int icycle, double tol, double* stats);
int aces_sym_set_lambda_diis_order(void* handle, int order);
int aces_sym_get_lambda(void* handle, double* d_l1, double* d_l2, int aces_norm);
int aces_sym_density(void* handle, double* d_dens);
int aces_sym_set_profiling(void* handle, int on);
int aces_sym_profile(void* handle, double* out, int reset);
long aces_sym_device_bytes(void* handle);
long aces_sym_packed_t2(void* handle);

/* ------------------------------------------------------------------------
 * Runtime, list store and record store (device-resident MOINTS/JOBARC).
 *   aces_init              acescore/aces_init.F:26 (lists + records, one stream)
 *   aces_set_symmetry      POP/VRT per irrep (COMMON /SYM/, records SYMPOPOA/VA)
 *   aces_irpdpd            IRPDPD(irrep, type) (acescore/aces_com_symloc.F:11-106)
 *   aces_isymoff           ISYMOFF(irrep_q, irrep_pq, type), 1-based
 *   aces_inipck            INIPCK (libr/inipck.F:3-33)
 *   aces_updmoi            UPDMOI (libr/updmoi.F:11)
 *   aces_getlst/putlst     GETLST/PUTLST (acescore/getlst.F:19, putlst.F:19)
 *   aces_getall/putall     GETALL/PUTALL (acesutil/getall.F)
 *   aces_sumsym2           SUMSYM2 (acesutil/sumsym2.F): list += add
 *   aces_getrec/putrec     GETREC/PUTREC (acescore/getrec.F:29-60, putrec.F:22),
 *                          lengths in 8-byte words, iflag semantics of GETREC
 * Fortran hosts link the same functionality under the reference's names and
 * ILP64 by-reference conventions: getlst_, putlst_, getall_, putall_,
 * sumsym2_, inipck_, updmoi_, getrec_, putrec_.
 * ---------------------------------------------------------------------- */
int aces_init(void* stream);
int aces_fin(void);
int aces_set_symmetry(int nirrep, const int* popa, const int* popb, const int* vrta, const int* vrtb);
long aces_irpdpd(int irrep, int type);
long aces_isymoff(int irrep_q, int irrep_pq, int type);
int aces_inipck(int irrepx, int typel, int typer, int list);
int aces_updmoi(long ncols, long nrows, int group, int list);
int aces_list_shape(int group, int list, long* nrows, long* ncols);
int aces_list_isytyp(int list, int* typel, int* typer);
int aces_getlst(double* dest, long ifirst, long ncols, int group, int list, int dest_on_device);
int aces_putlst(const double* src, long ifirst, long ncols, int group, int list, int src_on_device);
int aces_getall(double* dest, long length, int irrepx, int list, int on_device);
int aces_putall(const double* src, long length, int irrepx, int list, int on_device);
int aces_sumsym2(const double* add, int group, int list, int on_device);
double* aces_list_device_ptr(int group, int list);
/* A totally symmetric (irrepx 1) or other list of pair types (typel, typer) in
 * C1 order: to_c1 = 1 fills d_c1 (device, c1 rows x c1 columns, column major,
 * symmetry-forbidden elements zero) from the list's irrep groups; to_c1 = 0
 * writes the groups from d_c1.  The C1 order is the list's nirrep = 1 layout
 * over the irrep-sorted orbitals.  No reference counterpart (the reference
 * reads its lists by irrep, acescore/aces_com_symloc.F); the bridge the
 * irrep-blocked driver path uses, exposed for hosts that hold C1-ordered data. */
int aces_list_c1(int list, int typel, int typer, int irrepx, double* d_c1, int to_c1);
/* The job archive and list files on disk, ILP64 layout (8-byte integers, one
 * double per integer word), in directory dir (NULL or "": the working
 * directory, where the reference keeps them):
 *   aces_jobarc_read   JAINDX + JOBARC -> the record store (ACES_JA_INIT,
 *                      acescore/aces_ja_init.F: the logical records, in slot order)
 *   aces_jobarc_write  the record store -> JOBARC + JAINDX (ACES_JA_FIN,
 *                      acescore/aces_ja_fin.F; record placement of acescore/putrec.F)
 *   aces_lists_read    MOINTS/GAMLAM/MOABCD/DERINT/DERGAM -> the device list store,
 *                      placed by the MOIOVEC/MOIOWRD/MOIOSIZ/MOIODIS/MOIOFIL records,
 *                      pair types from ISYMTYP, physical records of IFLAGS(37)
 *                      integers, symmetry from COMPNIRR/SYMPOP* (ACES_IO_INIT,
 *                      acescore/aces_io_init.F, getlst.F); read the archive first
 *   aces_lists_write   every list -> the list files, placed as
 *                      acescore/aces_list_touch.F places them, and the MOIO*,
 *                      ISYMTYP, TOTRECMO, TOTWRDMO, COMPNIRR, SYMPOP* records
 *                      (ACES_IO_FIN, acescore/aces_io_fin.F); write the archive after */
/* The lists xintprc writes for the CC path, from factorised integrals on the
 * GPU (intprc/intproc.F; list 233 at intprc/intproc.F:240-266; denominators
 * intprc/dijab.F:178-185, 300-375): d_B (device) holds naux x n x n factors with
 * (pq|rs) = sum_P B[P][p][q] B[P][r][s], each B[P] symmetric (density fitting or
 * Cholesky over real orbitals), orbitals in list order (occupied then virtual,
 * irrep-sorted in each, symmetry-forbidden elements zero); eps (host) the n
 * orbital energies.  Under the symmetry set by aces_set_symmetry (closed shell,
 * 1, 2, 4 or 8 irreps) it writes lists 13, 16, 10, 30, 23 and 233 (C1 or
 * symmetry packed), the reciprocal denominators 1/D of lists 50 (AB), 48 (AA)
 * and 48 group 9 (T1), and record SCFEVALA.  Every integral block is an fp64
 * MFMA product over P. */
int aces_intprc_df(const double* d_B, int naux, const double* eps);
int aces_jobarc_read(const char* dir);
int aces_jobarc_write(const char* dir);
int aces_lists_read(const char* dir);
int aces_lists_write(const char* dir);
int aces_putrec(const char* name, long nwords, const void* src);
int aces_getrec(int iflag, const char* name, long* nwords, void* dest);

/* ------------------------------------------------------------------------
 * xvcc driver entry points over the list store (vcc/vcc.F:811-1085 order).
 *   aces_vcc_mp1           MP1 start (intprc/dfrmt2.F) into lists 46/44, T1 = 0
 *   aces_vcc_set_diis_order DIISLST (vcc/diislst.F): CC_EXPORDER vectors
 *   aces_genint            GENINT (vcc/genint.F:1) + the fused increment terms
 *   aces_dre3en            DRE3EN (vcc/dre3en.F): T2 increment -> lists 63/61
 *   aces_e4s               E4S (vcc/e4s.F): new T1 -> list 90 group 3
 *   aces_newt2             NEWT2 (libr2/newt2.F): list 63 <- 63/D, 61 <- ABTOAA
 *   aces_cmpeng            CMPENG (libr/cmpeng.F:60-188) on lists nlist2+1..+3 and
 *                          T1 group nlist1+1: ecorr3[0] = E(AA), ecorr3[2] = E(AB)
 *                          (ecorr3[1] untouched, as for RHF there), *etot =
 *                          2 E(AA) + E(AB), *etott2 = the same with T2 for tau
 *                          (any output pointer may be NULL)
 *   aces_dodiis0           DODIIS0 (vcc/dodiis0.F); *iconvg = 0 when converged
 *   aces_drmove            DRMOVE(..., 0, .TRUE.) (libr/drmove.f): T lists <- increment lists
 *   aces_drmove_lists      DRMOVE(..., nlist, sing1): lists 61/63 -> 44+nlist/46+nlist,
 *                          with sing1 also list 90 group 3 -> list 90+nlist group 1
 * Integral input lists (RHF): 13 <Mn|Ij> (14,14), 16 <Ab|Ij> (13,14),
 * 10 <Ij|Ka> (14,18), 30 <Ab|Ci> (13,11), 233 <Ab|Cd> (5,13) and
 * 23 <AJ||BI> (9,9); records SCFEVALA (orbital energies) or FOCKMOA (MO Fock),
 * orbitals in list order (occupied then virtual, each irrep-sorted).  With one
 * irrep the C1 engine runs on the lists' own memory; with 2, 4 or 8 irreps the
 * lists are symmetry packed (one group per irrep, DPD order) and the
 * irrep-blocked engine runs them (amplitude lists 44/46/61/63/90 in the same
 * packing, T1 = the irrep-1 (a,i) vector of pair type 9).
 * Fortran spellings with the reference's argument lists: genint_, feacont_,
 * fmicont_, fmecont_, t1int2a_, t1int2b_, t1int1_, initin_, initsn_, rnabij_,
 * dre3en_, e4s_, newt2_, cmpeng_, dodiis0_, drmove_.  Argument values they do
 * not implement (open-shell IUHF, DODIIS0 list sets other than xvcc's
 * 44,61,90,0,90,2,70, DAMP_PARAMETER != 1, RNABIJ type 'C') end the process
 * with status 1 after a message on stderr, the reference's fatal path
 * (aces_exit(1)); e4s_ returns SE4 = 0, as the reference does for CC methods.
 * ---------------------------------------------------------------------- */
int aces_vcc_mp1(void);
int aces_vcc_set_diis_order(int order);
/* prec as for aces_ccsd_cycle (0 fp64, 1 fp32, 2 bf16x3, 3 refinement); switching
 * to 3 records a new refinement anchor at the next GENINT. */
int aces_vcc_set_precision(int prec);
int aces_genint(void);
int aces_dre3en(void);
int aces_e4s(void);
int aces_newt2(void);
int aces_cmpeng(int nlist2, int nlist1, double* ecorr3, double* etot, double* etott2);
int aces_dodiis0(int icycle, double tol, int* iconvg, double* stats);
int aces_drmove(void);
int aces_drmove_lists(int nlist, int sing1);

#ifdef __cplusplus
}
#endif
#endif /* ACES2_AMD_H */
