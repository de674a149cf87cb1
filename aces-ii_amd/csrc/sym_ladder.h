// Particle-particle ladder on irrep-blocked storage (symmetric pair packing).
//
// Per pair irrep G the ladder is one MFMA product
//   Zp_G = Vp_G . Sp_G^T,  Zm_G = Vm_G . Dp_G^T
// with rows [(a<=b) | (k,b)] resp. [(a<b) | (k,b)] (the (k,b) rows carry the
// T1 dressing of <ab||cd>, vcc/ladab1.F, vcc/t1inw1.f), columns (c<=d)/(c<d),
// and Sp/Dp the symmetric/antisymmetric combinations of tau over (i<=j)/(i<j).
// The (k,b) compound follows the row order of an (O,V|..) layout.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "dpd.h"

namespace aces {
namespace dpd {

struct LadderDesc {
  int nirr;
  Space O, V;
  PairSpace VP, VM, OP, OM;
  long nkb[kMaxIrr];
  long kbsub[kMaxIrr][kMaxIrr];
  // block offsets (elements) of Vp, Vm, Sp, Dp, Zp, Zm
  long vp_off[kMaxIrr + 1], vm_off[kMaxIrr + 1], sp_off[kMaxIrr + 1], dp_off[kMaxIrr + 1];
  long zp_off[kMaxIrr + 1], zm_off[kMaxIrr + 1];
};

LadderDesc make_ladder(const Space& O, const Space& V);

// Vp/Vm from the dense compressed list 233 ([d][c][(a<=b)] holding <ab|cd>,
// intprc/intproc.F:250-257, orbitals irrep-sorted) and dense ovvv[k][c][b][d].
int build_ladder_sym(const LadderDesc* d_desc, const LadderDesc& h, const double* L233, const double* ovvv,
                     double* Vp, double* Vm, hipStream_t s);
// Sp/Dp from tau (layout tauL, device copy).
int pack_tau_sym(const LadderDesc* d_desc, const LadderDesc& h, const Layout* d_tauL, const double* tau, double* Sp,
                 double* Dp, hipStream_t s);
// Tiles of the inc2 side of the unpack: one per (box of the inc2 plan, value
// a, range of b, range of the box's occupied pairs in Z's column order).
struct UnpackTile {
  int box;
  short la, lb0, tb, tq;  // a; first b and b extent; pair extent
  int q0;                 // first pair (column offset inside the box's pair block)
};
constexpr int kUnpackLds = 2048;  // doubles per LDS image (Zp and Zm parts each)
constexpr int kUnpackBatch = 4;  // loads of a thread issued together in the tiled unpack
constexpr int kPackBatch = 4;    // elements of a thread per pass of the ladder-operand packing
void unpack_tiles(const Plan& inc2_plan, std::vector<UnpackTile>& tiles);
// inc2 += unpacked Z (tiles of the inc2 plan: Z rows read along their
// occupied pairs into LDS, inc2 written along b); Yb = (k,b) rows (walk of the
// Yb plan).
int unpack_ladder_sym(const LadderDesc* d_desc, const Plan& inc2_plan, const UnpackTile* tiles, long ntiles,
                      const Plan& yb_plan, const double* Zp, const double* Zm, double* inc2, double* Yb,
                      hipStream_t s);
// Adjoints: gZp/gZm from g_inc2 and gYb; gtau += pack adjoint (walk of the tau plan).
int unpack_ladder_sym_adjoint(const LadderDesc* d_desc, const LadderDesc& h, const Layout* d_inc2L,
                              const Layout* d_ybL, const double* g2, const double* gYb, double* gZp, double* gZm,
                              hipStream_t s);
int pack_tau_sym_adjoint(const LadderDesc* d_desc, const Plan& tau_plan, const double* gSp, const double* gDp,
                         double* gtau, hipStream_t s);

}  // namespace dpd
}  // namespace aces
