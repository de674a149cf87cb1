// Symmetry-packed (DPD) lists of the list store <-> the same lists in C1 order.
//
// A list with pair types (tl, tr) and overall irrep irrepx holds one group per
// irrep r of its right (column) pair, left pair irrep r ^ irrepx; a group is
// a column-major block nrows x ncols whose rows (columns) run over the left
// (right) pairs of that irrep in the reference's DPD order
// (acescore/aces_com_symloc.F:11-106): sub-blocks by the irrep of the pair's
// second orbital q, ascending; inside a sub-block p fastest, or for a packed
// type with both orbitals in one irrep the triangle q(q-1)/2 + p (p < q) or
// q(q+1)/2 + p (p <= q).  Orbitals are irrep-sorted inside each space.
//
// The C1 order of the same list is the nirrep = 1 case: one block whose row
// (column) index is p + N_p q (full) or the triangle index of the global
// orbital indices.  Elements that symmetry forbids are zero in C1 order.
#pragma once
#include <hip/hip_runtime.h>
#include "liststore.h"

namespace aces {

// Rows and columns of the C1 order of a list of pair types (tl, tr); tr = 0 is
// a single column (a vector list such as T1, list 90).
long c1_pairs(const DpdSym& sym, int type);

// dst (C1 order, c1_pairs(tl) x c1_pairs(tr), zero filled here) <- list.
// tl, tr: the list's pair types (tr = 0: single-column groups); irrepx 1-based.
int dpd_to_c1(const ListStore& L, int list, int tl, int tr, int irrepx, double* dst, hipStream_t s);
// list <- src (C1 order); the list's groups must exist with the DPD shapes.
int c1_to_dpd(ListStore& L, int list, int tl, int tr, int irrepx, const double* src, hipStream_t s);
// The same for one group of a vector list (T1: list 90 group g, pair type 9,
// totally symmetric, one column): dense t1[i][a] <-> the group.
int t1_to_dense(const ListStore& L, int group, int list, double* t1, hipStream_t s);
int dense_to_t1(ListStore& L, int group, int list, const double* t1, hipStream_t s);

}  // namespace aces
