// Integral and denominator lists from factorised integrals (see intprc_df.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "liststore.h"

namespace aces {

// B (device): naux x n x n factors, (pq|rs) = sum_P B[P][p][q] B[P][r][s],
// each B[P] symmetric, orbitals in list order; eps (host): the n orbital
// energies.  Writes lists 13, 16, 10, 30, 23, 233 (types of the list-fed
// driver), the reciprocal denominator lists 50 (AB), 48 (AA) and 48 group 9
// (T1), and record SCFEVALA, under the store's symmetry (closed shell).
int intprc_df(ListStore& L, RecordStore& R, const double* B, int naux, const double* eps, hipStream_t s);

}  // namespace aces
