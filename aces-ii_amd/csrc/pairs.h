// Packed pair indices of the reference's DPD storage (acescore/aces_com_symloc.F,
// libr/symexp4.f:36, libr/assym.f), 0-based, q outer and p inner:
//   'PCK2' (p <= q):  index = q*(q+1)/2 + p
//   'PACK' (p <  q):  index = q*(q-1)/2 + p
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace aces {

inline unsigned grid1(long n, int block = 256, long cap = 1L << 20) {
  long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Inverse of the packed pair index (q outer, p inner).
__device__ inline void unpack_le(long idx, int& p, int& q) {  // p <= q
  int qq = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
  while ((long)qq * (qq + 1) / 2 > idx) --qq;
  while ((long)(qq + 1) * (qq + 2) / 2 <= idx) ++qq;
  q = qq;
  p = (int)(idx - (long)qq * (qq + 1) / 2);
}
__device__ inline void unpack_lt(long idx, int& p, int& q) {  // p < q
  int qq = (int)((sqrt(8.0 * (double)idx + 1.0) + 1.0) * 0.5);
  while ((long)qq * (qq - 1) / 2 > idx) --qq;
  while ((long)(qq + 1) * qq / 2 <= idx) ++qq;
  q = qq;
  p = (int)(idx - (long)qq * (qq - 1) / 2);
}
__device__ inline long pk_le(int p, int q) { return (long)q * (q + 1) / 2 + p; }
__device__ inline long pk_lt(int p, int q) { return (long)q * (q - 1) / 2 + p; }

}  // namespace aces
