// Round-to-nearest-even fp32 -> bf16 and back, for the hi/lo operand splits.
#pragma once
#include <hip/hip_runtime.h>

namespace aces {

__device__ inline unsigned short bf16_rn(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

}  // namespace aces
