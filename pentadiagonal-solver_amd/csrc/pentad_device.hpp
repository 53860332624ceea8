// pentad_device.hpp -- device-side helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace pentad {

// Reciprocal of a pivot.  The hardware reciprocal (v_rcp_f32 / v_rcp_f64) is
// refined by Newton-Raphson steps e = 1 - d*r, r += r*e: one step brings the
// FP32 estimate to within an ulp, two steps take the FP64 estimate (about
// 2^-26 relative error) to the limit of double rounding.  Pivots of the
// systems this library targets are O(1) and far from the denormal range.
__device__ __forceinline__ float recip(float d) {
  float r = __builtin_amdgcn_rcpf(d);
  float e = __builtin_fmaf(-d, r, 1.0f);
  return __builtin_fmaf(r, e, r);
}

__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

template <typename T> __device__ __forceinline__ T fmad(T a, T b, T c) {
  return a * b + c; // contracted to v_fma_* by -ffp-contract=fast
}

} // namespace pentad
