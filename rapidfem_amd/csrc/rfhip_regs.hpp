// Device helpers for the compressed registers (an f64 value held as its f32
// rounding plus a bf16 of the remainder) and the KCL error norm, shared by
// the stage kernels, the streaming kernels and the Debye block.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rfhip {

// f32 -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One entry of the KCL controller's error norm (session.rs:160-171):
// (err / (atol + rtol max(|y_old|, |y_new|)))^2, +inf for a non-finite y_new.
__device__ __forceinline__ double werr_term(double yo, double yn, double e, double atol, double rtol) {
  if (!isfinite(yn)) return INFINITY;
  const double r = e / (atol + rtol * fmax(fabs(yo), fabs(yn)));
  return r * r;
}

}  // namespace rfhip
