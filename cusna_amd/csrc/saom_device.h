// saom_device.h -- device-only helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace saom {

// Approximate Gumbel noise log(-log u) in float32 for u = (x + 1/2) 2^-32,
// used only to discard alternatives that cannot win. a = -ln(u) is taken from
// u itself for u < 1/2, and from q = 1 - u (exact in integers) for u >= 1/2:
// q (1 + q/2) for q < 2^-8 (relative error below q^2/3), else -ln(1 - q).
// The three forms are computed without branches and selected. With
// v_log_f32 accurate to a few ulp, |result - log(-log u)| stays below 1e-4
// over every 32-bit x (checked on the device by a test); kGumbelSlack is 2^-8.
constexpr double kGumbelSlack = 1.0 / 256.0;

__device__ __forceinline__ float approx_loglog(uint32_t x) {
  constexpr float kLn2 = 0.69314718055994531f;
  constexpr float k2m32 = 2.3283064365386963e-10f;
  const uint32_t qi = 0xFFFFFFFFu - x;
  const float u = ((float)x + 0.5f) * k2m32;
  const float q = ((float)qi + 0.5f) * k2m32;
  const float lo = x < 0x80000000u ? u : 1.0f - q;  // argument of the logarithm
  const float a_log = -__builtin_amdgcn_logf(lo) * kLn2;
  const float a_ser = __builtin_fmaf(0.5f * q, q, q);
  const float a = (x >= 0x80000000u && qi < (1u << 24)) ? a_ser : a_log;
  return __builtin_amdgcn_logf(a) * kLn2;
}

}  // namespace saom
