// saom_device.h -- device-only helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace saom {

// Approximate Gumbel noise log(-log u) in float32 for u = (x + 1/2) 2^-32,
// used only to discard alternatives that cannot win. a = -ln(u) is taken from
// u itself for u < 1/2, and from q = 1 - u (exact in integers) for u >= 1/2:
// the series q + q^2/2 + q^3/3 + q^4/4 for q < 2^-8, else -ln(1 - q). With
// v_log_f32 accurate to a few ulp, |result - log(-log u)| stays below 1e-4
// over every 32-bit x (checked on the device by a test); kGumbelSlack is 2^-8.
constexpr double kGumbelSlack = 1.0 / 256.0;

__device__ __forceinline__ float approx_loglog(uint32_t x) {
  constexpr float kLn2 = 0.69314718055994531f;
  constexpr float k2m32 = 2.3283064365386963e-10f;
  float a;
  if (x < 0x80000000u) {
    const float u = ((float)x + 0.5f) * k2m32;
    a = -__builtin_amdgcn_logf(u) * kLn2;
  } else {
    const uint32_t qi = 0xFFFFFFFFu - x;
    const float q = ((float)qi + 0.5f) * k2m32;
    if (qi < (1u << 24)) a = q * (1.0f + q * (0.5f + q * (0.33333334f + q * 0.25f)));
    else a = -__builtin_amdgcn_logf(1.0f - q) * kLn2;
  }
  return __builtin_amdgcn_logf(a) * kLn2;
}

}  // namespace saom
