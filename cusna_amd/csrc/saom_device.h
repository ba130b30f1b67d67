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

// exp(x) for the softmax weights (x = f - shift <= 4, finite): x = (32 m + j)
// ln2/32 + r with |r| <= ln2/64, exp(x) = 2^m 2^(j/32) p(r), p the degree-6
// Taylor polynomial (truncation below 4e-18 relative), 2^(j/32) from a
// 32-entry table (in LDS in the period kernel). Within a few ulps of the
// correctly rounded value, like the library's exp (the weights enter only the
// scores, through ratios of their sums; checked against the host's exp by a
// test through cusna_exp_weight_eval); about half the library's instructions,
// with no range checks and without a branch per alternative.
constexpr int kExpTab = 32;
__device__ __forceinline__ double exp_table(double x, const double* tab) {
  const double kd = rint(x * 46.166241308446828);  // 32 / ln 2
  const int k = (int)kd;
  // ln2/32 = hi + lo, hi with 29 significant bits: k hi is exact for |k| < 2^24
  double r = fma(-kd, 0x1.62e42ff000000p-6, x);
  r = fma(-kd, -0x1.718432a1b0e26p-40, r);
  double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
  q = fma(r, q, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  q = fma(r, q, 1.0);
  q = fma(r, q, 1.0);
  return ldexp(tab[k & (kExpTab - 1)] * q, k >> 5);
}

}  // namespace saom
