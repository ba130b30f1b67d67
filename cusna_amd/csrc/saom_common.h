// saom_common.h -- arithmetic shared by the gfx950 kernels and by host code.
//
// Everything here is plain IEEE arithmetic (integer ops, +, *, /, sqrt, fma),
// so a host build and a device build of the same function return the same
// bits; host tests rely on that to check device-side formulas on the CPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define SAOM_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define SAOM_HD static inline
#endif

namespace saom {

constexpr int kMaxEff = 24;   // CUSNA_MAXEFF
constexpr int kMaxBeff = 6;   // CUSNA_MAXBEFF
constexpr int kMaxRate = 4;   // CUSNA_MAXRATE
constexpr long long kMaxSteps = 4000000;  // runaway cap per chain-period

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). The two multipliers and
// the two Weyl key increments are the published Random123 constants. The
// reference draws one block per ministep (counter = steps, chain, 0, 0) and
// one per alternative (counter = steps, chain, 1, j), keyed by the 64-bit seed
// (cran/cusna src/libcusna_saom_cpu.cpp:32-45, 167, 266, 381).
SAOM_HD void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                      uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint64_t m0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t y0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
    const uint32_t y2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)m1;
    c3 = (uint32_t)m0;
    c0 = y0;
    c2 = y2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// (x + 1/2) 2^-32, the open-interval uniform of the reference (:46).
SAOM_HD double unit_open(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

// x^1.5 for x >= 0, rounded once. s = sqrt(x) is correctly rounded and
// fma(-s, s, x) is its exact residual, so x*sqrt(x) = x*s + x*res/(2s) to
// about 2^-100 relative; p + (pe + tail) then rounds to the nearest double
// unless x^1.5 lies within that error of a rounding boundary. This replaces
// pow(x, 1.5) of the degree statistics (libcusna_saom_cpu.cpp:663, 667); a host
// test checks it against the C library's pow on every degree 0..8192.
SAOM_HD double pow_three_halves(double x) {
  if (x <= 0.0) return 0.0;
  const double s = sqrt(x);
  const double res = fma(-s, s, x);
  const double p = x * s;
  const double pe = fma(x, s, -p);
  const double tail = x * res / (2.0 * s);
  return p + (pe + tail);
}

// The reference's statistics call the C library's pow (glibc 2.35 on this
// platform), which rounds x^1.5 the other way from the correctly rounded value
// on exactly nine integers in 0..8192 (error 0.504 ulp). Degrees are integers
// below CUSNA_MAX_N, so stepping one ulp on those nine reproduces the library
// value on the whole domain. Sign = direction of the library's result.
SAOM_HD double pow_three_halves_libm(double x) {
  const double v = pow_three_halves(x);
  const int xi = (int)x;
  int dir = 0;
  switch (xi) {
    case 1335: case 2261: case 3151: case 3947: case 4072: case 4319: case 5340: case 5479:
      dir = -1;
      break;
    case 3617:
      dir = 1;
      break;
    default:
      break;
  }
  if (dir == 0 || (double)xi != x) return v;
  return nextafter(v, dir > 0 ? 1e300 : 0.0);
}

}  // namespace saom
