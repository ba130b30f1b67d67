// saom_launch.hip -- dispatch of the period kernel over its compiled shapes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "saom_kernels.h"

namespace saom {

#define SAOM_DECL(W) \
  int launch_period_w##W(const PeriodLaunch& L, int nw, bool replay, bool simple, void* stream);
SAOM_DECL(1)
SAOM_DECL(2)
SAOM_DECL(4)
SAOM_DECL(8)
SAOM_DECL(16)
SAOM_DECL(32)
SAOM_DECL(64)
SAOM_DECL(128)
#undef SAOM_DECL
size_t sim_lds_bytes_w1(const SimParams& P, int nw, int wmax);

size_t sim_lds_bytes(const SimParams& P, int nw, int wmax) {
  return sim_lds_bytes_w1(P, nw, wmax);
}

bool simple_effects(const int* t, int k) {
  for (int e = 0; e < k; ++e) {
    switch (t[e]) {
      case 5: case 6: case 22: case 24: case 26: case 27: case 28: case 29: case 30: case 32:
        return false;
      default:
        break;
    }
  }
  return true;
}

LaunchShape exact_shape(int W64, LaunchShape g, int k_free, size_t lds_generic) {
  int nw, R, wpe = 0;
  if (W64 <= 8) {
    nw = g.nw;
    R = (W64 + nw - 1) / nw;
    const int R2 = (W64 + 2 * nw - 1) / (2 * nw);
    const int chains_cu = lds_generic > 0 ? (int)((160 * 1024) / lds_generic) : 64;
    if (k_free > 0 && 2 * nw <= 8 && k_free * R2 <= 32 &&
        (k_free * R > 32 || (W64 >= 5 && chains_cu <= 4))) {
      if (k_free * R <= 32) wpe = 4;  // doubled for occupancy: the four-wave budget
      nw *= 2;
      R = R2;
    }
  } else if (W64 <= 12) {
    // three words per wave on four waves, one wave per SIMD per chain:
    // five or six waves of two words left one SIMD with two waves of the
    // chain, 1.7-1.8 x slower per chain on the 256-group batch
    R = 3;
    nw = 4;
  } else if (W64 <= 32) {
    R = 2;
    nw = (W64 + 1) / 2;
  } else {
    R = (W64 + 15) / 16;
    nw = (W64 + R - 1) / R;
  }
  return LaunchShape{nw, nw * R, wpe};
}

int launch_period_ex(const PeriodLaunch& L, LaunchShape s, bool replay, bool simple,
                     void* stream) {
  if (L.grid < 1) return 0;
  if (L.used) *L.used = false;
  // the specialised kernel (module API), in its exact shape when the caller
  // gave one; a launch the runtime refuses (for instance for its dynamic LDS
  // size) falls back to the generic kernel
  if (L.specialize && L.Ph) {
    bool ex = L.exact.nw > 0 && L.lds_exact <= 160 * 1024;
    LaunchShape rs = ex ? L.exact : s;
    size_t lds = ex ? L.lds_exact : L.lds;
    void* fn = nullptr;
    if (ex && rs.wpe > 2) {
      // the preferred shape needs rs.wpe waves per SIMD: compiled with the
      // default budget when it fits that occupancy's registers anyway (the
      // default budget schedules better: 113 against 117 ms at 500 actors),
      // else without the per-chain covariate registers (the rate-effect path
      // at 500 actors: 141 -> 127 VGPRs), else lean with the tighter budget,
      // unless that spills; then the alternative shape in the two-wave budget
      const int budget = 512 / rs.wpe;
      auto fits = [&](void* f) { return f && !uses_scratch(f) && kernel_vgprs(f) <= budget; };
      fn = rtc_period_kernel(*L.Ph, rs, replay, simple, L.same_n);
      if (!fits(fn)) fn = rtc_period_kernel(*L.Ph, rs, replay, simple, L.same_n, 0, true);
      if (!fits(fn)) {
        fn = rtc_period_kernel(*L.Ph, rs, replay, simple, L.same_n, rs.wpe, true);
        if (fn && uses_scratch(fn)) fn = nullptr;
      }
      if (!fn && L.exact_alt.nw > 0 && L.lds_exact_alt <= 160 * 1024) {
        rs = L.exact_alt;
        lds = L.lds_exact_alt;
      }
    }
    if (!fn) fn = rtc_period_kernel(*L.Ph, rs, replay, simple, L.same_n);
    if (fn) {
      PeriodLaunch Lr = L;
      if (ex) Lr.lds = lds;
      const int e = launch_period_rtc(fn, Lr, rs.nw, stream);
      if (e == 0) {
        if (L.used) *L.used = true;
        if (L.used_threads) *L.used_threads = 64 * rs.nw;
        return 0;
      }
      (void)hipGetLastError();
    }
  }
  if (L.used_threads) *L.used_threads = 64 * s.nw;
  switch (s.wmax) {
    case 1: return launch_period_w1(L, s.nw, replay, simple, stream);
    case 2: return launch_period_w2(L, s.nw, replay, simple, stream);
    case 4: return launch_period_w4(L, s.nw, replay, simple, stream);
    case 8: return launch_period_w8(L, s.nw, replay, simple, stream);
    case 16: return launch_period_w16(L, s.nw, replay, simple, stream);
    case 32: return launch_period_w32(L, s.nw, replay, simple, stream);
    case 64: return launch_period_w64(L, s.nw, replay, simple, stream);
    case 128: return launch_period_w128(L, s.nw, replay, simple, stream);
    default: return (int)hipErrorInvalidValue;
  }
}

#define SAOM_MDECL(W) \
  int launch_multi_w##W(const MultiParams* Pd, int grid, size_t lds, int nw, void* stream);
SAOM_MDECL(1)
SAOM_MDECL(2)
SAOM_MDECL(4)
SAOM_MDECL(8)
SAOM_MDECL(16)
SAOM_MDECL(32)
SAOM_MDECL(64)
SAOM_MDECL(128)
#undef SAOM_MDECL

int launch_multi(const MultiParams& P, const MultiParams* Pd, int nw, int wmax, void* stream) {
  if (P.M < 1) return 0;
  const size_t lds = multi_lds_bytes(P, nw, wmax);
  switch (wmax) {
    case 1: return launch_multi_w1(Pd, P.M, lds, nw, stream);
    case 2: return launch_multi_w2(Pd, P.M, lds, nw, stream);
    case 4: return launch_multi_w4(Pd, P.M, lds, nw, stream);
    case 8: return launch_multi_w8(Pd, P.M, lds, nw, stream);
    case 16: return launch_multi_w16(Pd, P.M, lds, nw, stream);
    case 32: return launch_multi_w32(Pd, P.M, lds, nw, stream);
    case 64: return launch_multi_w64(Pd, P.M, lds, nw, stream);
    case 128: return launch_multi_w128(Pd, P.M, lds, nw, stream);
    default: return (int)hipErrorInvalidValue;
  }
}

int launch_period(const SimParams& P, const SimParams* Pd, LaunchShape s, bool replay,
                  uint32_t k0, uint32_t k1, void* stream, bool specialize, bool* used,
                  bool free_shape, int* used_threads) {
  PeriodLaunch L;
  L.Pd = Pd;
  L.items = nullptr;
  L.grid = P.M;
  L.lds = sim_lds_bytes(P, s.nw, s.wmax);
  L.key0 = k0;
  L.key1 = k1;
  L.Ph = &P;
  L.specialize = specialize;
  L.used = used;
  L.used_threads = used_threads;
  L.exact = exact_shape(P.W64, s, free_shape ? P.k : 0, L.lds);
  L.lds_exact = sim_lds_bytes(P, L.exact.nw, L.exact.wmax);
  L.exact_alt = exact_shape(P.W64, s, free_shape ? P.k : 0);
  L.lds_exact_alt = sim_lds_bytes(P, L.exact_alt.nw, L.exact_alt.wmax);
  return launch_period_ex(L, s, replay, simple_effects(P.tid, P.k), stream);
}

}  // namespace saom
