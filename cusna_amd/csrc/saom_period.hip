// saom_period.hip -- the SAOM period simulator for gfx950 (MI355X).
//
// One workgroup of NW wave64s runs one chain of one period, i.e. the
// reference's simulate_one (cran/cusna src/libcusna_saom_cpu.cpp:142-303).
//
// State in LDS: the network as u64 bit rows stored word-major (word w of
// actor h at w * n + h, so a wave reading word w of 64 actors is conflict-free), in/out
// degrees, behaviour values. Alternative j of a network ministep belongs to
// lane (j & 63) of wave ((j >> 6) % NW) in round (j >> 6) / NW, so round r of
// wave w covers u64 word NW*r + w of every bit row: a lane's own j is always
// bit `lane` of the word the whole wave is looking at, and broadcasts of a
// row word serve all 64 lanes at once.
//
// Per network ministep (all uniform control flow):
//   1. Philox block (steps, chain, 0, 0): waiting time, net/behaviour, ego i.
//   2. row_i from LDS (broadcast), col_i from a ballot of bit i of each row.
//   3. Features per alternative j: tp_j = sum over out-neighbours h of i of
//      bit j of row_h (two-paths i->h->j); so_j = |row_i & row_j|;
//      cy_j = |row_j & col_i| (popcounts over W64 words).
//   4. f_j = sum_e theta_e d_e(j), effects in list order, no fused multiply-add
//      (same rounding as the reference's f += th[e] * d).
//   5. g_j = f_j - log(-log u_j) with u_j from Philox (steps, chain, 1, j);
//      block argmax with lowest-j ties (the reference's strict > scan from
//      gbest = -1e30, jstar = i).
//   6. Second pass: w_j = exp(f_j - lmax); S = sum w_j, T_e = sum w_j d_e(j);
//      score_e += d_e(j*) - T_e / S.
//   7. Toggle i->j* (and j*->i when symmetric), degrees, distance.
// Behaviour ministeps follow libcusna_saom_cpu.cpp:202-227.
//
// Specialised builds: runtime compilation (rtc.hip) defines SAOM_FIXED_K and
// the effect types / creation-endowment types / behaviour-dynamic flags of
// one effect list, and the word bound, block shape and variant of one
// launch. Effect loops then unroll over compile-time effect types (no
// per-effect dispatch) and, when they fit, the change statistics of the
// choice pass stay in registers for the score pass. Same arithmetic, same
// order: results are identical to the generic kernel's.
#ifndef SAOM_RTC
#include <hip/hip_runtime.h>

#include "saom_device.h"
#include "saom_kernels.h"
#endif

namespace saom {
namespace {

#ifdef SAOM_FIXED_K
constexpr int kFixedT[] = {SAOM_FIXED_TYPES};
constexpr int kFixedE[] = {SAOM_FIXED_ETYPES};
constexpr int kFixedD[] = {SAOM_FIXED_DYN};
// effect e's change statistic reads a static covariate at the alternative
// (altX, simX, sameX, higher, egoXaltX, egoPlusAltX, altRThresholdX): the
// values at the lane's actors are loaded once per chain
constexpr bool alt_cov(int e) {
#if defined(SAOM_LEAN) && SAOM_LEAN
  return e < 0;  // lean build: none (the values are read per ministep)
#else
  const int t = kFixedT[e];
  return (t == 11 || t == 12 || t == 13 || t == 20 || t == 21 || t == 25 || t == 31) &&
         kFixedD[e] == 0;
#endif
}
__device__ __forceinline__ int eff_type(const SimParams&, int e) { return kFixedT[e]; }
__device__ __forceinline__ int eff_ety(const SimParams&, int e) { return kFixedE[e]; }
__device__ __forceinline__ int eff_dyn(const SimParams&, int e) { return kFixedD[e]; }

// f(IntC<B>), ..., f(IntC<E-1>): a loop whose index is a compile-time
// constant in every iteration. The fixed effect list is walked this way so
// that each iteration holds only its own effect's code: a "#pragma unroll"
// loop is sized before the effect type folds, so with a dozen effects it
// exceeds the unroller's limit, stays a loop over every effect's code and
// indexes the cached change statistics dynamically, i.e. through scratch.
template <int V>
struct IntC {
  __device__ constexpr operator int() const { return V; }
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(IntC<B>{});
    static_for<B + 1, E>(f);
  }
}
#else
__device__ __forceinline__ int eff_type(const SimParams& P, int e) { return P.tid[e]; }
__device__ __forceinline__ int eff_ety(const SimParams& P, int e) { return P.ety[e]; }
__device__ __forceinline__ int eff_dyn(const SimParams& P, int e) { return P.dyn[e]; }
#endif

// Shape values a specialised launch may fix at compile time when every
// group of the launch shares them (the actor count, the behaviour-effect
// count, the feature flags): LDS offsets and loop bounds then fold into
// immediates instead of occupying scalar registers across the ministep loop.
__device__ __forceinline__ int shape_n(const SimParams& P) {
#ifdef SAOM_FIXED_N
  (void)P;
  return SAOM_FIXED_N;
#else
  return P.n;
#endif
}
__device__ __forceinline__ int shape_kb(const SimParams& P) {
#ifdef SAOM_FIXED_KB
  (void)P;
  return SAOM_FIXED_KB;
#else
  return P.kb;
#endif
}
__device__ __forceinline__ int shape_flags(const SimParams& P) {
#ifdef SAOM_FIXED_FLAGS
  (void)P;
  return SAOM_FIXED_FLAGS;
#else
  return P.flags;
#endif
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// A load at a wave-uniform address of memory that no kernel writes while the
// period runs (problem inputs): through the constant address space the
// compiler issues a scalar load, which it cannot prove safe for a generic
// pointer.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

// A wave-uniform 64-bit value moved to scalar registers (loops over its set
// bits then run on the scalar unit).
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Wave64 all-reductions: four DPP steps combine within each row of 16 lanes
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror), then
// the four row results are read into scalar registers and combined. Every lane
// ends with the same bits (each step combines a lane with a partner holding the
// same kind of partial, and a + b == b + a exactly).
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  // every lane's source lies in the wave for these patterns, so the value
  // kept for a lane without a source is never used (no initialising move)
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32<CTRL>((int)b);
  const int hi = dpp_i32<CTRL>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// A wave-uniform double moved to scalar registers.
__device__ __forceinline__ double uniform_f64(double v) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double wave_sum(double v) {
  v = v + dpp_f64<0xB1>(v);
  v = v + dpp_f64<0x4E>(v);
  v = v + dpp_f64<0x141>(v);
  v = v + dpp_f64<0x140>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
  // counts per lane fit in 32 bits
  int x = (int)v;
  x += dpp_i32<0xB1>(x);
  x += dpp_i32<0x4E>(x);
  x += dpp_i32<0x141>(x);
  x += dpp_i32<0x140>(x);
  return (long long)__builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16) +
         __builtin_amdgcn_readlane(x, 32) + __builtin_amdgcn_readlane(x, 48);
}

__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  v = fmax(v, dpp_f64<0x141>(v));
  v = fmax(v, dpp_f64<0x140>(v));
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
              fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

__device__ __forceinline__ float wave_max_f32(float v) {
  v = fmaxf(v, __int_as_float(dpp_i32<0xB1>(__float_as_int(v))));
  v = fmaxf(v, __int_as_float(dpp_i32<0x4E>(__float_as_int(v))));
  v = fmaxf(v, __int_as_float(dpp_i32<0x141>(__float_as_int(v))));
  v = fmaxf(v, __int_as_float(dpp_i32<0x140>(__float_as_int(v))));
  const float a = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16)));
  const float b = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)),
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48)));
  return fmaxf(a, b);
}

// Threshold of the choice's candidates for a float32 maximum Mw of the
// float64 bounds: Mw = fl32(Ma) of the largest bound Ma, |Mw - Ma| <=
// 2^-24 |Ma| < (|Mw| + 1) 2^-23, and every maximiser of the exact g lies
// within 2 eps of Ma, so every one of them reaches Mw - 2 eps - (|Mw| + 1) 2^-23.
__device__ __forceinline__ double choice_threshold(double Mw) {
  return Mw - 2.0 * kGumbelSlack - (fabs(Mw) + 1.0) * 0x1p-23;
}

// Lane exchange with the partner that differs in lane bit B: row_mirror
// (15 - l) for bit 3, row_half_mirror (7 - l) for bit 2, quad_perm xor 2 /
// xor 1 for bits 1 / 0, ds_bpermute xor 16 / xor 32 for bits 4 / 5.
template <int B>
__device__ __forceinline__ double lane_exchange(double v) {
  if constexpr (B == 3) return dpp_f64<0x140>(v);
  else if constexpr (B == 2) return dpp_f64<0x141>(v);
  else if constexpr (B == 1) return dpp_f64<0x4E>(v);
  else if constexpr (B == 0) return dpp_f64<0xB1>(v);
  else return __shfl_xor(v, 1 << B, 64);
}

// Pairwise sums across lane bit 5 (B = 5) or 4 (B = 4) of two values a, b:
// lanes with the bit clear get a(lane) + a(partner), lanes with it set
// b(partner) + b(lane). One v_permlane32_swap / v_permlane16_swap per 32-bit
// half exchanges the upper half-wave (odd rows of 16 lanes) of a with the
// lower half-wave (even rows) of b, which leaves both addends of every lane in
// the two registers without a select.
template <int B>
__device__ __forceinline__ double half_swap_sum(double a, double b) {
  const long long ab = __double_as_longlong(a), bb = __double_as_longlong(b);
  uint32_t a0 = (uint32_t)ab, a1 = (uint32_t)(ab >> 32);
  uint32_t b0 = (uint32_t)bb, b1 = (uint32_t)(bb >> 32);
  if constexpr (B == 5) {
    const auto s0 = __builtin_amdgcn_permlane32_swap(a0, b0, false, false);
    const auto s1 = __builtin_amdgcn_permlane32_swap(a1, b1, false, false);
    a0 = s0[0], b0 = s0[1], a1 = s1[0], b1 = s1[1];
  } else {
    static_assert(B == 4, "half_swap_sum: lane bit 4 or 5");
    const auto s0 = __builtin_amdgcn_permlane16_swap(a0, b0, false, false);
    const auto s1 = __builtin_amdgcn_permlane16_swap(a1, b1, false, false);
    a0 = s0[0], b0 = s0[1], a1 = s1[0], b1 = s1[1];
  }
  const double x = __longlong_as_double((long long)(((unsigned long long)a1 << 32) | a0));
  const double y = __longlong_as_double((long long)(((unsigned long long)b1 << 32) | b0));
  return x + y;
}

// Transposing all-reduction of KP (a power of two, <= 32) values per lane:
// level t exchanges half of the values still held with the partner across
// lane bit B[t] (B = 5, 4, 3, 2, 1, 0) and keeps the other half, so after
// log2(KP) levels each lane holds one value index; the remaining levels add
// the partner's copy. Returns the wave total of value index
// v(lane) = sum_t bit(lane, B[t]) << (log2(KP) - 1 - t); every lane whose
// remaining bits are zero holds a distinct index. The half-wave levels (bits
// 5 and 4) come first: their exchanges need no select (half_swap_sum), and
// they carry the most values.
template <int KP, int T = 0>
__device__ __forceinline__ double transpose_sum(double (&x)[KP], int lane) {
  constexpr int kBits[6] = {5, 4, 3, 2, 1, 0};
  if constexpr (T == 6) {
    return x[0];
  } else {
    constexpr int B = kBits[T];
    if constexpr (B >= 4) {
      constexpr int H = KP > 1 ? KP / 2 : 1;
      double y[H];
#pragma unroll
      for (int i = 0; i < H; ++i) y[i] = half_swap_sum<B>(x[i], x[KP > 1 ? i + H : i]);
      return transpose_sum<H, T + 1>(y, lane);
    } else if constexpr (KP > 1) {
      constexpr int H = KP / 2;
      const bool hi = (lane >> B) & 1;
      double y[H];
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const double send = hi ? x[i] : x[i + H];
        const double keep = hi ? x[i + H] : x[i];
        y[i] = keep + lane_exchange<B>(send);
      }
      return transpose_sum<H, T + 1>(y, lane);
    } else {
      double y[1] = {x[0] + lane_exchange<B>(x[0])};
      return transpose_sum<1, T + 1>(y, lane);
    }
  }
}

// Value index held by `lane` after transpose_sum<KP> (-1: a duplicate).
template <int KP>
__device__ __forceinline__ int transpose_index(int lane) {
  constexpr int kBits[6] = {5, 4, 3, 2, 1, 0};
  int v = 0, L = 0;
  while ((1 << L) < KP) ++L;
  for (int t = 0; t < 6; ++t) {
    const int b = (lane >> kBits[t]) & 1;
    if (t < L) v = (v << 1) | b;
    else if (b) return -1;
  }
  return v;
}

// lexicographic max of (g, -j): larger g wins, lower j breaks ties.
__device__ __forceinline__ void arg_combine(double& g, int& j, double g2, int j2) {
  if (g2 > g || (g2 == g && j2 < j)) {
    g = g2;
    j = j2;
  }
}

template <int CTRL>
__device__ __forceinline__ void argmax_step(double& g, int& j) {
  const double g2 = dpp_f64<CTRL>(g);
  const int j2 = dpp_i32<CTRL>(j);
  arg_combine(g, j, g2, j2);
}

__device__ __forceinline__ void wave_argmax(double& g, int& j) {
  argmax_step<0xB1>(g, j);
  argmax_step<0x4E>(g, j);
  argmax_step<0x141>(g, j);
  argmax_step<0x140>(g, j);
  double gb = readlane_f64(g, 0);
  int jb = __builtin_amdgcn_readlane(j, 0);
  arg_combine(gb, jb, readlane_f64(g, 16), __builtin_amdgcn_readlane(j, 16));
  arg_combine(gb, jb, readlane_f64(g, 32), __builtin_amdgcn_readlane(j, 32));
  arg_combine(gb, jb, readlane_f64(g, 48), __builtin_amdgcn_readlane(j, 48));
  g = gb;
  j = jb;
}

// Word 0 of the Philox4x32-10 blocks (c0, c1, 1, j[r]) of the lane's R
// alternatives. The rounds are unrolled with the R independent blocks
// interleaved: the key schedule and the parts of the first rounds that do not
// depend on j (c0, c1 and the key are wave-uniform) then stay in scalar
// registers, and no register rotation is needed between rounds.
template <int R>
__device__ __forceinline__ void philox_alternatives(uint32_t c0, uint32_t c1, const int (&j)[R],
                                                    uint32_t k0, uint32_t k1, uint32_t (&u)[R]) {
  uint32_t a[R], b[R], c[R], d[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    a[r] = c0;
    b[r] = c1;
    c[r] = 1u;
    d[r] = (uint32_t)j[r];
  }
#pragma unroll
  for (int round = 0; round < 9; ++round) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t m0 = (uint64_t)0xD2511F53u * a[r];
      const uint64_t m1 = (uint64_t)0xCD9E8D57u * c[r];
      const uint32_t y0 = (uint32_t)(m1 >> 32) ^ b[r] ^ k0;
      const uint32_t y2 = (uint32_t)(m0 >> 32) ^ d[r] ^ k1;
      b[r] = (uint32_t)m1;
      d[r] = (uint32_t)m0;
      a[r] = y0;
      c[r] = y2;
    }
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  // tenth round: only output word 0 is used, which needs one of the two products
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t m1 = (uint64_t)0xCD9E8D57u * c[r];
    u[r] = (uint32_t)(m1 >> 32) ^ b[r] ^ k0;
  }
}

// Per-alternative features of one lane (R alternatives). Alternative r of
// this lane is actor j = 64*(NW*r + wave) + lane; lanes past n read actor i
// (any valid row) and are masked out.
template <int R>
struct Alt {
  int j[R];         // alternative actor (clamped to i where j >= n)
  uint32_t in;      // bit r: j < n
  uint32_t ok;      // bit r: choosable (j == i always)
  uint32_t xij;     // bit r: tie i->j
  uint32_t xji;     // bit r: tie j->i
  int so[R], cy[R], tp[R];
  __device__ __forceinline__ int bij(int r) const { return (xij >> r) & 1; }
  __device__ __forceinline__ int bji(int r) const { return (xji >> r) & 1; }
  __device__ __forceinline__ double dl(int r) const { return 1.0 - 2.0 * (double)bij(r); }
  // dl(r) * y with the same bits (dl is +-1, the product a sign flip)
  __device__ __forceinline__ double sdl(int r, double y) const {
    return __longlong_as_double(
        (long long)((unsigned long long)__double_as_longlong(y) ^ ((unsigned long long)bij(r) << 63)));
  }
};

// Ego-level state shared by all alternatives of the ministep. Row and column
// of i sit in registers when the network is LDS-resident, word w in lane w
// (a word is read into scalar registers where it is used); a streamed network
// (rows in device memory, W64 up to 128 words) keeps them in LDS instead.
// (Reading them from LDS where they lie was measured slower also at four and
// eight waves per chain: 121 against 118 ms at 500 actors, 99 against 90 ms
// at 1000; every wave holding every word in registers, 82.0 against 81.5 ms at
// 500 actors.)
template <int WMAX, bool STREAM>
struct Ego {
  int i;
  int outdeg_i, indeg_i, rdeg_i;
  double zvi;
  uint64_t rl = 0, cl = 0;  // word `lane` of row i (out-ties) and column i (in-ties)
  const uint64_t* rowi_s;
  const uint64_t* coli_s;
  int W;  // words looped over: WMAX (zero-padded rows) or W64 (streamed)
  __device__ __forceinline__ uint64_t rowi(int w) const {
    if constexpr (STREAM) return rowi_s[w];
    else return readlane64(rl, w);
  }
  __device__ __forceinline__ uint64_t coli(int w) const {
    if constexpr (STREAM) return coli_s[w];
    else return readlane64(cl, w);
  }
  __device__ __forceinline__ int words() const {
    if constexpr (STREAM) return W;
    else return WMAX;
  }
};

// LDS views of the chain state. Word w of row x: rows[w * RS + x] (LDS, or
// device memory for a streamed network without partial residency); a
// streamed network with partial residency keeps words w < lw in LDS (rows)
// and the rest in device memory (rows_d, same layout, every word).
struct ChainLds {
  const uint64_t* rows;
  int RS, W64, NW, wv, lane;
  const int *indeg, *outdeg, *z, *tpall, *soall;
  const uint64_t *tpm0, *tpm1;
  const uint64_t* rows_d = nullptr;
  int lw = 1 << 30;
  template <bool STREAM>
  __device__ __forceinline__ uint64_t row(int w, int x) const {
    if constexpr (STREAM) {
      if (w >= lw) return rows_d[(size_t)w * RS + x];  // w is uniform: no divergence
    }
    return rows[(size_t)w * RS + x];
  }
};

// Change statistics of one non-interaction effect e for the lane's R
// alternatives (libcusna_saom_cpu.cpp:49-101), creation/endowment masking
// applied (:258-259). Covariate values: vmat row e, or the current behaviour
// minus zbar for behaviour-dynamic effects (:233, 244, 252-253).
template <int R, int WMAX, bool STREAM, bool SIMPLE, class EgoT>
__device__ __forceinline__ void eff_base(const SimParams& P, int e, const Alt<R>& A,
                                         const EgoT& U, const ChainLds& S, double (&d)[R],
                                         const double* vjr = nullptr) {
  auto ROW = [&](int w, int x) -> uint64_t { return S.template row<STREAM>(w, x); };
  const int* indeg = S.indeg;
  const int* outdeg = S.outdeg;
  const int* z = S.z;
  const int* tpall = S.tpall;
  const int* soall = S.soall;
  const int t = eff_type(P, e);
  const int n = shape_n(P);
  const bool dyn = eff_dyn(P, e) != 0;
  const double* vrow = P.vmat + (size_t)e * n;
  const double* w = P.wtab + (size_t)e * (n + 1);
  const double p1 = P.p1[e], p2 = P.p2[e];
  // vjr: the alternatives' covariate values held by the caller (a static
  // covariate's values at the lane's actors do not change within a period)
  auto VJ = [&](int r) -> double {
    if (vjr) return vjr[r];
    return dyn ? ((double)z[A.j[r]] - P.zbar) : vrow[A.j[r]];
  };
  const double vi = dyn ? U.zvi : load_uniform(vrow + U.i);
  switch (t) {
    case 0:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.dl(r);
      break;
    case 1:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)A.bji(r));
      break;
    case 2:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)(A.tp[r] + A.so[r]));
      break;
    case 3:
    case 23:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)A.cy[r]);
      break;
    case 4:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)(indeg[A.j[r]] + (1 - A.bij(r))));
      break;
    case 7:
#pragma unroll
      for (int r = 0; r < R; ++r)
        d[r] = A.sdl(r, sqrt((double)(indeg[A.j[r]] + (1 - A.bij(r)))));
      break;
    case 8:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)outdeg[A.j[r]]);
      break;
    case 9:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, sqrt((double)outdeg[A.j[r]]));
      break;
    case 10:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, vi);
      break;
    case 11:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, VJ(r));
      break;
    case 12:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (1.0 - fabs(vi - VJ(r)) * p1 - p2));
      break;
    case 13:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (fabs(vi - VJ(r)) < 1e-9 ? 1.0 : 0.0));
      break;
    case 14:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)U.indeg_i);
      break;
    case 15:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, sqrt((double)U.indeg_i));
      break;
    case 16:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (2.0 * (double)U.outdeg_i + A.dl(r)));
      break;
    case 17: {
      const double d0 = (double)U.outdeg_i;
      const double s0 = d0 * sqrt(d0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double d1 = d0 + A.dl(r);
        d[r] = d1 * sqrt(d1) - s0;
      }
      break;
    }
    case 18: {
      const double d0 = (double)U.outdeg_i;
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = fmin(d0 + A.dl(r), p1) - fmin(d0, p1);
      break;
    }
    case 19:
#pragma unroll
      for (int r = 0; r < R; ++r)
        d[r] = A.sdl(r, ((double)U.rdeg_i + (double)A.bji(r) * ((double)U.outdeg_i + A.dl(r))));
      break;
    case 20:
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double vj = VJ(r);
        d[r] = vi > vj + 1e-9 ? A.dl(r) : (fabs(vi - vj) <= 1e-9 ? A.dl(r) * 0.5 : 0.0);
      }
      break;
    case 21:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, vi * VJ(r));
      break;
    case 25:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (vi + VJ(r)));
      break;
    case 31:
#pragma unroll
      for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (VJ(r) >= p1 - 1e-12 ? 1.0 : 0.0));
      break;
    default:
      if constexpr (!SIMPLE) {
        switch (t) {
      case 5:  // transTies: |row_i & row_j & {h: tp_h == x_ij}| + [tp_j > 0]
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint64_t* tm = A.bij(r) ? S.tpm1 : S.tpm0;
          int c = 0;
#pragma unroll
          for (int wd = 0; wd < U.words(); ++wd)
            c += popc64(U.rowi(wd) & ROW(wd, A.j[r]) & tm[wd]);
          d[r] = A.sdl(r, (double)(c + (A.tp[r] > 0 ? 1 : 0)));
        }
        break;
      case 6:  // transRecTrip: x_ji tp_j + |row_i & col_i & row_j|
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int c = 0;
#pragma unroll
          for (int wd = 0; wd < U.words(); ++wd)
            c += popc64(U.rowi(wd) & U.coli(wd) & ROW(wd, A.j[r]));
          d[r] = A.sdl(r, (double)((A.bji(r) ? A.tp[r] : 0) + c));
        }
        break;
      case 22: {
        const int dc = P.dci[e];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double dyw =
              (dc >= 0 && P.dyc) ? P.dyc[((size_t)dc * n + U.i) * n + A.j[r]] : 0.0;
          d[r] = A.sdl(r, dyw);
        }
        break;
      }
      case 24:
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, w[A.tp[r]]);
        break;
      case 26:
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, w[A.so[r]]);
        break;
      case 27: {  // gwespBF: common in-neighbours #{h in in(i): h -> j}
        int c[R];
#pragma unroll
        for (int r = 0; r < R; ++r) c[r] = 0;
#pragma unroll
        for (int wd = 0; wd < U.words(); ++wd) {
          uint64_t m = U.coli(wd);
          while (m) {
            const int h = wd * 64 + __ffsll((long long)m) - 1;
            m &= m - 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int word = S.NW * r + S.wv;
              const uint64_t x = (!STREAM || word < S.W64) ? ROW(word, h) : 0ull;
              c[r] += (int)((x >> S.lane) & 1);
            }
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, w[c[r]]);
        break;
      }
      case 28:
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, w[A.cy[r]]);
        break;
      case 29:  // gwdspFF: out-neighbours h of j (h != i), in ascending order
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double s = 0.0;
          for (int wd = 0; wd < P.W64; ++wd) {
            uint64_t m = ROW(wd, A.j[r]);
            while (m) {
              const int h = wd * 64 + __ffsll((long long)m) - 1;
              m &= m - 1;
              if (h == U.i) continue;
              const int p = tpall[h];
              s += A.bij(r) ? (w[p] - w[p - 1]) : (w[p + 1] - w[p]);
            }
          }
          d[r] = A.sdl(r, s);
        }
        break;
      case 30:  // gwdspFB: in-neighbours h of j (h != i), in ascending order
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double s = 0.0;
          const int jw = A.j[r] >> 6, jb = A.j[r] & 63;
          for (int h = 0; h < n; ++h) {
            if (h == U.i || !((ROW(jw, h) >> jb) & 1)) continue;
            const int p = soall[h];
            s += A.bij(r) ? (w[p] - w[p - 1]) : (w[p + 1] - w[p]);
          }
          d[r] = A.sdl(r, s);
        }
        break;
          default:
#pragma unroll
            for (int r = 0; r < R; ++r) d[r] = 0.0;
            break;
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = 0.0;
      }
      break;
  }
  const int et = eff_ety(P, e);
  if (et != 0) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if ((et == 1 && A.bij(r) == 1) || (et == 2 && A.bij(r) == 0)) d[r] = 0.0;
  }
}

// Full change statistic of effect e, including interactions (type 32):
// d = (c2 < 0 ? delta : 1) * (d_c0 * d_c1) [* d_c2] (:246-251). Alternatives
// j == i get 0 (:238).
template <int R, int WMAX, bool STREAM, bool SIMPLE, class EgoT>
__device__ __forceinline__ void eff_full(const SimParams& P, int e, const Alt<R>& A,
                                         const EgoT& U, const ChainLds& S, double (&d)[R],
                                         bool mask = true, const double* vjr = nullptr) {
  if constexpr (SIMPLE) {
    eff_base<R, WMAX, STREAM, true>(P, e, A, U, S, d, vjr);
  } else {
    // One eff_base call site: a plain effect runs the loop once; an
    // interaction runs it over its two or three components.
    const bool inter = eff_type(P, e) == 32;
    const int c2 = P.c2[e];
    double prod[R];
    for (int q = 0; q < 3; ++q) {
      const int ce = !inter ? (q == 0 ? e : -1) : (q == 0 ? P.c0[e] : (q == 1 ? P.c1[e] : c2));
      if (ce < 0) break;
      double tmp[R];
      eff_base<R, WMAX, STREAM, false>(P, ce, A, U, S, tmp, inter ? nullptr : vjr);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (!inter) d[r] = tmp[r];
        else if (q == 0) prod[r] = tmp[r];
        else if (q == 1) d[r] = (c2 < 0 ? A.dl(r) : 1.0) * (prod[r] * tmp[r]);
        else d[r] = d[r] * tmp[r];
      }
    }
  }
  // the ego's own slot (no change) and slots past n: no change statistics
  // (callers that weight d themselves skip this)
  if (mask) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (A.j[r] == U.i || !((A.in >> r) & 1)) d[r] = 0.0;
  }
}

// Behaviour change statistic (libcusna_saom_cpu.cpp:103-113).
__device__ __forceinline__ double beh_d(int t, double dz, double zic, long long nb, double zsum,
                                        long long nH, long long nE, long long nL, double brange,
                                        double bsim) {
  switch (t) {
    case 0:
      return dz;
    case 1:
      return (2.0 * zic + dz) * dz;
    case 2:
      return nb > 0 ? dz * zsum / (double)nb : 0.0;
    case 3:
    case 4: {
      if (nb == 0) return 0.0;
      const double tc = dz > 0 ? (double)(nH - nE - nL) : (double)(nL - nE - nH);
      return t == 3 ? tc / (brange * (double)nb) : tc / brange - (double)nb * bsim;
    }
    default:
      return 0.0;
  }
}

struct LdsLayout {
  size_t rows, indeg, outdeg, z, tpall, soall, coli, rowi, tpm0, tpm1, red, tsl, dstar, sc, bsc,
      cnt, crate, ract, ru, rate, exptab, draw, dbatch, total;
};

// Word bounds above this stream the network from device memory.
constexpr int kLdsWordBound = 16;

__host__ __device__ inline size_t al8(size_t x) { return (x + 7) & ~(size_t)7; }

// Rows are padded to wmax words (zero above W64) so the unrolled word loops of
// the ministep need no bounds tests. Streamed networks (wmax above
// kLdsWordBound) keep their rows in device memory and row i in LDS.
__host__ __device__ inline LdsLayout lds_layout(int n, int k, int kb, int wmax, int RS, int nw,
                                                int flags, int lw) {
  const int W64 = wmax;
  const bool stream = wmax > kLdsWordBound;
  LdsLayout L;
  size_t o = 0;
  L.rows = o; o += (size_t)(stream ? lw : W64) * RS * 8;  // streamed: the leading lw words
  L.indeg = o; o += (size_t)n * 4;
  L.outdeg = o; o += (size_t)n * 4;
  L.z = o; if (flags & F_Z) o += (size_t)n * 4;
  L.tpall = o; if (flags & (F_TPALL | F_SOALL)) o += (size_t)n * 4;
  L.soall = o; if (flags & (F_TPALL | F_SOALL)) o += (size_t)n * 4;
  o = al8(o);
  L.coli = o; o += (size_t)W64 * 8;
  L.rowi = o; if (stream) o += (size_t)W64 * 8;
  L.tpm0 = o; if (flags & (F_TPALL | F_SOALL)) o += (size_t)W64 * 8;
  L.tpm1 = o; if (flags & (F_TPALL | F_SOALL)) o += (size_t)W64 * 8;
  L.red = o; o += (size_t)nw * 4 * 8;
  L.tsl = o; o += (size_t)nw * (k + 1) * 8;
  L.dstar = o; o += (size_t)k * 8;
  L.sc = o; o += (size_t)k * 8;
  L.bsc = o; o += (size_t)(kb > 0 ? kb : 1) * 8;
  L.cnt = o; o += (size_t)nw * 8 * 8;
  // rate effects: per-actor rates in act order (then their prefix sums),
  // the active list, the rate covariates (nrate rows), then S0, the
  // rate-weighted sums, the rate scores and rho
  const int nrate = (flags & F_RATE) ? (flags >> 16) & 7 : 0;
  L.crate = o; if (flags & F_RATE) o += (size_t)n * 8;
  L.ru = o; o += (size_t)nrate * n * 8;
  L.ract = o; if (flags & F_RATE) o += (size_t)n * 4;
  o = al8(o);
  L.rate = o; if (flags & F_RATE) o += (size_t)(2 * kMaxRate + 3) * 8;
  L.exptab = o; o += (size_t)kExpTab * 8;  // 2^(j / kExpTab), j < kExpTab (exp_table)
  // the ring of ministep draws (period_body): a double and three words per
  // entry, 64 entries, fewer where the full ring would cost a chain per CU
  L.draw = o;
  const size_t lds_cu = 160 * 1024, per_cu = lds_cu / (o + 20);
  L.dbatch = 64;
  while (L.dbatch > 1 && per_cu > 0 && (o + 20 * L.dbatch) * per_cu > lds_cu) L.dbatch >>= 1;
  o += 20 * L.dbatch;
  if (!stream) {
    // the tie walk of the ministep reads word h >> 6 (h < n) of the rows of
    // every lane's actor 64 (nw r + wave) + lane, which reaches 64 wmax - 1
    // past n (lanes that are masked out): keep those reads inside the block
    const size_t walk_end = ((size_t)((n + 63) / 64 - 1) * RS + 64 * (size_t)wmax) * 8;
    if (o < walk_end) o = walk_end;
  }
  L.total = o;
  return L;
}

// Rate-effect covariate of kind kd (0 covariate row ru, 1 out-degree,
// 2 in-degree) of actor a (libcusna_saom_cpu.cpp:175-176).
__device__ __forceinline__ double rate_cov(int kd, const double* ru, int a, const int* outdeg,
                                           const int* indeg) {
  return kd == 0 ? ru[a] : (kd == 1 ? (double)outdeg[a] : (double)indeg[a]);
}

// The period of one chain (the kernel body; see the file comment).
template <int NW, int R, int WMAX, bool REPLAY, bool SIMPLE>
__device__ __forceinline__ void period_body(const SimParams* __restrict__ Pp,
                                            const uint32_t* __restrict__ items,
                                            uint32_t key0_arg, uint32_t key1_arg) {
  // One group: block b runs chain b of Pp[0]. Several groups in one launch:
  // item b names the group g (high 12 bits, parameters Pp[g]) and the chain
  // (low 20 bits).
  const uint32_t it = items ? items[blockIdx.x] : (uint32_t)blockIdx.x;
  const int chain = items ? (int)(it & 0xFFFFFu) : (int)blockIdx.x;
  const SimParams& P = Pp[items ? (it >> 20) : 0u];
  // the Philox key comes from the launch (one group) or the group's
  // parameters, or from device memory when a captured graph replays the
  // launch with keys a device kernel has written
  // wave-uniform: keep the key in scalar registers
  const uint32_t key0 = __builtin_amdgcn_readfirstlane(
      P.keys ? P.keys[0] : (items ? P.key0 : key0_arg));
  const uint32_t key1 = __builtin_amdgcn_readfirstlane(
      P.keys ? P.keys[1] : (items ? P.key1 : key1_arg));
  constexpr int NT = 64 * NW;
  extern __shared__ __align__(16) unsigned char smem[];
  // word-major bit rows: word w of actor h at rows[w * RS + h], RS = n, so the
  // lanes of a wave reading word w of consecutive actors hit consecutive words
  const int n = shape_n(P), kb = shape_kb(P), pflags = shape_flags(P);
#ifdef SAOM_FIXED_K
  constexpr int k = SAOM_FIXED_K;
#else
  const int k = P.k;
#endif
#ifdef SAOM_FIXED_N
  constexpr int W64 = (SAOM_FIXED_N + 63) / 64, RS = SAOM_FIXED_N;
#elif defined(SAOM_FIXED_W64)
  constexpr int W64 = SAOM_FIXED_W64;  // every chain of the launch has W64 words
  const int RS = P.RS;
#else
  const int W64 = P.W64, RS = P.RS;
#endif
  // the wave index is uniform: scalar registers and scalar branches on it
  const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr bool STREAM = WMAX > kLdsWordBound;
  // streamed: this chain's rows in device memory (rows_d), same word-major
  // layout with W64 words (no padding), of which the leading lw words live
  // in LDS instead (partial residency: read and written there only)
  const int lw = STREAM ? P.lw : 0;
  const LdsLayout L = lds_layout(n, k, kb, WMAX, RS, NW, pflags, lw);
  uint64_t* rows = (uint64_t*)(smem + L.rows);
  const uint32_t* rows32 = (const uint32_t*)(smem + L.rows);  // halves of the LDS words
  // the lane's half of word wv of row 0 (two-path walk: word NW r + wv of row h
  // at a fixed offset from it)
  const uint32_t* tp_base = rows32 + 2 * (size_t)wv * RS + (lane >> 5);
  // word 0 of the row of the lane's actor 64 wv + lane, unclamped (the tie walk
  // reads bit h of row j as the half of word h >> 6 at a uniform offset from it)
  const uint32_t* so_base = rows32 + 2 * (size_t)(64 * wv + lane);
  uint64_t* rows_d = STREAM ? P.work + (size_t)chain * W64 * RS : nullptr;
  // word w of row x; w is uniform at every call, so the branch does not diverge
  auto RGET = [&](int w, int x) -> uint64_t {
    if constexpr (STREAM) {
      if (w >= lw) return rows_d[(size_t)w * RS + x];
    }
    return rows[(size_t)w * RS + x];
  };
  auto RPUT = [&](int w, int x, uint64_t v) {
    if constexpr (STREAM) {
      if (w >= lw) {
        rows_d[(size_t)w * RS + x] = v;
        return;
      }
    }
    rows[(size_t)w * RS + x] = v;
  };
  uint64_t* rowi_s = (uint64_t*)(smem + L.rowi);
  int* indeg = (int*)(smem + L.indeg);
  int* outdeg = (int*)(smem + L.outdeg);
  int* z = (int*)(smem + L.z);
  int* tpall = (int*)(smem + L.tpall);
  int* soall = (int*)(smem + L.soall);
  uint64_t* coli_s = (uint64_t*)(smem + L.coli);
  uint64_t* tpm0 = (uint64_t*)(smem + L.tpm0);
  uint64_t* tpm1 = (uint64_t*)(smem + L.tpm1);
  double* red = (double*)(smem + L.red);
  double* tsl = (double*)(smem + L.tsl);
  double* dstar = (double*)(smem + L.dstar);
  double* sc = (double*)(smem + L.sc);
  double* bsc = (double*)(smem + L.bsc);
  long long* cnt = (long long*)(smem + L.cnt);
  double* exptab = (double*)(smem + L.exptab);
  double* crate = (double*)(smem + L.crate);    // rates, then their prefix sums
  double* ru_s = (double*)(smem + L.ru);         // rate covariates, q * n + actor
  int* ract = (int*)(smem + L.ract);             // active actors
  double* rsum_s = (double*)(smem + L.rate);      // [0] S0, [1 + q] sum c * u_q
  double* rsc_s = rsum_s + kMaxRate + 1;          // rate scores
  double* rho_s = rsc_s + kMaxRate;               // integrated total rate
#if defined(SAOM_FIXED_K) && defined(SAOM_HAS_RATE)
  // specialised build: rate effects compiled in or out
  constexpr int kRateMask = SAOM_HAS_RATE ? ~0 : ~F_RATE;
#else
  constexpr int kRateMask = ~0;
#endif
  const int flags = (SIMPLE ? (pflags & (F_TP | F_Z | F_RDEG | F_RATE)) : pflags) & kRateMask;
  const ChainLds S{rows, RS, W64, NW, wv, lane, indeg, outdeg, z, tpall, soall, tpm0, tpm1,
                   rows_d, lw};

  // start network and distance mask (problem inputs, read every ministep)
  const uint64_t* const x0p = P.x0;
  const uint64_t* const kdp = P.kd;
#ifdef SAOM_FIXED_K
  // static covariate values at the lane's alternatives, per effect that reads
  // them (slots past n: 0, never chosen)
  double vjpre[SAOM_FIXED_K][R];
  static_for<0, SAOM_FIXED_K>([&](auto EC) {
    constexpr int e = EC;
    if constexpr (alt_cov(e)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = (NW * r + wv) * 64 + lane;
        vjpre[e][r] = j < n ? P.vmat[(size_t)e * n + j] : 0.0;
      }
    }
  });
#endif

  // ---- chain set-up (libcusna_saom_cpu.cpp:146-160)
  constexpr int WI = STREAM ? 1 : WMAX;  // words per row stored (x W64 when streamed)
  const int WS = STREAM ? W64 : WI;
  for (int q = tid; q < n * WS; q += NT) {
    const int h = q / WS, wd = q - h * WS;
    RPUT(wd, h, wd < W64 ? P.xs[(size_t)h * W64 + wd] : 0ull);
  }
  for (int q = tid; q < WMAX; q += NT) coli_s[q] = 0ull;
  for (int q = tid; q < n; q += NT) {
    indeg[q] = 0;
    if (flags & F_Z) z[q] = P.z0[q];
  }
  for (int q = tid; q < k; q += NT) sc[q] = 0.0;
  for (int q = tid; q < kb; q += NT) bsc[q] = 0.0;
  for (int q = tid; q < kExpTab; q += NT) exptab[q] = exp2((double)q / (double)kExpTab);
  if (flags & F_RATE) {
    if (tid == 0) {
      for (int q = 0; q < kMaxRate; ++q) rsc_s[q] = 0.0;
      *rho_s = 0.0;
    }
    for (int x = tid; x < n; x += NT) {
      ract[x] = P.act ? P.act[x] : x;
      for (int q = 0; q < P.nrate; ++q) ru_s[(size_t)q * n + x] = P.ru ? P.ru[(size_t)q * n + x] : 0.0;
    }
  }
  __syncthreads();
  for (int h = tid; h < n; h += NT) {
    int od = 0;
    for (int wd = 0; wd < W64; ++wd) {
      uint64_t m = RGET(wd, h);
      od += popc64(m);
      while (m) {
        const int j = wd * 64 + __ffsll((long long)m) - 1;
        m &= m - 1;
        atomicAdd(&indeg[j], 1);
      }
    }
    outdeg[h] = od;
  }
  __syncthreads();

  // uniform over the workgroup: scalar registers (the compiler cannot prove
  // a load from a kernel that stores to global memory uniform by itself)
  const double lam = uniform_f64(P.lam[(size_t)chain * P.lam_stride]);
  const double lamb = kb > 0 ? uniform_f64(P.lamb[(size_t)chain * P.lam_stride]) : 0.0;
  const double* th = P.theta + (size_t)chain * P.th_stride;
  const double* thb = kb > 0 ? P.thetab + (size_t)chain * P.thb_stride : nullptr;
  const int nact = P.nact;
  // rate effects: first position in act order whose running rate sum is to be
  // recomputed at the next draw (all at the first)
  int rp = 0;
  long long steps = 0, steps_beh = 0, dist = 0, bdist = 0;
  long long cap = kMaxSteps;
  if (REPLAY && P.smax < cap) cap = P.smax;
  double t = 0.0;
  // The draw of ministep s (Philox block (s, chain, 0, 0) and log u0, or,
  // with the basic rate, the waiting time -log u0 / (nact (lam + lamb)))
  // depends on s alone, not on the network. One wave computes a batch of
  // them at once, one per lane, into an LDS ring that every wave reads
  // where each value is used; the wave that fills a batch rotates with the
  // batch number, so the SIMDs share the work. A fill starts after the
  // previous ministep's closing barrier (every read of the old batch is
  // done) and ends with a barrier of its own.
  const int DB = (int)L.dbatch;  // a power of two
  double* dv = (double*)(smem + L.draw);
  uint32_t* dc1 = (uint32_t*)(dv + DB);
  uint32_t* dc2 = dc1 + DB;
  uint32_t* dc3 = dc2 + DB;
  const double rate_div = (double)nact * (lam + lamb);
  auto fill = [&](long long s0) {
    const int b = (int)(s0 / DB);
    if (NW == 1 || __builtin_amdgcn_readfirstlane(wv) == b % NW) {  // uniform branch
      const int x = lane;  // DB <= 64
      if (x < DB) {
        const long long s = s0 + x;
        uint32_t c0 = (uint32_t)s, c1 = (uint32_t)chain, c2 = 0u, c3 = 0u;
        philox10(c0, c1, c2, c3, key0, key1);
        double lu0;
        if (REPLAY) lu0 = s < cap ? P.lu0[(size_t)chain * P.smax + s] : 0.0;
        else lu0 = log(unit_open(c0));
        dv[x] = (flags & F_RATE) ? lu0 : -lu0 / rate_div;
        dc1[x] = c1;
        dc2[x] = c2;
        dc3[x] = c3;
      }
    }
    __syncthreads();
  };

  // problem scalars read every ministep, held in scalar registers for the
  // period (the compiler would otherwise reload them from the parameter block
  // at each use, an exposed scalar-memory wait for a lone wave)
  int p_sym = P.symflag, p_cond = P.cond, p_target = P.target, p_maxdeg = P.maxdeg;
  const uint64_t* p_pm = P.pm;
  const int* p_act = P.act;
  double p_zbar = P.zbar;
  asm volatile("" : "+s"(p_sym), "+s"(p_cond), "+s"(p_target), "+s"(p_maxdeg));
  asm volatile("" : "+s"(p_pm), "+s"(p_act), "+s"(p_zbar));
  for (;;) {
    // 1. ministep draw (:167, 194-200)
    if ((steps & (DB - 1)) == 0) fill(steps);
    const int di = (int)(steps & (DB - 1));
    const uint32_t c1 = dc1[di], c2 = dc2[di];
    int i;
    bool net;
    if (flags & F_RATE) {
      // rate effects (:170-193): per-actor rates in parallel. Lane 0 sums
      // the rates in act order, as the reference does, writing the running
      // sums as it goes (S0 and the running sums fix the waiting time and
      // the ego: the ego is the first actor whose running sum reaches
      // u * S0, found by counting the running sums below it in parallel).
      // The covariate-weighted sums sum_a c_a u_q(a), which only enter the
      // rate scores, are summed in parallel (per thread, then a fixed-order
      // lane and wave reduction): their rounding differs from the
      // reference's sequential sum by float64 ulps.
      //
      // Only the running sums from position rp on are recomputed: the rates
      // before it are those of the previous draw (rp is the first position in
      // act order whose actor's rate can have changed since, nact when none),
      // so the sums there, and S0 when rp == nact, are the ones already held.
      const int nr = P.nrate;
      double part[kMaxRate];
#pragma unroll
      for (int q = 0; q < kMaxRate; ++q) part[q] = 0.0;
      for (int x = tid; x < nact; x += NT) {
        const int a2 = ract[x];
        double sx = 0.0, uq[kMaxRate];
#pragma unroll
        for (int q = 0; q < kMaxRate; ++q) {
          uq[q] = 0.0;
          if (q < nr) {
            uq[q] = rate_cov(P.rkind[q], ru_s + (size_t)q * n, a2, outdeg, indeg);
            const double prod = P.ralpha[q] * uq[q];
            sx = sx + prod;
          }
        }
        if (sx > 30.0) sx = 30.0;
        else if (sx < -30.0) sx = -30.0;
        const double c = exp(sx);
        if (x >= rp) crate[x] = c;
#pragma unroll
        for (int q = 0; q < kMaxRate; ++q) part[q] += c * uq[q];
      }
      double* wsum = (double*)cnt;  // slots 0..3 of each wave's counters
#pragma unroll
      for (int q = 0; q < kMaxRate; ++q) {
        if (q < nr) {
          const double v = wave_sum(part[q]);
          if (lane == 0) wsum[wv * 8 + q] = v;
        }
      }
      __syncthreads();
      if (tid == 0) {
        // the loads of kB rates are in flight together before their sums
        int x = rp;
        double acc = x > 0 ? crate[x - 1] : 0.0;
        constexpr int kB = 16;
        for (; x + kB <= nact; x += kB) {
          double cx[kB];
#pragma unroll
          for (int m = 0; m < kB; ++m) cx[m] = crate[x + m];
#pragma unroll
          for (int m = 0; m < kB; ++m) {
            acc += cx[m];
            crate[x + m] = acc;
          }
        }
        for (; x < nact; ++x) {
          acc += crate[x];
          crate[x] = acc;
        }
        if (rp < nact) rsum_s[0] = acc;
        for (int q = 0; q < nr; ++q) {
          double t = 0.0;
          for (int w = 0; w < NW; ++w) t += wsum[w * 8 + q];
          rsum_s[1 + q] = t;
        }
      }
      __syncthreads();
      rp = nact;  // every running sum current
      const double S0 = rsum_s[0];
      const double tau = -dv[di] / (lam * S0);
      double tau_eff = tau;
      if (!p_cond && t + tau > 1.0) tau_eff = 1.0 - t;
      if (tid == 0) {
        *rho_s += tau_eff * S0;
        for (int q = 0; q < nr; ++q) {
          const double dec = tau_eff * lam * rsum_s[1 + q];
          rsc_s[q] -= dec;
        }
      }
      t = uniform_f64(t + tau);
      if (!p_cond && t > 1.0) break;
      const double tgt = unit_open(c2) * S0;
      long long below = 0;  // running sums are non-decreasing: count those below tgt
      for (int x = tid; x < nact; x += NT) below += crate[x] < tgt ? 1 : 0;
      below = wave_sum_ll(below);
      if (NW > 1) {
        if (lane == 0) cnt[wv * 8 + 5] = below;
        __syncthreads();
        below = 0;
        for (int q = 0; q < NW; ++q) below += cnt[q * 8 + 5];
      }
      i = ract[below < nact ? (int)below : nact - 1];
      net = true;
    } else {
      t = uniform_f64(t + dv[di]);  // the chain time lives in scalar registers
      if (!p_cond && t > 1.0) break;
      net = kb == 0 || unit_open(c1) < lam / (lam + lamb);
      long long ix = (long long)(unit_open(c2) * nact);
      if (ix >= nact) ix = nact - 1;
      i = p_act ? p_act[ix] : (int)ix;
    }
    i = __builtin_amdgcn_readfirstlane(i);  // uniform: scalar registers and loads
    auto ROWI = [&](int w) -> uint64_t { return RGET(w, i); };

    if (!net) {
      // ---- behaviour ministep (:202-227)
      const int zi = z[i];
      const uint32_t c3 = dc3[di];
      const double zic = (double)zi - p_zbar;
      long long nb = 0, nH = 0, nE = 0, nL = 0;
      double zs = 0.0;  // exact in any order: terms are multiples of ulp(float zbar)
      for (int q = tid; q < n; q += NT) {
        if ((ROWI(q >> 6) >> (q & 63)) & 1) {
          nb++;
          const int zj = z[q];
          zs += (double)zj - p_zbar;
          if (zj > zi) nH++;
          else if (zj == zi) nE++;
          else nL++;
        }
      }
      nb = wave_sum_ll(nb);
      nH = wave_sum_ll(nH);
      nE = wave_sum_ll(nE);
      nL = wave_sum_ll(nL);
      zs = wave_sum(zs);
      if (NW > 1) {
        if (lane == 0) {
          cnt[wv * 8 + 0] = nb;
          cnt[wv * 8 + 1] = nH;
          cnt[wv * 8 + 2] = nE;
          cnt[wv * 8 + 3] = nL;
          red[wv * 4 + 0] = zs;
        }
        __syncthreads();
        nb = nH = nE = nL = 0;
        zs = 0.0;
        for (int q = 0; q < NW; ++q) {
          nb += cnt[q * 8 + 0];
          nH += cnt[q * 8 + 1];
          nE += cnt[q * 8 + 2];
          nL += cnt[q * 8 + 3];
          zs += red[q * 4 + 0];
        }
      }
      const bool up = zi < P.zmax, dn = zi > P.zmin;
      double fu = 0.0, fd = 0.0;
      double cup[kMaxBeff], cdn[kMaxBeff];
#pragma unroll
      for (int e = 0; e < kMaxBeff; ++e) {
        cup[e] = 0.0;
        cdn[e] = 0.0;
        if (e < kb) {
          cup[e] = beh_d(P.bty[e], 1.0, zic, nb, zs, nH, nE, nL, P.brange, P.bsim);
          cdn[e] = beh_d(P.bty[e], -1.0, zic, nb, zs, nH, nE, nL, P.brange, P.bsim);
          const double pu_ = thb[e] * cup[e];
          const double pd_ = thb[e] * cdn[e];
          fu = fu + pu_;
          fd = fd + pd_;
        }
      }
      double mx = 0.0;
      if (up && fu > mx) mx = fu;
      if (dn && fd > mx) mx = fd;
      double pd = dn ? exp(fd - mx) : 0.0, p0 = exp(-mx), pu = up ? exp(fu - mx) : 0.0;
      const double ps = pd + p0 + pu;
      pd /= ps;
      p0 /= ps;
      pu /= ps;
      const double uc = unit_open(c3);
      const int dz = uc < pd ? -1 : (uc < pd + p0 ? 0 : 1);
      __syncthreads();
      if (tid == 0) {
#pragma unroll
        for (int e = 0; e < kMaxBeff; ++e) {
          if (e < kb) {
            double s = dz == 1 ? cup[e] : (dz == -1 ? cdn[e] : 0.0);
            if (up) s -= pu * cup[e];
            if (dn) s -= pd * cdn[e];
            bsc[e] += s;
          }
        }
        if (dz != 0) z[i] = zi + dz;
      }
      if (dz != 0 && P.bkd[i]) {
        const long long a = (long long)zi + dz - P.z0[i], b = (long long)zi - P.z0[i];
        bdist += (a < 0 ? -a : a) - (b < 0 ? -b : b);
      }
      __syncthreads();
      steps_beh++;
      steps++;
      if (p_cond && dist >= p_target) break;
      if (steps >= cap) break;
      continue;
    }

    // ---- network ministep (:229-288)
    // 2. row i (word w in lane w) and col i (ballot of bit i over the rows)
    Ego<WMAX, STREAM> U;
    U.i = i;
    if constexpr (STREAM) {
      for (int wd = tid; wd < W64; wd += NT) rowi_s[wd] = ROWI(wd);  // read after the barrier
      U.rowi_s = rowi_s;
      U.coli_s = coli_s;
      U.W = W64;
    }
    Alt<R> A;
    A.in = A.ok = A.xij = A.xji = 0u;
    const int iw = i >> 6, ib = i & 63;
    int xji[R], xij[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int word = NW * r + wv;
      const int j = word * 64 + lane;
      const bool in = j < n;
      A.j[r] = in ? j : i;
      A.in |= (uint32_t)in << r;
      if constexpr (STREAM) {
        xji[r] = in ? (int)((RGET(iw, A.j[r]) >> ib) & 1) : 0;
        const uint64_t rw = word < W64 ? ROWI(word) : 0ull;
        xij[r] = (int)((rw >> lane) & 1);
      } else {  // 32-bit halves: bit i of row j (unclamped) and bit j of row i
        const uint32_t cj = so_base[2 * ((size_t)iw * RS + 64 * NW * r) + (ib >> 5)];
        xji[r] = in ? (int)((cj >> (ib & 31)) & 1u) : 0;
        const uint32_t ri = rows32[2 * ((size_t)word * RS + i) + (lane >> 5)];
        xij[r] = (int)((ri >> (lane & 31)) & 1u);
      }
    }
    // Philox words of the lane's alternatives (counter steps, chain, 1, j) and
    // their Gumbel-bound logarithms: they depend on the step and j only, so they
    // are computed here, where they fill the waits of the loads above and of
    // the barrier, instead of lengthening the choice (not used in replay)
    uint32_t uj[R];
    float llb[R];  // float32 values of log(-log u_j) (approx_loglog) for the bounds
    if constexpr (!REPLAY) {
      philox_alternatives<R>(__builtin_amdgcn_readfirstlane((uint32_t)steps),
                             __builtin_amdgcn_readfirstlane((uint32_t)chain), A.j, key0, key1,
                             uj);
#pragma unroll
      for (int r = 0; r < R; ++r) llb[r] = approx_loglog(uj[r]);
      // keep them here (the compiler would otherwise sink them to their use)
#pragma unroll
      for (int r = 0; r < R; ++r) asm volatile("" : "+v"(uj[r]), "+v"(llb[r]));
    }
    // the ego's degrees and behaviour and row i with word w in lane w: read
    // before the barrier (the previous toggle precedes it), used after it
    U.outdeg_i = outdeg[i];
    U.indeg_i = indeg[i];
    U.zvi = (flags & F_Z) ? (double)z[i] - p_zbar : -p_zbar;
    const uint64_t rl = (!STREAM && lane < WMAX) ? ROWI(lane) : 0ull;
    U.rl = rl;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t bal = __ballot(xji[r]);
      if (lane == 0) coli_s[NW * r + wv] = bal;
      A.xji |= (uint32_t)xji[r] << r;
      A.xij |= (uint32_t)xij[r] << r;
    }
    __syncthreads();
    if constexpr (!STREAM) U.cl = lane < WMAX ? coli_s[lane] : 0ull;
    U.rdeg_i = 0;
    if (flags & F_RDEG) {
#pragma unroll
      for (int wd = 0; wd < U.words(); ++wd) U.rdeg_i += popc64(U.rowi(wd) & U.coli(wd));
    }

    // 3. features: tp_j (two-paths i->h->j), so_j = |row_i & row_j|,
    //    cy_j = |row_j & col_i|. When the ego has few ties, so and cy come
    //    from walking its out- and in-ties (bit h of row j per tie, three
    //    instructions) instead of popcounts over every row word (eight per
    //    word); the counts are the same either way. Streamed rows (device
    //    memory) are walked when the ego has fewer ties than a row has
    //    words: each is one load per alternative either way.
    const int deg_i = __builtin_amdgcn_readfirstlane(U.outdeg_i + U.indeg_i);
    const bool walk = STREAM ? deg_i < U.words() : deg_i * 3 < U.words() * 8;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      A.tp[r] = 0;
      A.so[r] = 0;
      A.cy[r] = 0;
    }
    auto bit_of_row_j = [&](int r, int h) -> int {  // x_{j h}
      if constexpr (!STREAM) {
        // the 32-bit half that holds bit h, at the actor's unclamped row:
        // lanes past n read another word of the block and are masked out
        const uint32_t x = so_base[2 * ((size_t)(h >> 6) * RS + 64 * NW * r) + ((h >> 5) & 1)];
        return (int)((x >> (h & 31)) & 1u);
      } else {
        const uint64_t x = RGET(h >> 6, A.j[r]);
        const uint32_t half = (h & 32) ? (uint32_t)(x >> 32) : (uint32_t)x;
        return (int)((half >> (h & 31)) & 1u);
      }
    };
    auto out_tie = [&](int h) {  // an out-tie i -> h of the ego
      if (flags & F_TP) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (STREAM) {
            const int word = NW * r + wv;
            const uint64_t x = word < W64 ? RGET(word, h) : 0ull;
            A.tp[r] += (int)((x >> lane) & 1);
          } else {  // the 32-bit half of the word that holds the lane's bit
            const uint32_t x = tp_base[2 * ((size_t)NW * r * RS + h)];
            A.tp[r] += (int)((x >> (lane & 31)) & 1u);
          }
        }
      }
      if (walk) {
#pragma unroll
        for (int r = 0; r < R; ++r) A.so[r] += bit_of_row_j(r, h);
      }
    };
    auto in_tie = [&](int h) {  // an in-tie h -> i of the ego
#pragma unroll
      for (int r = 0; r < R; ++r) A.cy[r] += bit_of_row_j(r, h);
    };
    if constexpr (!STREAM) {
      // scalar loops over the set bits of the ego's nonzero row and column
      // words: lane w holds word w (rl, cl), a ballot names the nonzero
      // words, and only those are read into scalar registers
      if ((flags & F_TP) || walk) {
        for (uint64_t nz = __ballot(rl != 0); nz; nz &= nz - 1) {
          const int wd = __ffsll((long long)nz) - 1;
          for (uint64_t m = readlane64(rl, wd); m; m &= m - 1)
            out_tie(wd * 64 + __ffsll((long long)m) - 1);
        }
      }
      if (walk) {
        for (uint64_t nz = __ballot(U.cl != 0); nz; nz &= nz - 1) {
          const int wd = __ffsll((long long)nz) - 1;
          for (uint64_t m = readlane64(U.cl, wd); m; m &= m - 1)
            in_tie(wd * 64 + __ffsll((long long)m) - 1);
        }
      }
    } else {
      if ((flags & F_TP) || walk) {
        for (int wd = 0; wd < U.words(); ++wd) {
          for (uint64_t m = uniform64(U.rowi(wd)); m; m &= m - 1)  // i's out-ties
            out_tie(wd * 64 + __ffsll((long long)m) - 1);
        }
      }
      if (walk) {
        for (int wd = 0; wd < U.words(); ++wd) {
          for (uint64_t m = uniform64(U.coli(wd)); m; m &= m - 1)  // i's in-ties
            in_tie(wd * 64 + __ffsll((long long)m) - 1);
        }
      }
    }
    if (!walk) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int so = 0, cy = 0;
#pragma unroll
        for (int wd = 0; wd < U.words(); ++wd) {
          const uint64_t x = RGET(wd, A.j[r]);
          so += popc64(U.rowi(wd) & x);
          cy += popc64(x & U.coli(wd));
        }
        A.so[r] = so;
        A.cy[r] = cy;
      }
    }
    if (flags & (F_TPALL | F_SOALL)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int word = NW * r + wv;
        const bool in = (A.in >> r) & 1;
        if (in) {
          tpall[A.j[r]] = A.tp[r];
          soall[A.j[r]] = A.so[r];
        }
        const uint64_t b0 = __ballot(in && A.tp[r] == 0);
        const uint64_t b1 = __ballot(in && A.tp[r] == 1);
        if (lane == 0) {
          tpm0[word] = b0;
          tpm1[word] = b1;
        }
      }
      __syncthreads();
    }
    // choosable alternatives (:238-241)
    const uint64_t* pmrow = p_pm + (size_t)i * W64;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int word = NW * r + wv;
      bool ok;
      if (!((A.in >> r) & 1)) ok = false;
      else if (A.j[r] == i) ok = true;
      else {
        ok = (load_uniform(pmrow + word) >> lane) & 1;
        if (ok && p_maxdeg > 0 && A.bij(r) == 0 && U.outdeg_i >= p_maxdeg) ok = false;
      }
      A.ok |= (uint32_t)ok << r;
    }

    // 4. f_j in effect order
    double f[R];
#pragma unroll
    for (int r = 0; r < R; ++r) f[r] = 0.0;
#ifdef SAOM_FIXED_K
    constexpr int KC = SAOM_CACHE_D ? SAOM_FIXED_K : 1;
    double dc[KC][R];
    // unmasked: f of the ego's own slot is set to 0 below, slots past n are
    // not choosable, and the score sums weight the cached values (the own
    // slot with weight 0)
    static_for<0, SAOM_FIXED_K>([&](auto EC) {
      constexpr int e = EC;
      double d[R];
      eff_full<R, WMAX, STREAM, SIMPLE>(P, e, A, U, S, d, !SAOM_CACHE_D,
                                        alt_cov(e) ? vjpre[e] : nullptr);
      const double te = load_uniform(th + e);
      // density and reciprocity change by -1, 0 or 1, so te * d is exact and
      // one fused operation rounds f + te * d as the separate ones do; the
      // first term starts the sum (0 + te * d differs only in a zero's sign)
      constexpr bool kUnitD = kFixedT[e] == 0 || kFixedT[e] == 1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (e == 0) {
          f[r] = te * d[r];
        } else if constexpr (kUnitD) {
          f[r] = fma(te, d[r], f[r]);
        } else {
          const double prod = te * d[r];
          f[r] = f[r] + prod;
        }
        if (SAOM_CACHE_D) dc[SAOM_CACHE_D ? e : 0][r] = d[r];
      }
    });
#else
    for (int e = 0; e < k; ++e) {
      double d[R];
      eff_full<R, WMAX, STREAM, SIMPLE>(P, e, A, U, S, d);
      const double te = load_uniform(th + e);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double prod = te * d[r];
        f[r] = f[r] + prod;
      }
    }
#endif
    // 5. Gumbel-max choice (:265-268) and lmax
    double gb = -1e30, lm = -1e30;
    int jb = i;
    double shift = 0.0;  // softmax shift: lmax (replay) or the noisy maximum's bound
    bool exact_choice = true;  // the choice needs the exact noise's arg-max
    if constexpr (REPLAY) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (A.j[r] == i) f[r] = 0.0;
        if (!((A.ok >> r) & 1)) continue;
        if (f[r] > lm) lm = f[r];
        const double ll = P.llu[((size_t)chain * P.smax + steps) * n + A.j[r]];
        arg_combine(gb, jb, f[r] - ll, A.j[r]);
      }
    } else {
      // Bound every g_j = f_j - log(-log u_j) in float, then evaluate the
      // exact float64 noise only where g_j can reach the maximum: every j
      // with the largest exact g (ties included) has a bound within 2 eps of
      // the largest bound. The largest bound is taken in float32 (a four-step
      // DPP maximum) with the threshold widened by its rounding
      // (choice_threshold); the candidates are found by ballot.
      double ga[R];
      float m32 = -INFINITY;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (A.j[r] == i) f[r] = 0.0;
        ga[r] = -INFINITY;
        if (!((A.ok >> r) & 1)) continue;
        ga[r] = f[r] - (double)llb[r];
        m32 = fmaxf(m32, (float)ga[r]);
      }
      double M = (double)wave_max_f32(m32);
      double thr = choice_threshold(M);
      // candidates of this wave and, when there is one, its index
      int near = 0, jm = i;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t b = __ballot(((A.ok >> r) & 1) && ga[r] >= thr);
        if (b) {
          near += __popcll(b);
          jm = 64 * (NW * r + wv) + __ffsll((long long)b) - 1;
        }
      }
      // A single candidate is the exact maximiser, so the exact noise is
      // evaluated only when there are several (rarely: the top gap of a
      // Gumbel maximum is below 2 eps in about 0.8 % of draws).
      bool unique = near == 1;
      if (NW > 1) {
        if (lane == 0) {
          red[wv * 4 + 3] = M;
          cnt[wv * 8 + 6] = near;
          cnt[wv * 8 + 7] = jm;
        }
        __syncthreads();
        int qm = 0;
        M = red[3];
        for (int q = 1; q < NW; ++q)
          if (red[q * 4 + 3] > M) {
            M = red[q * 4 + 3];
            qm = q;
          }
        thr = choice_threshold(M);  // the leading wave's own threshold
        unique = cnt[qm * 8 + 6] == 1;
        // another wave's bounds lie below its float32 maximum plus its rounding
        for (int q = 0; q < NW; ++q) {
          const double Mq = red[q * 4 + 3];
          if (q != qm && Mq + (fabs(Mq) + 1.0) * 0x1p-23 >= thr) unique = false;
        }
        jm = (int)cnt[qm * 8 + 7];
      }
      // M >= max_j f_j - 3.14 - 2^-23 (|M| + 1) (the noise -log(-log u) is at
      // least -log(33 ln 2) for every u >= 2^-33 the stream yields), so
      // exp(f_j - M) stays below e^3.15 and M serves as the softmax shift: the
      // scores are ratios of the weighted sums, invariant to the shift up to
      // rounding
      shift = M;
      if (unique) {
        jb = jm;
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (ga[r] >= thr) {
            const double ll = log(-log(unit_open(uj[r])));
            arg_combine(gb, jb, f[r] - ll, A.j[r]);
          }
        }
      }
      exact_choice = !unique;
    }
    if (exact_choice) {  // uniform over the workgroup
      wave_argmax(gb, jb);
      if constexpr (REPLAY) lm = wave_max(lm);
      if (NW > 1) {
        if (lane == 0) {
          red[wv * 4 + 0] = gb;
          red[wv * 4 + 1] = lm;
          red[wv * 4 + 2] = (double)jb;
        }
        __syncthreads();
        gb = red[0];
        lm = red[1];
        jb = (int)red[2];
        for (int q = 1; q < NW; ++q) {
          arg_combine(gb, jb, red[q * 4 + 0], (int)red[q * 4 + 2]);
          if constexpr (REPLAY) lm = fmax(lm, red[q * 4 + 1]);
        }
      }
    }
    if constexpr (REPLAY) shift = lm;
    // uniform over the workgroup: scalar registers (and scalar loads below)
    const int jstar = __builtin_amdgcn_readfirstlane(jb);

    // 6. softmax-weighted expected change statistics (:271-274)
    double wgt[R];
    double Sw = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {  // every alternative evaluated, no branch
      const bool okr = (A.ok >> r) & 1;
      const double e = exp_table(okr ? f[r] - shift : -1000.0, exptab);
      wgt[r] = okr ? e : 0.0;
      Sw += wgt[r];
    }
#if defined(SAOM_FIXED_K) && SAOM_CACHE_D
    {
      // all k + 1 sums in one transposing reduction
      constexpr int KP = SAOM_FIXED_K + 1 <= 2 ? 2 : SAOM_FIXED_K + 1 <= 4 ? 4
                       : SAOM_FIXED_K + 1 <= 8 ? 8 : SAOM_FIXED_K + 1 <= 16 ? 16 : 32;
      double x[KP];
#pragma unroll
      for (int e = 0; e < KP; ++e) x[e] = 0.0;
      // the cached change statistics are unmasked: the ego's own slot enters
      // the sums with weight 0 (slots past n have weight 0 already)
      double wt[R];
#pragma unroll
      for (int r = 0; r < R; ++r) wt[r] = A.j[r] == i ? 0.0 : wgt[r];
#pragma unroll
      for (int e = 0; e < SAOM_FIXED_K; ++e) {
        double T = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) T = fma(wt[r], dc[e][r], T);  // enters only the scores
        x[e] = T;
      }
      // the chosen alternative's change statistics, from the lane that holds
      // it (unmasked: the score update takes none when the ego keeps its ties)
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (((A.in >> r) & 1) && A.j[r] == jstar) {
#pragma unroll
          for (int e = 0; e < SAOM_FIXED_K; ++e) dstar[e] = dc[e][r];
        }
      x[SAOM_FIXED_K] = Sw;
      const double tot = transpose_sum<KP>(x, lane);
      const int v = transpose_index<KP>(lane);
      if (v >= 0 && v <= SAOM_FIXED_K) tsl[wv * (k + 1) + v] = tot;
    }
#else
    Sw = wave_sum(Sw);
    if (lane == 0) tsl[wv * (k + 1) + k] = Sw;
    auto expected = [&](auto ev) {  // ev: an int, or IntC<e> in a fixed list
      const int e = ev;
      double d[R];
      eff_full<R, WMAX, STREAM, SIMPLE>(P, e, A, U, S, d);
      double T = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        T += wgt[r] * d[r];
        if (((A.in >> r) & 1) && A.j[r] == jstar) dstar[e] = d[r];
      }
      T = wave_sum(T);
      if (lane == 0) tsl[wv * (k + 1) + e] = T;
    };
#ifdef SAOM_FIXED_K
    static_for<0, SAOM_FIXED_K>(expected);
#else
    for (int e = 0; e < k; ++e) expected(e);
#endif
#endif
    // 7. state of the chosen dyad before the toggle (:277-281), from row i as
    // read in this ministep (registers, or the LDS copy of a streamed row)
    int aij = 0, x0ij = 0, kdij = 0;
    uint64_t rw = 0;  // word jstar >> 6 of row i before the toggle
    if (jstar != i) {
      const int jw = jstar >> 6, jbit = jstar & 63;
      if constexpr (STREAM) rw = rowi_s[jw];
      else rw = readlane64(rl, jw);
      const uint64_t x0w = load_uniform(x0p + (size_t)i * W64 + jw);
      const uint64_t kdw = load_uniform(kdp + (size_t)i * W64 + jw);
      aij = (int)((rw >> jbit) & 1);
      x0ij = (int)((x0w >> jbit) & 1);
      kdij = (int)((kdw >> jbit) & 1);
    }
    auto toggle = [&]() {  // thread 0: the step's tie change and the rate scores
      if (jstar != i) {  // row i's word and i's degrees as read in this ministep
        const int dd = 1 - 2 * aij;
        RPUT(jstar >> 6, i, rw ^ (1ull << (jstar & 63)));
        indeg[jstar] += dd;
        outdeg[i] = U.outdeg_i + dd;
        if (p_sym) {
          RPUT(i >> 6, jstar, RGET(i >> 6, jstar) ^ (1ull << (i & 63)));
          indeg[i] = U.indeg_i + dd;
          outdeg[jstar] += dd;
        }
      }
      // rate-effect scores: covariate of the ego after the step (:283-284)
      if (flags & F_RATE)
        for (int q = 0; q < P.nrate; ++q)
          rsc_s[q] += rate_cov(P.rkind[q], ru_s + (size_t)q * n, i, outdeg, indeg);
    };
    // With the change statistics cached in registers nothing after the choice
    // reads the network, its degrees or the rate state, so the toggle goes
    // before the barrier that publishes the score sums, and the score update
    // (wave 0) needs no closing barrier: the next writes of the sums and of
    // dstar come after the next ministep's first two barriers, and the next
    // reads of the network after this barrier.
#if defined(SAOM_FIXED_K) && SAOM_CACHE_D
    constexpr bool kEarlyToggle = true;
#else
    constexpr bool kEarlyToggle = false;
#endif
    if (kEarlyToggle && tid == 0) toggle();
    __syncthreads();
    if (tid < k) {  // one lane per effect: score_e += d_e(j*) - T_e / S
      double St = 0.0, Tt = 0.0;
      for (int q = 0; q < NW; ++q) {
        St += tsl[q * (k + 1) + k];
        Tt += tsl[q * (k + 1) + tid];
      }
      // no change statistics when the ego keeps its ties (dstar holds the
      // unmasked values of its own slot in the cached form)
      // T / S from one Newton step on the hardware reciprocal (within 2 ulps
      // of the division; the scores are float32 outputs)
      double rs = __builtin_amdgcn_rcp(St);
      rs = fma(fma(-St, rs, 1.0), rs, rs);
      sc[tid] += (jstar == i ? 0.0 : dstar[tid]) - Tt * rs;
    }
    if (!kEarlyToggle && tid == 0) toggle();
    steps++;
    if (jstar != i && kdij) {
      const long long chg = p_sym ? 2 : 1;
      dist += (aij == x0ij) ? chg : -chg;
    }
    if ((flags & F_RATE) && jstar != i) {
      // degree-based rates of the actors whose degrees changed: their running
      // sums and all after them are recomputed at the next draw (all of them
      // when the active actors are a list, whose positions are not kept)
      for (int q = 0; q < P.nrate; ++q) {
        const int kd = P.rkind[q];
        if (kd == 1 || (kd == 2 && p_sym)) rp = min(rp, p_act ? 0 : i);
        if (kd == 2 || (kd == 1 && p_sym)) rp = min(rp, p_act ? 0 : jstar);
      }
    }
    if (!kEarlyToggle) __syncthreads();
    if (p_cond && dist >= p_target) break;
    if (steps >= cap) break;
  }

  // ---- period end (:290-302)
  __syncthreads();
  if (P.has_comp && P.inactw) {
    for (int h = tid; h < n; h += NT) {
      const bool hin = (P.inactw[h >> 6] >> (h & 63)) & 1;
      for (int wd = 0; wd < W64; ++wd) {
        const uint64_t x0 = P.x0[(size_t)h * W64 + wd];
        const uint64_t im = P.inactw[wd];
        const uint64_t v = RGET(wd, h);
        RPUT(wd, h, hin ? x0 : ((v & ~im) | (x0 & im)));
      }
    }
    __syncthreads();
  }
  long long dsum = 0;
  uint64_t* out_rows = P.end_rows + (size_t)chain * n * W64;
  for (int h = tid; h < n; h += NT) {
    int c = 0;
    for (int wd = 0; wd < W64; ++wd) {
      const uint64_t v = RGET(wd, h);
      out_rows[(size_t)h * W64 + wd] = v;
      c += popc64((v ^ P.x0[(size_t)h * W64 + wd]) & P.kd[(size_t)h * W64 + wd]);
    }
    if (P.actor_dist) P.actor_dist[(size_t)chain * n + h] = c;
    dsum += c;
    if (P.z_end) P.z_end[(size_t)chain * n + h] = z[h];
  }
  dsum = wave_sum_ll(dsum);
  if (lane == 0) cnt[wv * 8 + 4] = dsum;
  __syncthreads();
  if (tid == 0) {
    long long ds = 0;
    for (int q = 0; q < NW; ++q) ds += cnt[q * 8 + 4];
    P.dist[chain] = (int32_t)dist;
    P.dist_stat[chain] = (int32_t)ds;
    P.steps[chain] = (int32_t)steps;
    P.ntim[chain] = t;
    double* so = P.scores + (size_t)chain * (k + 1);
    so[0] = (flags & F_RATE) ? (double)(steps - steps_beh) / lam - *rho_s
                             : (double)(steps - steps_beh) / lam - (double)nact;
    if (flags & F_RATE)
      for (int q = 0; q < P.nrate; ++q) P.rscores[(size_t)chain * P.nrate + q] = rsc_s[q];
    for (int e = 0; e < k; ++e) so[1 + e] = sc[e];
    if (kb > 0) {
      P.bdist[chain] = (int32_t)bdist;
      double* bo = P.bscores + (size_t)chain * (kb + 1);
      bo[0] = (double)steps_beh / lamb - (double)nact;
      for (int e = 0; e < kb; ++e) bo[1 + e] = bsc[e];
    }
  }
}

#if defined(SAOM_HELPERS_ONLY)
// included by another kernel file for the helpers above (no period kernels)
}  // namespace
}  // namespace saom
#elif !defined(SAOM_RTC)
// Register budget: 256 VGPRs (two waves per SIMD) for every multi-wave
// shape. Measured at n = 500: a 12-effect list in four waves per chain ran
// 310 ms with three waves' budget (168 VGPRs, spills) and 284 ms with two,
// 527 ms with four (the LDS-bound occupancy); the batch of 256 groups of 30-1200 actors (up to 16
// waves per chain) 13.3 s and 13.0 s.
template <int NW, int R, int WMAX, bool REPLAY, bool SIMPLE>
__global__ void __launch_bounds__(64 * NW)
    __attribute__((amdgpu_waves_per_eu(NW == 1 ? 1 : 2)))
    saom_period_kernel(const SimParams* __restrict__ Pp, const uint32_t* __restrict__ items,
                       uint32_t key0_arg, uint32_t key1_arg) {
  period_body<NW, R, WMAX, REPLAY, SIMPLE>(Pp, items, key0_arg, key1_arg);
}

template <int NW, int WMAX, bool REPLAY, bool SIMPLE>
hipError_t launch_shape(const PeriodLaunch& L, hipStream_t st) {
  constexpr int R = WMAX / NW;
  auto kern = saom_period_kernel<NW, R, WMAX, REPLAY, SIMPLE>;
  // raise the dynamic LDS limit once per instantiation (not inside a graph
  // capture); 160 KB is the whole LDS of a gfx950 CU
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(kern, dim3(L.grid), dim3(64 * NW), L.lds, st, L.Pd, L.items, L.key0,
                     L.key1);
  return hipGetLastError();
}

template <int NW, int WMAX>
hipError_t launch_variant(const PeriodLaunch& L, bool replay, bool simple, hipStream_t st) {
  if (replay) return launch_shape<NW, WMAX, true, false>(L, st);
  if (simple) return launch_shape<NW, WMAX, false, true>(L, st);
  return launch_shape<NW, WMAX, false, false>(L, st);
}

}  // namespace

// This file is compiled once per word bound SAOM_WMAX (1, 2, 4, 8, 16) so the
// instantiations build in parallel; launch_period (saom_launch.hip) dispatches.
#ifndef SAOM_WMAX
#error "compile with -DSAOM_WMAX=<1|2|4|8|16|32|64|128>"
#endif
#define SAOM_CAT2(a, b) a##b
#define SAOM_CAT(a, b) SAOM_CAT2(a, b)

int SAOM_CAT(launch_period_w, SAOM_WMAX)(const PeriodLaunch& L, int nw, bool replay,
                                         bool simple, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  constexpr int WM = SAOM_WMAX;
  hipError_t e = hipErrorInvalidValue;
  if constexpr (WM <= kLdsWordBound) {
    if constexpr (WM <= 8) {
      if (nw == 1) e = launch_variant<1, WM>(L, replay, simple, st);
    }
    if constexpr (WM >= 2) {
      if (nw == 2) e = launch_variant<2, WM>(L, replay, simple, st);
    }
    if constexpr (WM >= 4) {
      if (nw == 4) e = launch_variant<4, WM>(L, replay, simple, st);
    }
    if constexpr (WM >= 8) {
      if (nw == 8) e = launch_variant<8, WM>(L, replay, simple, st);
    }
  }
  // streamed networks run one shape: 16 waves, WM / 16 words per lane
  if constexpr (WM >= 16) {
    if (nw == 16) e = launch_variant<16, WM>(L, replay, simple, st);
  }
  return (int)e;
}

size_t SAOM_CAT(sim_lds_bytes_w, SAOM_WMAX)(const SimParams& P, int nw, int wmax) {
  return lds_layout(P.n, P.k, P.kb, wmax, P.RS, nw, P.flags, P.lw).total;
}

}  // namespace saom
#else   // SAOM_RTC: one specialised kernel, shape and effect list from macros
}  // namespace

// Register budget: the launch's occupancy in waves per SIMD (set by the
// host from the chain's LDS size; two waves, as the generic kernel, when the
// tighter budget would spill).
#ifndef SAOM_WAVES_PER_EU
#define SAOM_WAVES_PER_EU (SAOM_NW == 1 ? 1 : 2)
#endif
extern "C" __global__ void __launch_bounds__(64 * SAOM_NW)
    __attribute__((amdgpu_waves_per_eu(SAOM_WAVES_PER_EU)))
    saom_period_rtc(const SimParams* __restrict__ Pp, const uint32_t* __restrict__ items,
                    uint32_t key0_arg, uint32_t key1_arg) {
  period_body<SAOM_NW, SAOM_WMAX / SAOM_NW, SAOM_WMAX, SAOM_REPLAY != 0, SAOM_SIMPLE != 0>(
      Pp, items, key0_arg, key1_arg);
}

}  // namespace saom
#endif
