// saom_multi.hip -- multi-network SAOM periods for gfx950 (MI355X).
//
// D one-mode networks (2..4) co-evolve over one actor set: the reference's
// simulate_one_multi (cran/cusna src/libcusna_saom_cpu.cpp:460-510). One
// workgroup of NW wave64s runs one chain; the mapping of alternatives to
// lanes, the per-alternative features, the float32-bounded Gumbel-max choice
// and the score pass are those of the single-network kernel
// (saom_period.hip), applied to the network the ministep picks.
//
// Per ministep:
//   1. Philox block (steps, chain, 0, 0): waiting time -log(u0) / (nact *
//      sum_d lam_d); network d with probability lam_d / sum lam (cumulative
//      scan, :472); active actor i (uniform over act[0..nact)).
//   2. Row i / column i and the features of network d; alternatives are the
//      permitted toggles of d (perm) plus j = i.
//   3. f_j = sum over the effects of network d (netof[e] == d), in list
//      order, of theta_e * d_e(j). Ids 0-31 use the single-network change
//      statistics on d, with reciprocated degree 0 as the reference passes
//      (:488); ids 33-37 are cross-network effects reading source network
//      netref[e] in its current state (:430-443): crprod w_ij, crprodRecip
//      w_ji, crprodMutual w_ij w_ji, from |out_w(i) & out_w(j)| (square root
//      when parm >= 2), to #{h : w_ih x_hj}. Creation/endowment masks apply
//      to both (:493-494).
//   4. Gumbel-max choice, softmax-weighted scores of d's effects, toggle of
//      d (both cells when d is symmetric), per-network step counts.
// Period end: the end networks (no composition restore: the reference's
// multi-network engine has none), ministeps, time, rate scores
// steps_d / lam_d - nact (:505) and effect scores.
//
// Chain state: the D networks as word-major bit rows (word w of actor h at
// w * n + h) in LDS when they fit, else in a per-chain device buffer (rows
// padded to WMAX words for WMAX <= 16; streamed layout with W64 words above);
// in/out degrees of every network in LDS.
//
// Runtime specialisation (rtc.hip) compiles this file with SAOM_RTC and the
// effect list fixed (SAOM_FIXED_* as for the single-network kernel, plus
// SAOM_FIXED_NETOF / SAOM_FIXED_NETREF): the effect loops unroll and, when
// k x alternatives-per-lane <= 32, the change statistics stay in registers
// between the choice and score passes and the score sums take one transposing
// reduction. Same arithmetic in the same order as the generic kernel.
#ifndef SAOM_RTC
#define SAOM_HELPERS_ONLY
#include "saom_period.hip"
#undef SAOM_HELPERS_ONLY
#endif

namespace saom {
namespace {

#if defined(SAOM_FIXED_K) && defined(SAOM_FIXED_NETOF)
constexpr int kFixedNetof[] = {SAOM_FIXED_NETOF};
constexpr int kFixedNetref[] = {SAOM_FIXED_NETREF};
__device__ __forceinline__ int eff_netof(const MultiParams&, int e) { return kFixedNetof[e]; }
__device__ __forceinline__ int eff_netref(const MultiParams&, int e) { return kFixedNetref[e]; }
#else
__device__ __forceinline__ int eff_netof(const MultiParams& Q, int e) { return Q.netof[e]; }
__device__ __forceinline__ int eff_netref(const MultiParams& Q, int e) { return Q.netref[e]; }
#endif

struct MultiLds {
  size_t rows, indeg, outdeg, tpall, soall, coli, rowi, tpm0, tpm1, red, tsl, dstar, sc, sd,
      dc, exptab, total;
};

__host__ __device__ inline MultiLds multi_layout(int n, int k, int D, int wmax, int RS, int nw,
                                                 int flags, bool rows_global) {
  const bool stream = wmax > kLdsWordBound;
  const bool all = (flags & (F_TPALL | F_SOALL)) != 0;
  MultiLds L;
  size_t o = 0;
  L.rows = o; if (!rows_global) o += (size_t)D * wmax * RS * 8;
  L.indeg = o; o += (size_t)D * n * 4;
  L.outdeg = o; o += (size_t)D * n * 4;
  L.tpall = o; if (all) o += (size_t)n * 4;
  L.soall = o; if (all) o += (size_t)n * 4;
  o = al8(o);
  L.coli = o; o += (size_t)wmax * 8;
  L.rowi = o; if (stream) o += (size_t)wmax * 8;
  L.tpm0 = o; if (all) o += (size_t)wmax * 8;
  L.tpm1 = o; if (all) o += (size_t)wmax * 8;
  L.red = o; o += (size_t)nw * 6 * 8;  // per wave: exact-noise arg-max, bound maximum, candidates
  L.tsl = o; o += (size_t)nw * (k + 1) * 8;
  L.dstar = o; o += (size_t)k * 8;
  L.sc = o; o += (size_t)k * 8;
  L.sd = o; o += (size_t)kMaxNet * 8;
  L.dc = o; o += (size_t)kMaxNet * 4;
  o = al8(o);
  L.exptab = o; o += (size_t)kExpTab * 8;  // 2^(j / kExpTab) (exp_table)
  L.total = al8(o);
  return L;
}

// Change statistic of effect e (of the ministep's network) for the lane's R
// alternatives: cross-network ids 33-37 here, ids 0-31 by eff_base.
template <int R, int WMAX, bool STREAM>
__device__ __forceinline__ void multi_eff(const MultiParams& Q, int e, const Alt<R>& A,
                                          const Ego<WMAX, STREAM>& U, const ChainLds& S,
                                          const uint64_t* rows_all, size_t netw,
                                          double (&d)[R], bool mask = true) {
  const SimParams& P = Q.base;
  const int t = eff_type(P, e);
  if (t < 33 || t > 37) {
    eff_base<R, WMAX, STREAM, false>(P, e, A, U, S, d);
  } else {
    const uint64_t* rw = rows_all + (size_t)eff_netref(Q, e) * netw;
    const int RS = S.RS, i = U.i, iw = i >> 6, ib = i & 63;
    auto wij = [&](int r) -> int {  // bit j of row i of the source network
      const int word = S.NW * r + S.wv;
      const uint64_t x = (!STREAM || word < S.W64) ? rw[(size_t)word * RS + i] : 0ull;
      return (int)((x >> S.lane) & 1);
    };
    auto wji = [&](int r) -> int { return (int)((rw[(size_t)iw * RS + A.j[r]] >> ib) & 1); };
    switch (t) {
      case 33:
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)wij(r));
        break;
      case 34:
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)wji(r));
        break;
      case 35:
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)(wij(r) & wji(r)));
        break;
      case 36: {
        const bool root = P.p1[e] >= 2.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int c = 0;
          for (int wd = 0; wd < U.words(); ++wd)
            c += popc64(rw[(size_t)wd * RS + i] & rw[(size_t)wd * RS + A.j[r]]);
          d[r] = A.sdl(r, (root ? sqrt((double)c) : (double)c));
        }
        break;
      }
      default: {  // 37: out-neighbours h of i in the source network, bit j of row h here
        int c[R];
#pragma unroll
        for (int r = 0; r < R; ++r) c[r] = 0;
        for (int wd = 0; wd < U.words(); ++wd) {
          uint64_t m = rw[(size_t)wd * RS + i];
          while (m) {
            const int h = wd * 64 + __ffsll((long long)m) - 1;
            m &= m - 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int word = S.NW * r + S.wv;
              const uint64_t x =
                  (!STREAM || word < S.W64) ? S.rows[(size_t)word * RS + h] : 0ull;
              c[r] += (int)((x >> S.lane) & 1);
            }
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) d[r] = A.sdl(r, (double)c[r]);
        break;
      }
    }
    const int et = eff_ety(P, e);
    if (et != 0) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if ((et == 1 && A.bij(r) == 1) || (et == 2 && A.bij(r) == 0)) d[r] = 0.0;
    }
  }
  // the ego's own slot and slots past n: none (callers that weight d
  // themselves skip this)
  if (mask) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (A.j[r] == U.i || !((A.in >> r) & 1)) d[r] = 0.0;
  }
}

template <int NW, int R, int WMAX>
__device__ __forceinline__ void multi_body(const MultiParams* __restrict__ Qp) {
  const MultiParams& Q = *Qp;
  const SimParams& P = Q.base;
  constexpr int NT = 64 * NW;
  constexpr bool STREAM = WMAX > kLdsWordBound;
  extern __shared__ __align__(16) unsigned char smem[];
  const int chain = blockIdx.x;
  const int n = P.n, k = P.k, D = Q.D, W64 = P.W64, RS = P.RS;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const bool rg = STREAM || Q.rows_global;
  const MultiLds L = multi_layout(n, k, D, WMAX, RS, NW, P.flags, rg);
  const int WS = STREAM ? W64 : WMAX;  // words stored per row
  const size_t netw = (size_t)WS * RS;   // words per network
  uint64_t* rows_all =
      rg ? Q.work + (size_t)chain * D * netw : (uint64_t*)(smem + L.rows);
  int* indeg_all = (int*)(smem + L.indeg);
  int* outdeg_all = (int*)(smem + L.outdeg);
  int* tpall = (int*)(smem + L.tpall);
  int* soall = (int*)(smem + L.soall);
  uint64_t* coli_s = (uint64_t*)(smem + L.coli);
  uint64_t* rowi_s = (uint64_t*)(smem + L.rowi);
  uint64_t* tpm0 = (uint64_t*)(smem + L.tpm0);
  uint64_t* tpm1 = (uint64_t*)(smem + L.tpm1);
  double* red = (double*)(smem + L.red);
  double* tsl = (double*)(smem + L.tsl);
  double* dstar = (double*)(smem + L.dstar);
  double* sc = (double*)(smem + L.sc);
  long long* sd = (long long*)(smem + L.sd);
  double* exptab = (double*)(smem + L.exptab);
  int* dc = (int*)(smem + L.dc);
  const int flags = P.flags;

  // ---- chain set-up (:465-478)
  for (int q = tid; q < D * n * WS; q += NT) {
    const int d = q / (n * WS);
    const int rem = q - d * n * WS;
    const int h = rem / WS, wd = rem - h * WS;
    rows_all[(size_t)d * netw + (size_t)wd * RS + h] =
        wd < W64 ? Q.x0[((size_t)d * n + h) * W64 + wd] : 0ull;
  }
  for (int q = tid; q < WMAX; q += NT) coli_s[q] = 0ull;
  for (int q = tid; q < D * n; q += NT) indeg_all[q] = 0;
  for (int q = tid; q < k; q += NT) sc[q] = 0.0;
  for (int q = tid; q < kExpTab; q += NT) exptab[q] = exp2((double)q / (double)kExpTab);
  for (int q = tid; q < kMaxNet; q += NT) {
    sd[q] = 0;
    dc[q] = 0;
  }
  __syncthreads();
  for (int q = tid; q < D * n; q += NT) {
    const int d = q / n, h = q - d * n;
    const uint64_t* rd = rows_all + (size_t)d * netw;
    int od = 0;
    for (int wd = 0; wd < W64; ++wd) {
      uint64_t m = rd[(size_t)wd * RS + h];
      od += popc64(m);
      while (m) {
        const int j = wd * 64 + __ffsll((long long)m) - 1;
        m &= m - 1;
        atomicAdd(&indeg_all[d * n + j], 1);
      }
    }
    outdeg_all[q] = od;
  }
  __syncthreads();

  const double* th = Q.theta + (size_t)chain * k;
  const double* lam = Q.lam + (size_t)chain * D;
  double lam_tot = 0.0;
  for (int d = 0; d < D; ++d) lam_tot += lam[d];
  const int nact = P.nact;
  const uint32_t key0 = __builtin_amdgcn_readfirstlane(P.key0);
  const uint32_t key1 = __builtin_amdgcn_readfirstlane(P.key1);
  long long steps = 0;
  double t = 0.0;

  for (;;) {
    // 1. waiting time, network, actor (:467-476)
    uint32_t c0 = (uint32_t)steps, c1 = (uint32_t)chain, c2 = 0u, c3 = 0u;
    philox10(c0, c1, c2, c3, key0, key1);
    t += -log(unit_open(c0)) / ((double)nact * lam_tot);
    if (t > 1.0) break;
    const double uv = unit_open(c1) * lam_tot;
    int d = 0;
    double acc = lam[0];
    while (d < D - 1 && uv >= acc) {
      d += 1;
      acc += lam[d];
    }
    long long ix = (long long)(unit_open(c2) * nact);
    if (ix >= nact) ix = nact - 1;
    // uniform: scalar registers and loads
    const int i = __builtin_amdgcn_readfirstlane(P.act ? P.act[ix] : (int)ix);
    uint64_t* rows = rows_all + (size_t)d * netw;
    int* indeg = indeg_all + d * n;
    int* outdeg = outdeg_all + d * n;
    const ChainLds S{rows, RS, W64, NW, wv, lane, indeg, outdeg, nullptr, tpall, soall, tpm0,
                     tpm1};
    auto ROWI = [&](int w) -> uint64_t { return rows[(size_t)w * RS + i]; };

    // 2. row i, col i, features of network d
    Ego<WMAX, STREAM> U;
    U.i = i;
    if constexpr (STREAM) {
      for (int wd = tid; wd < W64; wd += NT) rowi_s[wd] = ROWI(wd);
      U.rowi_s = rowi_s;
      U.coli_s = coli_s;
      U.W = W64;
    } else {
      U.rl = lane < WMAX ? ROWI(lane) : 0ull;  // word `lane` of row i
    }
    Alt<R> A;
    A.in = A.ok = A.xij = A.xji = 0u;
    const int iw = i >> 6, ib = i & 63;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int word = NW * r + wv;
      const int j = word * 64 + lane;
      const bool in = j < n;
      A.j[r] = in ? j : i;
      const int xji = in ? (int)((rows[(size_t)iw * RS + A.j[r]] >> ib) & 1) : 0;
      const uint64_t bal = __ballot(xji);
      if (lane == 0) coli_s[word] = bal;
      const uint64_t rw = (!STREAM || word < W64) ? ROWI(word) : 0ull;
      const int xij = (int)((rw >> lane) & 1);
      A.in |= (uint32_t)in << r;
      A.xji |= (uint32_t)xji << r;
      A.xij |= (uint32_t)xij << r;
    }
    // Philox words of the lane's alternatives and their float32 Gumbel-bound
    // logarithms: they depend on the step and j only, so they fill the waits
    // of the loads above and of the barrier (kept here by an empty asm operand)
    uint32_t uj[R];
    float llb[R];
    philox_alternatives<R>(__builtin_amdgcn_readfirstlane((uint32_t)steps),
                           __builtin_amdgcn_readfirstlane((uint32_t)chain), A.j, key0, key1, uj);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      llb[r] = approx_loglog(uj[r]);
      asm volatile("" : "+v"(uj[r]), "+v"(llb[r]));
    }
    __syncthreads();
    if constexpr (!STREAM) U.cl = lane < WMAX ? coli_s[lane] : 0ull;  // word `lane` of column i
    U.outdeg_i = outdeg[i];
    U.indeg_i = indeg[i];
    U.zvi = 0.0;
    U.rdeg_i = 0;
    // features of network d: tp, so and cy as the single-network kernel
    // computes them, with the same tie walk for egos with few ties.
    const int deg_i = __builtin_amdgcn_readfirstlane(U.outdeg_i + U.indeg_i);
    const bool walk = STREAM ? deg_i < U.words() : deg_i * 3 < U.words() * 8;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      A.tp[r] = 0;
      A.so[r] = 0;
      A.cy[r] = 0;
    }
    auto bit_of_row_j = [&](int r, int h) -> int {  // x_{j h}
      const uint64_t x = rows[(size_t)(h >> 6) * RS + A.j[r]];
      const uint32_t half = (h & 32) ? (uint32_t)(x >> 32) : (uint32_t)x;
      return (int)((half >> (h & 31)) & 1u);
    };
    if ((flags & F_TP) || walk) {
#pragma unroll
      for (int wd = 0; wd < U.words(); ++wd) {
        uint64_t m = uniform64(U.rowi(wd));  // scalar loop over i's out-ties
        while (m) {
          const int h = wd * 64 + __ffsll((long long)m) - 1;
          m &= m - 1;
          if (flags & F_TP) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int word = NW * r + wv;
              const uint64_t x =
                  (!STREAM || word < W64) ? rows[(size_t)(word) * RS + h] : 0ull;
              A.tp[r] += (int)((x >> lane) & 1);
            }
          }
          if (walk) {
#pragma unroll
            for (int r = 0; r < R; ++r) A.so[r] += bit_of_row_j(r, h);
          }
        }
      }
    }
    if (walk) {
#pragma unroll
      for (int wd = 0; wd < U.words(); ++wd) {
        uint64_t m = uniform64(U.coli(wd));  // scalar loop over i's in-ties
        while (m) {
          const int h = wd * 64 + __ffsll((long long)m) - 1;
          m &= m - 1;
#pragma unroll
          for (int r = 0; r < R; ++r) A.cy[r] += bit_of_row_j(r, h);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int so = 0, cy = 0;
#pragma unroll
        for (int wd = 0; wd < U.words(); ++wd) {
          const uint64_t x = rows[(size_t)(wd) * RS + A.j[r]];
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
    // permitted toggles of network d, and j = i (:481-483)
    const uint64_t* pmrow = Q.pm + ((size_t)d * n + i) * W64;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int word = NW * r + wv;
      bool ok;
      if (!((A.in >> r) & 1)) ok = false;
      else if (A.j[r] == i) ok = true;
      else ok = (pmrow[word] >> lane) & 1;
      A.ok |= (uint32_t)ok << r;
    }

    // 3. f_j over the effects of network d, in list order
    double f[R];
#pragma unroll
    for (int r = 0; r < R; ++r) f[r] = 0.0;
#ifdef SAOM_FIXED_K
    constexpr int KC = SAOM_CACHE_D ? SAOM_FIXED_K : 1;
    double dc[KC][R];
    static_for<0, SAOM_FIXED_K>([&](auto EC) {
      constexpr int e = EC;
      if (SAOM_CACHE_D) {
#pragma unroll
        for (int r = 0; r < R; ++r) dc[SAOM_CACHE_D ? e : 0][r] = 0.0;
      }
      if (eff_netof(Q, e) != d) return;
      // unmasked: f of the ego's own slot is set to 0 below, slots past n are
      // not choosable, and the score sums weight the cached values (the own
      // slot with weight 0)
      double dv[R];
      multi_eff<R, WMAX, STREAM>(Q, e, A, U, S, rows_all, netw, dv, !SAOM_CACHE_D);
      const double te = load_uniform(th + e);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double prod = te * dv[r];
        f[r] = f[r] + prod;
        if (SAOM_CACHE_D) dc[SAOM_CACHE_D ? e : 0][r] = dv[r];
      }
    });
#else
    for (int e = 0; e < k; ++e) {
      if (eff_netof(Q, e) != d) continue;
      double dv[R];
      multi_eff<R, WMAX, STREAM>(Q, e, A, U, S, rows_all, netw, dv);
      const double te = load_uniform(th + e);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double prod = te * dv[r];
        f[r] = f[r] + prod;
      }
    }
#endif
    // 4. Gumbel-max choice with the float32 bound, as the single-network
    // kernel: the largest bound as a float32 DPP maximum, the candidates
    // (bounds within the slack of it, widened by its rounding) by ballot, and
    // the exact float64 noise only when there are several
    double gb = -1e30, shift = 0.0;
    int jb = i;
    bool exact_choice;
    {
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
      int near = 0, jm = i;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint64_t b = __ballot(((A.ok >> r) & 1) && ga[r] >= thr);
        if (b) {
          near += __popcll(b);
          jm = 64 * (NW * r + wv) + __ffsll((long long)b) - 1;
        }
      }
      bool unique = near == 1;
      if (NW > 1) {
        if (lane == 0) {
          red[wv * 6 + 3] = M;
          red[wv * 6 + 4] = (double)near;
          red[wv * 6 + 5] = (double)jm;
        }
        __syncthreads();
        int qm = 0;
        M = red[3];
        for (int q = 1; q < NW; ++q)
          if (red[q * 6 + 3] > M) {
            M = red[q * 6 + 3];
            qm = q;
          }
        thr = choice_threshold(M);
        unique = red[qm * 6 + 4] == 1.0;
        for (int q = 0; q < NW; ++q) {
          const double Mq = red[q * 6 + 3];
          if (q != qm && Mq + (fabs(Mq) + 1.0) * 0x1p-23 >= thr) unique = false;
        }
        jm = (int)red[qm * 6 + 5];
      }
      // the bound's maximum is within 3.15 of max_j f_j: the softmax shift
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
      if (NW > 1) {
        if (lane == 0) {
          red[wv * 6 + 0] = gb;
          red[wv * 6 + 2] = (double)jb;
        }
        __syncthreads();
        gb = red[0];
        jb = (int)red[2];
        for (int q = 1; q < NW; ++q) arg_combine(gb, jb, red[q * 6 + 0], (int)red[q * 6 + 2]);
      }
    }
    // uniform over the workgroup: scalar registers
    const int jstar = __builtin_amdgcn_readfirstlane(jb);

    // softmax-weighted expected change statistics of d's effects (:499-502)
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
      // all k + 1 sums in one transposing reduction (zeros for the other
      // networks' effects, which the score update skips)
      constexpr int KP = SAOM_FIXED_K + 1 <= 2 ? 2 : SAOM_FIXED_K + 1 <= 4 ? 4
                       : SAOM_FIXED_K + 1 <= 8 ? 8 : SAOM_FIXED_K + 1 <= 16 ? 16 : 32;
      double x[KP];
#pragma unroll
      for (int e = 0; e < KP; ++e) x[e] = 0.0;
      // the ego's own slot enters the sums with weight 0 (cached values are
      // unmasked; slots past n have weight 0 already)
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
      if (eff_netof(Q, e) != d) return;
      double dv[R];
      multi_eff<R, WMAX, STREAM>(Q, e, A, U, S, rows_all, netw, dv);
      double T = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        T += wgt[r] * dv[r];
        if (((A.in >> r) & 1) && A.j[r] == jstar) dstar[e] = dv[r];
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
    auto toggle = [&]() {  // thread 0: the step's tie change in network d
      sd[d] += 1;
      if (jstar != i) {
        const int aij = (int)((ROWI(jstar >> 6) >> (jstar & 63)) & 1);
        const int dd = 1 - 2 * aij;
        rows[(size_t)(jstar >> 6) * RS + i] ^= (1ull << (jstar & 63));
        indeg[jstar] += dd;
        outdeg[i] += dd;
        if (Q.symf[d]) {
          rows[(size_t)(i >> 6) * RS + jstar] ^= (1ull << (i & 63));
          indeg[i] += dd;
          outdeg[jstar] += dd;
        }
      }
    };
    // With the change statistics cached in registers nothing after the choice
    // reads the networks, so the toggle goes before the barrier that
    // publishes the score sums and the ministep needs no closing barrier (the
    // next writes of the sums and of dstar follow the next ministep's first
    // barriers; its reads of the networks follow this one).
#if defined(SAOM_FIXED_K) && SAOM_CACHE_D
    constexpr bool kEarlyToggle = true;
#else
    constexpr bool kEarlyToggle = false;
#endif
    if (kEarlyToggle && tid == 0) toggle();
    __syncthreads();
    if (tid < k && Q.netof[tid] == d) {  // (netof of a lane's own effect)
      double St = 0.0, Tt = 0.0;
      for (int q = 0; q < NW; ++q) {
        St += tsl[q * (k + 1) + k];
        Tt += tsl[q * (k + 1) + tid];
      }
      // no change statistics when the ego keeps its ties (dstar holds the
      // unmasked values of its own slot in the cached form)
      sc[tid] += (jstar == i ? 0.0 : dstar[tid]) - Tt / St;
    }
    if (!kEarlyToggle && tid == 0) toggle();
    steps++;
    if (!kEarlyToggle) __syncthreads();
    if (steps >= kMaxSteps) break;
  }

  // ---- period end (:503-510)
  __syncthreads();
  for (int q = tid; q < D * n; q += NT) {
    const int d = q / n, h = q - d * n;
    const uint64_t* rd = rows_all + (size_t)d * netw;
    uint64_t* out = Q.end_rows + (((size_t)d * Q.M + chain) * n + h) * W64;
    const uint64_t* x0 = Q.x0 + ((size_t)d * n + h) * W64;
    const uint64_t* kd = Q.kd ? Q.kd + ((size_t)d * n + h) * W64 : nullptr;
    int c = 0;
    for (int wd = 0; wd < W64; ++wd) {
      const uint64_t v = rd[(size_t)wd * RS + h];
      out[wd] = v;
      if (kd) c += popc64((v ^ x0[wd]) & kd[wd]);
    }
    if (c) atomicAdd(&dc[d], c);
  }
  __syncthreads();
  if (tid == 0) {
    Q.steps[chain] = (int32_t)steps;
    Q.ntim[chain] = t;
    for (int d = 0; d < D; ++d) {
      Q.rscores[(size_t)chain * D + d] = (double)sd[d] / lam[d] - (double)nact;
      if (Q.dist_stat) Q.dist_stat[(size_t)d * Q.M + chain] = dc[d];
    }
    for (int e = 0; e < k; ++e) Q.scores[(size_t)chain * k + e] = sc[e];
  }
}

#ifndef SAOM_RTC
template <int NW, int R, int WMAX>
__global__ void __launch_bounds__(64 * NW)
    saom_multi_kernel(const MultiParams* __restrict__ Qp) {
  multi_body<NW, R, WMAX>(Qp);
}

template <int NW, int WMAX>
hipError_t launch_multi_shape(const MultiParams* Pd, int grid, size_t lds, hipStream_t st) {
  auto kern = saom_multi_kernel<NW, WMAX / NW, WMAX>;
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * NW), lds, st, Pd);
  return hipGetLastError();
}

}  // namespace

#ifndef SAOM_WMAX
#error "compile with -DSAOM_WMAX=<1|2|4|8|16|32|64|128>"
#endif
#define SAOM_MCAT2(a, b) a##b
#define SAOM_MCAT(a, b) SAOM_MCAT2(a, b)

// Block shapes: 1, 2 or 4 waves per chain (at most WMAX) for word bounds up
// to 16; 16 waves for the streamed word bounds.
int SAOM_MCAT(launch_multi_w, SAOM_WMAX)(const MultiParams* Pd, int grid, size_t lds, int nw,
                                         void* stream) {
  hipStream_t st = (hipStream_t)stream;
  constexpr int WM = SAOM_WMAX;
  hipError_t e = hipErrorInvalidValue;
  if constexpr (WM <= kLdsWordBound) {
    if (nw == 1) e = launch_multi_shape<1, WM>(Pd, grid, lds, st);
    if constexpr (WM >= 2) {
      if (nw == 2) e = launch_multi_shape<2, WM>(Pd, grid, lds, st);
    }
    if constexpr (WM >= 4) {
      if (nw == 4) e = launch_multi_shape<4, WM>(Pd, grid, lds, st);
    }
  } else {
    if (nw == 16) e = launch_multi_shape<16, WM>(Pd, grid, lds, st);
  }
  return (int)e;
}

#if SAOM_WMAX == 1
size_t multi_lds_bytes(const MultiParams& P, int nw, int wmax) {
  return multi_layout(P.base.n, P.base.k, P.D, wmax, P.base.RS, nw, P.base.flags,
                      wmax > kLdsWordBound || P.rows_global)
      .total;
}
#endif

}  // namespace saom
#else   // SAOM_RTC: one specialised kernel, shape and effect list from macros
}  // namespace

extern "C" __global__ void __launch_bounds__(64 * SAOM_NW)
    saom_multi_rtc(const MultiParams* __restrict__ Qp) {
  multi_body<SAOM_NW, SAOM_WMAX / SAOM_NW, SAOM_WMAX>(Qp);
}

}  // namespace saom
#endif
