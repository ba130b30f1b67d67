// saom_stats.hip -- statistics of packed networks and behaviour vectors (gfx950).
//
// net_stats_kernel restates cusna_network_stats (cran/cusna
// src/libcusna_saom_cpu.cpp:606-689) on u64 bit rows, one workgroup per
// network. Integer-valued statistics (density, recip, transTrip, cycle3,
// transTies, transRecTrip, inPop, outAct, reciAct, transTriads) are exact
// int64 counts from popcounts and out-neighbour walks, which equal the
// reference's double sums of integers. Statistics that sum non-integers have
// their terms computed by the workgroup in windows and added by one lane in the
// reference's order (actors ascending; ties row-major), so they round exactly
// as the reference does.
//
// beh_stats_kernel restates cusna_behavior_stats (:767-817).
#include <hip/hip_runtime.h>

#include "saom_device.h"
#include "saom_kernels.h"

namespace saom {
namespace {

constexpr int kStatThreads = 256;
// term window of the ordered sums, in doubles (LDS rows / rows in scratch)
constexpr int kTermCapLds = 512, kTermCapScratch = 4096;

__device__ __forceinline__ int bit_of(const uint64_t* row, int j) {
  return (int)((row[j >> 6] >> (j & 63)) & 1);
}

// Degree statistic term of outPop/inAct families (:608-616).
__device__ __forceinline__ double degree_term(int t, double pm, double cur, double start) {
  if (t == 8 || t == 9) {
    if (pm >= 1.0) return t == 9 ? sqrt(cur) : cur;
    return pm == 0.0 ? start : cur + start;
  }
  if (pm >= 1.0) return t == 15 ? sqrt(cur) : cur;
  const double s = pm == 0.0 ? start : cur + start;
  return t == 15 ? sqrt(s) : s;
}

__device__ __forceinline__ long long block_sum_ll(long long v, long long* slots) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) slots[wv] = v;
  __syncthreads();
  long long s = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += slots[q];
  __syncthreads();
  return s;
}

// Exclusive prefix sum of v over the workgroup; total receives the sum.
__device__ __forceinline__ int block_exscan(int v, int* slots, int& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) slots[wv] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
    const int sq = slots[q];
    before += q < wv ? sq : 0;
    all += sq;
  }
  __syncthreads();
  total = all;
  return before + x - v;
}

// s += buf[0], s += buf[1], ... in order (one lane); eight loads in flight.
__device__ __forceinline__ void ordered_add(double& s, const double* buf, int len) {
  int x = 0;
  for (; x + 8 <= len; x += 8) {
    double c[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) c[m] = buf[x + m];
#pragma unroll
    for (int m = 0; m < 8; ++m) s += c[m];
  }
  for (; x < len; ++x) s += buf[x];
}

enum : int {
  I_E = 0, I_RECIP, I_TT, I_C3, I_TTIES, I_TRT, I_INPOP, I_OUTACT, I_RECIACT, I_N
};

// GROWS: the masked rows (and columns) live in device scratch instead of LDS,
// for networks above the LDS capacity; the arithmetic is the same.
template <bool GROWS>
__global__ void __launch_bounds__(kStatThreads) net_stats_kernel(const StatParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = P.n, W64 = P.W64, k = P.k;
  const int tid = threadIdx.x, NT = blockDim.x;
  const int net = blockIdx.x;
  bool need_cols = false;
  for (int e = 0; e < k; ++e)
    if (P.tid[e] == 27 || P.tid[e] == 28 || P.tid[e] == 30) need_cols = true;
  size_t o = 0;
  uint64_t* rows;
  uint64_t* cols;
  if constexpr (GROWS) {
    rows = P.scratch + (size_t)net * n * W64 * (need_cols ? 2 : 1);
    cols = rows + (size_t)n * W64;
  } else {
    rows = (uint64_t*)(smem + o);
    o += (size_t)n * W64 * 8;
    cols = (uint64_t*)(smem + o);
    if (need_cols) o += (size_t)n * W64 * 8;
  }
  int* indeg = (int*)(smem + o);
  o += (size_t)n * 4;
  int* outdeg = (int*)(smem + o);
  o += (size_t)n * 4;
  o = (o + 7) & ~(size_t)7;
  long long* slots = (long long*)(smem + o);
  double* tbuf = (double*)(smem + o + 64 * 8);

  const uint64_t* src = P.rows + (size_t)net * n * W64;
  for (int q = tid; q < n * W64; q += NT) {
    uint64_t v = src[q];
    if (P.keep) v &= P.keep[q];
    if (P.force) v |= P.force[q];
    rows[q] = v;
    if (need_cols) cols[q] = 0ull;
  }
  for (int q = tid; q < n; q += NT) indeg[q] = 0;
  __syncthreads();
  for (int h = tid; h < n; h += NT) {
    int od = 0;
    for (int wd = 0; wd < W64; ++wd) {
      uint64_t m = rows[(size_t)h * W64 + wd];
      od += __popcll(m);
      while (m) {
        const int j = wd * 64 + __ffsll((long long)m) - 1;
        m &= m - 1;
        atomicAdd(&indeg[j], 1);
        if (need_cols) atomicOr(&cols[(size_t)j * W64 + (h >> 6)], 1ull << (h & 63));
      }
    }
    outdeg[h] = od;
  }
  __syncthreads();

  // integer statistics, actors split over the workgroup
  long long acc[I_N];
  for (int q = 0; q < I_N; ++q) acc[q] = 0;
  for (int i = tid; i < n; i += NT) {
    const uint64_t* ri = rows + (size_t)i * W64;
    const long long od = outdeg[i], idg = indeg[i];
    acc[I_E] += od;
    acc[I_INPOP] += idg * idg;
    acc[I_OUTACT] += od * od;
    long long rd = 0;
    for (int wd = 0; wd < W64; ++wd) {
      uint64_t m = ri[wd];
      while (m) {
        const int h = wd * 64 + __ffsll((long long)m) - 1;
        m &= m - 1;
        const uint64_t* rh = rows + (size_t)h * W64;
        const int hi = bit_of(rh, i);  // a_hi
        rd += hi;
        int tt = 0, c3 = 0, trt = 0;
        for (int w2 = 0; w2 < W64; ++w2) tt += __popcll(ri[w2] & rh[w2]);
        // cycle3: j in out(h) with a_ji; transRecTrip: j in out(h) & out(i) with a_ji
        for (int w2 = 0; w2 < W64; ++w2) {
          uint64_t m2 = rh[w2];
          while (m2) {
            const int j = w2 * 64 + __ffsll((long long)m2) - 1;
            m2 &= m2 - 1;
            const int ji = bit_of(rows + (size_t)j * W64, i);
            c3 += ji;
            if (ji && bit_of(ri, j)) trt++;
          }
        }
        acc[I_TT] += tt;
        acc[I_C3] += c3;
        acc[I_TRT] += trt;
      }
    }
    acc[I_RECIP] += rd;
    acc[I_RECIACT] += od * rd;
    // transTies: j in out(i) reachable by a two-path from i
    int tties = 0;
    for (int wd = 0; wd < W64; ++wd) {
      uint64_t reach = 0;
      uint64_t m = ri[wd];
      (void)m;
      for (int w2 = 0; w2 < W64; ++w2) {
        uint64_t mm = ri[w2];
        while (mm) {
          const int h = w2 * 64 + __ffsll((long long)mm) - 1;
          mm &= mm - 1;
          reach |= rows[(size_t)h * W64 + wd];
        }
      }
      tties += __popcll(ri[wd] & reach);
    }
    acc[I_TTIES] += tties;
  }
  long long tot[I_N];
  for (int q = 0; q < I_N; ++q) tot[q] = block_sum_ll(acc[q], slots);

  // Per effect. A statistic that sums non-integers keeps the reference's
  // order (actors ascending; ties row-major): the workgroup writes a window of
  // terms in that order into tbuf, then lane 0 adds them to its running sum,
  // so the sum rounds exactly as the reference's loop. A term the reference
  // skips is written as +0.0: the running sum starts at +0.0 and can never
  // become -0.0, so adding +0.0 leaves it bit-identical.
  constexpr int kCap = GROWS ? kTermCapScratch : kTermCapLds;
  int* islots = (int*)slots;
  double* out = P.out + (size_t)net * k;
  for (int e = 0; e < k; ++e) {
    const int t = P.tid[e];
    const double pm = P.p1[e], q2 = P.p2[e];
    const double* wr = P.wtab + (size_t)e * (n + 1);
    const bool dyn = P.dyn[e] != 0;
    const double* vrow = P.vmat + (size_t)e * n;
    const uint8_t* mrow = P.vmiss ? P.vmiss + (size_t)e * n : nullptr;
    auto V = [&](int i) -> double { return dyn ? (P.vb ? P.vb[i] : 0.0) : vrow[i]; };
    auto MISS = [&](int i) -> bool { return mrow && mrow[i]; };
    double s = 0.0;  // lane 0's running sum
    auto add_window = [&](int len) {
      __syncthreads();
      if (tid == 0) ordered_add(s, tbuf, len);
      __syncthreads();
    };
    // term of actor i in the per-actor sums (:663-671)
    auto actor_term = [&](int i) -> double {
      switch (t) {
        case 7:
          return pow_three_halves_libm((double)indeg[i]);
        case 17:
          return pow_three_halves_libm((double)outdeg[i]);
        case 8:
        case 9:
          return (double)indeg[i] *
                 degree_term(t, pm, (double)outdeg[i], P.od0 ? (double)P.od0[i] : 0.0);
        case 14:
        case 15:
          return (double)outdeg[i] *
                 degree_term(t, pm, (double)indeg[i], P.id0 ? (double)P.id0[i] : 0.0);
        case 18:
          return fmin((double)outdeg[i], pm);
        case 10:
          return MISS(i) ? 0.0 : (double)outdeg[i] * V(i);
        default:  // 11
          return MISS(i) ? 0.0 : (double)indeg[i] * V(i);
      }
    };
    // pairs (i, j) of word wd that the tie sums visit: the ties of i, or for
    // 29/30 every pair (wr[0] != 0) or the pairs with a two-path i->h->j (29)
    // or a shared out-neighbour h (30)
    const bool all_pairs = (t == 29 || t == 30);
    auto pair_mask = [&](const uint64_t* ri, int wd) -> uint64_t {
      if (!all_pairs) return ri[wd];
      if (wr[0] != 0.0) {
        uint64_t m = ~0ull;
        if (wd == W64 - 1 && (n & 63)) m &= (1ull << (n & 63)) - 1;
        return m;
      }
      const uint64_t* base = t == 29 ? rows : cols;
      uint64_t m = 0;
      for (int w2 = 0; w2 < W64; ++w2) {
        uint64_t mm = ri[w2];
        while (mm) {
          const int h = w2 * 64 + __ffsll((long long)mm) - 1;
          mm &= mm - 1;
          m |= base[(size_t)h * W64 + wd];
        }
      }
      return m;
    };
    // term of pair (i, j) in the tie sums (:672-683)
    auto tie_term = [&](int i, const uint64_t* ri, int j) -> double {
      const uint64_t* rj = rows + (size_t)j * W64;
      switch (t) {
        case 12:
          return (MISS(i) || MISS(j)) ? 0.0 : 1.0 - fabs(V(i) - V(j)) * pm - q2;
        case 13:
          return (MISS(i) || MISS(j)) ? 0.0 : (fabs(V(i) - V(j)) < 1e-9 ? 1.0 : 0.0);
        case 20: {
          if (MISS(i) || MISS(j)) return 0.5;
          const double dd = V(i) - V(j);
          return (dd > 1e-9 ? 1.0 : 0.0) + 0.5 * (fabs(dd) <= 1e-9 ? 1.0 : 0.0);
        }
        case 21:
          return (MISS(i) || MISS(j)) ? 0.0 : V(i) * V(j);
        case 25:
          return (MISS(i) || MISS(j)) ? 0.0 : V(i) + V(j);
        case 31:
          return (!MISS(j) && V(j) >= pm - 1e-12) ? 1.0 : 0.0;
        case 24:
        case 29: {
          if (t == 29 && i == j) return 0.0;
          int c = 0;  // two-paths i->h->j
          for (int w2 = 0; w2 < W64; ++w2) {
            uint64_t mm = ri[w2];
            while (mm) {
              const int h = w2 * 64 + __ffsll((long long)mm) - 1;
              mm &= mm - 1;
              c += bit_of(rows + (size_t)h * W64, j);
            }
          }
          return wr[c];
        }
        case 26:
        case 30: {
          if (t == 30 && i == j) return 0.0;
          int c = 0;
          for (int w2 = 0; w2 < W64; ++w2) c += __popcll(ri[w2] & rj[w2]);
          return wr[c];
        }
        case 27: {
          int c = 0;
          for (int w2 = 0; w2 < W64; ++w2)
            c += __popcll(cols[(size_t)i * W64 + w2] & cols[(size_t)j * W64 + w2]);
          return wr[c];
        }
        default: {  // 28
          int c = 0;
          for (int w2 = 0; w2 < W64; ++w2) c += __popcll(cols[(size_t)i * W64 + w2] & rj[w2]);
          return wr[c];
        }
      }
    };
    switch (t) {
      case 0:
        s = (double)tot[I_E];
        if (P.symflag) s *= 0.5;
        break;
      case 1:
        s = (double)tot[I_RECIP];
        break;
      case 2:
        s = (double)tot[I_TT];
        break;
      case 3:
        s = (double)tot[I_C3] / 3.0;
        break;
      case 23:
        s = (double)tot[I_TT] / 6.0;
        break;
      case 5:
        s = (double)tot[I_TTIES];
        break;
      case 6:
        s = (double)tot[I_TRT];
        break;
      case 4:
        s = (double)tot[I_INPOP];
        break;
      case 16:
        s = (double)tot[I_OUTACT];
        break;
      case 19:
        s = (double)tot[I_RECIACT];
        break;
      case 7: case 17: case 8: case 9: case 14: case 15: case 18: case 10: case 11:
        // windows of NT actors (kCap >= NT)
        for (int b = 0; b < n; b += NT) {
          if (b + tid < n) tbuf[tid] = actor_term(b + tid);
          add_window(min(NT, n - b));
        }
        break;
      case 12: case 13: case 20: case 21: case 24: case 25: case 26: case 27: case 28:
      case 29: case 30: case 31:
        // windows of NT actors; a window's pairs go through tbuf kCap at a time
        for (int b = 0; b < n; b += NT) {
          const int i = b + tid;
          const bool act = i < n;
          const uint64_t* ri = rows + (size_t)(act ? i : 0) * W64;
          int cnt = 0;
          if (act) {
            if (!all_pairs) cnt = outdeg[i];
            else
              for (int wd = 0; wd < W64; ++wd) cnt += __popcll(pair_mask(ri, wd));
          }
          int wtot;
          const int off = block_exscan(cnt, islots, wtot);
          for (int lo = 0; lo < wtot; lo += kCap) {
            const int hi = min(lo + kCap, wtot);
            if (act && off < hi && off + cnt > lo) {
              int kq = off;  // window index of the next pair of i
              for (int wd = 0; wd < W64 && kq < hi; ++wd) {
                uint64_t m = pair_mask(ri, wd);
                const int c = __popcll(m);
                if (kq + c <= lo) {
                  kq += c;
                  continue;
                }
                while (m && kq < hi) {
                  const int j = wd * 64 + __ffsll((long long)m) - 1;
                  m &= m - 1;
                  if (kq >= lo) tbuf[kq - lo] = tie_term(i, ri, j);
                  ++kq;
                }
              }
            }
            add_window(hi - lo);
          }
        }
        break;
      default:
        break;  // X (22) and interactions (32) have no statistic (:684)
    }
    if (tid == 0) out[e] = s;
  }
}

// Behaviour statistics (:767-817), one workgroup per behaviour vector.
template <bool GMEM>
__global__ void __launch_bounds__(kStatThreads) beh_stats_kernel(const BehStatParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = P.n, W64 = P.W64;
  const int tid = threadIdx.x, NT = blockDim.x;
  const int c = blockIdx.x;
  double* u;
  if constexpr (GMEM) u = P.scratch + (size_t)c * 5 * n;
  else u = (double*)smem;
  double* term = u + n;
  double* okd = term + n;
  double* od = okd + n;
  double* nb = od + n;
  const int32_t* zc = P.z + (size_t)c * n;
  for (int i = tid; i < n; i += NT) {
    okd[i] = (P.bkdist && !P.bkdist[i]) ? 0.0 : 1.0;
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    u[i] = okd[i] != 0.0 ? ((double)zc[i] - P.zbar) : 0.0;
    double d = 0.0, b = 0.0;
    const uint64_t* xr = P.xnet + (size_t)i * W64;
    for (int wd = 0; wd < W64; ++wd) {
      uint64_t m = xr[wd];
      while (m) {
        const int j = wd * 64 + __ffsll((long long)m) - 1;
        m &= m - 1;
        d += 1.0;
        b += okd[j];
      }
    }
    od[i] = d;
    nb[i] = b;
  }
  __syncthreads();
  for (int e = 0; e < P.kb; ++e) {
    const int t = P.bty[e];
    if (t == 2 || t == 3 || t == 4) {
      for (int i = tid; i < n; i += NT) {
        const uint64_t* xr = P.xnet + (size_t)i * W64;
        double x = 0.0;
        bool skip;
        if (t == 2) {
          skip = od[i] <= 0.0;
          if (!skip) {
            for (int wd = 0; wd < W64; ++wd) {
              uint64_t m = xr[wd];
              while (m) {
                const int j = wd * 64 + __ffsll((long long)m) - 1;
                m &= m - 1;
                x += u[j];
              }
            }
            x = u[i] * (x / od[i]);
          }
        } else {
          skip = okd[i] == 0.0;
          if (!skip) {
            for (int wd = 0; wd < W64; ++wd) {
              uint64_t m = xr[wd];
              while (m) {
                const int j = wd * 64 + __ffsll((long long)m) - 1;
                m &= m - 1;
                if (okd[j] != 0.0) x += 1.0 - fabs(u[i] - u[j]) / P.brange;
              }
            }
            x -= nb[i] * P.bsim;
            if (t == 3) x = nb[i] > 0.0 ? x / nb[i] : 0.0;
          }
        }
        // A skipped actor adds nothing; adding +0.0 instead is bit-identical
        // because the running sum starts at +0.0 and can never become -0.0.
        term[i] = skip ? 0.0 : x;
      }
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      if (t == 0) {
        ordered_add(s, u, n);
      } else if (t == 1) {
        for (int i = 0; i < n; ++i) s += u[i] * u[i];
      } else if (t == 2 || t == 3 || t == 4) {
        ordered_add(s, term, n);
      }
      P.out[(size_t)c * P.kb + e] = s;
    }
    __syncthreads();
  }
}

__global__ void expand_dense_kernel(const uint64_t* rows, int n, int W64, long long total,
                                    uint8_t* dense) {
  const long long nn = (long long)n * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const long long c = q / nn;
    const long long rem = q - c * nn;
    const int i = (int)(rem / n), j = (int)(rem - (long long)i * n);
    dense[q] = (uint8_t)((rows[(c * n + i) * W64 + (j >> 6)] >> (j & 63)) & 1);
  }
}

__global__ void philox_kernel(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    philox10(c0, c1, c2, c3, key[0], key[1]);
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
  }
}

// max |approx_loglog(x) - log(-log u(x))| over x = x0 + q*stride, q < count;
// the float64 reference is the same device log() the period kernel uses.
__global__ void gumbel_check_kernel(uint32_t x0, uint32_t stride, uint32_t count,
                                    unsigned long long* max_bits) {
  double m = 0.0;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count;
       q += gridDim.x * blockDim.x) {
    const uint32_t x = x0 + q * stride;
    const double exact = log(-log(unit_open(x)));
    const double err = fabs((double)approx_loglog(x) - exact);
    m = fmax(m, err);
  }
  atomicMax(max_bits, (unsigned long long)__double_as_longlong(m));
}

// out[q] = exp_table(x[q]) with the period kernel's table.
__global__ void exp_weight_kernel(const double* x, int count, double* out) {
  __shared__ double tab[kExpTab];
  if (threadIdx.x < kExpTab) tab[threadIdx.x] = exp2((double)threadIdx.x / (double)kExpTab);
  __syncthreads();
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x)
    out[q] = exp_table(x[q], tab);
}

}  // namespace

int launch_exp_weight(const double* x, int count, double* out, void* stream) {
  hipLaunchKernelGGL(exp_weight_kernel, dim3(256), dim3(256), 0, (hipStream_t)stream, x, count,
                     out);
  return (int)hipGetLastError();
}

int launch_gumbel_check(uint32_t x0, uint32_t stride, uint32_t count,
                        unsigned long long* max_bits, void* stream) {
  hipLaunchKernelGGL(gumbel_check_kernel, dim3(1024), dim3(256), 0, (hipStream_t)stream, x0,
                     stride, count, max_bits);
  return (int)hipGetLastError();
}

bool stats_need_cols(const int* t, int k) {
  for (int e = 0; e < k; ++e)
    if (t[e] == 27 || t[e] == 28 || t[e] == 30) return true;
  return false;
}

int launch_net_stats(const StatParams& P, void* stream) {
  if (P.M < 1) return 0;
  const bool need_cols = stats_need_cols(P.tid, P.k);
  const bool grows = net_stats_scratch_words(P.n, P.W64, need_cols) > 0;
  if (grows && !P.scratch) return (int)hipErrorInvalidValue;
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)net_stats_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
      160 * 1024);
  if (attr != hipSuccess) return (int)attr;
  static const hipError_t attr_g = hipFuncSetAttribute(
      (const void*)net_stats_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
      160 * 1024);
  if (attr_g != hipSuccess) return (int)attr_g;
  if (grows) {
    const size_t lds = (size_t)P.n * 8 + 8 + 64 * 8 + (size_t)kTermCapScratch * 8;
    hipLaunchKernelGGL(net_stats_kernel<true>, dim3(P.M), dim3(kStatThreads), lds,
                       (hipStream_t)stream, P);
  } else {
    const size_t lds = net_stats_lds(P.n, P.W64, need_cols);
    hipLaunchKernelGGL(net_stats_kernel<false>, dim3(P.M), dim3(kStatThreads), lds,
                       (hipStream_t)stream, P);
  }
  return (int)hipGetLastError();
}

size_t net_stats_lds(int n, int W64, bool need_cols) {
  return (size_t)n * W64 * 8 * (need_cols ? 2 : 1) + (size_t)n * 8 + 8 + 64 * 8 +
         (size_t)kTermCapLds * 8;
}

size_t net_stats_scratch_words(int n, int W64, bool need_cols) {
  if (net_stats_lds(n, W64, need_cols) <= 160 * 1024) return 0;
  return (size_t)n * W64 * (need_cols ? 2 : 1);
}

size_t beh_stats_scratch_doubles(int n) {
  return (size_t)n * 5 * 8 > 160 * 1024 ? (size_t)n * 5 : 0;
}

int launch_beh_stats(const BehStatParams& P, void* stream) {
  if (P.M < 1) return 0;
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)beh_stats_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
      160 * 1024);
  if (attr != hipSuccess) return (int)attr;
  if (beh_stats_scratch_doubles(P.n) > 0) {
    if (!P.scratch) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(beh_stats_kernel<true>, dim3(P.M), dim3(kStatThreads), 0,
                       (hipStream_t)stream, P);
  } else {
    hipLaunchKernelGGL(beh_stats_kernel<false>, dim3(P.M), dim3(kStatThreads),
                       (size_t)P.n * 5 * 8, (hipStream_t)stream, P);
  }
  return (int)hipGetLastError();
}

// ------------------------------------------------------------------ GoF

__device__ __forceinline__ uint64_t gof_word(const GofParams& P, const uint64_t* src, int i,
                                             int w) {
  uint64_t v = src[(size_t)i * P.W64 + w];
  if (P.keep) v &= P.keep[(size_t)i * P.W64 + w];
  if (P.force) v |= P.force[(size_t)i * P.W64 + w];
  // no bits beyond actor n - 1 (every column index below stays in range)
  if (w == P.W64 - 1 && (P.n & 63)) v &= (1ull << (P.n & 63)) - 1ull;
  return v;
}

// Degree distributions: one workgroup per network, degrees in LDS.
__global__ void __launch_bounds__(kStatThreads) gof_degree_kernel(const GofParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  int* deg = (int*)smem;
  const int n = P.n, net = blockIdx.x;
  const uint64_t* src = P.rows + (size_t)net * n * P.W64;
  for (int v = threadIdx.x; v < n; v += blockDim.x) deg[v] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int od = 0;
    for (int w = 0; w < P.W64; ++w) {
      uint64_t m = gof_word(P, src, i, w);
      od += __popcll(m);
      if (P.kind == 0)
        while (m) {
          const int j = w * 64 + __ffsll((long long)m) - 1;
          m &= m - 1;
          atomicAdd(&deg[j], 1);
        }
    }
    if (P.kind == 1) deg[i] = od;
  }
  __syncthreads();
  for (int L = 0; L < P.nlev; ++L) {
    const int lev = P.levls[L];
    int c = 0;
    for (int v = threadIdx.x; v < n; v += blockDim.x) c += deg[v] <= lev ? 1 : 0;
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ int part[kStatThreads / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += part[q];
      P.out[(size_t)net * P.nlev + L] = (double)t;
    }
    __syncthreads();
  }
}

// Geodesic distribution: one workgroup per network, masked rows in LDS; each
// thread runs a breadth-first expansion from its source actors with the
// reached set and frontier as bit rows in registers (word bound WMAX), adding
// the number of newly reached actors at each distance into LDS counters.
// reached_k = (reached_{k-1} | OR of the rows of the frontier) without the
// source, which equals the reference's boolean matrix powers with the diagonal
// cleared after each step.
template <int WMAX>
__global__ void __launch_bounds__(kStatThreads) gof_geodesic_kernel(const GofParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = P.n, W64 = P.W64, net = blockIdx.x, kmax = P.kmax;
  uint64_t* rows = (uint64_t*)smem;
  unsigned long long* delta = (unsigned long long*)(rows + (size_t)n * W64);
  const uint64_t* src = P.rows + (size_t)net * n * W64;
  for (int q = threadIdx.x; q < n * W64; q += blockDim.x)
    rows[q] = gof_word(P, src, q / W64, q % W64);
  for (int k = threadIdx.x; k <= kmax; k += blockDim.x) delta[k] = 0ull;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    uint64_t R[WMAX], F[WMAX];
    const int iw = i >> 6;
    const uint64_t ib = 1ull << (i & 63);
    int cnt = 0;
#pragma unroll
    for (int w = 0; w < WMAX; ++w) {
      R[w] = w < W64 ? rows[(size_t)i * W64 + w] : 0ull;
      F[w] = R[w];
      cnt += __popcll(w == iw ? (R[w] & ~ib) : R[w]);
    }
    if (kmax >= 1 && cnt) atomicAdd(&delta[1], (unsigned long long)cnt);
    for (int k = 2; k <= kmax; ++k) {
      uint64_t N[WMAX];
#pragma unroll
      for (int w = 0; w < WMAX; ++w) N[w] = 0ull;
#pragma unroll
      for (int fw = 0; fw < WMAX; ++fw) {
        uint64_t m = F[fw];
        while (m) {
          const int h = fw * 64 + __ffsll((long long)m) - 1;
          m &= m - 1;
#pragma unroll
          for (int w = 0; w < WMAX; ++w)
            if (w < W64) N[w] |= rows[(size_t)h * W64 + w];
        }
      }
      int grow = 0, any = 0;
#pragma unroll
      for (int w = 0; w < WMAX; ++w) {
        uint64_t rn = R[w] | N[w];
        if (w == iw) rn &= ~ib;
        const uint64_t old = w == iw ? (R[w] & ~ib) : R[w];
        grow += __popcll(rn) - __popcll(old);
        F[w] = rn & ~R[w];
        any |= F[w] != 0ull;
        R[w] = rn;
      }
      if (grow) atomicAdd(&delta[k], (unsigned long long)grow);
      if (!any) break;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // counts[k] = sum of deltas up to k; levels beyond kmax saturate
    double* out = P.out + (size_t)net * (P.nlev + 1);
    for (int k = 2; k <= kmax; ++k) delta[k] += delta[k - 1];
    for (int L = 0; L < P.nlev; ++L) {
      const int lev = P.levls[L] < kmax ? P.levls[L] : kmax;
      out[L] = lev >= 1 ? (double)delta[lev] : 0.0;
    }
    const unsigned long long reach = kmax >= 1 ? delta[kmax] : 0ull;
    out[P.nlev] = (double)((long long)n * (n - 1) - (long long)reach);
  }
}

// Geodesic distribution of networks whose rows exceed the LDS: one wave per
// (network, source actor). The reached set and the frontier are bit rows
// spread over the wave's lanes (word w in lane w % 64, slot w / 64, WPL slots:
// W64 <= 64 WPL); the frontier's actors are walked in order from scalar
// copies of its words, and the rows of four of them at a time are read from
// device memory (masked as gof_word), each lane ORing its words. The same
// reached/frontier recurrence as gof_geodesic_kernel; newly reached counts go
// to the network's counters (integer atomics: order-free).
template <int WPL>
__global__ void __launch_bounds__(256) gof_geodesic_stream_kernel(const GofParams P) {
  const int lane = threadIdx.x & 63;
  const long long gw = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int n = P.n, W64 = P.W64, kmax = P.kmax;
  if (gw >= (long long)P.M * n) return;  // whole waves
  const int net = (int)(gw / n), i = (int)(gw - (long long)net * n);
  const uint64_t* src = P.rows + (size_t)net * n * W64;
  unsigned long long* delta = P.counters + (size_t)net * (kmax + 1);
  const int iw = i >> 6;
  const uint64_t ib = 1ull << (i & 63);
  uint64_t R[WPL], F[WPL];
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < WPL; ++q) {
    const int w = lane + 64 * q;
    R[q] = w < W64 ? gof_word(P, src, i, w) : 0ull;
    F[q] = R[q];
    cnt += __popcll(w == iw ? (R[q] & ~ib) : R[q]);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if (kmax >= 1 && cnt && lane == 0) atomicAdd(&delta[1], (unsigned long long)cnt);
  for (int k = 2; k <= kmax; ++k) {
    uint64_t N[WPL];
#pragma unroll
    for (int q = 0; q < WPL; ++q) N[q] = 0ull;
#pragma unroll
    for (int q = 0; q < WPL; ++q) {
      uint64_t lanes = __ballot(F[q] != 0ull);  // lanes holding frontier words in slot q
      while (lanes) {
        const int ln = __ffsll((long long)lanes) - 1;
        lanes &= lanes - 1;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)F[q], ln);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(F[q] >> 32), ln);
        uint64_t word = ((uint64_t)hi << 32) | lo;
        const int base = (ln + 64 * q) * 64;
        while (word) {
          int h[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            h[t] = word ? base + __ffsll((long long)word) - 1 : -1;
            word &= word ? word - 1 : 0ull;
          }
          uint64_t got[4][WPL];
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q2 = 0; q2 < WPL; ++q2) {
              const int w2 = lane + 64 * q2;
              got[t][q2] = (h[t] >= 0 && w2 < W64) ? gof_word(P, src, h[t], w2) : 0ull;
            }
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q2 = 0; q2 < WPL; ++q2) N[q2] |= got[t][q2];
        }
      }
    }
    int grow = 0, any = 0;
#pragma unroll
    for (int q = 0; q < WPL; ++q) {
      const int w = lane + 64 * q;
      uint64_t rn = R[q] | N[q];
      if (w == iw) rn &= ~ib;
      const uint64_t old = w == iw ? (R[q] & ~ib) : R[q];
      grow += __popcll(rn) - __popcll(old);
      F[q] = rn & ~R[q];
      any |= F[q] != 0ull;
      R[q] = rn;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) grow += __shfl_xor(grow, o, 64);
    if (grow && lane == 0) atomicAdd(&delta[k], (unsigned long long)grow);
    if (__ballot(any) == 0ull) break;
  }
}

// Clears the reach counters of the device-memory geodesic form.
__global__ void gof_counters_clear_kernel(unsigned long long* c, size_t len) {
  const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < len) c[x] = 0ull;
}

// counts[k] = newly reached pairs up to distance k; outputs as the LDS form.
__global__ void gof_geodesic_finish_kernel(const GofParams P) {
  const int net = blockIdx.x * blockDim.x + threadIdx.x;
  if (net >= P.M) return;
  const int kmax = P.kmax;
  unsigned long long* delta = P.counters + (size_t)net * (kmax + 1);
  for (int k = 2; k <= kmax; ++k) delta[k] += delta[k - 1];
  double* out = P.out + (size_t)net * (P.nlev + 1);
  for (int L = 0; L < P.nlev; ++L) {
    const int lev = P.levls[L] < kmax ? P.levls[L] : kmax;
    out[L] = lev >= 1 ? (double)delta[lev] : 0.0;
  }
  const unsigned long long reach = kmax >= 1 ? delta[kmax] : 0ull;
  out[P.nlev] = (double)((long long)P.n * (P.n - 1) - (long long)reach);
}

size_t gof_geodesic_lds(int n, int W64, int kmax) {
  if (W64 > 16) return 0;
  const size_t b = (size_t)n * W64 * 8 + (size_t)(kmax + 1) * 8;
  return b <= 160 * 1024 ? b : 0;
}

int launch_gof(const GofParams& P, void* stream) {
  if (P.M < 1) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (P.kind == 0 || P.kind == 1) {
    hipLaunchKernelGGL(gof_degree_kernel, dim3(P.M), dim3(kStatThreads), (size_t)P.n * 4, st, P);
    return (int)hipGetLastError();
  }
  const size_t lds = gof_geodesic_lds(P.n, P.W64, P.kmax);
  if (lds == 0 || P.stream) {
    if (!P.counters || P.W64 > 128) return (int)hipErrorInvalidValue;
    const size_t len = (size_t)P.M * (P.kmax + 1);
    hipLaunchKernelGGL(gof_counters_clear_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256),
                       0, st, P.counters, len);
    const long long waves = (long long)P.M * P.n;
    const dim3 grid((unsigned)((waves + 3) / 4));
    if (P.W64 <= 64)
      hipLaunchKernelGGL(gof_geodesic_stream_kernel<1>, grid, dim3(256), 0, st, P);
    else
      hipLaunchKernelGGL(gof_geodesic_stream_kernel<2>, grid, dim3(256), 0, st, P);
    hipLaunchKernelGGL(gof_geodesic_finish_kernel, dim3((P.M + 63) / 64), dim3(64), 0, st, P);
    return (int)hipGetLastError();
  }
  int wmax = 1;
  while (wmax < P.W64) wmax <<= 1;
  auto go = [&](const void* k, void (*kern)(const GofParams)) -> int {
    const hipError_t attr =
        hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (attr != hipSuccess) return (int)attr;
    hipLaunchKernelGGL(kern, dim3(P.M), dim3(kStatThreads), lds, st, P);
    return (int)hipGetLastError();
  };
  switch (wmax) {
    case 1: return go((const void*)gof_geodesic_kernel<1>, gof_geodesic_kernel<1>);
    case 2: return go((const void*)gof_geodesic_kernel<2>, gof_geodesic_kernel<2>);
    case 4: return go((const void*)gof_geodesic_kernel<4>, gof_geodesic_kernel<4>);
    case 8: return go((const void*)gof_geodesic_kernel<8>, gof_geodesic_kernel<8>);
    case 16: return go((const void*)gof_geodesic_kernel<16>, gof_geodesic_kernel<16>);
    default: return (int)hipErrorInvalidValue;
  }
}

// ------------------------------------------------------- cross statistics

__device__ __forceinline__ uint64_t cross_row(const CrossParams& P, const uint64_t* src, int i,
                                              int w) {
  uint64_t v = src[(size_t)i * P.W64 + w];
  if (P.keep) v &= P.keep[(size_t)i * P.W64 + w];
  if (P.force) v |= P.force[(size_t)i * P.W64 + w];
  if (w == P.W64 - 1 && (P.n & 63)) v &= (1ull << (P.n & 63)) - 1ull;
  return v;
}

// One workgroup per network. Integer-valued statistics are exact int64 sums
// in any order; for `from` with the square root (parm >= 2) the workgroup
// computes the terms a window at a time and one lane adds them in the
// reference's order (ties row-major), so it rounds identically.
__global__ void __launch_bounds__(kStatThreads) cross_stats_kernel(const CrossParams P) {
  const int n = P.n, W64 = P.W64, net = blockIdx.x, t = P.t;
  const uint64_t* src = P.rows + (size_t)net * n * W64;
  double* out = P.out + (size_t)net * P.out_stride;
  if (t == 36 && P.parm >= 2.0) {
    __shared__ double tbuf[kTermCapLds];
    __shared__ int islots[kStatThreads / 64];
    const int tid = threadIdx.x, NT = blockDim.x;
    double acc = 0.0;  // lane 0's running sum
    // windows of NT actors; a window's ties go through tbuf kTermCapLds at a time
    for (int b = 0; b < n; b += NT) {
      const int i = b + tid;
      const bool act = i < n;
      int cnt = 0;
      if (act)
        for (int w = 0; w < W64; ++w) cnt += __popcll(cross_row(P, src, i, w));
      int wtot;
      const int off = block_exscan(cnt, islots, wtot);
      for (int lo = 0; lo < wtot; lo += kTermCapLds) {
        const int hi = min(lo + kTermCapLds, wtot);
        if (act && off < hi && off + cnt > lo) {
          int kq = off;  // window index of the next tie of i
          for (int w = 0; w < W64 && kq < hi; ++w) {
            uint64_t m = cross_row(P, src, i, w);
            const int c = __popcll(m);
            if (kq + c <= lo) {
              kq += c;
              continue;
            }
            while (m && kq < hi) {
              const int j = w * 64 + __ffsll((long long)m) - 1;
              m &= m - 1;
              if (kq >= lo) {
                int cc = 0;
                for (int q = 0; q < W64; ++q)
                  cc += __popcll(P.wp[(size_t)i * W64 + q] & P.wp[(size_t)j * W64 + q]);
                tbuf[kq - lo] = sqrt((double)cc);
              }
              ++kq;
            }
          }
        }
        __syncthreads();
        if (tid == 0) ordered_add(acc, tbuf, hi - lo);
        __syncthreads();
      }
    }
    if (tid == 0) *out = acc;
    return;
  }
  long long c = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    for (int w = 0; w < W64; ++w) {
      const uint64_t a = cross_row(P, src, i, w);
      const uint64_t wi = P.wp[(size_t)i * W64 + w], wt = P.wpT[(size_t)i * W64 + w];
      switch (t) {
        case 33: c += __popcll(a & wi); break;
        case 34: c += __popcll(a & wt); break;
        case 35: c += __popcll(a & wi & wt); break;
        case 36: {  // sum over ties i->j of |out_wp(i) & out_wp(j)|
          uint64_t m = a;
          while (m) {
            const int j = w * 64 + __ffsll((long long)m) - 1;
            m &= m - 1;
            for (int q = 0; q < W64; ++q)
              c += __popcll(P.wp[(size_t)i * W64 + q] & P.wp[(size_t)j * W64 + q]);
          }
          break;
        }
        default: {  // 37: sum over h in out_wp(i) of |out_a(h) & out_a(i)|
          uint64_t m = wi;
          while (m) {
            const int h = w * 64 + __ffsll((long long)m) - 1;
            m &= m - 1;
            for (int q = 0; q < W64; ++q)
              c += __popcll(cross_row(P, src, h, q) & cross_row(P, src, i, q));
          }
          break;
        }
      }
    }
  }
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
  __shared__ long long part[kStatThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long tot = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) tot += part[q];
    *out = (double)tot;
  }
}

int launch_cross_stats(const CrossParams& P, void* stream) {
  if (P.M < 1) return 0;
  if (P.t < 33 || P.t > 37) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(cross_stats_kernel, dim3(P.M), dim3(kStatThreads), 0, (hipStream_t)stream,
                     P);
  return (int)hipGetLastError();
}

int launch_expand_dense(const uint64_t* rows, int n, int W64, int M, uint8_t* dense,
                        void* stream) {
  const long long total = (long long)M * n * n;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(expand_dense_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, rows, n, W64, total, dense);
  return (int)hipGetLastError();
}

int launch_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, void* stream) {
  hipLaunchKernelGGL(philox_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, ctr, key, out);
  return (int)hipGetLastError();
}

}  // namespace saom
