// saom_kernels.h -- launch parameters of the gfx950 SAOM kernels (host + device).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "saom_common.h"

namespace saom {

// Feature bits: which per-alternative counts the effect list needs.
enum : int {
  F_TP = 1 << 0,     // two-paths i->h->j per alternative (tp)
  F_TPALL = 1 << 1,  // tp of every actor in LDS (transTies, gwdspFF)
  F_TPIN = 1 << 2,   // common in-neighbours of i and j (gwespBF)
  F_RR = 1 << 3,     // |row_i & col_i & row_j| (transRecTrip)
  F_TT = 1 << 4,     // transTies masked popcount
  F_SOALL = 1 << 5,  // shared out-ties of every actor with i in LDS (gwdspFB)
  F_Z = 1 << 6,      // behaviour values live in LDS
  F_RDEG = 1 << 7,   // reciprocated degree of the ego (reciAct)
  F_RATE = 1 << 8,   // non-basic rate effects: per-actor rates in LDS (count in bits 16-18)
};

// One SAOM period for a batch of chains (see saom_period.hip). Rows are u64
// bit rows: word w of actor h holds ties h->64w..64w+63 (bit j&63).
struct SimParams {
  int n, k, kb, M, W64, RS;
  int nact, has_comp, symflag, maxdeg, cond, target;
  int zmin, zmax;
  int smax;  // replay: recorded ministeps per chain (0 = free running)
  int lw;    // streamed rows: leading words of every row kept in LDS (0 .. W64)
  int flags;
  uint32_t key0, key1;
  double zbar, brange, bsim;
  int tid[kMaxEff], ety[kMaxEff], dyn[kMaxEff], dci[kMaxEff];
  int c0[kMaxEff], c1[kMaxEff], c2[kMaxEff];
  double p1[kMaxEff], p2[kMaxEff];
  int bty[kMaxBeff];
  // non-basic rate effects (libcusna_saom_cpu.cpp:170-193): actor a's rate is
  // lam * exp(clamp(sum_q ralpha[q] u_q(a), -30, 30)), u_q = ru[q*n + a]
  // (rkind 0), out-degree (1) or in-degree (2)
  int nrate;
  int rkind[kMaxRate];
  double ralpha[kMaxRate];
  const double* ru;    // nrate*n
  double* rscores;     // M*nrate
  const uint64_t *x0, *xs, *kd, *pm;  // n*W64 words each
  const uint64_t *inactw;             // W64 words or null
  const double *theta, *lam, *thetab, *lamb;  // chain c reads row c * stride
  int th_stride, thb_stride, lam_stride;      // k, kb, 1 per chain; 0 = shared row
  const uint32_t* keys;                       // [opt] Philox key in device memory
  const double *vmat, *wtab, *dyc;            // k*n, k*(n+1), nd*n*n (or null)
  const int32_t *z0, *bkd, *act;              // n each (act: nact used)
  const double *lu0, *llu;                    // replay inputs (or null)
  uint64_t *end_rows;                         // M*n*W64 end networks
  uint64_t *work;                             // streamed: M*W64*n working rows
  double *ntim, *scores, *bscores;            // M, M*(k+1), M*(kb+1)
  int32_t *dist, *dist_stat, *steps, *bdist, *z_end, *actor_dist;
};

// Launch shape chosen for one call.
struct LaunchShape {
  int nw;    // waves per chain (workgroup = 64*nw threads)
  int wmax;  // compile-time bound on W64 (1 .. 16 LDS-resident; 32, 64, 128 streamed)
  int wpe = 0;  // specialised kernel's register budget in waves per SIMD (0: two)
};

// LDS bytes one chain needs with the given shape.
size_t sim_lds_bytes(const SimParams& P, int nw, int wmax);

// Effects whose change statistics need no per-ministep arrays beyond the
// two-path counts (no transTies, transRecTrip, dyadic X, gwesp/gwdsp family,
// interactions): they run in the SIMPLE specialisation of the kernel.
bool simple_effects(const int* type_ids, int k);

// Launch the period kernel. P is the host copy (shape, LDS size); Pd the same
// struct in device memory, read by the kernel. Returns 0 or a hipError_t.
// *used_threads (optional): threads per chain of the launch that ran.
int launch_period(const SimParams& P, const SimParams* Pd, LaunchShape s, bool replay,
                  uint32_t key0, uint32_t key1, void* stream, bool specialize = false,
                  bool* used = nullptr, bool free_shape = false, int* used_threads = nullptr);

// General form. One group: Pd, grid = chains, key0/key1 the Philox key.
// Several groups: Pd[g] are group g's device parameters (its Philox key in
// their key0/key1 or keys), items[b] = (g << 20) | chain names block b's
// work; grid = number of items; lds = the largest of the groups' LDS sizes.
// All groups of a launch share the word bound and the block shape.
// specialize: use the runtime-specialised kernel for the effect list of Ph
// (host copy of the parameters; every group of the launch must share it)
// when it compiles and fits, else the generic kernel. *used (optional) tells
// which ran.
struct PeriodLaunch {
  const SimParams* Pd;
  const uint32_t* items;
  int grid;
  size_t lds;
  uint32_t key0, key1;
  const SimParams* Ph = nullptr;
  bool specialize = false;
  bool* used = nullptr;
  int* used_threads = nullptr;  // threads per chain of the launch that ran
  // shape and LDS bytes of the specialised kernel (exact_shape); nw = 0: the
  // generic shape. exact_alt: the shape to use instead when the preferred
  // one's register budget (exact.wpe) would make the compiler spill.
  LaunchShape exact{0, 0};
  size_t lds_exact = 0;
  LaunchShape exact_alt{0, 0};
  size_t lds_exact_alt = 0;
  // every group of the launch has the actor count of Ph
  bool same_n = true;
};
// Word bound and block shape of a specialised kernel for networks of W64
// words: the generic shape rounds the word bound up to a power of two; the
// specialised kernel is compiled for any bound, so it pads only to a multiple
// of the waves per chain (R words per lane): R = ceil(W64 / nw) with the
// generic waves up to 8 words, three words per lane on four waves at 9..12
// words (one wave per SIMD per chain), two words per lane from 13 words
// (more waves, less LDS padding), and at most 16 waves for streamed
// networks. k_free > 0
// (the caller left the block shape to the library; k_free effects) doubles
// the waves per chain up to 8 words when that brings k x R within the 32
// change statistics per lane that the specialised kernel keeps in registers
// between its choice and score passes (without them it computes every change
// statistic twice), and also when a CU holds at most four chains of
// lds_generic bytes (rows of 5..8 words, 385..512 actors): four waves per
// SIMD then hide latency that two cannot, given a four-wave register budget
// (wpe = 4; at 500 actors 118 against 124 ms per 4096-chain period).
LaunchShape exact_shape(int W64, LaunchShape generic, int k_free = 0, size_t lds_generic = 0);
int launch_period_ex(const PeriodLaunch& L, LaunchShape s, bool replay, bool simple,
                     void* stream);

// Runtime specialisation (rtc.hip): the compiled kernel for P's effect list
// and this launch shape (null when unavailable), and its launch.
// fixed_n: every group of the launch has P's actor count (compiled in); the
// word count W64 is compiled in either way (every launch holds networks of
// one word count: a plan's chains, one bucket of a groups call, the periods
// of a MoM fit);
// waves_per_eu > 0: the register budget in waves per SIMD (0: two waves);
// lean: without the registers that hold the alternatives' static covariate
// values for the whole period (they are loaded per ministep instead).
void* rtc_period_kernel(const SimParams& P, LaunchShape s, bool replay, bool simple,
                        bool fixed_n, int waves_per_eu = 0, bool lean = false);
// Whether a compiled kernel uses scratch memory (register spills), and its
// vector registers per lane.
bool uses_scratch(void* fn);
int kernel_vgprs(void* fn);
int launch_period_rtc(void* fn, const PeriodLaunch& L, int nw, void* stream);
// Launch policy: CUSNA_SPECIALIZE=0 never, =1 always, unset: when the launch
// has at least this many chains x actors.
constexpr long long kSpecializeWork = 1ll << 18;
bool specialize_policy(long long chains_x_actors);
// Same effect list, behaviour-effect count and feature flags (what the
// specialised kernel bakes in).
bool same_effects(const SimParams& a, const SimParams& b);

// Multi-network period (saom_multi.hip): D networks over one actor set
// (libcusna_saom_cpu.cpp:427-510). `base` carries the shared shape and the
// per-effect descriptors read by the single-network change statistics
// (n, k, W64, RS, nact, flags, tid/ety/dci/p1/p2, vmat, wtab, dyc, act);
// its dyn flags are zero.
constexpr int kMaxNet = 4;  // CUSNA_MAXNET
struct MultiParams {
  SimParams base;
  int D;
  int rows_global;  // 1: working rows in `work` (chain state above the LDS)
  int netof[kMaxEff], netref[kMaxEff];
  int symf[kMaxNet];
  const uint64_t *x0, *pm, *kd;  // D*n*W64 each (kd: distance dyads, or null)
  const double *theta, *lam;     // M*k, M*D
  uint64_t* work;                // rows_global: M*D*WS*n words (WS = row words stored)
  uint64_t* end_rows;            // D*M*n*W64: network d of chain c at (d*M + c)*n*W64
  double *scores, *rscores, *ntim;  // M*k, M*D, M
  int32_t* steps;                   // M
  int32_t* dist_stat;               // [opt] D*M: |end xor start| over kd, per network
  int M;
};
// Launch the multi-network period kernel (one workgroup of 64*nw threads
// per chain). Pd is the device copy of P. Returns 0 or a hipError_t.
int launch_multi(const MultiParams& P, const MultiParams* Pd, int nw, int wmax, void* stream);
size_t multi_lds_bytes(const MultiParams& P, int nw, int wmax);
// Runtime-specialised multi-network kernel for P's effect list and this
// shape (null when unavailable), and its launch.
void* rtc_multi_kernel(const MultiParams& P, int nw, int wmax);
int launch_multi_rtc(void* fn, const MultiParams* Pd, int grid, size_t lds, int nw,
                     void* stream);

// Statistics of M packed networks (rows u64, n*W64 per network).
struct StatParams {
  int n, k, M, W64, symflag;
  int tid[kMaxEff], dyn[kMaxEff];
  double p1[kMaxEff], p2[kMaxEff];
  const uint64_t* rows;   // M*n*W64
  const uint64_t *keep, *force;  // n*W64 or null
  const double *vmat, *vb, *wtab;
  const uint8_t* vmiss;   // k*n or null
  const int32_t *od0, *id0;  // n or null
  double* out;  // M*k
  // networks whose rows (and columns) do not fit the LDS: M times
  // net_stats_scratch_words u64 of device memory for them, else null
  uint64_t* scratch;
};
int launch_net_stats(const StatParams& P, void* stream);
size_t net_stats_lds(int n, int W64, bool need_cols);
// u64 words of device scratch one network needs (0: LDS-resident).
size_t net_stats_scratch_words(int n, int W64, bool need_cols);
bool stats_need_cols(const int* type_ids, int k);

struct BehStatParams {
  int n, M, kb, W64;
  int bty[kMaxBeff];
  const int32_t* z;        // M*n
  const uint64_t* xnet;    // n*W64
  const int32_t* bkdist;   // n or null
  double zbar, brange, bsim;
  double* out;             // M*kb
  double* scratch;         // M*beh_stats_scratch_doubles(n) when above the LDS, else null
};
int launch_beh_stats(const BehStatParams& P, void* stream);
size_t beh_stats_scratch_doubles(int n);

// sienaGOF auxiliary distributions of M packed networks
// (libcusna_saom_cpu.cpp:716-765): kind 0 indegree / 1 outdegree counts of
// actors with degree <= levls[L]; kind 2 geodesic counts of ordered pairs within
// distance levls[L] plus the unreachable pairs. Networks are (rows & keep) |
// force when the masks are given.
struct GofParams {
  int n, M, W64, kind, nlev;
  int kmax;                      // geodesic: largest level, capped at n
  const uint64_t* rows;          // M*n*W64
  const uint64_t *keep, *force;  // n*W64 or null
  const int32_t* levls;          // nlev (device)
  double* out;                   // M*(nlev + (kind == 2))
  // geodesic on rows read from device memory (gof_geodesic_lds() == 0 or
  // stream set): per-network reach counters, M*(kmax + 1), cleared by the launch
  unsigned long long* counters;
  int stream;                    // 1: the device-memory form even where the LDS one fits
};
int launch_gof(const GofParams& P, void* stream);
// LDS bytes of the geodesic kernel (0: the network is too large for it and
// the rows are read from device memory).
size_t gof_geodesic_lds(int n, int W64, int kmax);

// Cross-network statistic of M packed networks against a predictor network
// (libcusna_saom_cpu.cpp:819-857). Networks are (rows & keep) | force.
struct CrossParams {
  int n, M, W64, t;
  double parm;
  const uint64_t* rows;          // network c at rows + c * n * W64
  const uint64_t *keep, *force;  // n*W64 or null
  const uint64_t *wp, *wpT;      // predictor rows and their transpose, n*W64
  double* out;                   // out[c * out_stride]
  int out_stride;
};
int launch_cross_stats(const CrossParams& P, void* stream);

// Expand packed end rows (M*n*W64) to dense uint8 (M*n*n).
int launch_expand_dense(const uint64_t* rows, int n, int W64, int M, uint8_t* dense,
                        void* stream);

int launch_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, void* stream);

// The period kernel's softmax-weight exponential on count values (device).
int launch_exp_weight(const double* x, int count, double* out, void* stream);
int launch_gumbel_check(uint32_t x0, uint32_t stride, uint32_t count,
                        unsigned long long* max_bits, void* stream);

}  // namespace saom
