/*
 * libcusna.h -- C ABI of the MI355X SAOM period simulator (libcusna_amd.so).
 *
 * This header declares the entry points of the drop-in library. The structs
 * and the v1 functions keep the layout and meaning of the reference ABI
 * (cran/cusna src/libcusna.h) so that the R glue
 * (src/cusna_saom_glue.cpp:57-189, src/cusna_glue.cpp:55-89) can be relinked
 * against this library unchanged except for the symbol it calls. Every
 * declaration cites the reference declaration it replaces.
 *
 * Conventions (reference libcusna.h:28-35):
 *   - extern "C", fixed-width integers, plain-old-data structs;
 *   - the caller owns every buffer; NULL skips an optional output;
 *   - matrices are row-major; a packed bit row is W = ceil(n/32) uint32 words
 *     and bit (j & 31) of word (j >> 5) is the tie to actor j.
 *
 * All compute runs on the GPU (gfx950). There is no CPU fallback: without a
 * usable device the compute entry points return CUSNA_ERR_NO_CUDA.
 */
#ifndef CUSNA_AMD_LIBCUSNA_H
#define CUSNA_AMD_LIBCUSNA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUSNA_API __attribute__((visibility("default")))

/* Capacity limits (reference libcusna.h:61-65). */
#define CUSNA_MAXEFF 24
#define CUSNA_MAXBEFF 6
#define CUSNA_MAXRATE 4
#define CUSNA_MAXNET 4
#define CUSNA_MAX_N 8192

/* ABI version (reference libcusna.h:69). */
#define CUSNA_ABI_VERSION 1

/* Status codes (reference libcusna.h:72-80). */
typedef enum cusna_status {
  CUSNA_OK = 0,
  CUSNA_ERR_NULL_ARG = 1,    /* a required pointer is NULL                   */
  CUSNA_ERR_SHAPE = 2,       /* n/W/k/M out of range, or struct version      */
  CUSNA_ERR_NO_CUDA = 3,     /* no usable GPU (name kept from the ABI)       */
  CUSNA_ERR_CUDA = 4,        /* a HIP runtime call failed                    */
  CUSNA_ERR_UNSUPPORTED = 5, /* option not implemented                       */
  CUSNA_ERR_INTERNAL = 6
} cusna_status;

/*
 * One SAOM period for M chains (reference libcusna.h:90-180). Field order and
 * types are the reference's, byte for byte; the comments give the meaning
 * this implementation honours.
 */
typedef struct cusna_saom_problem {
  int32_t struct_version; /* CUSNA_ABI_VERSION                                */

  int32_t n;       /* actors, 1..CUSNA_MAX_N                                  */
  int32_t W;       /* (n + 31) / 32                                           */
  int32_t k;       /* network effects, 1..CUSNA_MAXEFF                        */
  int32_t M;       /* chains                                                  */
  int32_t threads; /* 0 = let the library choose the launch shape             */

  const uint32_t *x0_out;  /* data start network, packed out-rows (n*W)       */
  const uint32_t *x0_in;   /* packed in-rows; accepted, not needed            */
  const uint32_t *sim_out; /* [opt] simulation start network; NULL = x0_out   */
  const uint32_t *sim_in;  /* [opt] accepted, not needed                      */

  const float *lam;   /* [M] basic rate per chain                             */
  const float *theta; /* [M*k] evaluation parameters per chain                */

  const int32_t *type_ids;  /* [k] effect ids (0 density ... 32 interaction) */
  const int32_t *etypes;    /* [k] 0 eval, 1 creation, 2 endowment; NULL = 0 */
  const float *wtab;        /* [k*(n+1)] gwesp/gwdsp weight tables           */
  const float *p1;          /* [k] effect parameter 1                        */
  const float *p2;          /* [k] effect parameter 2                        */
  const float *vmat;        /* [k*n] covariate values                        */
  const uint8_t *vmiss;     /* [k*n] covariate missing flags (statistics)    */
  const int32_t *dyc_idx;   /* [k] index into dyc or -1; NULL = all -1       */
  const float *dyc;         /* [opt] nd*n*n dyadic covariates; NULL = zeros  */
  const int32_t *int0;      /* [k] interaction component 0 or -1             */
  const int32_t *int1;      /* [k] interaction component 1 or -1             */
  const int32_t *int2;      /* [k] interaction component 2 or -1             */
  const int32_t *dyn_flags; /* [k] covariate is the co-evolving behaviour    */

  const uint32_t *kdist;     /* [opt] dyads in the period distance (n*W)     */
  const uint32_t *keep_out;  /* [opt] statistic network (A & keep) | force   */
  const uint32_t *keep_in;   /* accepted, not needed                         */
  const uint32_t *force_out; /* [opt]                                        */
  const uint32_t *force_in;  /* accepted, not needed                         */
  const uint32_t *perm;      /* [opt] toggleable dyads (n*W)                 */

  const int32_t *od0; /* [n] start out-degrees (degree effects, p1 <= 0)      */
  const int32_t *id0; /* [n] start in-degrees                                 */

  int32_t maxdegree;   /* > 0 forbids new ties of egos at the bound           */
  int32_t need_mask;   /* accepted, not needed                                */
  int32_t conditional; /* 1 = stop at target_dist                             */
  int32_t target_dist;
  int32_t symflag; /* symmetric network: both cells flip                      */
  uint64_t seed;   /* Philox key                                              */

  int32_t has_comp;       /* restore inactive actors' ties at period end      */
  int32_t nact;           /* active actors (with act)                         */
  const int32_t *act;     /* [opt] active actor ids, first nact used          */
  const uint32_t *inactw; /* [opt] W words, inactive-actor bit mask           */

  int32_t kb;            /* behaviour effects, 0 = network only               */
  const float *lamb;     /* [M] behaviour rate per chain                      */
  const float *thetab;   /* [M*kb] behaviour parameters per chain             */
  const int32_t *bty;    /* [kb] behaviour effect ids (0 linear ... 4 totSim) */
  const int32_t *z0;     /* [n] start behaviour                               */
  const float *vb;       /* [n] centred start behaviour                       */
  const int32_t *bkdist; /* [n] actor counted in the behaviour distance       */
  int32_t zmin, zmax;
  float zbar, brange, bsimmean;

  int32_t nrate;        /* non-basic rate effects, 0..CUSNA_MAXRATE: actor a's */
                        /* rate is lam * exp(sum_q ralpha[q] u_q(a)), clamped */
                        /* to exp(+-30); the result's rscores and actor_dist */
                        /* are filled when nrate > 0                          */
  const int32_t *rkind; /* [nrate] 0 covariate ru row, 1 outdegree, 2 indegree */
  const float *ralpha;  /* [nrate] rate-effect parameters                     */
  const float *ru;      /* [nrate*n] covariate rows (rows of kind 1, 2 unread) */

  /* float64 descriptors; when given they replace the float32 ones above. */
  const double *vmat64;
  const double *wtab64;
  const double *vb64;
  const double *p1_64;
  const double *p2_64;
} cusna_saom_problem;

/* Outputs (reference libcusna.h:185-208). */
typedef struct cusna_saom_result {
  int32_t struct_version;

  double *stats;      /* [M*k] statistics of the masked end network          */
  double *ntim;       /* [M] simulated time                                  */
  float *scores;      /* [M*(k+1)] column 0 rate score, then the k effects   */
  int32_t *dist;      /* [M] signed simulated distance                       */
  int32_t *dist_stat; /* [M] |x_end xor x0| over kdist                       */
  int32_t *steps;     /* [M] ministeps                                       */

  int32_t *bdist;   /* [opt] [M]                                             */
  float *bscores;   /* [opt] [M*(kb+1)]                                      */
  int32_t *z_end;   /* [opt] [M*n]                                           */

  float *rscores;      /* [opt] [M*nrate]                                    */
  int32_t *actor_dist; /* [opt] [M*n] per-ego distance over kdist            */

  int32_t want_end_nets;
  uint8_t *a_ends; /* [opt] [M*n*n] dense end networks, [chain][i][j]        */
} cusna_saom_result;

/*
 * M chains of one period on the GPU. Replaces cusna_saom_simulate_period
 * (reference libcusna.h:212-213). Synchronous: results are in the caller's
 * host buffers on return. Unlike the reference CPU mirror
 * (libcusna_saom_cpu.cpp:334-336, 369) every chain c uses its own theta row
 * and lam[c], as the ABI comment at libcusna.h:107-109 specifies.
 */
CUSNA_API cusna_status cusna_saom_simulate_period(const cusna_saom_problem *problem,
                                                  cusna_saom_result *result);

/*
 * cusna_saom_simulate_period without device allocation. No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). Same
 * arguments, results and status codes as cusna_saom_simulate_period; the
 * device buffers of the call (the problem's device copy, chain state, packed
 * and dense end networks, outputs before their copy to the caller) are placed
 * in one static device array of CUSNA_NOALLOC_SCRATCH_BYTES per GPU, which
 * the GPU holds from the library's load until the process ends. A call whose
 * buffers need more returns CUSNA_ERR_SHAPE (cusna_last_error names the
 * size). Runs the generic period kernel (no runtime specialisation: loading
 * a new module allocates) on the null stream. Calls on one GPU share the
 * array and are serialised by a lock within the process; calls from two
 * processes on one GPU use separate copies. The array keeps the last call's
 * contents; every buffer a call reads is written by that call first.
 */
#define CUSNA_NOALLOC_SCRATCH_BYTES (64u << 20)
CUSNA_API cusna_status cusna_saom_simulate_period_noalloc(const cusna_saom_problem *problem,
                                                          cusna_saom_result *result);

/*
 * Multi-network co-evolution: D one-mode networks (2..CUSNA_MAXNET) over one
 * actor set. Layout byte-identical to the reference (libcusna.h:238-278).
 * Each ministep picks a network in proportion to its rate, an active actor,
 * and one toggle of that network. Effect e belongs to network netof[e]; type
 * ids 33-37 are cross-network effects (crprod, crprodRecip, crprodMutual,
 * from, to) reading the current state of network netref[e].
 */
typedef struct cusna_saom_multi_problem {
  int32_t struct_version; /* CUSNA_ABI_VERSION                               */
  int32_t n;              /* actors, shared by the D networks                */
  int32_t W;              /* (n + 31) / 32                                   */
  int32_t k;              /* effects over all networks (<= CUSNA_MAXEFF)     */
  int32_t M;              /* chains                                          */
  int32_t D;              /* networks, 2..CUSNA_MAXNET                       */
  int32_t threads;        /* block size hint; 0 = automatic                  */

  const uint32_t *x0_out; /* D*n*W packed out-rows of the start networks     */
  const uint32_t *x0_in;  /* D*n*W packed in-rows                            */
  const float *lam;       /* M*D rate of network d in chain c                */
  const float *theta;     /* M*k parameters of chain c                       */

  const int32_t *type_ids; /* k effect type ids                              */
  const int32_t *etypes;   /* k: 0 evaluation, 1 creation, 2 endowment       */
  const int32_t *netof;    /* k: network of effect e                         */
  const int32_t *netref;   /* k: source network of a cross effect            */
  const float *wtab;       /* k*(n+1) geometric weight tables                */
  const float *p1;         /* k                                              */
  const float *p2;         /* k                                              */
  const float *vmat;       /* k*n covariate values                           */
  const float *dyc;        /* [opt] dyadic covariates, nd*n*n                */
  const int32_t *dyc_idx;  /* k: dyadic covariate of effect e, or -1         */
  const uint32_t *perm;    /* [opt] D*n*W permitted toggles per network      */
  const int32_t *symf;     /* D: symmetric network (paired flips)            */
  uint64_t seed;           /* Philox key                                     */
  const int32_t *act;      /* [opt] n: active actors first                   */
  int32_t nact;            /* active actors (n without composition change)   */
} cusna_saom_multi_problem;

/* Outputs, caller-allocated; NULL skips one (reference libcusna.h:271-278). */
typedef struct cusna_saom_multi_result {
  int32_t struct_version;
  float *rscores;  /* [opt] [M*D] rate score of each network                 */
  float *scores;   /* [opt] [M*k] effect scores                              */
  double *ntim;    /* [opt] [M] simulated time                               */
  int32_t *steps;  /* [opt] [M] ministeps                                    */
  uint8_t *a_ends; /* [opt] [M*D*n*n] dense end networks, [chain][d][i][j]   */
} cusna_saom_multi_result;

/*
 * M chains of one multi-network period on the GPU (reference libcusna.h:
 * 284-285; the CPU mirror it must equal is cusna_saom_simulate_multi_cpu,
 * libcusna_saom_cpu.cpp:548-604): same trajectories, end networks, steps,
 * ntim; scores to float32. As in that mirror, a type-19 (reciAct) effect sees
 * reciprocated degree 0 and there is no composition restore. Dyadic covariate
 * q is read at dyc + q*n*n as the ABI documents (the mirror copies only the
 * first n*n values; with dyc_idx in {-1, 0} both agree).
 */
CUSNA_API cusna_status cusna_saom_simulate_multi(const cusna_saom_multi_problem *problem,
                                                 cusna_saom_multi_result *result);

/*
 * Cross-network statistic of M dense networks a_batch (M*n*n) against one
 * predictor network wp (n*n) on the GPU (reference cusna_cross_stats,
 * libcusna_saom_cpu.cpp:819-857): t = 33 sum a*wp, 34 sum a*wp^T,
 * 35 sum a*wp*wp^T, 36 sum over ties of |out(i) & out(j)| in wp (its square
 * root when parm >= 2), 37 sum over ties of (wp a)_ij. out[M].
 */
CUSNA_API cusna_status cusna_cross_stats(const uint8_t *a_batch, const uint8_t *wp, int32_t n,
                                         int32_t M, int32_t t, double parm, double *out);

/*
 * Device-resident multi-network problem (no counterpart in the reference,
 * whose estimator calls the one-shot function per batch, R/mom.R:352-364):
 * create uploads the problem once (and its theta/lam rows); set_params
 * replaces the per-chain rows (M*k theta, M*D lam); run simulates one period
 * with the given Philox key; fetch copies outputs like the one-shot call.
 * kdist ([opt] D*n*W packed rows): dyads counted in each network's period
 * distance; cusna_multi_plan_dist then returns D*M distances (network-major)
 * of the last run.
 */
typedef struct cusna_multi_plan cusna_multi_plan;
CUSNA_API cusna_status cusna_multi_plan_create(const cusna_saom_multi_problem *problem,
                                               int32_t want_end_nets, const uint32_t *kdist,
                                               cusna_multi_plan **plan);
CUSNA_API cusna_status cusna_multi_plan_set_params(cusna_multi_plan *plan, const float *theta,
                                                   const float *lam);
CUSNA_API cusna_status cusna_multi_plan_run(cusna_multi_plan *plan, uint64_t seed);
CUSNA_API cusna_status cusna_multi_plan_fetch(cusna_multi_plan *plan,
                                              cusna_saom_multi_result *result);
CUSNA_API cusna_status cusna_multi_plan_dist(cusna_multi_plan *plan, int32_t *dist_stat);
CUSNA_API cusna_status cusna_multi_plan_destroy(cusna_multi_plan *plan);
/* 1 when the last run launched the kernel specialised (hipRTC) to the plan's
 * effect list: launches of at least 2^18 chains x actors, CUSNA_SPECIALIZE=0/1
 * overriding; results are identical either way. */
CUSNA_API int32_t cusna_multi_plan_specialized(const cusna_multi_plan *plan);

/*
 * Statistics of network d's end states for the multi-network estimator, on
 * the device (the reference computes them on the host from dense a_ends,
 * R/mom.R:280-308, 371-391): spec lists network d's effects in list order;
 * ids 0-31 use the network statistics (as cusna_network_stats), ids 33-37 the
 * cross statistics (as cusna_cross_stats) against predictor cross_wp[q]; all
 * on the statistic network (A & keep) | force. cusna_multi_plan_stats then
 * writes M*k statistics of the last run, column e = effect e of the plan.
 */
typedef struct cusna_multi_stat_spec {
  int32_t struct_version;
  int32_t k;                /* effects of network d                         */
  const int32_t *type_ids;  /* k                                            */
  const double *p1, *p2;    /* k                                            */
  const double *vmat;       /* k*n                                          */
  const uint8_t *vmiss;     /* [opt] k*n                                    */
  const double *wtab;       /* [opt] k*(n+1)                                */
  const int32_t *od0, *id0; /* [opt] n start degrees                        */
  int32_t symflag;
  const uint32_t *keep;     /* [opt] n*W; NULL = off-diagonal               */
  const uint32_t *force;    /* [opt] n*W; NULL = none                       */
  const uint32_t *cross_wp; /* [opt] k*n*W predictor rows (ids 33-37)       */
} cusna_multi_stat_spec;
CUSNA_API cusna_status cusna_multi_plan_set_stats(cusna_multi_plan *plan, int32_t d,
                                                  const cusna_multi_stat_spec *spec);
CUSNA_API cusna_status cusna_multi_plan_stats(cusna_multi_plan *plan, double *out);

/* Pack dense n*n uint8 into out/in bit rows (reference libcusna.h:299-301). */
CUSNA_API cusna_status cusna_pack_bits(const uint8_t *dense, int32_t n, uint32_t *out_rows,
                                       uint32_t *in_rows, int32_t *W_out);

/* One Philox4x32-10 block, evaluated on the GPU (reference libcusna.h:304-305). */
CUSNA_API cusna_status cusna_philox4x32(const uint32_t counter[4], const uint32_t key[2],
                                        uint32_t out[4]);

/*
 * Statistics of ONE dense network (reference libcusna.h:313-318), computed on
 * the GPU by the same kernel that scores simulated end networks.
 */
CUSNA_API cusna_status cusna_network_stats(const uint8_t *a, int32_t n, int32_t k,
                                           const int32_t *type_ids, const double *p1,
                                           const double *p2, const double *vmat,
                                           const uint8_t *vmiss, const int32_t *dyn_flags,
                                           const double *vb, const int32_t *od0,
                                           const int32_t *id0, const double *wtab,
                                           int32_t symflag, double *out);

/* Behaviour statistics over M behaviour vectors (reference libcusna.h:359-362). */
CUSNA_API cusna_status cusna_behavior_stats(const int32_t *z_batch, int32_t n, int32_t M,
                                            int32_t kb, const int32_t *bty,
                                            const uint8_t *xnet, const int32_t *bkdist,
                                            double zbar, double brange, double bsimmean,
                                            double *out);

/* Introspection (reference libcusna.h:527-530). cusna_has_cuda keeps its ABI
 * name and reports whether a HIP device is usable. */
/* sienaGOF auxiliary distributions over a batch of end networks (reference
 * libcusna.h:334-343, libcusna_saom_cpu.cpp:716-765). a_batch is M*n*n dense;
 * kind 0 indegree / 1 outdegree: out[c][L] = actors with degree <= levls[L]
 * (M*nlev); kind 2 geodesic: ordered pairs i != j at distance <= levls[L],
 * then the unreachable pairs (M*(nlev+1)); geodesic levels must be >= 0.
 * Any n <= CUSNA_MAX_N: networks whose rows exceed the LDS (n > ~1000) are
 * walked from device memory. Status codes as the reference's. */
CUSNA_API cusna_status cusna_gof_distribution(const uint8_t *a_batch, int32_t n, int32_t M,
                                              int32_t kind, const int32_t *levls, int32_t nlev,
                                              double *out);

CUSNA_API int32_t cusna_abi_version(void);
CUSNA_API const char *cusna_version_string(void);
CUSNA_API int32_t cusna_has_cuda(void);
CUSNA_API const char *cusna_last_error(void);

/* ------------------------------------------------------------------------
 * Extensions above v1. None of them changes a v1 struct.
 * ------------------------------------------------------------------------ */

/*
 * Replay variant of cusna_saom_simulate_period, for verification. No
 * counterpart in the reference: it exists so that trajectories can be checked
 * bit for bit against the CPU path independently of how the device and the
 * host C library round log(). Instead of evaluating log() on its Philox
 * uniforms, chain c at ministep s reads
 *   log_u0[(c*smax + s)]           = log(u0)        (waiting time), and
 *   loglog_u[(c*smax + s)*n + j]   = log(-log(u_j)) (Gumbel noise of j),
 * as recorded by a host. A chain that reaches smax ministeps stops there.
 */
CUSNA_API cusna_status cusna_saom_simulate_period_replay(const cusna_saom_problem *problem,
                                                         cusna_saom_result *result,
                                                         const double *log_u0,
                                                         const double *loglog_u, int32_t smax);

/*
 * Statistics of M packed networks in one launch. No counterpart in the
 * reference (which scores one dense network per call, libcusna.h:313-318):
 * rows[M*n*W] are packed out-rows; keep/force ([opt], n*W packed) give the
 * statistic network (A & keep) | force as in libcusna_saom_cpu.cpp:415-417.
 * Descriptor meaning is that of cusna_network_stats. out is [M*k].
 */
CUSNA_API cusna_status cusna_network_stats_packed(const uint32_t *rows, int32_t n, int32_t M,
                                                  int32_t k, const int32_t *type_ids,
                                                  const double *p1, const double *p2,
                                                  const double *vmat, const uint8_t *vmiss,
                                                  const int32_t *dyn_flags, const double *vb,
                                                  const int32_t *od0, const int32_t *id0,
                                                  const double *wtab, int32_t symflag,
                                                  const uint32_t *keep, const uint32_t *force,
                                                  double *out);

/*
 * Verification hook, no counterpart in the reference. The period kernel
 * discards alternatives whose float32 bound on the Gumbel noise log(-log u)
 * cannot reach the maximum, and evaluates the float64 noise only for the
 * rest; that is exact as long as the float32 value is within 2^-8 of the
 * float64 one. This returns the largest difference over the uniforms
 * u = (x + 1/2) 2^-32, x = x0 + q*stride (mod 2^32), q < count.
 */
CUSNA_API cusna_status cusna_gumbel_bound_check(uint32_t x0, uint32_t stride, uint32_t count,
                                               double *max_abs_err);

/*
 * Verification hook, no counterpart in the reference. The period kernel
 * weighs the alternatives of the score sums by exp(f_j - shift) with an
 * exponential of its own (a 32-entry table of 2^(j/32) and a degree-6
 * polynomial); this evaluates it on count device-side copies of x.
 */
CUSNA_API cusna_status cusna_exp_weight_eval(const double *x, int32_t count, double *out);

/*
 * Device-resident period plans. No counterpart in the reference, whose period
 * call re-marshals everything every time (cusna_saom_glue.cpp:57-189); the
 * reserved Layer-2 estimator types (libcusna.h:541-542) are where the
 * reference meant such state to live. A plan uploads one problem's
 * chain-invariant inputs once; each run launches the period kernels on the
 * plan's own HIP stream and leaves the outputs in device memory until fetched.
 *
 *   create        validate + upload problem (problem->theta/lam used as the
 *                 initial per-chain parameters); want_end_nets reserves the
 *                 dense a_ends buffer
 *   set_params    replace per-chain theta[M*k], lam[M] (thetab/lamb with kb>0)
 *   run           launch one period with the given Philox seed (asynchronous)
 *   sync / fetch  wait; copy outputs into a cusna_saom_result
 *   kernel_ms     duration of the last period kernel (HIP events)
 *   shape         threads per chain of the last run's launch (the specialised
 *                 kernel may run its own block shape); before the first run,
 *                 and after autotune, the plan's block shape
 */
typedef struct cusna_period_plan cusna_period_plan;
CUSNA_API cusna_status cusna_period_plan_create(const cusna_saom_problem *problem,
                                                int32_t want_end_nets, cusna_period_plan **plan);
CUSNA_API cusna_status cusna_period_plan_set_params(cusna_period_plan *plan, const float *theta,
                                                    const float *lam, const float *thetab,
                                                    const float *lamb);
CUSNA_API cusna_status cusna_period_plan_run(cusna_period_plan *plan, uint64_t seed,
                                             int32_t want_stats);
CUSNA_API cusna_status cusna_period_plan_sync(cusna_period_plan *plan);
CUSNA_API cusna_status cusna_period_plan_fetch(cusna_period_plan *plan, cusna_saom_result *result);
CUSNA_API double cusna_period_plan_kernel_ms(cusna_period_plan *plan);
CUSNA_API int32_t cusna_period_plan_shape(const cusna_period_plan *plan);
CUSNA_API void cusna_period_plan_destroy(cusna_period_plan *plan);

/*
 * Block-shape autotuner (no counterpart in the reference, whose CPU engine
 * has one shape: an OpenMP loop over chains, libcusna_saom_cpu.cpp:384-387).
 * Times the period kernel in each compiled block shape (64 to 1024 threads
 * per chain) on the plan's first probe_chains chains (<= 0: min(M, 2048)),
 * keeps the fastest for this plan and for later plans of the same network
 * size and effect set created with threads = 0, and returns it in
 * *best_threads. The probe runs overwrite the plan's outputs. Trajectories,
 * statistics and end networks do not depend on the shape; scores can differ
 * in the last float64 bits (their sums run in a shape-dependent order).
 */
/*
 * Runtime specialisation of the period kernel (no counterpart in the
 * reference). The library embeds the kernel source and can compile, with
 * hipRTC, a kernel for one effect list and launch shape: effect dispatch
 * resolved at compile time, change statistics kept in registers between the
 * choice and score passes. Results are identical to the generic kernel's.
 * Code objects are cached in memory and on disk (CUSNA_RTC_CACHE, else
 * $XDG_CACHE_HOME/cusna_amd, else $HOME/.cache/cusna_amd). Plans use it when
 * M x n >= 2^18 unless CUSNA_SPECIALIZE=0 (never) or =1 (always) is set.
 *   specialize    mode 1 on, 0 off, -1 by that policy (replay plans: off)
 *   specialized   1 when the plan's last run launched the specialised kernel
 */
/* GoF distributions (as cusna_gof_distribution) of a plan's end networks,
 * read where they are in device memory after the last run (no dense copy);
 * masked != 0 applies the plan's keep/force masks (the statistic network). */
CUSNA_API cusna_status cusna_period_plan_gof(cusna_period_plan *plan, int32_t masked,
                                             int32_t kind, const int32_t *levls, int32_t nlev,
                                             double *out);

CUSNA_API cusna_status cusna_period_plan_specialize(cusna_period_plan *plan, int32_t mode);
CUSNA_API int32_t cusna_period_plan_specialized(const cusna_period_plan *plan);

CUSNA_API cusna_status cusna_period_plan_autotune(cusna_period_plan *plan, int32_t probe_chains,
                                                  int32_t *best_threads);

/*
 * Several independent period problems ("groups": networks of different
 * sizes, periods, parameter sets) in one call: results[g] is what
 * cusna_saom_simulate_period(problems[g], results[g]) returns. The groups run
 * in one period-kernel launch per distinct launch shape, items ordered
 * longest-first. Up to 4096 groups of up to 2^20 chains each. No counterpart
 * in the reference (whose R layer calls the period function once per group
 * and period, mom.R:143-148).
 */
CUSNA_API cusna_status cusna_saom_simulate_groups(const cusna_saom_problem *const *problems,
                                                  int32_t ngroups,
                                                  cusna_saom_result *const *results);

/*
 * Method-of-moments estimator (the reference's Layer-2 estimator, reserved as
 * opaque types at libcusna.h:541-542 and implemented in R at R/mom.R:91-269).
 * The caller prepares one cusna_saom_problem per period (inputs only; M,
 * theta, lam and seed are ignored) and the theta-ordered target vector
 * (R/targets.R:99-126). Parameter layouts, by mode:
 *   0 unconditional  [rate x P, eval x k]
 *   1 conditional    [eval x k]                 (internal rate 1)
 *   2 co-evolution   [net rate x P, beh rate x P, eval x k, beh x kb]
 * Phase 1 (n1 chains), phase 2 (nsub sub-phases of nD derivative chains and
 * n2[s] Robbins-Monro iterations of batch2 chains, gain firstg * 0.5^s) and
 * phase 3 (n3 chains) follow R/mom.R:218-269; every simulation and moment
 * reduction runs on the device, phase-2 iterations replayed from a HIP graph
 * when use_graph is set. With max_repeats > 0, a phase-3 result with
 * max|tconv| > conv_tol is followed by one more last-gain sub-phase (with
 * the surrogate: one Newton step from that phase 3's means and score-function
 * derivative) and a new phase 3, up to max_repeats times (no counterpart in
 * the reference).
 */
typedef struct cusna_mom_options {
  int32_t struct_version; /* CUSNA_ABI_VERSION */
  int32_t mode;           /* 0 unconditional, 1 conditional, 2 co-evolution */
  int32_t P;              /* periods, 1..16 */
  int32_t nsub;           /* phase-2 sub-phases, 0..8 */
  int32_t n2[8];          /* iterations per sub-phase */
  int32_t n1, nD, batch2, n3;
  double firstg;          /* 0.4 in the reference (mom.R:26) */
  double diagonalize;     /* 0.2 */
  uint64_t seed;          /* Philox keys of round r, period q derive from (seed, r, q) */
  int32_t use_graph;      /* replay phase-2 iterations from a captured HIP graph */
  int32_t max_repeats;    /* extra sub-phase + phase 3 rounds until converged */
  double conv_tol;        /* max|tconv| bound for max_repeats (0.1) */
  /* co-evolution: the float64 behaviour mean, range and similarity mean the
   * behaviour statistics take (mom.R:114-115); the period problems carry
   * their float32 roundings for the simulator */
  double beh_zbar, beh_brange, beh_bsimmean;
  /* Surrogate phase 2 (no counterpart in the reference) in place of the
   * Robbins-Monro sub-phases: per round, 2^surrogate_pairs_log2 chains (0:
   * 2^13) with per-chain theta drawn uniformly from a box around the phase-1
   * Newton point, the surrogate perceptron trained on them for
   * surrogate_steps Adam steps (0: 200), E s(theta) = target solved on its
   * mean; the box re-centres on the solution and shrinks to two standard
   * errors over surrogate_rounds rounds (0: 1). Then up to
   * surrogate_corrections (negative: 6) Newton steps with residual and
   * score-function derivative from nD plain chains, stopping once every
   * residual is below conv_tol moment standard deviations. Phase 3 as usual,
   * its repeats as above. */
  int32_t surrogate;
  int32_t surrogate_pairs_log2;
  int32_t surrogate_steps;
  int32_t surrogate_rounds;
  int32_t surrogate_corrections;
  /* period kernel specialised at run time to the effect list (hipRTC): 1 on,
   * -1 off, 0 when a simulation round has >= 2^18 chains x actors or the
   * whole schedule >= 2^22 (the ~1 s compile, cached on disk, amortised) */
  int32_t specialize;
  /* progress lines on stderr, as the R estimator's verbose = TRUE prints them
   * (mom.R:230, 246): theta after phase 1 and after each sub-phase, then
   * max|tconv| of each phase 3 */
  int32_t verbose;
  /* surrogate: plain Robbins-Monro sub-phases run before it (< 0: none, the
   * box then comes from phase 1; 0: one when phase 1's moments lie more than
   * 50 standard deviations from the targets, else none); with some, the box
   * is centred on the Newton point of nD chains at their theta. Whatever the
   * count, when the corrections leave a residual above 3 moment standard
   * deviations the
   * remaining plain sub-phases run from the theta the surrogate started at
   * (the plain estimator is the fallback), and a phase-3 Newton repeat moves each
   * parameter by at most two of that phase 3's standard errors. */
  int32_t surrogate_warmup;
} cusna_mom_options;

typedef struct cusna_mom_result {
  int32_t struct_version; /* CUSNA_ABI_VERSION */
  int32_t p;              /* out: parameter count */
  /* caller-owned outputs, each optional (NULL = not wanted) */
  double *theta, *se, *tconv, *sim_means; /* p */
  double *D, *msf;                        /* p*p, row-major */
  double *rate_est, *rate_sd;             /* P, conditional mode (mom.R:260, 265) */
  double *phase3_stats;                   /* n3*p moment rows of phase 3 */
  double *phase3_sp, *phase3_scp;         /* P*n3*p per-period moment / score rows */
  double tconv_max;                       /* sqrt(dev' msf^-1 dev) */
  double max_abs_tconv;
  int64_t n_sims;                         /* chains simulated (each over all periods) */
  int32_t repeats;
  int32_t reserved;
  double phase_seconds[3];
  double wall_seconds;
} cusna_mom_result;

/*
 * Surrogate perceptron (no counterpart in the reference). Four layers
 * x (p_in) -> width -> width -> width -> [mu (p_out), log-variance (p_out)],
 * SiLU activations, trained by Adam on the Gaussian negative log-likelihood
 * mean_b sum_a 0.5 [(y - mu)^2 exp(-lv) + lv]; products on the BF16 matrix
 * cores with float32 accumulation (weights and activations rounded to
 * bfloat16 as they enter a product), float32 master weights.
 *
 *   create        p_in, p_out in 1..64; width a multiple of 64 up to 256,
 *                 batch a multiple of 64 (a training step is two launches:
 *                 the batch in 16-row slices through the whole network, then
 *                 the weight gradients with Adam per 64 x 64 weight tile)
 *   param_count   floats in the flat parameter vector: for each layer l,
 *                 W_l [K_l][N_l] row-major then b_l [N_l] (layer_shape);
 *                 K_0 = p_in and N_3 = 2 p_out padded to multiples of 64
 *   forward_backward  one batch (batch x p_in, batch x p_out, host float32):
 *                 outputs (batch x N_3), gradients (flat), mean loss
 *   train         `steps` Adam steps on minibatches drawn with replacement
 *                 from n host samples, learning rate decaying linearly from lr
 *                 to lr / 10; last minibatch's mean loss
 *   predict       outputs [mu, lv] (n x 2 p_out) for n inputs
 */
typedef struct cusna_mlp cusna_mlp;
CUSNA_API cusna_status cusna_mlp_create(int32_t p_in, int32_t p_out, int32_t width, int32_t batch,
                                        uint64_t seed, cusna_mlp **mlp);
CUSNA_API int64_t cusna_mlp_param_count(const cusna_mlp *mlp);
CUSNA_API cusna_status cusna_mlp_layer_shape(const cusna_mlp *mlp, int32_t layer, int32_t *K,
                                             int32_t *N);
CUSNA_API cusna_status cusna_mlp_get_params(cusna_mlp *mlp, float *flat);
CUSNA_API cusna_status cusna_mlp_set_params(cusna_mlp *mlp, const float *flat);
CUSNA_API cusna_status cusna_mlp_forward_backward(cusna_mlp *mlp, const float *x, const float *y,
                                                  float *out, float *grads, double *loss);
CUSNA_API cusna_status cusna_mlp_train(cusna_mlp *mlp, const float *x_all, const float *y_all,
                                       int64_t n, int32_t steps, float lr, uint64_t seed,
                                       double *last_loss);
CUSNA_API cusna_status cusna_mlp_predict(cusna_mlp *mlp, const float *x, int64_t n, float *out);
CUSNA_API void cusna_mlp_destroy(cusna_mlp *mlp);

CUSNA_API cusna_status cusna_mom_estimate(const cusna_saom_problem *const *periods,
                                          const uint8_t *const *xnet, const double *targets,
                                          const double *rate_consts, const double *theta0,
                                          const cusna_mom_options *options,
                                          cusna_mom_result *result);

#ifdef __cplusplus
}
#endif
#endif /* CUSNA_AMD_LIBCUSNA_H */
