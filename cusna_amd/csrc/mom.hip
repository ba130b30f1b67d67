// mom.hip -- method-of-moments (Robbins-Monro) estimator on device-resident
// period plans.
//
// Restates the reference's R estimator (cran/cusna R/mom.R:91-269): phase 1
// (n1 chains, derivative, first step), phase 2 (nsub sub-phases: nD chains for
// the derivative, then n2[sub] Robbins-Monro iterations of batch2 chains with
// gain firstg * 0.5^(sub-1), theta averaged), phase 3 (n3 chains: derivative,
// moment covariance, convergence t-ratios, standard errors). Parameter layouts
// (R/mom.R:39, 131-138): unconditional [rate x P, eval x k]; conditional
// [eval x k] (internal rate 1); co-evolution [net rate x P, beh rate x P,
// eval x k, beh x kb].
//
// MI355X layout: every simulation runs on the device; the moment vector,
// its column means, the score-covariance derivative D and the moment
// covariance are reduced on the device; a phase-2 iteration (seed draw,
// P period kernels, statistics, moment reduction, guarded theta update) is
// captured once per sub-phase in a hipGraph and replayed n2 times with no host
// round trip. The host only inverts the p x p matrices between sub-phases.
// theta is rounded to float32 before the kernels read it, as the R glue
// rounds it for every period call (cusna_saom_glue.cpp:75-78).
//
// Surrogate phase 2 (no counterpart in the reference): instead of the
// Robbins-Monro iterations, simulate 2^q chains with per-chain theta drawn
// uniformly from a box around the phase-1 Newton point, fit the BF16 MFMA
// perceptron (theta -> mean and log-variance of the moments) to them, and
// solve E s(theta) = target on the fitted mean by damped Newton with its
// exact Jacobian; then correct the solution by a few Newton steps whose
// residual and derivative come from nD plain chains. Phase 3 then runs on
// the plain simulator, as in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "mlp.h"
#include "plan.h"

namespace {

using namespace cusna_impl;

constexpr int kMomThreads = 256;

struct PeriodOut {  // device pointers of one period's outputs
  const int32_t* dist_stat;
  const double* scores;  // M*(k+1)
  const double* stats;   // M*k
  const int32_t* bdist;
  const double* bscores;  // M*(kb+1)
  const double* bstats;   // M*kb
};

struct GatherArgs {
  int M, P, p, k, kb, mode;
  const PeriodOut* per;      // P entries (device)
  const double* rate_consts;  // P
  double* S;                  // M*p
  double* SP;                 // P*M*p
  double* SC;                 // P*M*p
};

__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Philox keys of this simulation round: one per period, from (seed, round).
__global__ void mom_seed_kernel(uint32_t* keys, unsigned long long* counter, uint64_t seed,
                                int P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long it = counter[0];
  for (int pp = 0; pp < P; ++pp) {
    const uint64_t x = splitmix64(seed ^ splitmix64((uint64_t)it * 64u + (uint64_t)pp));
    keys[2 * pp] = (uint32_t)x;
    keys[2 * pp + 1] = (uint32_t)(x >> 32);
  }
  counter[0] = it + 1;
}

// Per chain: the period moment rows sp, score rows scp (R/mom.R:152-166) and
// their sum over periods s. Scores pass through float32, the type of the ABI's
// score outputs (libcusna.h:185-208) that the R estimator reads.
__global__ void mom_gather_kernel(const GatherArgs G) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.M) return;
  const int P = G.P, p = G.p, k = G.k, kb = G.kb;
  const int off = G.mode == 2 ? 2 * P : (G.mode == 1 ? 0 : P);
  double* s = G.S + (size_t)c * p;
  for (int a = 0; a < p; ++a) s[a] = 0.0;
  for (int pp = 0; pp < P; ++pp) {
    const PeriodOut o = G.per[pp];
    double* sp = G.SP + ((size_t)pp * G.M + c) * p;
    double* sc = G.SC + ((size_t)pp * G.M + c) * p;
    for (int a = 0; a < p; ++a) {
      sp[a] = 0.0;
      sc[a] = 0.0;
    }
    if (G.mode != 1) {
      sp[pp] = (double)o.dist_stat[c] + G.rate_consts[pp];
      sc[pp] = (double)(float)o.scores[(size_t)c * (k + 1)];
    }
    if (G.mode == 2) {
      sp[P + pp] = (double)o.bdist[c];
      sc[P + pp] = (double)(float)o.bscores[(size_t)c * (kb + 1)];
      for (int e = 0; e < kb; ++e) {
        sp[off + k + e] = o.bstats[(size_t)c * kb + e];
        sc[off + k + e] = (double)(float)o.bscores[(size_t)c * (kb + 1) + 1 + e];
      }
    }
    for (int e = 0; e < k; ++e) {
      sp[off + e] = o.stats[(size_t)c * k + e];
      sc[off + e] = (double)(float)o.scores[(size_t)c * (k + 1) + 1 + e];
    }
    for (int a = 0; a < p; ++a) s[a] += sp[a];
  }
}

__device__ inline double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Column means of an M x p matrix, one block per column (fixed tree order).
__global__ void mom_mean_kernel(const double* X, int M, int p, double* mean) {
  __shared__ double sh[kMomThreads];
  const int a = blockIdx.x;
  double v = 0.0;
  for (int c = threadIdx.x; c < M; c += blockDim.x) v += X[(size_t)c * p + a];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) mean[a] = v / (double)M;
}

// Column variances (divisor M - 1) of an M x p matrix, one block per column.
__global__ void mom_var_kernel(const double* X, int M, int p, const double* mean, double* var) {
  __shared__ double sh[kMomThreads];
  const int a = blockIdx.x;
  double v = 0.0;
  for (int c = threadIdx.x; c < M; c += blockDim.x) {
    const double d = X[(size_t)c * p + a] - mean[a];
    v += d * d;
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) var[a] = v / (double)(M - 1);
}

// Surrogate training inputs: chain i gets theta_a = c_a + h_a u (u uniform
// on (-1, 1) from a counter hash of (seed, round, i, a)), rates kept at or
// above a tenth of the box centre; the kernels read the float32 rounding
// (row i of th_sim, slot p = internal rate 1), the perceptron the
// normalised x = (theta - c) / h of that rounding.
__global__ void sample_theta_kernel(const double* c, const double* h, int p, int M,
                                    uint64_t seed, unsigned round, int n_rates, float* x,
                                    double* th_sim) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  for (int a = 0; a < p; ++a) {
    const uint64_t r = splitmix64(seed ^ splitmix64(((uint64_t)round << 40) ^
                                                    ((uint64_t)i << 7) ^ (uint64_t)a));
    const double u = 2.0 * ((double)(r >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    double th = c[a] + h[a] * u;
    if (a < n_rates) th = fmax(th, 0.1 * fabs(c[a]));
    const float t32 = (float)th;
    x[(size_t)i * p + a] = (float)(((double)t32 - c[a]) / h[a]);
    th_sim[(size_t)i * (p + 1) + a] = (double)t32;
  }
  th_sim[(size_t)i * (p + 1) + p] = 1.0;
}

// y = (S - mean) / sd per column, float32 (perceptron targets).
__global__ void normalize_kernel(const double* S, int M, int p, const double* mean,
                                 const double* var, float* y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)M * p) return;
  const int a = (int)(i % p);
  const double sd = var[a] > 0.0 ? sqrt(var[a]) : 1.0;
  y[i] = (float)((S[i] - mean[a]) / sd);
}

// out[a][b] (+)= sum_c (A[c][a] - mA[a]) (B[c][b] - mB[b]) / div, one block
// per entry.
__global__ void mom_cov_kernel(const double* A, const double* mA, const double* B,
                               const double* mB, int M, int p, double div, int accumulate,
                               double* out) {
  __shared__ double sh[kMomThreads];
  const int a = blockIdx.x / p, b = blockIdx.x - a * p;
  double v = 0.0;
  for (int c = threadIdx.x; c < M; c += blockDim.x)
    v += (A[(size_t)c * p + a] - mA[a]) * (B[(size_t)c * p + b] - mB[b]);
  v = block_sum(v, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = (accumulate ? out[blockIdx.x] : 0.0) + v / div;
}

// Guarded Robbins-Monro step (R/mom.R:189-197, 236-243):
// theta -= guard(gain * dinv (mean - target)).
__global__ void mom_update_kernel(double* theta, double* theta_sim, double* thav,
                                  const double* dinv, const double* mean, const double* target,
                                  int p, double gain, int n_rates, int accumulate) {
  __shared__ double dev[64];
  const int a = threadIdx.x;
  if (a < p) dev[a] = mean[a] - target[a];
  __syncthreads();
  if (a >= p) return;
  double st = 0.0;
  for (int b = 0; b < p; ++b) st += dinv[a * p + b] * dev[b];
  st *= gain;
  const double th = theta[a];
  const double cap = fmax(1.0, 0.5 * fabs(th));
  st = fmin(fmax(st, -cap), cap);
  if (a < n_rates && th - st < 0.1 * th) st = 0.5 * th;
  const double nt = th - st;
  theta[a] = nt;
  theta_sim[a] = (double)(float)nt;
  if (accumulate) thav[a] += nt;
}

// End of a sub-phase (R/mom.R:245): theta = thav / n, thav reset, the
// float32 copy the kernels read refreshed; slot p holds the internal rate 1
// of conditional runs.
__global__ void mom_average_kernel(double* theta, double* theta_sim, double* thav, int p,
                                   double n) {
  const int a = threadIdx.x;
  if (a < p) {
    theta[a] = thav[a] / n;
    theta_sim[a] = (double)(float)theta[a];
    thav[a] = 0.0;
  }
  if (a == 0) theta_sim[p] = 1.0;
}

// ---------------------------------------------------------------- host algebra

using Mat = std::vector<double>;

// Inverse by Gauss-Jordan with partial pivoting; false when (numerically)
// singular: reciprocal condition (1-norm) below machine epsilon, the point at
// which R's solve() refuses (mom.R:20 then falls back to the pseudo-inverse).
bool invert(const Mat& A, int n, Mat& inv) {
  Mat a = A;
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  double anorm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(A[(size_t)i * n + j]);
    anorm = std::max(anorm, s);
  }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (a[(size_t)piv * n + c] == 0.0) return false;
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        std::swap(a[(size_t)c * n + j], a[(size_t)piv * n + j]);
        std::swap(inv[(size_t)c * n + j], inv[(size_t)piv * n + j]);
      }
    const double d = a[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) {
      a[(size_t)c * n + j] /= d;
      inv[(size_t)c * n + j] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
        inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j];
      }
    }
  }
  double inorm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(inv[(size_t)i * n + j]);
    inorm = std::max(inorm, s);
  }
  if (!(anorm * inorm < 1.0 / 2.220446049250313e-16)) return false;
  return true;
}

// Moore-Penrose pseudo-inverse by one-sided Jacobi SVD, singular values below
// sqrt(eps) * max dropped (mom.R:15-19).
Mat pinv(const Mat& A, int n) {
  Mat U = A, V((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < n; ++i) {
          const double up = U[(size_t)i * n + p], uq = U[(size_t)i * n + q];
          al += up * up;
          be += uq * uq;
          ga += up * uq;
        }
        if (ga == 0.0) continue;
        off = std::max(off, std::fabs(ga) / std::sqrt(al * be));
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < n; ++i) {
          const double up = U[(size_t)i * n + p], uq = U[(size_t)i * n + q];
          U[(size_t)i * n + p] = cs * up - sn * uq;
          U[(size_t)i * n + q] = sn * up + cs * uq;
          const double vp = V[(size_t)i * n + p], vq = V[(size_t)i * n + q];
          V[(size_t)i * n + p] = cs * vp - sn * vq;
          V[(size_t)i * n + q] = sn * vp + cs * vq;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> sv(n);
  double smax = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += U[(size_t)i * n + j] * U[(size_t)i * n + j];
    sv[j] = std::sqrt(s);
    smax = std::max(smax, sv[j]);
  }
  const double tol = std::sqrt(2.220446049250313e-16) * smax;
  Mat out((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    if (!(sv[j] > tol)) continue;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c)
        out[(size_t)r * n + c] +=
            V[(size_t)r * n + j] * (U[(size_t)c * n + j] / sv[j]) / sv[j];
  }
  return out;
}

Mat safe_solve(const Mat& A, int n) {
  Mat inv;
  if (invert(A, n, inv)) return inv;
  return pinv(A, n);
}

// ------------------------------------------------- surrogate on the host

// bfloat16 rounding (nearest even) of a float32, as the device rounds
// weights entering its products.
double bf16_round(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16;
  float r;
  std::memcpy(&r, &u, 4);
  return (double)r;
}

// The trained perceptron in float64 with its bfloat16 weights: mean and
// log-variance of the normalised moments at normalised x, and the exact
// Jacobian d mean / d x by forward-mode differentiation.
struct HostNet {
  int p = 0, width = 0;
  int K[kMlpLayers], N[kMlpLayers];
  std::vector<double> W[kMlpLayers], b[kMlpLayers];

  void load(const Mlp& net, const std::vector<float>& flat, int p_, int width_) {
    p = p_;
    width = width_;
    for (int l = 0; l < kMlpLayers; ++l) {
      K[l] = net.K(l);
      N[l] = net.N(l);
      W[l].resize((size_t)K[l] * N[l]);
      b[l].resize(N[l]);
      for (size_t q = 0; q < W[l].size(); ++q) W[l][q] = bf16_round(flat[net.w_offset(l) + q]);
      for (int q = 0; q < N[l]; ++q) b[l][q] = flat[net.b_offset(l) + q];
    }
  }

  void eval(const double* x, double* mu, double* lv, double* J) const {
    std::vector<double> h(x, x + p), T((size_t)p * p, 0.0);  // T[i][d] = dh_i/dx_d
    for (int i = 0; i < p; ++i) T[(size_t)i * p + i] = 1.0;
    int hin = p;
    for (int l = 0; l < kMlpLayers; ++l) {
      const int Nl = (l == kMlpLayers - 1) ? 2 * p : width;
      std::vector<double> z(b[l].begin(), b[l].begin() + Nl), Tz((size_t)Nl * p, 0.0);
      for (int i = 0; i < hin; ++i) {
        const double* wr = &W[l][(size_t)i * N[l]];
        const double hi = h[i];
        const double* ti = &T[(size_t)i * p];
        for (int nn = 0; nn < Nl; ++nn) {
          const double w = wr[nn];
          if (w == 0.0) continue;
          z[nn] += hi * w;
          double* tz = &Tz[(size_t)nn * p];
          for (int d = 0; d < p; ++d) tz[d] += ti[d] * w;
        }
      }
      if (l < kMlpLayers - 1) {
        h.assign(Nl, 0.0);
        T.assign((size_t)Nl * p, 0.0);
        for (int nn = 0; nn < Nl; ++nn) {
          const double sg = 1.0 / (1.0 + std::exp(-z[nn]));
          h[nn] = z[nn] * sg;
          const double ds = sg * (1.0 + z[nn] * (1.0 - sg));
          for (int d = 0; d < p; ++d) T[(size_t)nn * p + d] = Tz[(size_t)nn * p + d] * ds;
        }
        hin = Nl;
      } else {
        for (int a = 0; a < p; ++a) {
          mu[a] = z[a];
          lv[a] = z[p + a];
          for (int d = 0; d < p; ++d) J[(size_t)a * p + d] = Tz[(size_t)a * p + d];
        }
      }
    }
  }
};

// Damped Newton for mu(x) = t inside the box |x_a| <= 1: full steps from
// x = 0, halved until the residual norm falls; stops when the step is tiny
// or no halving helps. Returns the final residual norm.
double surrogate_newton(const HostNet& H, const std::vector<double>& t, std::vector<double>& x,
                        std::vector<double>& J, std::vector<double>& lv) {
  const int p = H.p;
  std::vector<double> mu(p), r(p), xn(p), mun(p), lvn(p), Jn((size_t)p * p);
  x.assign(p, 0.0);
  J.assign((size_t)p * p, 0.0);
  lv.assign(p, 0.0);
  H.eval(x.data(), mu.data(), lv.data(), J.data());
  auto norm = [&](const std::vector<double>& m) {
    double s2 = 0.0;
    for (int a = 0; a < p; ++a) s2 += (m[a] - t[a]) * (m[a] - t[a]);
    return std::sqrt(s2);
  };
  double nr = norm(mu);
  for (int it = 0; it < 60 && nr > 1e-9; ++it) {
    for (int a = 0; a < p; ++a) r[a] = mu[a] - t[a];
    const Mat Ji = safe_solve(J, p);
    std::vector<double> dx(p, 0.0);
    for (int a = 0; a < p; ++a)
      for (int d = 0; d < p; ++d) dx[a] += Ji[(size_t)a * p + d] * r[d];
    double alpha = 1.0, nn = nr;
    bool moved = false;
    for (int half = 0; half < 12; ++half, alpha *= 0.5) {
      for (int a = 0; a < p; ++a) xn[a] = std::min(1.0, std::max(-1.0, x[a] - alpha * dx[a]));
      H.eval(xn.data(), mun.data(), lvn.data(), Jn.data());
      nn = norm(mun);
      if (nn < nr) {
        moved = true;
        break;
      }
    }
    if (!moved) break;
    double step = 0.0;
    for (int a = 0; a < p; ++a) step = std::max(step, std::fabs(xn[a] - x[a]));
    x = xn;
    mu = mun;
    lv = lvn;
    J = Jn;
    nr = nn;
    if (step < 1e-10) break;
  }
  return nr;
}

}  // namespace

// ------------------------------------------------------------------ engine

// One set of period plans simulated together: all periods of a round in one
// launch (item b = period b % P, chain b / P), then statistics, moment rows
// and their column means.
struct SimSet {
  std::vector<std::unique_ptr<cusna_period_plan>> plans;
  DeviceBlock D;
  int M = 0;  // chains allocated
  saom::SimParams* groups = nullptr;
  uint32_t* items = nullptr;
  PeriodOut* per = nullptr;
  double *S = nullptr, *SP = nullptr, *SC = nullptr, *bstats = nullptr, *mean = nullptr;
  bool specialize = false;  // runtime-specialised period kernel (same effects in every period)
};

struct cusna_mom_engine {
  int P = 0, k = 0, kb = 0, p = 0, mode = 0, n = 0, Mmax = 0, n_rates = 0, W64 = 0;
  std::chrono::steady_clock::time_point t0;  // start of the estimate (progress lines)
  SimSet main;
  DeviceBlock D;
  hipStream_t s = nullptr;
  double *theta = nullptr, *theta_sim = nullptr, *thav = nullptr, *dinv = nullptr;
  double *targets = nullptr, *rate_consts = nullptr;
  double *mSP = nullptr, *mSC = nullptr, *Dm = nullptr, *msf = nullptr, *var = nullptr;
  double* bscratch = nullptr;  // behaviour statistics above the LDS (else null)
  size_t bscratch_doubles = 0;
  uint64_t* xnet = nullptr;
  int32_t* bkdist = nullptr;
  uint32_t* keys = nullptr;
  unsigned long long* counter = nullptr;
  uint64_t seed = 0;
  long long n_sims = 0;
  int specialize = -1;  // 1 on, 0 off, -1 by the launch-size policy
  long long work = 0;   // chains x periods x actors of the whole schedule
  double beh_zbar = 0.0, beh_brange = 1.0, beh_bsim = 0.0;
  ~cusna_mom_engine() {
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

// Plans of every period for M chains. theta_base: the float32-rounded
// parameter rows the kernels read (one shared row when row_stride is 0, else
// row c at c * row_stride), laid out as the theta vector plus the constant
// internal rate 1 in slot p.
cusna_status build_simset(cusna_mom_engine* E, SimSet& X, const cusna_saom_problem* const* periods,
                          int M, double* theta_base, int row_stride) {
  const int P = E->P, p = E->p, k = E->k, kb = E->kb, mode = E->mode;
  hipStream_t s = E->s;
  X.M = M;
  DeviceBlock& D = X.D;
  const size_t o_S = D.add((size_t)M * p * 8), o_SP = D.add((size_t)P * M * p * 8);
  const size_t o_SC = D.add((size_t)P * M * p * 8), o_mean = D.add(p * 8);
  const size_t o_bst = D.add((size_t)P * M * (kb > 0 ? kb : 1) * 8);
  const size_t o_per = D.add(P * sizeof(PeriodOut));
  const size_t o_grp = D.add(P * sizeof(saom::SimParams)), o_items = D.add((size_t)P * M * 4);
  HIP_TRY(D.alloc());
  X.S = D.at<double>(o_S);
  X.SP = D.at<double>(o_SP);
  X.SC = D.at<double>(o_SC);
  X.mean = D.at<double>(o_mean);
  X.bstats = D.at<double>(o_bst);
  X.per = D.at<PeriodOut>(o_per);
  X.groups = D.at<saom::SimParams>(o_grp);
  X.items = D.at<uint32_t>(o_items);
  std::vector<PeriodOut> per(P);
  std::vector<saom::SimParams> grp(P);
  for (int pp = 0; pp < P; ++pp) {
    const cusna_saom_problem* pb = periods[pp];
    cusna_saom_problem q = *pb;
    q.M = M;
    std::vector<float> thf((size_t)M * k, 0.0f), lmf(M, 1.0f),
        thbf((size_t)M * (kb > 0 ? kb : 1), 0.0f), lmbf(M, 1.0f);
    q.theta = thf.data();
    q.lam = lmf.data();
    if (kb > 0) {
      q.thetab = thbf.data();
      q.lamb = lmbf.data();
    }
    q.conditional = mode == 1 ? 1 : 0;
    cusna_period_plan* raw = nullptr;
    cusna_status st = plan_create(&q, false, nullptr, nullptr, 0, s, &raw);
    if (st != CUSNA_OK) return st;
    X.plans.emplace_back(raw);
    saom::SimParams& SPm = raw->P;
    const int off = mode == 2 ? 2 * P : (mode == 1 ? 0 : P);
    SPm.theta = theta_base + off;
    SPm.th_stride = row_stride;
    SPm.lam = mode == 1 ? theta_base + p : theta_base + pp;
    SPm.lam_stride = row_stride;
    if (mode == 2) {
      SPm.lamb = theta_base + P + pp;
      SPm.thetab = theta_base + 2 * P + k;
      SPm.thb_stride = row_stride;
    }
    SPm.keys = E->keys + 2 * pp;
    HIP_TRY(hipMemcpyAsync(raw->D.at<char>(raw->o_params), &SPm, sizeof(SPm),
                           hipMemcpyHostToDevice, s));
    per[pp].dist_stat = SPm.dist_stat;
    per[pp].scores = SPm.scores;
    per[pp].stats = raw->D.at<double>(raw->o_stats);
    per[pp].bdist = SPm.bdist;
    per[pp].bscores = SPm.bscores;
    per[pp].bstats = X.bstats + (size_t)pp * M * (kb > 0 ? kb : 1);
    if (raw->nw != X.plans[0]->nw || raw->wmax != X.plans[0]->wmax) {
      set_error("cusna_mom_estimate: periods need the same launch shape");
      return CUSNA_ERR_SHAPE;
    }
    grp[pp] = SPm;
  }
  std::vector<uint32_t> items((size_t)P * M);
  for (size_t b = 0; b < items.size(); ++b)
    items[b] = ((uint32_t)(b % P) << 20) | (uint32_t)(b / P);
  HIP_TRY(hipMemcpyAsync(X.per, per.data(), P * sizeof(PeriodOut), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(X.groups, grp.data(), P * sizeof(saom::SimParams),
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(X.items, items.data(), items.size() * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  bool uniform = true;
  for (int pp = 1; pp < P; ++pp) uniform = uniform && saom::same_effects(grp[pp], grp[0]);
  X.specialize = uniform && (E->specialize > 0 ||
                             (E->specialize < 0 &&
                              saom::specialize_policy(std::max((long long)P * M * E->n,
                                                               E->work / 16))));
  return CUSNA_OK;
}

// One simulation round of M chains over every period of X: fresh keys,
// one period launch, statistics, moment rows (X.S, X.SP, X.SC), means.
cusna_status engine_simulate(cusna_mom_engine* E, SimSet& X, int M) {
  hipStream_t s = E->s;
  hipLaunchKernelGGL(mom_seed_kernel, dim3(1), dim3(64), 0, s, E->keys, E->counter, E->seed,
                     E->P);
  {
    cusna_period_plan* p0 = X.plans[0].get();
    saom::PeriodLaunch L;
    L.Pd = X.groups;
    L.items = X.items;
    L.grid = E->P * M;
    L.lds = saom::sim_lds_bytes(p0->P, p0->nw, p0->wmax);
    L.key0 = L.key1 = 0u;
    L.Ph = &p0->P;
    L.specialize = X.specialize;
    const saom::LaunchShape gen{p0->nw, p0->wmax};
    const int kf = p0->free_shape ? p0->P.k : 0;
    L.exact = saom::exact_shape(p0->W64, gen, kf, L.lds);
    L.lds_exact = saom::sim_lds_bytes(p0->P, L.exact.nw, L.exact.wmax);
    L.exact_alt = saom::exact_shape(p0->W64, gen, kf);
    L.lds_exact_alt = saom::sim_lds_bytes(p0->P, L.exact_alt.nw, L.exact_alt.wmax);
    for (int pp = 1; pp < E->P; ++pp) L.same_n = L.same_n && X.plans[pp]->n == p0->n;
    const int le = saom::launch_period_ex(L, saom::LaunchShape{p0->nw, p0->wmax}, false,
                                          saom::simple_effects(p0->P.tid, p0->P.k), s);
    if (le != 0) {
      set_error(std::string("period kernel launch: ") + hipGetErrorString((hipError_t)le));
      return CUSNA_ERR_CUDA;
    }
  }
  for (int pp = 0; pp < E->P; ++pp) {
    cusna_period_plan* pl = X.plans[pp].get();
    cusna_status st = plan_stats(pl, M, s);
    if (st != CUSNA_OK) return st;
    if (E->mode == 2) {
      saom::BehStatParams BP;
      std::memset(&BP, 0, sizeof(BP));
      BP.n = E->n;
      BP.M = M;
      BP.kb = E->kb;
      BP.W64 = E->W64;
      for (int e = 0; e < E->kb; ++e) BP.bty[e] = pl->P.bty[e];
      BP.z = pl->P.z_end;
      BP.xnet = E->xnet + (size_t)pp * E->n * E->W64;
      BP.bkdist = E->bkdist + (size_t)pp * E->n;
      BP.zbar = E->beh_zbar;
      BP.brange = E->beh_brange;
      BP.bsim = E->beh_bsim;
      BP.out = X.bstats + (size_t)pp * X.M * E->kb;
      BP.scratch = E->bscratch;
      if (E->bscratch_doubles > 0 && M > E->Mmax) {
        set_error("cusna_mom_estimate: behaviour scratch smaller than the round");
        return CUSNA_ERR_INTERNAL;
      }
      if (saom::launch_beh_stats(BP, s)) return CUSNA_ERR_CUDA;
    }
  }
  GatherArgs G{M, E->P, E->p, E->k, E->kb, E->mode, X.per, E->rate_consts, X.S, X.SP, X.SC};
  hipLaunchKernelGGL(mom_gather_kernel, dim3((M + 127) / 128), dim3(128), 0, s, G);
  hipLaunchKernelGGL(mom_mean_kernel, dim3(E->p), dim3(kMomThreads), 0, s, X.S, M, E->p, X.mean);
  E->n_sims += M;
  return hipGetLastError() == hipSuccess ? CUSNA_OK : CUSNA_ERR_CUDA;
}

// D = sum over periods of crossprod(centred sp, centred scp) / M (mom.R:173-182)
void engine_derivative(cusna_mom_engine* E, SimSet& X, int M) {
  hipStream_t s = E->s;
  const int p = E->p;
  for (int pp = 0; pp < E->P; ++pp) {
    const double* sp = X.SP + (size_t)pp * M * p;  // gather's layout: [period][chain][a]
    const double* sc = X.SC + (size_t)pp * M * p;
    hipLaunchKernelGGL(mom_mean_kernel, dim3(p), dim3(kMomThreads), 0, s, sp, M, p, E->mSP);
    hipLaunchKernelGGL(mom_mean_kernel, dim3(p), dim3(kMomThreads), 0, s, sc, M, p, E->mSC);
    hipLaunchKernelGGL(mom_cov_kernel, dim3(p * p), dim3(kMomThreads), 0, s, sp, E->mSP, sc,
                       E->mSC, M, p, (double)M, pp > 0 ? 1 : 0, E->Dm);
  }
}

struct GraphHolder {  // a captured phase-2 iteration and its executable
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  ~GraphHolder() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
  }
};

cusna_status fetch(cusna_mom_engine* E, double* dst, const double* src, size_t n) {
  HIP_TRY(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, E->s));
  HIP_TRY(hipStreamSynchronize(E->s));
  return CUSNA_OK;
}

Mat dinv_of(const Mat& Dm, int p, double diag) {
  Mat mix(Dm);
  for (int a = 0; a < p; ++a)
    for (int b = 0; b < p; ++b)
      mix[(size_t)a * p + b] =
          (1.0 - diag) * Dm[(size_t)a * p + b] + (a == b ? diag * Dm[(size_t)a * p + b] : 0.0);
  return safe_solve(mix, p);
}

// sqrt(|diag(A^-1 S A^-T)|)
std::vector<double> sandwich_se(const Mat& A, const Mat& S, int p) {
  const Mat Ai = safe_solve(A, p);
  std::vector<double> se(p);
  for (int a = 0; a < p; ++a) {
    double v = 0.0;
    for (int b = 0; b < p; ++b)
      for (int c = 0; c < p; ++c)
        v += Ai[(size_t)a * p + b] * S[(size_t)b * p + c] * Ai[(size_t)a * p + c];
    se[a] = std::sqrt(std::fabs(v));
  }
  return se;
}

// Surrogate phase 2. Box centre c and half-widths h on the host. Each
// training round simulates the box, trains the perceptron and solves
// E s(theta) = target on its mean (the next round's box re-centres there and,
// when the solution is inside, shrinks to two standard errors). Then up to
// `corrections` simulation-checked Newton steps: nD chains at theta give the
// residual mean(s) - target and the score-function derivative, the step is
// guarded as the Robbins-Monro steps are; they stop once every |residual| is
// below conv_tol moment standard deviations. Leaves theta in the engine.
// Seconds since the estimate started (prefix of the progress lines).
double elapsed(const cusna_mom_engine* E) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - E->t0).count();
}

// A progress line on stderr (cusna_mom_options.verbose): the tag and the
// current theta, as the R estimator's verbose output (mom.R:230, 246).
cusna_status log_theta(cusna_mom_engine* E, const char* tag) {
  std::vector<double> th(E->p);
  cusna_status st = fetch(E, th.data(), E->theta, E->p);
  if (st != CUSNA_OK) return st;
  std::fprintf(stderr, "[%.2f s] %s theta =", elapsed(E), tag);
  for (double v : th) std::fprintf(stderr, " %.3f", v);
  std::fprintf(stderr, "\n");
  std::fflush(stderr);
  return CUSNA_OK;
}

// theta on the device: float64 and the float32-rounded copy the kernels read
// (with the trailing 1 of the per-chain layout).
cusna_status put_theta(cusna_mom_engine* E, const std::vector<double>& theta) {
  const int p = E->p;
  std::vector<double> ts(p + 1);
  for (int a = 0; a < p; ++a) ts[a] = (double)(float)theta[a];
  ts[p] = 1.0;
  HIP_TRY(hipMemcpyAsync(E->theta, theta.data(), p * 8, hipMemcpyHostToDevice, E->s));
  HIP_TRY(hipMemcpyAsync(E->theta_sim, ts.data(), (p + 1) * 8, hipMemcpyHostToDevice, E->s));
  HIP_TRY(hipStreamSynchronize(E->s));  // the host vectors go out of scope
  return CUSNA_OK;
}

// One Newton step theta -= D^-1 (sbar - t), each coordinate capped at
// max(1, |theta| / 2) and, given standard errors, at two of them; rates kept
// above a tenth of their value.
void newton_step(std::vector<double>& theta, const Mat& Dm, const std::vector<double>& sbar,
                 const std::vector<double>& tgt, int p, int n_rates,
                 const std::vector<double>* se = nullptr) {
  const Mat Di = safe_solve(Dm, p);
  for (int a = 0; a < p; ++a) {
    double stp = 0.0;
    for (int b = 0; b < p; ++b) stp += Di[(size_t)a * p + b] * (sbar[b] - tgt[b]);
    double cap = std::max(1.0, 0.5 * std::fabs(theta[a]));
    if (se && (*se)[a] > 0.0) cap = std::min(cap, 2.0 * (*se)[a]);
    stp = std::min(cap, std::max(-cap, stp));
    if (a < n_rates && theta[a] - stp < 0.1 * theta[a]) stp = 0.5 * theta[a];
    theta[a] -= stp;
  }
}

// Surrogate phase 2; *worst_out: the smallest largest residual (in moment
// standard deviations) the corrections measured (the caller falls back to
// the plain sub-phases above kSurrogateFallback).
constexpr double kSurrogateFallback = 3.0;
// Phase 1 further than this (largest moment deviation, in standard deviations)
// from the targets: its linearisation cannot centre the surrogate's box, and
// the automatic warm start runs one plain sub-phase first.
constexpr double kSurrogateWarmDev = 50.0;

cusna_status surrogate_phase(cusna_mom_engine* E, const cusna_saom_problem* const* periods,
                             const cusna_mom_options* opt, std::vector<double> c,
                             std::vector<double> h, const std::vector<double>& tgt,
                             double* worst_out) {
  const int p = E->p;
  hipStream_t s = E->s;
  const int q = opt->surrogate_pairs_log2 > 0 ? opt->surrogate_pairs_log2 : 13;
  if (q > 20 || ((long long)E->P << q) > (1ll << 30)) {
    set_error("cusna_mom_estimate: surrogate_pairs_log2 too large");
    return CUSNA_ERR_SHAPE;
  }
  const int Mtr = 1 << q;
  const int steps = opt->surrogate_steps > 0 ? opt->surrogate_steps : 200;
  const int rounds = opt->surrogate_rounds > 0 ? opt->surrogate_rounds : 1;
  const int corrections = opt->surrogate_corrections >= 0 ? opt->surrogate_corrections : 6;
  if (E->bscratch_doubles > 0 && Mtr > E->Mmax) {
    set_error("cusna_mom_estimate: surrogate rounds above the behaviour scratch");
    return CUSNA_ERR_UNSUPPORTED;
  }
  DeviceBlock D;
  const size_t o_th = D.add((size_t)Mtr * (p + 1) * 8), o_x = D.add((size_t)Mtr * p * 4);
  const size_t o_y = D.add((size_t)Mtr * p * 4), o_c = D.add(p * 8), o_h = D.add(p * 8);
  HIP_TRY(D.alloc());
  double* th_tr = D.at<double>(o_th);
  float* x_tr = D.at<float>(o_x);
  float* y_tr = D.at<float>(o_y);
  SimSet T;
  cusna_status st = build_simset(E, T, periods, Mtr, th_tr, p + 1);
  if (st != CUSNA_OK) return st;
  Mlp net;
  if ((st = net.init(p, p, 256, 1024, opt->seed ^ 0x5u, s)) != CUSNA_OK) return st;
  std::vector<float> flat(net.param_count());
  HostNet H;
  std::vector<double> mean(p), var(p), sd(p), tn(p), x, Jn, lv;
  for (int r = 0; r < rounds; ++r) {
    HIP_TRY(hipMemcpyAsync(D.at<char>(o_c), c.data(), p * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(D.at<char>(o_h), h.data(), p * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(sample_theta_kernel, dim3((Mtr + 255) / 256), dim3(256), 0, s,
                       D.at<double>(o_c), D.at<double>(o_h), p, Mtr, opt->seed, (unsigned)r,
                       E->n_rates, x_tr, th_tr);
    if ((st = engine_simulate(E, T, Mtr)) != CUSNA_OK) return st;
    hipLaunchKernelGGL(mom_var_kernel, dim3(p), dim3(kMomThreads), 0, s, T.S, Mtr, p, T.mean,
                       E->var);
    hipLaunchKernelGGL(normalize_kernel, dim3((unsigned)(((size_t)Mtr * p + 255) / 256)),
                       dim3(256), 0, s, T.S, Mtr, p, T.mean, E->var, y_tr);
    if ((st = fetch(E, mean.data(), T.mean, p)) != CUSNA_OK) return st;
    if ((st = fetch(E, var.data(), E->var, p)) != CUSNA_OK) return st;
    if ((st = net.reset(opt->seed + 17u * (uint64_t)r)) != CUSNA_OK) return st;
    if ((st = net.train(x_tr, y_tr, Mtr, steps, 2e-3f, opt->seed + 1009u * (uint64_t)r)) !=
        CUSNA_OK)
      return st;
    if ((st = net.get_params(flat.data())) != CUSNA_OK) return st;
    H.load(net, flat, p, 256);
    for (int a = 0; a < p; ++a) {
      sd[a] = var[a] > 0.0 ? std::sqrt(var[a]) : 1.0;
      tn[a] = (tgt[a] - mean[a]) / sd[a];
    }
    surrogate_newton(H, tn, x, Jn, lv);
    if (opt->verbose) {
      std::fprintf(stderr, "[%.2f s] [surrogate %d] %d pairs simulated, %d steps trained\n",
                   elapsed(E), r + 1, Mtr, steps);
      std::fflush(stderr);
    }
    bool interior = true;
    for (int a = 0; a < p; ++a) {
      c[a] = c[a] + h[a] * x[a];
      if (std::fabs(x[a]) > 0.9) interior = false;
    }
    if (interior && r + 1 < rounds) {
      // J in theta units and the predicted moment variances (diagonal)
      Mat J((size_t)p * p), Sg((size_t)p * p, 0.0);
      for (int a = 0; a < p; ++a) {
        for (int d = 0; d < p; ++d) J[(size_t)a * p + d] = sd[a] * Jn[(size_t)a * p + d] / h[d];
        Sg[(size_t)a * p + a] = std::exp(std::min(12.0, std::max(-12.0, lv[a]))) * var[a];
      }
      const std::vector<double> se = sandwich_se(J, Sg, p);
      for (int a = 0; a < p; ++a)
        h[a] = std::min(h[a], std::max(2.0 * se[a], 1e-3 * (1.0 + std::fabs(c[a]))));
    }
  }
  // simulation-checked Newton corrections: residual and score-function
  // derivative from nD plain chains at theta
  std::vector<double> theta = c, sbar(p), msf((size_t)p * p), Dm((size_t)p * p);
  if ((st = put_theta(E, theta)) != CUSNA_OK) return st;
  SimSet& X = E->main;
  double best = 1e300;
  for (int it = 0; it < corrections; ++it) {
    if ((st = engine_simulate(E, X, opt->nD)) != CUSNA_OK) return st;
    engine_derivative(E, X, opt->nD);
    hipLaunchKernelGGL(mom_cov_kernel, dim3(p * p), dim3(kMomThreads), 0, s, X.S, X.mean, X.S,
                       X.mean, opt->nD, p, (double)(opt->nD - 1), 0, E->msf);
    if ((st = fetch(E, sbar.data(), X.mean, p)) != CUSNA_OK) return st;
    if ((st = fetch(E, msf.data(), E->msf, (size_t)p * p)) != CUSNA_OK) return st;
    if ((st = fetch(E, Dm.data(), E->Dm, (size_t)p * p)) != CUSNA_OK) return st;
    double worst = 0.0;
    for (int a = 0; a < p; ++a) {
      const double v = msf[(size_t)a * p + a];
      worst = std::max(worst, std::fabs(sbar[a] - tgt[a]) / (v > 0.0 ? std::sqrt(v) : 1.0));
    }
    if (opt->verbose) {
      std::fprintf(stderr, "[%.2f s] [correction %d] worst residual %.4f sd\n", elapsed(E),
                   it + 1, worst);
      std::fflush(stderr);
    }
    best = std::min(best, worst);
    if (worst <= opt->conv_tol) break;
    newton_step(theta, Dm, sbar, tgt, p, E->n_rates);
    if ((st = put_theta(E, theta)) != CUSNA_OK) return st;
  }
  if (corrections == 0) best = 0.0;  // nothing measured: no fallback
  *worst_out = best;
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

}  // namespace

extern "C" {

CUSNA_API cusna_status cusna_mom_estimate(const cusna_saom_problem* const* periods,
                                          const uint8_t* const* xnet, const double* targets,
                                          const double* rate_consts, const double* theta0,
                                          const cusna_mom_options* opt, cusna_mom_result* res) {
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  if (!periods || !targets || !theta0 || !opt || !res) return CUSNA_ERR_NULL_ARG;
  if (opt->struct_version != CUSNA_ABI_VERSION || res->struct_version != CUSNA_ABI_VERSION)
    return CUSNA_ERR_SHAPE;
  const int P = opt->P, mode = opt->mode;
  if (P < 1 || P > 16 || mode < 0 || mode > 2 || opt->nsub < 0 || opt->nsub > 8)
    return CUSNA_ERR_SHAPE;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  const cusna_saom_problem* p0 = periods[0];
  if (!p0) return CUSNA_ERR_NULL_ARG;
  const int k = p0->k, kb = mode == 2 ? p0->kb : 0, n = p0->n;
  const int p = mode == 2 ? 2 * P + k + kb : (mode == 1 ? k : P + k);
  if (p > 64) {
    set_error("cusna_mom_estimate: more than 64 parameters");
    return CUSNA_ERR_SHAPE;
  }
  if (mode == 2 && (kb < 1 || !xnet)) return CUSNA_ERR_NULL_ARG;
  if (opt->n1 < 2 || opt->nD < 2 || opt->batch2 < 1 || opt->n3 < 2) {
    set_error("cusna_mom_estimate: n1, nD and n3 must be at least 2, batch2 at least 1");
    return CUSNA_ERR_SHAPE;
  }
  for (int sub = 0; sub < opt->nsub; ++sub)
    if (opt->n2[sub] < 0) {
      set_error("cusna_mom_estimate: negative n2");
      return CUSNA_ERR_SHAPE;
    }
  const int Mmax = std::max(std::max(opt->n1, opt->n3), std::max(opt->nD, opt->batch2));
  if (Mmax > (1 << 20)) {
    set_error("cusna_mom_estimate: at most 2^20 chains per simulation round");
    return CUSNA_ERR_SHAPE;
  }
  for (int pp = 0; pp < P; ++pp) {
    const cusna_saom_problem* pb = periods[pp];
    if (!pb || pb->n != n || pb->k != k || (mode == 2 && pb->kb != kb)) return CUSNA_ERR_SHAPE;
  }

  std::unique_ptr<cusna_mom_engine> E(new cusna_mom_engine());
  E->t0 = t_start;
  E->P = P;
  E->k = k;
  E->kb = kb;
  E->p = p;
  E->mode = mode;
  E->n = n;
  E->W64 = (n + 63) / 64;
  E->Mmax = Mmax;
  E->n_rates = mode == 1 ? 0 : (mode == 2 ? 2 * P : P);
  E->seed = opt->seed;
  E->specialize = opt->specialize > 0 ? 1 : (opt->specialize < 0 ? 0 : -1);
  {
    long long chains = (long long)opt->n1 + opt->n3;
    for (int sub = 0; sub < opt->nsub; ++sub)
      chains += opt->nD + (long long)opt->n2[sub] * opt->batch2;
    E->work = chains * P * n;
  }
  E->beh_zbar = opt->beh_zbar;
  E->beh_brange = opt->beh_brange;
  E->beh_bsim = opt->beh_bsimmean;
  HIP_TRY(hipStreamCreateWithFlags(&E->s, hipStreamNonBlocking));
  const int W64 = E->W64;
  DeviceBlock& D = E->D;
  const size_t o_theta = D.add(p * 8), o_tsim = D.add((p + 1) * 8), o_thav = D.add(p * 8);
  const size_t o_dinv = D.add((size_t)p * p * 8), o_tg = D.add(p * 8), o_rc = D.add(P * 8);
  const size_t o_mSP = D.add(p * 8), o_mSC = D.add(p * 8), o_Dm = D.add((size_t)p * p * 8);
  const size_t o_msf = D.add((size_t)p * p * 8), o_var = D.add(p * 8);
  const size_t o_xn = D.add((size_t)P * n * W64 * 8), o_bk = D.add((size_t)P * n * 4);
  const size_t o_keys = D.add(P * 8), o_cnt = D.add(8);
  E->bscratch_doubles = mode == 2 ? saom::beh_stats_scratch_doubles(n) : 0;
  const size_t o_bscr = D.add((size_t)Mmax * E->bscratch_doubles * 8);
  HIP_TRY(D.alloc());
  E->theta = D.at<double>(o_theta);
  E->theta_sim = D.at<double>(o_tsim);
  E->thav = D.at<double>(o_thav);
  E->dinv = D.at<double>(o_dinv);
  E->targets = D.at<double>(o_tg);
  E->rate_consts = D.at<double>(o_rc);
  E->mSP = D.at<double>(o_mSP);
  E->mSC = D.at<double>(o_mSC);
  E->Dm = D.at<double>(o_Dm);
  E->msf = D.at<double>(o_msf);
  E->var = D.at<double>(o_var);
  E->xnet = D.at<uint64_t>(o_xn);
  E->bkdist = D.at<int32_t>(o_bk);
  E->keys = D.at<uint32_t>(o_keys);
  E->counter = D.at<unsigned long long>(o_cnt);
  E->bscratch = E->bscratch_doubles ? D.at<double>(o_bscr) : nullptr;
  hipStream_t s = E->s;

  std::vector<double> rc(P, 0.0);
  if (rate_consts)
    for (int q = 0; q < P; ++q) rc[q] = rate_consts[q];
  std::vector<double> th0(theta0, theta0 + p), zeros((size_t)p * p, 0.0);
  std::vector<double> tsim(p + 1);
  for (int a = 0; a < p; ++a) tsim[a] = (double)(float)th0[a];
  tsim[p] = 1.0;
  const unsigned long long zero = 0ull;
  HIP_TRY(hipMemcpyAsync(E->theta, th0.data(), p * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(E->theta_sim, tsim.data(), (p + 1) * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(E->thav, zeros.data(), p * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(E->targets, targets, p * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(E->rate_consts, rc.data(), P * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(E->counter, &zero, 8, hipMemcpyHostToDevice, s));
  if (mode == 2) {
    for (int pp = 0; pp < P; ++pp) {
      std::vector<uint64_t> xr((size_t)n * W64, 0ull);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          if (xnet[pp][(size_t)i * n + j]) xr[(size_t)i * W64 + (j >> 6)] |= 1ull << (j & 63);
      HIP_TRY(hipMemcpyAsync(E->xnet + (size_t)pp * n * W64, xr.data(), xr.size() * 8,
                             hipMemcpyHostToDevice, s));
      std::vector<int32_t> bk(n, 1);
      if (periods[pp]->bkdist)
        for (int i = 0; i < n; ++i) bk[i] = periods[pp]->bkdist[i];
      HIP_TRY(hipMemcpyAsync(E->bkdist + (size_t)pp * n, bk.data(), n * 4, hipMemcpyHostToDevice,
                             s));
      HIP_TRY(hipStreamSynchronize(s));
    }
  }
  // the estimation rounds read the engine's float32-rounded theta (one shared
  // row) and the Philox keys from device memory
  cusna_status st = build_simset(E.get(), E->main, periods, Mmax, E->theta_sim, 0);
  if (st != CUSNA_OK) return st;
  SimSet& X = E->main;

  Mat Dm((size_t)p * p);
  std::vector<double> tgt(targets, targets + p);
  auto upload_dinv = [&](const Mat& di) -> cusna_status {
    HIP_TRY(hipMemcpyAsync(E->dinv, di.data(), (size_t)p * p * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return CUSNA_OK;
  };
  auto t_phase = clk::now();
  double phase_t[3] = {0, 0, 0};

  // ---- phase 1 (mom.R:224-229)
  if ((st = engine_simulate(E.get(), X, opt->n1)) != CUSNA_OK) return st;
  engine_derivative(E.get(), X, opt->n1);
  if ((st = fetch(E.get(), Dm.data(), E->Dm, (size_t)p * p)) != CUSNA_OK) return st;
  if ((st = upload_dinv(dinv_of(Dm, p, opt->diagonalize))) != CUSNA_OK) return st;
  hipLaunchKernelGGL(mom_update_kernel, dim3(1), dim3(64), 0, s, E->theta, E->theta_sim,
                     E->thav, E->dinv, X.mean, E->targets, p, 0.5 * opt->firstg, E->n_rates, 0);
  phase_t[0] = std::chrono::duration<double>(clk::now() - t_phase).count();
  if (opt->verbose && (st = log_theta(E.get(), "[phase 1]")) != CUSNA_OK) return st;
  t_phase = clk::now();

  // ---- phase 2 (mom.R:232-248), sub-phases replayed from a captured graph
  auto subphase = [&](double gain, int n2) -> cusna_status {
    cusna_status st2 = engine_simulate(E.get(), X, opt->nD);
    if (st2 != CUSNA_OK) return st2;
    engine_derivative(E.get(), X, opt->nD);
    if ((st2 = fetch(E.get(), Dm.data(), E->Dm, (size_t)p * p)) != CUSNA_OK) return st2;
    if ((st2 = upload_dinv(dinv_of(Dm, p, opt->diagonalize))) != CUSNA_OK) return st2;
    if (n2 < 1) return CUSNA_OK;
    auto iteration = [&]() -> cusna_status {
      cusna_status st3 = engine_simulate(E.get(), X, opt->batch2);
      if (st3 != CUSNA_OK) return st3;
      hipLaunchKernelGGL(mom_update_kernel, dim3(1), dim3(64), 0, s, E->theta, E->theta_sim,
                         E->thav, E->dinv, X.mean, E->targets, p, gain, E->n_rates, 1);
      return hipGetLastError() == hipSuccess ? CUSNA_OK : CUSNA_ERR_CUDA;
    };
    if (opt->use_graph) {
      GraphHolder G;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const long long before = E->n_sims;
      cusna_status cst = iteration();
      hipError_t ce = hipStreamEndCapture(s, &G.g);
      if (cst != CUSNA_OK) return cst;
      if (ce != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        return CUSNA_ERR_CUDA;
      }
      E->n_sims = before;
      HIP_TRY(hipGraphInstantiate(&G.ge, G.g, nullptr, nullptr, 0));
      for (int it = 0; it < n2; ++it) {
        HIP_TRY(hipGraphLaunch(G.ge, s));
        E->n_sims += opt->batch2;
      }
      HIP_TRY(hipStreamSynchronize(s));
    } else {
      for (int it = 0; it < n2; ++it) {
        cusna_status st3 = iteration();
        if (st3 != CUSNA_OK) return st3;
      }
    }
    hipLaunchKernelGGL(mom_average_kernel, dim3(1), dim3(64), 0, s, E->theta, E->theta_sim,
                       E->thav, p, (double)n2);
    return CUSNA_OK;
  };
  auto run_sub = [&](int sub) -> cusna_status {
    const double gain = opt->firstg * std::pow(0.5, sub);
    cusna_status sst = subphase(gain, opt->n2[sub]);
    if (sst != CUSNA_OK) return sst;
    if (opt->verbose) {
      char tag[64];
      std::snprintf(tag, sizeof tag, "[phase 2.%d] gain=%.3f", sub + 1, gain);
      return log_theta(E.get(), tag);
    }
    return CUSNA_OK;
  };
  if (opt->surrogate) {
    // plain sub-phases first when asked; then the box: Newton point of the
    // moments and derivative at the current theta (phase 1's, or nD chains at
    // the sub-phases' theta), half-widths covering the step and three
    // standard errors
    // phase 1's moments and their covariance (the rows phase 1 left in X)
    std::vector<double> m1(p), msf1((size_t)p * p);
    auto moments = [&](int nb) -> cusna_status {
      hipLaunchKernelGGL(mom_cov_kernel, dim3(p * p), dim3(kMomThreads), 0, s, X.S, X.mean,
                         X.S, X.mean, nb, p, (double)(nb - 1), 0, E->msf);
      cusna_status mst = fetch(E.get(), m1.data(), X.mean, p);
      if (mst != CUSNA_OK) return mst;
      return fetch(E.get(), msf1.data(), E->msf, (size_t)p * p);
    };
    if ((st = moments(opt->n1)) != CUSNA_OK) return st;
    double dev1 = 0.0;  // largest deviation from the targets, in moment standard deviations
    for (int a = 0; a < p; ++a) {
      const double v = msf1[(size_t)a * p + a];
      dev1 = std::max(dev1, std::fabs(m1[a] - tgt[a]) / (v > 0.0 ? std::sqrt(v) : 1.0));
    }
    int warm = opt->surrogate_warmup;
    if (warm == 0) warm = dev1 > kSurrogateWarmDev ? 1 : 0;  // automatic
    warm = std::min(std::max(warm, 0), opt->nsub);
    if (opt->verbose) {
      std::fprintf(stderr, "[%.2f s] [surrogate] phase 1 %.1f sd from the targets: %d plain "
                   "sub-phase(s) first\n", elapsed(E.get()), dev1, warm);
      std::fflush(stderr);
    }
    for (int sub = 0; sub < warm; ++sub)
      if ((st = run_sub(sub)) != CUSNA_OK) return st;
    std::vector<double> base = th0;
    if (warm > 0) {
      if ((st = fetch(E.get(), base.data(), E->theta, p)) != CUSNA_OK) return st;
      if ((st = engine_simulate(E.get(), X, opt->nD)) != CUSNA_OK) return st;
      engine_derivative(E.get(), X, opt->nD);
      if ((st = fetch(E.get(), Dm.data(), E->Dm, (size_t)p * p)) != CUSNA_OK) return st;
      if ((st = moments(opt->nD)) != CUSNA_OK) return st;
    }
    const Mat D1i = safe_solve(Dm, p);
    std::vector<double> c(p), h(p);
    const std::vector<double> se1 = sandwich_se(Dm, msf1, p);
    for (int a = 0; a < p; ++a) {
      double step = 0.0;
      for (int b = 0; b < p; ++b) step += D1i[(size_t)a * p + b] * (m1[b] - tgt[b]);
      const double cap = 2.0 * std::max(1.0, std::fabs(base[a]));
      step = std::min(cap, std::max(-cap, step));
      c[a] = base[a] - step;
      if (a < E->n_rates && c[a] < 0.2 * base[a]) c[a] = 0.2 * base[a];
      h[a] = std::max(std::max(2.0 * std::fabs(c[a] - base[a]), 3.0 * se1[a]),
                      1e-3 * (1.0 + std::fabs(c[a])));
    }
    std::vector<double> start(p);
    if ((st = fetch(E.get(), start.data(), E->theta, p)) != CUSNA_OK) return st;
    double worst = 0.0;
    if ((st = surrogate_phase(E.get(), periods, opt, c, h, tgt, &worst)) != CUSNA_OK) return st;
    if (worst > kSurrogateFallback) {
      // the plain estimator is the fallback, from the theta the surrogate
      // started at (a failed surrogate's points can be far off)
      if ((st = put_theta(E.get(), start)) != CUSNA_OK) return st;
      if (opt->verbose) {
        std::fprintf(stderr, "[%.2f s] [surrogate] residual %.2f sd: plain sub-phases %d..%d\n",
                     elapsed(E.get()), worst, warm + 1, opt->nsub);
        std::fflush(stderr);
      }
      for (int sub = warm; sub < opt->nsub; ++sub)
        if ((st = run_sub(sub)) != CUSNA_OK) return st;
    }
  } else {
    for (int sub = 0; sub < opt->nsub; ++sub)
      if ((st = run_sub(sub)) != CUSNA_OK) return st;
  }
  phase_t[1] = std::chrono::duration<double>(clk::now() - t_phase).count();

  // ---- phase 3 (mom.R:250-258), optionally repeated until max |t| <= tol
  std::vector<double> sim_means(p), tconv(p), se(p);
  double tconv_max = 0.0, max_abs_t = 0.0;
  Mat msf((size_t)p * p);
  int repeats = 0;
  std::vector<double> theta_h(p);
  for (;;) {
    t_phase = clk::now();
    if ((st = engine_simulate(E.get(), X, opt->n3)) != CUSNA_OK) return st;
    engine_derivative(E.get(), X, opt->n3);
    hipLaunchKernelGGL(mom_cov_kernel, dim3(p * p), dim3(kMomThreads), 0, s, X.S, X.mean, X.S,
                       X.mean, opt->n3, p, (double)(opt->n3 - 1), 0, E->msf);
    if ((st = fetch(E.get(), Dm.data(), E->Dm, (size_t)p * p)) != CUSNA_OK) return st;
    if ((st = fetch(E.get(), msf.data(), E->msf, (size_t)p * p)) != CUSNA_OK) return st;
    if ((st = fetch(E.get(), sim_means.data(), X.mean, p)) != CUSNA_OK) return st;
    if ((st = fetch(E.get(), theta_h.data(), E->theta, p)) != CUSNA_OK) return st;
    std::vector<double> dev(p);
    max_abs_t = 0.0;
    for (int a = 0; a < p; ++a) {
      dev[a] = sim_means[a] - tgt[a];
      tconv[a] = dev[a] / std::sqrt(msf[(size_t)a * p + a]);
      max_abs_t = std::max(max_abs_t, std::fabs(tconv[a]));
    }
    const Mat msfi = safe_solve(msf, p);
    double q = 0.0;
    for (int a = 0; a < p; ++a)
      for (int b = 0; b < p; ++b) q += dev[a] * msfi[(size_t)a * p + b] * dev[b];
    tconv_max = std::sqrt(q);
    se = sandwich_se(Dm, msf, p);  // diag(D^-1 msf D^-T)
    phase_t[2] += std::chrono::duration<double>(clk::now() - t_phase).count();
    if (opt->verbose) {
      std::fprintf(stderr, "[%.2f s] [phase 3] max|tconv| = %.4f (repeat %d)\n",
                   elapsed(E.get()), max_abs_t, repeats);
      std::fflush(stderr);
    }
    if (max_abs_t <= opt->conv_tol || repeats >= opt->max_repeats) break;
    ++repeats;
    t_phase = clk::now();
    if (opt->surrogate) {
      // surrogate path: one more simulation-checked Newton step, from this
      // phase 3's own means and score-function derivative
      newton_step(theta_h, Dm, sim_means, tgt, p, E->n_rates, &se);
      if ((st = put_theta(E.get(), theta_h)) != CUSNA_OK) return st;
    } else {
      const int last = opt->nsub > 0 ? opt->nsub - 1 : 0;
      const double gain = opt->firstg * std::pow(0.5, last);
      if ((st = subphase(gain, opt->nsub > 0 ? opt->n2[last] : 10)) != CUSNA_OK) return st;
    }
    phase_t[1] += std::chrono::duration<double>(clk::now() - t_phase).count();
  }

  // conditional runs estimate the rates from the simulated times (mom.R:260, 265)
  if (mode == 1 && (res->rate_est || res->rate_sd)) {
    for (int pp = 0; pp < P; ++pp) {
      std::vector<double> nt(opt->n3);
      if ((st = fetch(E.get(), nt.data(), X.plans[pp]->P.ntim, opt->n3)) != CUSNA_OK) return st;
      double m = 0.0;
      for (double x : nt) m += x;
      m /= opt->n3;
      double v = 0.0;
      for (double x : nt) v += (x - m) * (x - m);
      if (res->rate_est) res->rate_est[pp] = m;
      if (res->rate_sd) res->rate_sd[pp] = std::sqrt(v / (opt->n3 - 1));
    }
  }
  if (res->phase3_stats) {
    if ((st = fetch(E.get(), res->phase3_stats, X.S, (size_t)opt->n3 * p)) != CUSNA_OK) return st;
  }
  if (res->phase3_sp) {
    if ((st = fetch(E.get(), res->phase3_sp, X.SP, (size_t)P * opt->n3 * p)) != CUSNA_OK)
      return st;
  }
  if (res->phase3_scp) {
    if ((st = fetch(E.get(), res->phase3_scp, X.SC, (size_t)P * opt->n3 * p)) != CUSNA_OK)
      return st;
  }
  res->p = p;
  for (int a = 0; a < p; ++a) {
    if (res->theta) res->theta[a] = theta_h[a];
    if (res->se) res->se[a] = se[a];
    if (res->tconv) res->tconv[a] = tconv[a];
    if (res->sim_means) res->sim_means[a] = sim_means[a];
  }
  if (res->D) std::copy(Dm.begin(), Dm.end(), res->D);
  if (res->msf) std::copy(msf.begin(), msf.end(), res->msf);
  res->tconv_max = tconv_max;
  res->max_abs_tconv = max_abs_t;
  res->n_sims = E->n_sims;
  res->repeats = repeats;
  for (int q = 0; q < 3; ++q) res->phase_seconds[q] = phase_t[q];
  res->wall_seconds = std::chrono::duration<double>(clk::now() - t_start).count();
  return CUSNA_OK;
}

}  // extern "C"
