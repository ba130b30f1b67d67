/*
 * saom_oracle.c -- CPU oracle of the SAOM period simulator.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the GPU library is
 * compared against, and the CPU leg that the benchmark times. The product
 * library (cusna_amd) never links, loads or calls it.
 *
 * A plain-C restatement of the reference CPU engine
 * (cran/cusna src/libcusna_saom_cpu.cpp), written from its semantics:
 *   oracle_philox4x32     <- philox4, u01               (:32-46)
 *   ego_change            <- ego_ds                     (:49-101)
 *   beh_change            <- beh_ds                     (:103-113)
 *   run_chain             <- simulate_one               (:142-303)
 *   unpack_rows           <- unpack                     (:306-311)
 *   oracle_saom_simulate_period <- cusna_saom_simulate_period_cpu (:315-425)
 *   oracle_network_stats  <- cusna_network_stats        (:606-689)
 *   oracle_behavior_stats <- cusna_behavior_stats       (:767-817)
 *   oracle_gof_distribution <- cusna_gof_distribution   (:716-765)
 *   cross_change          <- cross_ds                   (:430-443)
 *   run_chain_multi       <- simulate_one_multi         (:460-510)
 *   oracle_saom_simulate_multi <- cusna_saom_simulate_multi_cpu (:548-604)
 *   oracle_cross_stats    <- cusna_cross_stats          (:819-857)
 * Summation orders, float32->double promotions and the Philox stream layout
 * are the reference's. Build with -ffp-contract=off so that no a*b+c is fused.
 *
 * Two switches extend the reference behaviour for checking the GPU path; both
 * are off by default, and with both off the output equals the reference's:
 *   per_chain  chain c uses theta row c and lam[c] (the reference uses row 0
 *              for every chain, libcusna_saom_cpu.cpp:334-336, 369);
 *   replay     log(u0) and log(-log(u_j)) are read from recorded arrays
 *              (filled by oracle_replay_noise) instead of being evaluated.
 * A dyadic covariate `dyc`, when the caller provides one, is honoured; the
 * reference zero-fills it (libcusna_saom_cpu.cpp:354) and the R glue never
 * passes one, so with dyc == NULL both agree.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/libcusna.h"

#define ORACLE_API __attribute__((visibility("default")))
#define ORACLE_MAX_STEPS 4000000L /* runaway cap, libcusna_saom_cpu.cpp:115 */

typedef struct oracle_opts {
  int32_t per_chain;
  const double *log_u0;   /* [M*smax] or NULL */
  const double *loglog_u; /* [M*smax*n] or NULL */
  int32_t smax;
} oracle_opts;

/* ---------------------------------------------------------------- Philox */

/* Philox4x32 with 10 rounds (Salmon et al., SC'11); multipliers and Weyl
 * key increments are the published Random123 constants. */
static void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t m0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t m1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t y0 = (uint32_t)(m1 >> 32) ^ c[1] ^ k0;
    uint32_t y1 = (uint32_t)m1;
    uint32_t y2 = (uint32_t)(m0 >> 32) ^ c[3] ^ k1;
    uint32_t y3 = (uint32_t)m0;
    c[0] = y0;
    c[1] = y1;
    c[2] = y2;
    c[3] = y3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

static void draw4(uint32_t a, uint32_t b, uint32_t s, uint32_t d, uint32_t k0, uint32_t k1,
                  uint32_t out[4]) {
  out[0] = a;
  out[1] = b;
  out[2] = s;
  out[3] = d;
  philox10(out, k0, k1);
}

/* (x + 1/2) / 2^32: open interval (0, 1). */
static double unit_open(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

ORACLE_API int oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  if (!ctr || !key || !out) return CUSNA_ERR_NULL_ARG;
  draw4(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
  return CUSNA_OK;
}

/* Recorded transcendental noise for replay: log(u0) of stream 0 and
 * log(-log(u_j)) of stream 1, for chains 0..M-1 and ministeps 0..smax-1. */
ORACLE_API int oracle_replay_noise(uint64_t seed, int32_t M, int32_t n, int32_t smax,
                                   double *log_u0, double *loglog_u) {
  if (!log_u0 || !loglog_u || M < 1 || n < 1 || smax < 1) return CUSNA_ERR_SHAPE;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t c = 0; c < M; ++c) {
    uint32_t r[4];
    for (int32_t s = 0; s < smax; ++s) {
      draw4((uint32_t)s, (uint32_t)c, 0u, 0u, k0, k1, r);
      log_u0[(size_t)c * smax + s] = log(unit_open(r[0]));
      for (int32_t j = 0; j < n; ++j) {
        draw4((uint32_t)s, (uint32_t)c, 1u, (uint32_t)j, k0, k1, r);
        loglog_u[((size_t)c * smax + s) * n + j] = log(-log(unit_open(r[0])));
      }
    }
  }
  return CUSNA_OK;
}

/* --------------------------------------------------------- change stats */

/* Dense state of one chain: a[i*n+j] = tie i->j. */
typedef struct chain_view {
  int n;
  const uint8_t *a;
  const int64_t *indeg, *outdeg;
  const int *tp; /* two-paths from the ego: tp[h] = sum_x a[i,x] a[x,h] */
  int64_t rdeg_i;
} chain_view;

#define A(v, r, c) ((v)->a[(size_t)(r) * (v)->n + (c)])

/* Ego-level change statistic of effect type t when ego i toggles i->j
 * (dl = +1 adds, -1 removes). libcusna_saom_cpu.cpp:49-101. */
static double ego_change(int t, const chain_view *v, int i, int j, double dl, double vi, double vj,
                         double dyw, double p1, double p2, const double *w) {
  const int n = v->n;
  switch (t) {
    case 0: /* density */
      return dl;
    case 1: /* recip */
      return dl * (double)A(v, j, i);
    case 2: { /* transTrip: two-paths i->h->j plus shared out-ties */
      int64_t c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, i, h)) c += (int64_t)A(v, h, j) + (int64_t)A(v, j, h);
      return dl * (double)c;
    }
    case 3:    /* cycle3 */
    case 23: { /* transTriads (same ego-level count) */
      int64_t c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, j, h) && A(v, h, i)) ++c;
      return dl * (double)c;
    }
    case 4: /* inPop */
      return dl * (double)(v->indeg[j] + (1 - A(v, i, j)));
    case 5: { /* transTies */
      const int want = A(v, i, j) ? 1 : 0;
      int64_t c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, i, h) && A(v, j, h) && v->tp[h] == want) ++c;
      if (v->tp[j] > 0) ++c;
      return dl * (double)c;
    }
    case 6: { /* transRecTrip */
      int64_t c = 0;
      if (A(v, j, i))
        for (int h = 0; h < n; ++h)
          if (A(v, i, h) && A(v, h, j)) ++c;
      for (int h = 0; h < n; ++h)
        if (A(v, i, h) && A(v, h, i) && A(v, j, h)) ++c;
      return dl * (double)c;
    }
    case 7: /* inPopSqrt */
      return dl * sqrt((double)(v->indeg[j] + (1 - A(v, i, j))));
    case 8: /* outPop */
      return dl * (double)v->outdeg[j];
    case 9: /* outPopSqrt */
      return dl * sqrt((double)v->outdeg[j]);
    case 10: /* egoX */
      return dl * vi;
    case 11: /* altX */
      return dl * vj;
    case 12: /* simX */
      return dl * (1.0 - fabs(vi - vj) * p1 - p2);
    case 13: /* sameX */
      return dl * (fabs(vi - vj) < 1e-9 ? 1.0 : 0.0);
    case 14: /* inAct */
      return dl * (double)v->indeg[i];
    case 15: /* inActSqrt */
      return dl * sqrt((double)v->indeg[i]);
    case 16: /* outAct */
      return dl * (2.0 * (double)v->outdeg[i] + dl);
    case 17: { /* outActSqrt */
      const double d0 = (double)v->outdeg[i], d1 = d0 + dl;
      return d1 * sqrt(d1) - d0 * sqrt(d0);
    }
    case 18: { /* outTrunc */
      const double d0 = (double)v->outdeg[i];
      return fmin(d0 + dl, p1) - fmin(d0, p1);
    }
    case 19: /* reciAct */
      return dl * ((double)v->rdeg_i + (double)A(v, j, i) * ((double)v->outdeg[i] + dl));
    case 20: /* higher */
      if (vi > vj + 1e-9) return dl;
      if (fabs(vi - vj) <= 1e-9) return dl * 0.5;
      return 0.0;
    case 21: /* egoXaltX */
      return dl * vi * vj;
    case 22: /* X (dyadic) */
      return dl * dyw;
    case 24: { /* gwesp (FF) */
      int c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, i, h) && A(v, h, j)) ++c;
      return dl * w[c];
    }
    case 25: /* egoPlusAltX */
      return dl * (vi + vj);
    case 26: { /* gwespFB */
      int c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, i, h) && A(v, j, h)) ++c;
      return dl * w[c];
    }
    case 27: { /* gwespBF */
      int c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, h, i) && A(v, h, j)) ++c;
      return dl * w[c];
    }
    case 28: { /* gwespBB */
      int c = 0;
      for (int h = 0; h < n; ++h)
        if (A(v, h, i) && A(v, j, h)) ++c;
      return dl * w[c];
    }
    case 29: { /* gwdspFF */
      const int xij = A(v, i, j);
      double s = 0.0;
      for (int h = 0; h < n; ++h) {
        if (!A(v, j, h) || h == i) continue;
        int p = 0;
        for (int m = 0; m < n; ++m)
          if (A(v, i, m) && A(v, m, h)) ++p;
        s += xij ? (w[p] - w[p - 1]) : (w[p + 1] - w[p]);
      }
      return dl * s;
    }
    case 30: { /* gwdspFB */
      const int xij = A(v, i, j);
      double s = 0.0;
      for (int h = 0; h < n; ++h) {
        if (!A(v, h, j) || h == i) continue;
        int p = 0;
        for (int m = 0; m < n; ++m)
          if (A(v, i, m) && A(v, h, m)) ++p;
        s += xij ? (w[p] - w[p - 1]) : (w[p + 1] - w[p]);
      }
      return dl * s;
    }
    case 31: /* altRThresholdX */
      return dl * (vj >= p1 - 1e-12 ? 1.0 : 0.0);
    default:
      return 0.0;
  }
}

/* Behaviour change statistic for a step dz (+1/-1). libcusna_saom_cpu.cpp:103-113. */
static double beh_change(int t, double dz, double zic, int64_t nb, double zsum, int64_t nH,
                         int64_t nE, int64_t nL, double brange, double bsim) {
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

/* ------------------------------------------------------------ one chain */

typedef struct period_ctx {
  int n, k, kb, nrate, M;
  const uint8_t *x0, *xs, *kd, *pm; /* dense n*n: data start, sim start, kdist, perm */
  const double *th, *thb, *lam, *lamb; /* th[M*k] (row per chain), lam[M] */
  const int32_t *tid, *ety, *dyn, *dci, *c0, *c1, *c2, *bty, *rkind;
  const double *wtab, *p1, *p2, *vmat, *dyc, *ralpha, *ru;
  const int32_t *z0, *bkd, *act;
  int nact, has_comp, symflag, maxdeg, cond, target;
  const uint8_t *inact;
  int zmin, zmax;
  double zbar, brange, bsim;
  uint32_t k0, k1;
  int need_tp, need_rdeg;
  const double *rep_lu0, *rep_llu;
  int32_t smax;
} period_ctx;

typedef struct chain_out {
  int64_t dist, dist_stat, steps, bdist;
  double t, rate_score, brate_score;
} chain_out;

/* One chain of one period. libcusna_saom_cpu.cpp:142-303. */
static void run_chain(const period_ctx *C, uint32_t c, int row, uint8_t *a, int32_t *z,
                      double *sc, double *bsc, double *rsc, int64_t *adist, chain_out *o) {
  const int n = C->n, k = C->k, kb = C->kb, nrate = C->nrate;
  const double *th = C->th + (size_t)row * k;
  const double *thb = C->thb + (size_t)row * (kb ? kb : 1);
  const double lam = C->lam[row], lamb = C->lamb[row];
  int64_t *indeg = calloc((size_t)n, sizeof(int64_t));
  int64_t *outdeg = calloc((size_t)n, sizeof(int64_t));
  double *fj = malloc(sizeof(double) * (size_t)n);
  double *ds = malloc(sizeof(double) * (size_t)n * (size_t)k);
  double *wsum = malloc(sizeof(double) * (size_t)k);
  double *crate = malloc(sizeof(double) * (size_t)n);
  double *cup = malloc(sizeof(double) * (size_t)(kb ? kb : 1));
  double *cdn = malloc(sizeof(double) * (size_t)(kb ? kb : 1));
  double *ub = malloc(sizeof(double) * (size_t)(nrate ? nrate : 1));
  double *rsum = malloc(sizeof(double) * (size_t)(nrate ? nrate : 1));
  uint8_t *okj = malloc((size_t)n);
  int *tp = calloc((size_t)n, sizeof(int));

  for (size_t q = 0; q < (size_t)n * n; ++q) a[q] = C->xs[q];
  for (int r = 0; r < n; ++r)
    for (int s = 0; s < n; ++s)
      if (a[(size_t)r * n + s]) {
        outdeg[r]++;
        indeg[s]++;
      }
  for (int r = 0; r < n; ++r) {
    z[r] = C->z0[r];
    crate[r] = 1.0;
  }
  for (int e = 0; e < k; ++e) sc[e] = 0.0;
  for (int e = 0; e < kb; ++e) bsc[e] = 0.0;
  for (int q = 0; q < nrate; ++q) rsc[q] = 0.0;

  int64_t dist = 0, bdist = 0, steps = 0, steps_beh = 0;
  const int64_t cap = (C->rep_lu0 && C->smax < ORACLE_MAX_STEPS) ? C->smax : ORACLE_MAX_STEPS;
  double t = 0.0, rho = 0.0;
  uint32_t r4[4];
  chain_view V = {n, a, indeg, outdeg, tp, 0};

  for (;;) {
    draw4((uint32_t)steps, c, 0u, 0u, C->k0, C->k1, r4);
    const double lu0 =
        C->rep_lu0 ? C->rep_lu0[(size_t)c * C->smax + steps] : log(unit_open(r4[0]));
    int i;
    int net_step;
    if (nrate > 0) { /* rate effects: libcusna_saom_cpu.cpp:170-193 */
      double S0 = 0.0;
      for (int q = 0; q < nrate; ++q) rsum[q] = 0.0;
      for (int x = 0; x < C->nact; ++x) {
        const int a2 = C->act[x];
        double s = 0.0;
        for (int q = 0; q < nrate; ++q) {
          const double u = C->rkind[q] == 0   ? C->ru[(size_t)q * n + a2]
                           : C->rkind[q] == 1 ? (double)outdeg[a2]
                                              : (double)indeg[a2];
          ub[q] = u;
          s += C->ralpha[q] * u;
        }
        if (s > 30.0)
          s = 30.0;
        else if (s < -30.0)
          s = -30.0;
        const double cr = exp(s);
        crate[a2] = cr;
        S0 += cr;
        for (int q = 0; q < nrate; ++q) rsum[q] += cr * ub[q];
      }
      const double tau = -lu0 / (lam * S0);
      double tau_eff = tau;
      if (!C->cond && t + tau > 1.0) tau_eff = 1.0 - t;
      rho += tau_eff * S0;
      for (int q = 0; q < nrate; ++q) rsc[q] -= tau_eff * lam * rsum[q];
      t += tau;
      if (!C->cond && t > 1.0) break;
      net_step = 1;
      const double tgt = unit_open(r4[2]) * S0;
      double acc = 0.0;
      i = C->act[C->nact - 1];
      for (int x = 0; x < C->nact; ++x) {
        acc += crate[C->act[x]];
        if (acc >= tgt) {
          i = C->act[x];
          break;
        }
      }
    } else {
      t += -lu0 / ((double)C->nact * (lam + lamb));
      if (!C->cond && t > 1.0) break;
      net_step = kb == 0 || unit_open(r4[1]) < lam / (lam + lamb);
      int64_t x = (int64_t)(unit_open(r4[2]) * C->nact);
      if (x >= C->nact) x = C->nact - 1;
      i = C->act[x];
    }

    if (!net_step) { /* behaviour ministep: libcusna_saom_cpu.cpp:202-227 */
      const int zi = z[i];
      const double zic = (double)zi - C->zbar;
      int64_t nb = 0, nH = 0, nE = 0, nL = 0;
      double zsum = 0.0;
      for (int j = 0; j < n; ++j) {
        if (!a[(size_t)i * n + j]) continue;
        nb++;
        zsum += (double)z[j] - C->zbar;
        if (z[j] > zi)
          nH++;
        else if (z[j] == zi)
          nE++;
        else
          nL++;
      }
      const int up = zi < C->zmax, dn = zi > C->zmin;
      double fu = 0.0, fd = 0.0;
      for (int e = 0; e < kb; ++e) {
        cup[e] = beh_change(C->bty[e], 1.0, zic, nb, zsum, nH, nE, nL, C->brange, C->bsim);
        cdn[e] = beh_change(C->bty[e], -1.0, zic, nb, zsum, nH, nE, nL, C->brange, C->bsim);
        fu += thb[e] * cup[e];
        fd += thb[e] * cdn[e];
      }
      double mx = 0.0;
      if (up && fu > mx) mx = fu;
      if (dn && fd > mx) mx = fd;
      double pd = dn ? exp(fd - mx) : 0.0, p0 = exp(-mx), pu = up ? exp(fu - mx) : 0.0;
      const double ps = pd + p0 + pu;
      pd /= ps;
      p0 /= ps;
      pu /= ps;
      const double uc = unit_open(r4[3]);
      const int d = uc < pd ? -1 : (uc < pd + p0 ? 0 : 1);
      for (int e = 0; e < kb; ++e) {
        double s = d == 1 ? cup[e] : (d == -1 ? cdn[e] : 0.0);
        if (up) s -= pu * cup[e];
        if (dn) s -= pd * cdn[e];
        bsc[e] += s;
      }
      if (d != 0) {
        if (C->bkd[i]) bdist += llabs((int64_t)zi + d - C->z0[i]) - llabs((int64_t)zi - C->z0[i]);
        z[i] = zi + d;
      }
      steps_beh++;
      steps++;
      if (C->cond && dist >= C->target) break;
      if (steps >= cap) break;
      continue;
    }

    /* network ministep: libcusna_saom_cpu.cpp:229-288 */
    const uint8_t *ai = a + (size_t)i * n;
    if (C->need_tp) {
      for (int h = 0; h < n; ++h) tp[h] = 0;
      for (int x = 0; x < n; ++x)
        if (ai[x])
          for (int h = 0; h < n; ++h) tp[h] += a[(size_t)x * n + h];
    }
    V.rdeg_i = 0;
    if (C->need_rdeg)
      for (int h = 0; h < n; ++h)
        if (ai[h] && a[(size_t)h * n + i]) V.rdeg_i++;
    const double zvi = (double)z[i] - C->zbar;

    double gbest = -1e30, lmax = -1e30;
    int jstar = i;
    for (int j = 0; j < n; ++j) {
      double f = 0.0;
      int ok = 1;
      double *dj = ds + (size_t)j * k;
      if (j == i) {
        for (int e = 0; e < k; ++e) dj[e] = 0.0;
      } else {
        if (!C->pm[(size_t)i * n + j])
          ok = 0;
        else if (C->maxdeg > 0 && ai[j] == 0 && outdeg[i] >= C->maxdeg)
          ok = 0;
        if (ok) {
          const double dl = 1.0 - 2.0 * ai[j];
          const double zvj = (double)z[j] - C->zbar;
          for (int e = 0; e < k; ++e) {
            double d;
            if (C->tid[e] == 32) { /* interaction of earlier components */
              d = C->c2[e] < 0 ? dl : 1.0;
              d *= dj[C->c0[e]] * dj[C->c1[e]];
              if (C->c2[e] >= 0) d *= dj[C->c2[e]];
            } else {
              const double vi = C->dyn[e] ? zvi : C->vmat[(size_t)e * n + i];
              const double vj = C->dyn[e] ? zvj : C->vmat[(size_t)e * n + j];
              const double dyw =
                  (C->dci[e] >= 0 && C->dyc) ? C->dyc[((size_t)C->dci[e] * n + i) * n + j] : 0.0;
              d = ego_change(C->tid[e], &V, i, j, dl, vi, vj, dyw, C->p1[e], C->p2[e],
                             C->wtab + (size_t)e * (n + 1));
              if ((C->ety[e] == 1 && ai[j] == 1) || (C->ety[e] == 2 && ai[j] == 0)) d = 0.0;
            }
            dj[e] = d;
            f += th[e] * d;
          }
        }
      }
      okj[j] = (uint8_t)ok;
      fj[j] = f;
      if (ok && f > lmax) lmax = f;
      if (ok) {
        double ll;
        if (C->rep_llu) {
          ll = C->rep_llu[((size_t)c * C->smax + steps) * n + j];
        } else {
          draw4((uint32_t)steps, c, 1u, (uint32_t)j, C->k0, C->k1, r4);
          ll = log(-log(unit_open(r4[0])));
        }
        const double g = f - ll;
        if (g > gbest || (g == gbest && j < jstar)) {
          gbest = g;
          jstar = j;
        }
      }
    }

    double sume = 0.0;
    for (int e = 0; e < k; ++e) wsum[e] = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!okj[j]) continue;
      const double w = exp(fj[j] - lmax);
      sume += w;
      for (int e = 0; e < k; ++e) wsum[e] += w * ds[(size_t)j * k + e];
    }
    for (int e = 0; e < k; ++e) sc[e] += ds[(size_t)jstar * k + e] - wsum[e] / sume;
    steps++;

    if (jstar != i) {
      const int j = jstar;
      const int64_t dl = 1 - 2 * (int64_t)a[(size_t)i * n + j];
      const int64_t chg = C->symflag ? 2 : 1;
      if (C->kd[(size_t)i * n + j])
        dist += (a[(size_t)i * n + j] == C->x0[(size_t)i * n + j]) ? chg : -chg;
      indeg[j] += dl;
      outdeg[i] += dl;
      a[(size_t)i * n + j] ^= 1u;
      if (C->symflag) {
        indeg[i] += dl;
        outdeg[j] += dl;
        a[(size_t)j * n + i] ^= 1u;
      }
    }
    if (nrate > 0)
      for (int q = 0; q < nrate; ++q)
        rsc[q] += C->rkind[q] == 0   ? C->ru[(size_t)q * n + i]
                  : C->rkind[q] == 1 ? (double)outdeg[i]
                                     : (double)indeg[i];
    if (C->cond && dist >= C->target) break;
    if (steps >= cap) break;
  }

  if (C->has_comp)
    for (int r = 0; r < n; ++r)
      if (C->inact[r])
        for (int s = 0; s < n; ++s) {
          a[(size_t)r * n + s] = C->x0[(size_t)r * n + s];
          a[(size_t)s * n + r] = C->x0[(size_t)s * n + r];
        }

  int64_t dsum = 0;
  for (int r = 0; r < n; ++r) {
    int64_t cnt = 0;
    for (int s = 0; s < n; ++s)
      if (C->kd[(size_t)r * n + s] && a[(size_t)r * n + s] != C->x0[(size_t)r * n + s]) cnt++;
    adist[r] = cnt;
    dsum += cnt;
  }
  o->dist = dist;
  o->dist_stat = dsum;
  o->steps = steps;
  o->t = t;
  o->bdist = bdist;
  o->rate_score = nrate > 0 ? (double)(steps - steps_beh) / lam - rho
                            : (double)(steps - steps_beh) / lam - (double)C->nact;
  o->brate_score = kb > 0 ? (double)steps_beh / lamb - (double)C->nact : 0.0;

  free(indeg);
  free(outdeg);
  free(fj);
  free(ds);
  free(wsum);
  free(crate);
  free(cup);
  free(cdn);
  free(ub);
  free(rsum);
  free(okj);
  free(tp);
}

static void unpack_rows(const uint32_t *rows, int n, int W, uint8_t *dense) {
  for (int r = 0; r < n; ++r)
    for (int s = 0; s < n; ++s)
      dense[(size_t)r * n + s] = (uint8_t)((rows[(size_t)r * W + (s >> 5)] >> (s & 31)) & 1u);
}

static void offdiag(int n, uint8_t *dense) {
  for (int r = 0; r < n; ++r)
    for (int s = 0; s < n; ++s) dense[(size_t)r * n + s] = (uint8_t)(r != s);
}

ORACLE_API int oracle_network_stats(const uint8_t *a, int32_t n, int32_t k, const int32_t *type_ids,
                                    const double *p1, const double *p2, const double *vmat,
                                    const uint8_t *vmiss, const int32_t *dyn_flags,
                                    const double *vb, const int32_t *od0, const int32_t *id0,
                                    const double *wtab, int32_t symflag, double *out);

/* M chains of one period. libcusna_saom_cpu.cpp:315-425. */
ORACLE_API int oracle_saom_simulate_period(const cusna_saom_problem *pb, cusna_saom_result *r,
                                           const oracle_opts *opt) {
  if (!pb || !r) return CUSNA_ERR_NULL_ARG;
  if (pb->struct_version != CUSNA_ABI_VERSION) return CUSNA_ERR_SHAPE;
  const int n = pb->n, W = pb->W, k = pb->k, M = pb->M, kb = pb->kb, nrate = pb->nrate;
  if (n < 1 || n > CUSNA_MAX_N || W != (n + 31) / 32 || k < 1 || M < 1) return CUSNA_ERR_SHAPE;
  const int per_chain = opt && opt->per_chain;
  const int rows_th = per_chain ? M : 1;
  const size_t nn = (size_t)n * n;

  uint8_t *x0 = malloc(nn), *xs = malloc(nn), *kd = malloc(nn), *pm = malloc(nn);
  uint8_t *keep = malloc(nn), *force = calloc(nn, 1);
  unpack_rows(pb->x0_out, n, W, x0);
  unpack_rows(pb->sim_out ? pb->sim_out : pb->x0_out, n, W, xs);
  if (pb->kdist) unpack_rows(pb->kdist, n, W, kd); else offdiag(n, kd);
  if (pb->perm) unpack_rows(pb->perm, n, W, pm); else offdiag(n, pm);
  if (pb->keep_out) unpack_rows(pb->keep_out, n, W, keep); else offdiag(n, keep);
  if (pb->force_out) unpack_rows(pb->force_out, n, W, force);

  double *th = malloc(sizeof(double) * (size_t)rows_th * k);
  double *thb = calloc((size_t)rows_th * (kb ? kb : 1), sizeof(double));
  double *lam = malloc(sizeof(double) * (size_t)rows_th);
  double *lamb = malloc(sizeof(double) * (size_t)rows_th);
  for (int c = 0; c < rows_th; ++c) {
    for (int e = 0; e < k; ++e) th[(size_t)c * k + e] = pb->theta[(size_t)c * k + e];
    for (int e = 0; e < kb; ++e)
      thb[(size_t)c * kb + e] = pb->thetab ? pb->thetab[(size_t)c * kb + e] : 0.0;
    lam[c] = pb->lam ? pb->lam[c] : 1.0;
    lamb[c] = pb->lamb ? pb->lamb[c] : 0.0;
  }
  double *wtab = malloc(sizeof(double) * (size_t)k * (n + 1));
  for (size_t q = 0; q < (size_t)k * (n + 1); ++q)
    wtab[q] = pb->wtab64 ? pb->wtab64[q] : (pb->wtab ? pb->wtab[q] : 0.0);
  double *p1 = malloc(sizeof(double) * k), *p2 = malloc(sizeof(double) * k);
  for (int e = 0; e < k; ++e) {
    p1[e] = pb->p1_64 ? pb->p1_64[e] : (pb->p1 ? pb->p1[e] : 0.0);
    p2[e] = pb->p2_64 ? pb->p2_64[e] : (pb->p2 ? pb->p2[e] : 0.0);
  }
  double *vmat = malloc(sizeof(double) * (size_t)k * n);
  for (size_t q = 0; q < (size_t)k * n; ++q)
    vmat[q] = pb->vmat64 ? pb->vmat64[q] : (pb->vmat ? pb->vmat[q] : 0.0);
  double *vb = malloc(sizeof(double) * n);
  for (int q = 0; q < n; ++q) vb[q] = pb->vb64 ? pb->vb64[q] : (pb->vb ? pb->vb[q] : 0.0);
  double *ralpha = calloc((size_t)(nrate ? nrate : 1), sizeof(double));
  double *ru = calloc((size_t)(nrate ? nrate : 1) * n, sizeof(double));
  for (int q = 0; q < nrate; ++q) ralpha[q] = pb->ralpha[q];
  for (size_t q = 0; nrate && q < (size_t)nrate * n; ++q) ru[q] = pb->ru[q];
  double *dyc = NULL;
  if (pb->dyc && pb->dyc_idx) {
    int nd = 0;
    for (int e = 0; e < k; ++e)
      if (pb->dyc_idx[e] + 1 > nd) nd = pb->dyc_idx[e] + 1;
    if (nd > 0) {
      dyc = malloc(sizeof(double) * (size_t)nd * nn);
      for (size_t q = 0; q < (size_t)nd * nn; ++q) dyc[q] = pb->dyc[q];
    }
  }
  int32_t *zero_k = calloc((size_t)k, sizeof(int32_t));
  int32_t *neg_k = malloc(sizeof(int32_t) * k);
  for (int e = 0; e < k; ++e) neg_k[e] = -1;
  int32_t *z0 = calloc((size_t)n, sizeof(int32_t)), *bkd = malloc(sizeof(int32_t) * n);
  int32_t *act = malloc(sizeof(int32_t) * n);
  uint8_t *inact = calloc((size_t)n, 1);
  for (int q = 0; q < n; ++q) {
    if (pb->z0) z0[q] = pb->z0[q];
    bkd[q] = pb->bkdist ? pb->bkdist[q] : 1;
    act[q] = pb->act ? pb->act[q] : q;
    if (pb->has_comp && pb->inactw) inact[q] = (uint8_t)((pb->inactw[q >> 5] >> (q & 31)) & 1u);
  }

  period_ctx C;
  memset(&C, 0, sizeof(C));
  C.n = n; C.k = k; C.kb = kb; C.nrate = nrate; C.M = M;
  C.x0 = x0; C.xs = xs; C.kd = kd; C.pm = pm;
  C.th = th; C.thb = thb; C.lam = lam; C.lamb = lamb;
  C.tid = pb->type_ids;
  C.ety = pb->etypes ? pb->etypes : zero_k;
  C.dyn = pb->dyn_flags ? pb->dyn_flags : zero_k;
  C.dci = pb->dyc_idx ? pb->dyc_idx : neg_k;
  C.c0 = pb->int0 ? pb->int0 : neg_k;
  C.c1 = pb->int1 ? pb->int1 : neg_k;
  C.c2 = pb->int2 ? pb->int2 : neg_k;
  C.bty = pb->bty; C.rkind = pb->rkind;
  C.wtab = wtab; C.p1 = p1; C.p2 = p2; C.vmat = vmat; C.dyc = dyc;
  C.ralpha = ralpha; C.ru = ru;
  C.z0 = z0; C.bkd = bkd; C.act = act; C.inact = inact;
  C.nact = pb->act ? pb->nact : n;
  C.has_comp = pb->has_comp; C.symflag = pb->symflag; C.maxdeg = pb->maxdegree;
  C.cond = pb->conditional; C.target = pb->target_dist;
  C.zmin = pb->zmin; C.zmax = pb->zmax;
  C.zbar = pb->zbar; C.brange = pb->brange; C.bsim = pb->bsimmean;
  C.k0 = (uint32_t)(pb->seed & 0xFFFFFFFFu); C.k1 = (uint32_t)(pb->seed >> 32);
  for (int e = 0; e < k; ++e) {
    if (pb->type_ids[e] == 5) C.need_tp = 1;
    if (pb->type_ids[e] == 19) C.need_rdeg = 1;
  }
  if (opt && opt->log_u0 && opt->loglog_u && opt->smax > 0) {
    C.rep_lu0 = opt->log_u0;
    C.rep_llu = opt->loglog_u;
    C.smax = opt->smax;
  }

#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int c = 0; c < M; ++c) {
    uint8_t *a = malloc(nn);
    int32_t *z = malloc(sizeof(int32_t) * n);
    double *sc = malloc(sizeof(double) * k);
    double *bsc = malloc(sizeof(double) * (kb ? kb : 1));
    double *rs = malloc(sizeof(double) * (nrate ? nrate : 1));
    int64_t *ad = malloc(sizeof(int64_t) * n);
    chain_out o;
    run_chain(&C, (uint32_t)c, per_chain ? c : 0, a, z, sc, bsc, rs, ad, &o);
    if (r->dist) r->dist[c] = (int32_t)o.dist;
    if (r->dist_stat) r->dist_stat[c] = (int32_t)o.dist_stat;
    if (r->steps) r->steps[c] = (int32_t)o.steps;
    if (r->ntim) r->ntim[c] = o.t;
    if (r->scores) {
      r->scores[(size_t)c * (k + 1)] = (float)o.rate_score;
      for (int e = 0; e < k; ++e) r->scores[(size_t)c * (k + 1) + 1 + e] = (float)sc[e];
    }
    if (kb > 0) {
      if (r->bdist) r->bdist[c] = (int32_t)o.bdist;
      if (r->z_end)
        for (int q = 0; q < n; ++q) r->z_end[(size_t)c * n + q] = z[q];
      if (r->bscores) {
        r->bscores[(size_t)c * (kb + 1)] = (float)o.brate_score;
        for (int e = 0; e < kb; ++e) r->bscores[(size_t)c * (kb + 1) + 1 + e] = (float)bsc[e];
      }
    }
    if (nrate > 0) {
      if (r->rscores)
        for (int q = 0; q < nrate; ++q) r->rscores[(size_t)c * nrate + q] = (float)rs[q];
      if (r->actor_dist)
        for (int q = 0; q < n; ++q) r->actor_dist[(size_t)c * n + q] = (int32_t)ad[q];
    }
    if (r->want_end_nets && r->a_ends) memcpy(r->a_ends + (size_t)c * nn, a, nn);
    if (r->stats) {
      uint8_t *mk = malloc(nn);
      for (size_t q = 0; q < nn; ++q) mk[q] = (uint8_t)((a[q] & keep[q]) | force[q]);
      oracle_network_stats(mk, n, k, pb->type_ids, p1, p2, vmat, pb->vmiss, pb->dyn_flags, vb,
                           pb->od0, pb->id0, wtab, pb->symflag, r->stats + (size_t)c * k);
      free(mk);
    }
    free(a);
    free(z);
    free(sc);
    free(bsc);
    free(rs);
    free(ad);
  }

  free(x0); free(xs); free(kd); free(pm); free(keep); free(force);
  free(th); free(thb); free(lam); free(lamb); free(wtab); free(p1); free(p2);
  free(vmat); free(vb); free(ralpha); free(ru); free(dyc); free(zero_k); free(neg_k);
  free(z0); free(bkd); free(act); free(inact);
  return CUSNA_OK;
}

/* ------------------------------------------------------------ statistics */

/* Degree statistic helper for outPop/inAct families. libcusna_saom_cpu.cpp:608-616. */
static double degree_term(int t, double pm, double cur, double start) {
  if (t == 8 || t == 9) {
    if (pm >= 1.0) return t == 9 ? sqrt(cur) : cur;
    return pm == 0.0 ? start : cur + start;
  }
  if (pm >= 1.0) return t == 15 ? sqrt(cur) : cur;
  const double s = pm == 0.0 ? start : cur + start;
  return t == 15 ? sqrt(s) : s;
}

/* Two-path style products: M[i][j] = sum_h L(i,h) R(h,j) with optional
 * transposes (rowT: use a[h][i]; colT: use a[j][h]). */
static int *pair_counts(const uint8_t *a, int n, int rowT, int colT) {
  int *m = malloc(sizeof(int) * (size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      int s = 0;
      for (int h = 0; h < n; ++h) {
        const int l = rowT ? a[(size_t)h * n + i] : a[(size_t)i * n + h];
        const int rr = colT ? a[(size_t)j * n + h] : a[(size_t)h * n + j];
        s += l & rr;
      }
      m[(size_t)i * n + j] = s;
    }
  return m;
}

/* Statistics of one network. libcusna_saom_cpu.cpp:619-689. */
ORACLE_API int oracle_network_stats(const uint8_t *a, int32_t n, int32_t k, const int32_t *type_ids,
                                    const double *p1, const double *p2, const double *vmat,
                                    const uint8_t *vmiss, const int32_t *dyn_flags,
                                    const double *vb, const int32_t *od0, const int32_t *id0,
                                    const double *wtab, int32_t symflag, double *out) {
  if (!a || !type_ids || !out || n < 1 || n > CUSNA_MAX_N || k < 1) return CUSNA_ERR_NULL_ARG;
  double *ind = calloc((size_t)n, sizeof(double)), *outd = calloc((size_t)n, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (a[(size_t)i * n + j]) {
        outd[i] += 1.0;
        ind[j] += 1.0;
      }
  int need_aa = 0, need_fb = 0, need_bf = 0, need_bb = 0;
  for (int e = 0; e < k; ++e) {
    const int t = type_ids[e];
    if (t == 2 || t == 3 || t == 5 || t == 6 || t == 23 || t == 24 || t == 29) need_aa = 1;
    if (t == 26 || t == 30) need_fb = 1;
    if (t == 27) need_bf = 1;
    if (t == 28) need_bb = 1;
  }
  int *aa = need_aa ? pair_counts(a, n, 0, 0) : NULL;
  int *fb = need_fb ? pair_counts(a, n, 0, 1) : NULL;
  int *bf = need_bf ? pair_counts(a, n, 1, 0) : NULL;
  int *bb = need_bb ? pair_counts(a, n, 1, 1) : NULL;
  const size_t nn = (size_t)n * n;
#define AIJ(i, j) a[(size_t)(i) * n + (j)]
#define VAL(i) (dyn_flags && dyn_flags[e] ? (vb ? vb[i] : 0.0) : (vmat ? vmat[(size_t)e * n + (i)] : 0.0))
#define MISS(i) (vmiss && vmiss[(size_t)e * n + (i)])
  for (int e = 0; e < k; ++e) {
    const int t = type_ids[e];
    const double *wr = wtab ? wtab + (size_t)e * (n + 1) : NULL;
    const double pm = p1 ? p1[e] : 0.0, q2 = p2 ? p2[e] : 0.0;
    double s = 0.0;
    switch (t) {
      case 0:
        for (int i = 0; i < n; ++i) s += outd[i];
        if (symflag) s *= 0.5;
        break;
      case 1:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) s += (AIJ(i, j) && AIJ(j, i));
        break;
      case 2:
        for (size_t q = 0; q < nn; ++q)
          if (a[q]) s += aa[q];
        break;
      case 3:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(j, i)) s += aa[(size_t)i * n + j];
        s /= 3.0;
        break;
      case 23:
        for (size_t q = 0; q < nn; ++q)
          if (a[q]) s += aa[q];
        s /= 6.0;
        break;
      case 5:
        for (size_t q = 0; q < nn; ++q)
          if (a[q] && aa[q] > 0) s += 1.0;
        break;
      case 6:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j) && AIJ(j, i)) s += aa[(size_t)i * n + j];
        break;
      case 4:
        for (int j = 0; j < n; ++j) s += ind[j] * ind[j];
        break;
      case 7:
        for (int j = 0; j < n; ++j) s += pow(ind[j], 1.5);
        break;
      case 8:
      case 9:
        for (int i = 0; i < n; ++i) s += ind[i] * degree_term(t, pm, outd[i], od0 ? od0[i] : 0.0);
        break;
      case 14:
      case 15:
        for (int i = 0; i < n; ++i) s += outd[i] * degree_term(t, pm, ind[i], id0 ? id0[i] : 0.0);
        break;
      case 16:
        for (int i = 0; i < n; ++i) s += outd[i] * outd[i];
        break;
      case 17:
        for (int i = 0; i < n; ++i) s += pow(outd[i], 1.5);
        break;
      case 18:
        for (int i = 0; i < n; ++i) s += fmin(outd[i], pm);
        break;
      case 19:
        for (int i = 0; i < n; ++i) {
          double rd = 0.0;
          for (int j = 0; j < n; ++j) rd += (AIJ(i, j) && AIJ(j, i));
          s += outd[i] * rd;
        }
        break;
      case 10:
        for (int i = 0; i < n; ++i)
          if (!MISS(i)) s += outd[i] * VAL(i);
        break;
      case 11:
        for (int i = 0; i < n; ++i)
          if (!MISS(i)) s += ind[i] * VAL(i);
        break;
      case 12:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j) && !MISS(i) && !MISS(j)) s += 1.0 - fabs(VAL(i) - VAL(j)) * pm - q2;
        break;
      case 13:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j) && !MISS(i) && !MISS(j)) s += fabs(VAL(i) - VAL(j)) < 1e-9 ? 1.0 : 0.0;
        break;
      case 20:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j)) {
              double x;
              if (MISS(i) || MISS(j)) {
                x = 0.5;
              } else {
                const double d = VAL(i) - VAL(j);
                x = (d > 1e-9 ? 1.0 : 0.0) + 0.5 * (fabs(d) <= 1e-9 ? 1.0 : 0.0);
              }
              s += x;
            }
        break;
      case 21:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j) && !MISS(i) && !MISS(j)) s += VAL(i) * VAL(j);
        break;
      case 25:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j) && !MISS(i) && !MISS(j)) s += VAL(i) + VAL(j);
        break;
      case 31:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j) && !MISS(j) && VAL(j) >= pm - 1e-12) s += 1.0;
        break;
      case 24:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j)) s += wr[aa[(size_t)i * n + j]];
        break;
      case 26:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j)) s += wr[fb[(size_t)i * n + j]];
        break;
      case 27:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j)) s += wr[bf[(size_t)i * n + j]];
        break;
      case 28:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (AIJ(i, j)) s += wr[bb[(size_t)i * n + j]];
        break;
      case 29:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (i != j) s += wr[aa[(size_t)i * n + j]];
        break;
      case 30:
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (i != j) s += wr[fb[(size_t)i * n + j]];
        break;
      default:
        break; /* X (22) and interaction (32) have no end statistic here */
    }
    out[e] = s;
  }
#undef AIJ
#undef VAL
#undef MISS
  free(ind);
  free(outd);
  free(aa);
  free(fb);
  free(bf);
  free(bb);
  return CUSNA_OK;
}

/* Behaviour statistics over M behaviour vectors. libcusna_saom_cpu.cpp:767-817. */
ORACLE_API int oracle_behavior_stats(const int32_t *zb, int32_t n, int32_t M, int32_t kb,
                                     const int32_t *bty, const uint8_t *xnet,
                                     const int32_t *bkdist, double zbar, double brange,
                                     double bsim, double *out) {
  if (!zb || !bty || !xnet || !out || n < 1 || n > CUSNA_MAX_N || M < 1 || kb < 1)
    return CUSNA_ERR_NULL_ARG;
  double *ok = malloc(sizeof(double) * n), *od = malloc(sizeof(double) * n);
  double *nb = malloc(sizeof(double) * n), *u = malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) ok[i] = (bkdist && !bkdist[i]) ? 0.0 : 1.0;
  for (int i = 0; i < n; ++i) {
    double d = 0.0, b = 0.0;
    for (int j = 0; j < n; ++j)
      if (xnet[(size_t)i * n + j]) {
        d += 1.0;
        b += ok[j];
      }
    od[i] = d;
    nb[i] = b;
  }
  for (int c = 0; c < M; ++c) {
    const int32_t *z = zb + (size_t)c * n;
    for (int i = 0; i < n; ++i) u[i] = ok[i] != 0.0 ? ((double)z[i] - zbar) : 0.0;
    for (int e = 0; e < kb; ++e) {
      const int t = bty[e];
      double s = 0.0;
      if (t == 0) {
        for (int i = 0; i < n; ++i) s += u[i];
      } else if (t == 1) {
        for (int i = 0; i < n; ++i) s += u[i] * u[i];
      } else if (t == 2) {
        for (int i = 0; i < n; ++i) {
          if (od[i] <= 0.0) continue;
          double x = 0.0;
          for (int j = 0; j < n; ++j)
            if (xnet[(size_t)i * n + j]) x += u[j];
          s += u[i] * (x / od[i]);
        }
      } else if (t == 3 || t == 4) {
        for (int i = 0; i < n; ++i) {
          if (ok[i] == 0.0) continue;
          double x = 0.0;
          for (int j = 0; j < n; ++j)
            if (xnet[(size_t)i * n + j] && ok[j] != 0.0) x += 1.0 - fabs(u[i] - u[j]) / brange;
          x -= nb[i] * bsim;
          if (t == 3) x = nb[i] > 0.0 ? x / nb[i] : 0.0;
          s += x;
        }
      }
      out[(size_t)c * kb + e] = s;
    }
  }
  free(ok);
  free(od);
  free(nb);
  free(u);
  return CUSNA_OK;
}

ORACLE_API int oracle_set_threads(int t) {
#ifdef _OPENMP
  const int prev = omp_get_max_threads();
  omp_set_num_threads(t < 1 ? 1 : t);
  return prev;
#else
  (void)t;
  return 0;
#endif
}

/* sienaGOF auxiliary distributions (libcusna_saom_cpu.cpp:716-765): kind 0
 * indegree / 1 outdegree -- actors with degree <= levls[L]; kind 2 geodesic --
 * ordered pairs i != j within distance levls[L] (boolean powers of the
 * adjacency, diagonal cleared after each step), plus the unreachable pairs.
 * The argument checks and their status codes are the reference's. */
ORACLE_API int oracle_gof_distribution(const uint8_t *a_batch, int n, int M, int kind,
                                       const int32_t *levls, int nlev, double *out) {
  if (!a_batch || !levls || !out || n < 1 || M < 1 || nlev < 1) return CUSNA_ERR_NULL_ARG;
  const size_t nn = (size_t)n * n;
  if (kind == 0 || kind == 1) {
    int *deg = (int *)malloc(sizeof(int) * n);
    for (int c = 0; c < M; ++c) {
      const uint8_t *a = a_batch + (size_t)c * nn;
      for (int v = 0; v < n; ++v) deg[v] = 0;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          if (a[(size_t)i * n + j]) deg[kind == 0 ? j : i]++;
      for (int L = 0; L < nlev; ++L) {
        long cnt = 0;
        for (int v = 0; v < n; ++v)
          if (deg[v] <= levls[L]) cnt++;
        out[(size_t)c * nlev + L] = (double)cnt;
      }
    }
    free(deg);
    return CUSNA_OK;
  }
  if (kind == 2) {
    int maxk = 0;
    for (int L = 0; L < nlev; ++L) {
      if (levls[L] < 0) return CUSNA_ERR_SHAPE;
      if (levls[L] > maxk) maxk = levls[L];
    }
    uint8_t *cur = (uint8_t *)malloc(nn), *nxt = (uint8_t *)malloc(nn);
    long *counts = (long *)calloc((size_t)maxk + 1, sizeof(long));
    for (int c = 0; c < M; ++c) {
      const uint8_t *reach = a_batch + (size_t)c * nn;
      for (size_t q = 0; q < nn; ++q) cur[q] = reach[q] ? 1 : 0;
      for (int k = 1; k <= maxk; ++k) {
        long cn = 0;
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j)
            if (i != j && cur[(size_t)i * n + j]) cn++;
        counts[k] = cn;
        if (k < maxk) {
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
              uint8_t r = cur[(size_t)i * n + j];
              if (!r)
                for (int m = 0; m < n; ++m)
                  if (cur[(size_t)i * n + m] && reach[(size_t)m * n + j]) {
                    r = 1;
                    break;
                  }
              nxt[(size_t)i * n + j] = r;
            }
          for (int i = 0; i < n; ++i) nxt[(size_t)i * n + i] = 0;
          uint8_t *t = cur;
          cur = nxt;
          nxt = t;
        }
      }
      for (int L = 0; L < nlev; ++L) out[(size_t)c * (nlev + 1) + L] = (double)counts[levls[L]];
      out[(size_t)c * (nlev + 1) + nlev] = (double)((long)n * (n - 1) - counts[maxk]);
    }
    free(cur);
    free(nxt);
    free(counts);
    return CUSNA_OK;
  }
  return CUSNA_ERR_UNSUPPORTED;
}

/* ------------------------------------------------- multi-network periods */

/* Cross-network change statistic: ego i toggles i->j in its own network x
 * (dl = +1 adds, -1 removes) while source network w is read in its current
 * state. libcusna_saom_cpu.cpp:430-443. */
static double cross_change(int t, const uint8_t *w, const uint8_t *x, int n, int i, int j, double dl,
                           double parm) {
  switch (t) {
    case 33: /* crprod: w_ij */
      return dl * (double)w[(size_t)i * n + j];
    case 34: /* crprodRecip: w_ji */
      return dl * (double)w[(size_t)j * n + i];
    case 35: /* crprodMutual: w_ij w_ji */
      return dl * (double)(w[(size_t)i * n + j] & w[(size_t)j * n + i]);
    case 36: { /* from: shared out-ties of i and j in w */
      int64_t c = 0;
      for (int h = 0; h < n; ++h) c += w[(size_t)i * n + h] & w[(size_t)j * n + h];
      return dl * (parm >= 2.0 ? sqrt((double)c) : (double)c);
    }
    case 37: { /* to: two-paths i -w-> h -x-> j */
      int64_t c = 0;
      for (int h = 0; h < n; ++h) c += w[(size_t)i * n + h] & x[(size_t)h * n + j];
      return dl * (double)c;
    }
  }
  return 0.0;
}

typedef struct multi_ctx {
  int n, k, D, nact, need_tp, nd;
  const uint8_t *x0, *pm; /* D*n*n dense start networks and permitted toggles */
  const double *lam, *th; /* M*D, M*k */
  const int32_t *tid, *ety, *netof, *netref, *dci, *symf, *act;
  const double *wtab, *p1, *p2, *vmat, *dyc;
  uint32_t k0, k1;
} multi_ctx;

/* One chain of a multi-network period: end networks a (D*n*n), effect scores
 * sc (k), rate scores rsc (D). libcusna_saom_cpu.cpp:460-510. */
static void run_chain_multi(const multi_ctx *C, uint32_t c, uint8_t *a, double *sc, double *rsc,
                            double *ntim, int64_t *nsteps) {
  const int n = C->n, k = C->k, D = C->D;
  const size_t nn = (size_t)n * n;
  const double *lam = C->lam + (size_t)c * D;
  const double *th = C->th + (size_t)c * k;
  int64_t *indeg = calloc((size_t)D * n, sizeof(int64_t));
  int64_t *outdeg = calloc((size_t)D * n, sizeof(int64_t));
  int64_t *steps_d = calloc((size_t)D, sizeof(int64_t));
  int *tp = calloc((size_t)n, sizeof(int));
  double *fj = malloc(sizeof(double) * (size_t)n);
  double *ds = malloc(sizeof(double) * (size_t)n * (size_t)k);
  double *wsum = malloc(sizeof(double) * (size_t)k);
  uint8_t *okj = malloc((size_t)n);

  memcpy(a, C->x0, (size_t)D * nn);
  for (int d = 0; d < D; ++d)
    for (int r = 0; r < n; ++r)
      for (int s = 0; s < n; ++s)
        if (a[d * nn + (size_t)r * n + s]) {
          outdeg[(size_t)d * n + r]++;
          indeg[(size_t)d * n + s]++;
        }
  double lam_tot = 0.0;
  for (int d = 0; d < D; ++d) lam_tot += lam[d];
  for (int e = 0; e < k; ++e) sc[e] = 0.0;

  int64_t steps = 0;
  double t = 0.0;
  uint32_t r4[4];
  for (;;) {
    draw4((uint32_t)steps, c, 0u, 0u, C->k0, C->k1, r4);
    t += -log(unit_open(r4[0])) / ((double)C->nact * lam_tot);
    if (t > 1.0) break;
    /* network in proportion to its rate, then an active actor */
    const double uv = unit_open(r4[1]) * lam_tot;
    int d = 0;
    double acc = lam[0];
    while (d < D - 1 && uv >= acc) acc += lam[++d];
    int64_t x = (int64_t)(unit_open(r4[2]) * C->nact);
    if (x >= C->nact) x = C->nact - 1;
    const int i = C->act[x];
    uint8_t *ad = a + (size_t)d * nn;
    chain_view V = {n, ad, indeg + (size_t)d * n, outdeg + (size_t)d * n, tp, 0};
    if (C->need_tp) {
      for (int h = 0; h < n; ++h) tp[h] = 0;
      for (int y = 0; y < n; ++y)
        if (ad[(size_t)i * n + y])
          for (int h = 0; h < n; ++h) tp[h] += ad[(size_t)y * n + h];
    }
    double gbest = -1e30, lmax = -1e30;
    int jstar = i;
    for (int j = 0; j < n; ++j) {
      double f = 0.0;
      int ok = 1;
      if (j == i) {
        for (int e = 0; e < k; ++e) ds[(size_t)j * k + e] = 0.0;
      } else if (!C->pm[(size_t)d * nn + (size_t)i * n + j]) {
        ok = 0;
      } else {
        const int aij = ad[(size_t)i * n + j];
        const double dl = 1.0 - 2.0 * aij;
        for (int e = 0; e < k; ++e) {
          if (C->netof[e] != d) {
            ds[(size_t)j * k + e] = 0.0;
            continue;
          }
          double v;
          if (C->tid[e] >= 33) {
            v = cross_change(C->tid[e], a + (size_t)C->netref[e] * nn, ad, n, i, j, dl, C->p1[e]);
          } else {
            const int dc = C->dci[e];
            const double dyw = (dc >= 0 && C->dyc) ? C->dyc[((size_t)dc * n + i) * n + j] : 0.0;
            v = ego_change(C->tid[e], &V, i, j, dl, C->vmat[(size_t)e * n + i],
                           C->vmat[(size_t)e * n + j], dyw, C->p1[e], C->p2[e],
                           C->wtab + (size_t)e * (n + 1));
          }
          if ((C->ety[e] == 1 && aij == 1) || (C->ety[e] == 2 && aij == 0)) v = 0.0;
          ds[(size_t)j * k + e] = v;
          const double prod = th[e] * v;
          f = f + prod;
        }
      }
      okj[j] = (uint8_t)ok;
      fj[j] = f;
      if (ok && f > lmax) lmax = f;
      draw4((uint32_t)steps, c, 1u, (uint32_t)j, C->k0, C->k1, r4);
      if (ok) {
        const double g = f - log(-log(unit_open(r4[0])));
        if (g > gbest || (g == gbest && j < jstar)) {
          gbest = g;
          jstar = j;
        }
      }
    }
    double se = 0.0;
    for (int e = 0; e < k; ++e) wsum[e] = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!okj[j]) continue;
      const double w = exp(fj[j] - lmax);
      se += w;
      for (int e = 0; e < k; ++e) wsum[e] += w * ds[(size_t)j * k + e];
    }
    for (int e = 0; e < k; ++e)
      if (C->netof[e] == d) sc[e] += ds[(size_t)jstar * k + e] - wsum[e] / se;
    steps++;
    steps_d[d]++;
    if (jstar != i) {
      const int j = jstar;
      const int64_t dd = 1 - 2 * (int64_t)ad[(size_t)i * n + j];
      indeg[(size_t)d * n + j] += dd;
      outdeg[(size_t)d * n + i] += dd;
      ad[(size_t)i * n + j] ^= 1u;
      if (C->symf[d]) {
        indeg[(size_t)d * n + i] += dd;
        outdeg[(size_t)d * n + j] += dd;
        ad[(size_t)j * n + i] ^= 1u;
      }
    }
    if (steps >= ORACLE_MAX_STEPS) break;
  }
  for (int d = 0; d < D; ++d) rsc[d] = (double)steps_d[d] / lam[d] - (double)C->nact;
  *ntim = t;
  *nsteps = steps;
  free(indeg);
  free(outdeg);
  free(steps_d);
  free(tp);
  free(fj);
  free(ds);
  free(wsum);
  free(okj);
}

/* M chains of a multi-network period. libcusna_saom_cpu.cpp:548-604, except
 * that dyadic covariate q of nd is read at dyc + q*n*n as the ABI documents
 * (the reference copies only the first n*n values; with dyc_idx in {-1, 0}
 * both agree). */
ORACLE_API int oracle_saom_simulate_multi(const cusna_saom_multi_problem *pb,
                                          cusna_saom_multi_result *r, int32_t nd) {
  if (!pb || !r) return CUSNA_ERR_NULL_ARG;
  if (pb->struct_version != CUSNA_ABI_VERSION) return CUSNA_ERR_SHAPE;
  const int n = pb->n, W = pb->W, k = pb->k, M = pb->M, D = pb->D;
  if (n < 1 || n > CUSNA_MAX_N || W != (n + 31) / 32 || k < 1 || k > CUSNA_MAXEFF || M < 1 ||
      D < 2 || D > CUSNA_MAXNET)
    return CUSNA_ERR_SHAPE;
  const size_t nn = (size_t)n * n;
  uint8_t *x0 = malloc((size_t)D * nn), *pm = malloc((size_t)D * nn);
  for (int d = 0; d < D; ++d) {
    unpack_rows(pb->x0_out + (size_t)d * n * W, n, W, x0 + d * nn);
    if (pb->perm) unpack_rows(pb->perm + (size_t)d * n * W, n, W, pm + d * nn);
    else offdiag(n, pm + d * nn);
  }
  double *lam = malloc(sizeof(double) * (size_t)M * D), *th = malloc(sizeof(double) * (size_t)M * k);
  double *wtab = calloc((size_t)k * (n + 1), sizeof(double)), *p1 = calloc((size_t)k, 8),
         *p2 = calloc((size_t)k, 8), *vmat = calloc((size_t)k * n, 8);
  for (size_t q = 0; q < (size_t)M * D; ++q) lam[q] = pb->lam[q];
  for (size_t q = 0; q < (size_t)M * k; ++q) th[q] = pb->theta[q];
  for (int e = 0; e < k; ++e) {
    if (pb->p1) p1[e] = pb->p1[e];
    if (pb->p2) p2[e] = pb->p2[e];
  }
  if (pb->wtab)
    for (size_t q = 0; q < (size_t)k * (n + 1); ++q) wtab[q] = pb->wtab[q];
  if (pb->vmat)
    for (size_t q = 0; q < (size_t)k * n; ++q) vmat[q] = pb->vmat[q];
  double *dyc = NULL;
  if (pb->dyc && nd > 0) {
    dyc = malloc(sizeof(double) * nd * nn);
    for (size_t q = 0; q < (size_t)nd * nn; ++q) dyc[q] = pb->dyc[q];
  }
  int32_t *act = malloc(sizeof(int32_t) * (size_t)n);
  int32_t *dci = malloc(sizeof(int32_t) * (size_t)k);
  const int nact = pb->act ? pb->nact : n;
  for (int q = 0; q < n; ++q) act[q] = pb->act ? pb->act[q] : q;
  for (int e = 0; e < k; ++e) dci[e] = pb->dyc_idx ? pb->dyc_idx[e] : -1;
  int need_tp = 0;
  for (int e = 0; e < k; ++e) need_tp |= pb->type_ids[e] == 5;

  multi_ctx C = {n, k, D, nact, need_tp, nd, x0, pm, lam, th, pb->type_ids, pb->etypes,
                 pb->netof, pb->netref, dci, pb->symf, act, wtab, p1, p2, vmat, dyc,
                 (uint32_t)(pb->seed & 0xFFFFFFFFu), (uint32_t)(pb->seed >> 32)};
#pragma omp parallel for schedule(static)
  for (int c = 0; c < M; ++c) {
    uint8_t *a = malloc((size_t)D * nn);
    double *sc = malloc(sizeof(double) * (size_t)k), *rs = malloc(sizeof(double) * (size_t)D);
    double nt;
    int64_t st;
    run_chain_multi(&C, (uint32_t)c, a, sc, rs, &nt, &st);
    if (r->ntim) r->ntim[c] = nt;
    if (r->steps) r->steps[c] = (int32_t)st;
    if (r->rscores)
      for (int d = 0; d < D; ++d) r->rscores[(size_t)c * D + d] = (float)rs[d];
    if (r->scores)
      for (int e = 0; e < k; ++e) r->scores[(size_t)c * k + e] = (float)sc[e];
    if (r->a_ends) memcpy(r->a_ends + (size_t)c * D * nn, a, (size_t)D * nn);
    free(a);
    free(sc);
    free(rs);
  }
  free(x0);
  free(pm);
  free(lam);
  free(th);
  free(wtab);
  free(p1);
  free(p2);
  free(vmat);
  free(dyc);
  free(act);
  free(dci);
  return CUSNA_OK;
}

/* Cross-network statistic of M networks against predictor wp.
 * libcusna_saom_cpu.cpp:819-857. */
ORACLE_API int oracle_cross_stats(const uint8_t *a_batch, const uint8_t *wp, int32_t n, int32_t M,
                                  int32_t t, double parm, double *out) {
  if (!a_batch || !wp || !out || n < 1 || n > CUSNA_MAX_N || M < 1) return CUSNA_ERR_NULL_ARG;
  if (t < 33 || t > 37) return CUSNA_ERR_UNSUPPORTED;
  const size_t nn = (size_t)n * n;
  for (int c = 0; c < M; ++c) {
    const uint8_t *a = a_batch + (size_t)c * nn;
    double acc = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const uint8_t aij = a[(size_t)i * n + j];
        switch (t) {
          case 33: acc += aij & wp[(size_t)i * n + j]; break;
          case 34: acc += aij & wp[(size_t)j * n + i]; break;
          case 35: acc += aij & wp[(size_t)i * n + j] & wp[(size_t)j * n + i]; break;
          case 36:
            if (aij) {
              int64_t s = 0;
              for (int h = 0; h < n; ++h) s += wp[(size_t)i * n + h] & wp[(size_t)j * n + h];
              acc += parm >= 2.0 ? sqrt((double)s) : (double)s;
            }
            break;
          case 37:
            if (aij) {
              int64_t s = 0;
              for (int h = 0; h < n; ++h) s += wp[(size_t)i * n + h] & a[(size_t)h * n + j];
              acc += (double)s;
            }
            break;
        }
      }
    out[c] = acc;
  }
  return CUSNA_OK;
}
