"""CPU oracle of the method-of-moments algebra -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's R estimator algebra (cran/cusna
R/mom.R), used by the tests to check what the device engine reduces and
solves; nothing in the library imports it.

  pinv             R/mom.R:15-19   (SVD pseudo-inverse, tol sqrt(eps) * d_max)
  safe_solve       R/mom.R:20      (solve, else pinv)
  derivative       R/mom.R:173-182 (sum over periods of crossprod of centred
                                    moment and score rows / M)
  dinv             R/mom.R:184-187 ((1-d) D + d diag(D), inverted)
  guard_step       R/mom.R:189-197
  phase3_summary   R/mom.R:250-258 (sim means, msf, tconv, tconv_max, se)
  rm_update        R/mom.R:236-243 (one Robbins-Monro step)

Pinned by tests/test_mom_oracle.py against closed forms, brute-force loops
and numpy.linalg (the R code is not runnable here).
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(float).eps


def pinv(m: np.ndarray, tol: float = np.sqrt(EPS)) -> np.ndarray:
    """R/mom.R:15-19."""
    u, d, vt = np.linalg.svd(m)
    keep = d > max(tol * d[0], 0.0)
    return vt[keep].T @ (u[:, keep].T / d[keep][:, None])


def safe_solve(m: np.ndarray) -> np.ndarray:
    """R/mom.R:20: solve(m), falling back to pinv when R's solve() refuses
    (exactly singular or reciprocal condition number below eps)."""
    try:
        if np.linalg.cond(m, 1) * EPS >= 1.0:
            raise np.linalg.LinAlgError
        return np.linalg.inv(m)
    except np.linalg.LinAlgError:
        return pinv(m)


def derivative(sp: np.ndarray, scp: np.ndarray) -> np.ndarray:
    """R/mom.R:173-182; sp, scp are (P, M, p)."""
    P, M, p = sp.shape
    D = np.zeros((p, p))
    for q in range(P):
        s_c = sp[q] - sp[q].mean(0)
        sc_c = scp[q] - scp[q].mean(0)
        D += s_c.T @ sc_c / M
    return D


def dinv(D: np.ndarray, diagonalize: float) -> np.ndarray:
    """R/mom.R:184-187."""
    mix = (1 - diagonalize) * D + diagonalize * np.diag(np.diag(D))
    return safe_solve(mix)


def guard_step(theta: np.ndarray, step: np.ndarray, n_rates: int) -> np.ndarray:
    """R/mom.R:189-197: cap |step| at max(1, |theta|/2); a rate may not fall
    below a tenth of its value (else halve it)."""
    cap = np.maximum(1.0, 0.5 * np.abs(theta))
    step = np.minimum(np.maximum(step, -cap), cap)
    for m in range(n_rates):
        if theta[m] - step[m] < 0.1 * theta[m]:
            step[m] = 0.5 * theta[m]
    return step


def rm_update(theta: np.ndarray, stats: np.ndarray, targets: np.ndarray, dinv_m: np.ndarray,
              gain: float, n_rates: int) -> np.ndarray:
    """R/mom.R:236-243: theta - guard(gain * dinv (colMeans(stats) - targets))."""
    dev = stats.mean(0) - targets
    return theta - guard_step(theta, gain * (dinv_m @ dev), n_rates)


def phase3_summary(stats: np.ndarray, sp: np.ndarray, scp: np.ndarray,
                   targets: np.ndarray) -> dict:
    """R/mom.R:250-258."""
    D = derivative(sp, scp)
    di = safe_solve(D)
    msf = np.cov(stats, rowvar=False)
    dev = stats.mean(0) - targets
    tconv = dev / np.sqrt(np.diag(msf))
    tconv_max = float(np.sqrt(dev @ safe_solve(msf) @ dev))
    se = np.sqrt(np.abs(np.diag(di @ msf @ di.T)))
    return dict(D=D, msf=msf, sim_means=stats.mean(0), tconv=tconv, tconv_max=tconv_max,
                se=se)


def _period_rows(r: dict, q: int, P: int, p: int, k: int, kb: int, mode: int,
                 rate_const: float) -> tuple:
    """Moment and score rows of one period (R/mom.R:152-166); scores pass
    through float32 as the ABI returns them."""
    M = r["stats"].shape[0]
    off = 2 * P if mode == 2 else (0 if mode == 1 else P)
    sp, scp = np.zeros((M, p)), np.zeros((M, p))
    sc = np.asarray(r["scores"], np.float32).astype(float).reshape(M, k + 1)
    if mode != 1:
        sp[:, q] = r["dist_stat"] + rate_const
        scp[:, q] = sc[:, 0]
    if mode == 2:
        bsc = np.asarray(r["bscores"], np.float32).astype(float).reshape(M, kb + 1)
        sp[:, P + q] = r["bdist"]
        scp[:, P + q] = bsc[:, 0]
        sp[:, off + k:] = r["bstats"]
        scp[:, off + k:] = bsc[:, 1:]
    sp[:, off:off + k] = r["stats"]
    scp[:, off:off + k] = sc[:, 1:]
    return sp, scp


def mom_run(simulate_rows, P: int, p: int, targets: np.ndarray, theta0: np.ndarray, mode: int,
            n1: int, nD: int, batch2: int, n3: int, n2, firstg: float = 0.4,
            diagonalize: float = 0.2, seed: int = 1234) -> dict:
    """CPU driver of the three phases (R/mom.R:218-269). ``simulate_rows(q,
    theta, M, seed)`` runs period q for M chains and returns (sp, scp) of
    that period (R/mom.R:143-168 per period); seeds come from numpy instead of
    R's sample.int."""
    rng = np.random.default_rng(seed)
    n_rates = 0 if mode == 1 else (2 * P if mode == 2 else P)
    state = {"n_sims": 0}

    def simulate(theta, M):
        sp = np.zeros((P, M, p))
        scp = np.zeros((P, M, p))
        for q in range(P):
            sp[q], scp[q] = simulate_rows(q, theta, M, int(rng.integers(1, 2**31 - 1)))
        state["n_sims"] += M
        return sp.sum(0), sp, scp

    theta = np.asarray(theta0, float).copy()
    s, sp, scp = simulate(theta, n1)
    di = dinv(derivative(sp, scp), diagonalize)
    theta = rm_update(theta, s, targets, di, 0.5 * firstg, n_rates)
    for sub, nit in enumerate(n2):
        gain = firstg * 0.5 ** sub
        s, sp, scp = simulate(theta, nD)
        di = dinv(derivative(sp, scp), diagonalize)
        thav = np.zeros(p)
        for _ in range(int(nit)):
            s, _, _ = simulate(theta, batch2)
            theta = rm_update(theta, s, targets, di, gain, n_rates)
            thav += theta
        theta = thav / int(nit)
    s, sp, scp = simulate(theta, n3)
    out = phase3_summary(s, sp, scp, targets)
    out.update(theta=theta, n_sims=state["n_sims"])
    return out


def reference_rows(est, simulate, behavior_stats=None):
    """``simulate_rows`` for ``mom_run`` over an estimator context ``est``
    (its per-period problems, layout and, for co-evolution, start networks):
    each call copies the period's problem with M chains, the float32-rounded
    parameters of theta and the seed, and runs ``simulate`` on it (the
    reference engine or the C oracle)."""
    import copy

    P, p, k, kb, mode = est.P, est.p, est.k, est.kb, est.mode

    def rows(q, theta, M, seed):
        pr = copy.copy(est.problems[q])
        pr._keep = []
        pr.M, pr.seed = int(M), int(seed)
        if mode == 2:
            pr.lam, pr.lamb = float(theta[q]), float(theta[P + q])
            pr.theta, pr.thetab = theta[2 * P:2 * P + k], theta[2 * P + k:]
        elif mode == 1:
            pr.lam, pr.theta = 1.0, theta
        else:
            pr.lam, pr.theta = float(theta[q]), theta[P:]
        r = simulate(pr)
        if mode == 2:
            z = np.asarray(r["z_end"]).reshape(M, est.n)
            r["bstats"] = behavior_stats(z, pr.bty, est.xnets[q], pr.bkdist, pr.zbar,
                                         pr.brange, pr.bsimmean)
        return _period_rows(r, q, P, p, k, kb, mode, float(est.rate_consts[q]))

    return rows
