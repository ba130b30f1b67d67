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
