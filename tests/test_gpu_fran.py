"""GPU: the FRAN closure for RSiena's siena07() (R/fran.R:27-63), the cases of
tests/testthat/test-fran.R that need no RSiena, on the 50-actor three-wave
panel of the reference's estimate fixture.

Bars: the closure's rows are the period pieces of the device simulator, and
those pieces match the oracle run on the same seeds -- statistics and the
rate statistic bit-exact, ntim to 1e-12 relative, scores to
1e-5 * max(1, |ref|) (float32 outputs of float64 sums)."""
from dataclasses import replace

import numpy as np
import pytest

from mom_cases import estimate_fixture

pytestmark = pytest.mark.gpu

EFFS = ["density", "recip", "transTrip"]
SCORE_TOL = 1e-5


def _waves():
    g = estimate_fixture()
    return [np.array(w, float) for w in g["inputs"]["s50_waves"]]


def _fran(names, **kw):
    from cusna_amd.fran import cusna_fran

    return cusna_fran(_waves(), names, **kw)


def test_fra_and_sc_are_periods_by_parameters_conditional(gpu_lib):
    fran = _fran(EFFS)
    np.random.seed(1)
    r = fran({"theta": [-2.5, 2.4, 0.6], "cconditional": True})
    assert r["OK"] and r["feasible"]
    assert r["fra"].shape == (2, 3) and r["sc"].shape == (2, 3)
    assert r["ntim0"].shape == (2,) and np.all(r["ntim0"] > 0)
    assert r["fra"].sum(axis=0).shape == (3,)


def test_unconditional_rows_carry_each_periods_rate_statistic(gpu_lib):
    fran = _fran(["density", "recip"], conditional=False)
    np.random.seed(2)
    r = fran({"theta": [6.0, 5.0, -2.5, 2.4], "cconditional": False})
    assert r["fra"].shape == (2, 4) and r["sc"].shape == (2, 4)
    assert r["fra"][0, 1] == 0 and r["fra"][1, 0] == 0
    assert r["fra"][0, 0] > 0 and r["fra"][1, 1] > 0
    assert r["sc"][0, 1] == 0 and r["sc"][1, 0] == 0


@pytest.mark.parametrize("conditional,theta", [(True, [-2.5, 2.4, 0.6]),
                                               (False, [6.0, 5.0, -2.5, 2.4, 0.6])])
def test_rows_are_the_period_pieces_and_match_the_oracle(gpu_lib, oracle_lib, conditional, theta):
    fran = _fran(EFFS, conditional=conditional)
    est = fran.est
    np.random.seed(3)
    r = fran({"theta": theta})
    np.random.seed(3)
    s = est.simulate(np.asarray(theta), 1)
    for q in range(est.P):
        np.testing.assert_array_equal(r["fra"][q], s["pieces"][q]["sp"][0])
        np.testing.assert_array_equal(r["sc"][q], s["pieces"][q]["scp"][0])
    np.testing.assert_array_equal(r["fra"].sum(axis=0), s["stats"][0])
    np.testing.assert_array_equal(r["ntim0"], s["ntim"][0])

    # each period's piece against the oracle simulator on that period's seed
    k, P = est.k, est.P
    off = 0 if conditional else P
    th = np.asarray(theta[off:], float)
    for q, pc in enumerate(s["pieces"]):
        lam = 1.0 if conditional else float(theta[q])
        ref = oracle_lib.simulate(replace(est.problems[q], theta=th.copy(), lam=lam, M=1,
                                          seed=pc["seed"]))
        np.testing.assert_array_equal(pc["sp"][0, off:off + k], np.asarray(ref["stats"])[0])
        rsc = np.asarray(ref["scores"], np.float64).reshape(1, k + 1)
        gsc = pc["scp"][0]
        assert np.all(np.abs(gsc[off:off + k] - rsc[0, 1:]) <= SCORE_TOL *
                      np.maximum(1.0, np.abs(rsc[0, 1:])))
        np.testing.assert_allclose(s["ntim"][0, q], np.asarray(ref["ntim"])[0], rtol=1e-12)
        if not conditional:
            assert pc["sp"][0, q] == np.asarray(ref["dist_stat"])[0] + est.rate_consts[q]
            assert abs(gsc[q] - rsc[0, 0]) <= SCORE_TOL * max(1.0, abs(rsc[0, 0]))


def test_closure_counts_calls_and_leaves_the_stream_to_the_caller(gpu_lib):
    fran = _fran(EFFS)
    z = {"theta": [-2.5, 2.4, 0.6]}
    np.random.seed(4)
    r1 = fran(z)
    r2 = fran(z)
    assert fran.counter == 2
    assert not np.array_equal(r1["fra"], r2["fra"])   # successive calls draw new seeds
    np.random.seed(4)
    r1b = fran(z)
    np.testing.assert_array_equal(r1["fra"], r1b["fra"])   # ... reproducibly
    np.testing.assert_array_equal(r1["sc"], r1b["sc"])
    # an explicit generator gives the same stream handling without the global one
    a = fran(z, rng=np.random.default_rng(9))
    b = fran(z, rng=np.random.default_rng(9))
    np.testing.assert_array_equal(a["fra"], b["fra"])


def test_robbins_monro_on_the_closure_solves_the_moment_equation(gpu_lib):
    """siena07()'s phase-2 form driven through the closure, one chain per
    call: theta <- theta - a * diag(D)^-1 (colSums(fra) - target), with D the
    score-function derivative of 200 chains at the estimator's default start
    (sum over periods of cov(sp, scp), R/mom.R:186-195) and the gain halved
    every 100 calls. The average of the last 100 iterates must solve the
    moment equation: at it, 1000 chains give t-ratios
    (mean - target) / sd below 0.3 for every effect."""
    fran = _fran(EFFS)
    est = fran.est
    theta = est.default_theta0()
    np.random.seed(5)
    s = est.simulate(theta, 200)
    D = np.zeros((est.p, est.p))
    for pc in s["pieces"]:
        sc = pc["sp"] - pc["sp"].mean(axis=0)
        cc = pc["scp"] - pc["scp"].mean(axis=0)
        D += sc.T @ cc / sc.shape[0]
    dinv = 1.0 / np.diag(D)
    assert np.all(dinv > 0), np.diag(D)
    a, tail = 0.2, []
    for step in range(400):
        r = fran({"theta": theta})
        theta = theta - a * dinv * (r["fra"].sum(axis=0) - est.targets)
        if step % 100 == 99:
            a /= 2
        if step >= 300:
            tail.append(theta.copy())
    assert fran.counter == 400
    th = np.mean(tail, axis=0)
    assert np.all(np.isfinite(th))
    chk = est.simulate(th, 1000)["stats"]
    t = (chk.mean(axis=0) - est.targets) / chk.std(axis=0)
    assert np.all(np.abs(t) < 0.3), (th, t)
