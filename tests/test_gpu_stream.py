"""GPU: networks above the LDS (n > 1024): rows streamed from device memory.

The streamed kernel shapes (word bounds 32, 64, 128) and the statistics
kernels with device scratch are checked against the oracle at the parity
tests' bar, across effect families, conditional runs and co-evolution.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_period_equal

pytestmark = pytest.mark.gpu


def _net(n, deg, seed):
    from cusna_amd.workloads import bernoulli_network

    return bernoulli_network(n, deg, np.random.default_rng(seed))


def _problem(n, names, theta, lam, M, seed, **kw):
    from cusna_amd.effects import Effect
    from cusna_amd.workloads import effect_problem

    rng = np.random.default_rng(seed + 7)
    cov = rng.integers(1, 6, n).astype(float)
    eff = [Effect(x, cov) if x in ("egoX", "altX", "simX") else Effect(x) for x in names]
    return effect_problem(_net(n, 3.0, seed), eff, theta, lam, M, seed, **kw)


CASES = {
    "n1100_simple": (1100, ["density", "recip", "transTrip", "cycle3", "egoX", "simX"],
                     [-3.0, 1.8, 0.3, -0.1, 0.1, 0.5], 1.5, 6),
    "n1100_general": (1100, ["density", "recip", "transTies", "gwesp", "inPopSqrt"],
                      [-3.0, 1.5, 0.2, 0.4, -0.1], 1.0, 4),
    "n2100_degree": (2100, ["density", "recip", "inPop", "outAct", "altX"],
                     [-3.2, 1.6, 0.02, -0.02, 0.1], 0.6, 3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_streamed_period_matches_oracle(case, gpu_lib, oracle_lib):
    n, names, theta, lam, M = CASES[case]
    pr = _problem(n, names, theta, lam, M, 31)
    got = gpu_lib.simulate_period(pr)
    ref = oracle_lib.simulate(pr)
    assert_period_equal(got, ref, label=case)
    assert np.all(np.asarray(got["steps"]) > 100)


def test_streamed_conditional(gpu_lib, oracle_lib):
    pr = _problem(1300, ["density", "recip", "transTrip"], [-3.0, 1.8, 0.3], 1.0, 4, 41,
                  conditional=True, target_dist=150)
    got = gpu_lib.simulate_period(pr)
    ref = oracle_lib.simulate(pr)
    assert_period_equal(got, ref, label="conditional n1300")
    assert np.all(np.asarray(got["dist"]) >= 150)


def test_streamed_coevolution(gpu_lib, oracle_lib):
    from cusna_amd.data import build_saom_data
    from cusna_amd.effects import BehEffect, Effect
    from cusna_amd.workloads import effect_problem

    n = 1100
    rng = np.random.default_rng(5)
    x = _net(n, 3.0, 5)
    z = rng.binomial(4, 0.4, size=(n, 2)) + 1
    d = build_saom_data([x, x], behavior=z.astype(float))
    eff = [Effect("density"), Effect("recip"), Effect("transTrip"), Effect("simX", dyn=True)]
    bef = [BehEffect("linear"), BehEffect("quad"), BehEffect("avAlt")]
    pr = effect_problem(x, eff, [-3.0, 1.6, 0.3, 0.8], 1.0, 4, 55, beh=d.beh,
                        z0=d.beh.imputed[:, 0], beh_effects=bef, thetab=[-0.2, -0.3, 0.8],
                        lamb=1.0)
    got = gpu_lib.simulate_period(pr)
    ref = oracle_lib.simulate(pr)
    assert_period_equal(got, ref, label="coevol n1100")


def test_statistics_above_lds(gpu_lib, oracle_lib):
    from cusna_amd.effects import Effect, prepare_effects

    n = 1200  # 19 words per row: rows alone exceed the LDS
    rng = np.random.default_rng(9)
    cov = rng.normal(size=n)
    eff = [Effect("density"), Effect("recip"), Effect("transTrip"), Effect("cycle3"),
           Effect("gwespBF"), Effect("gwespBB"), Effect("inPop"), Effect("simX", cov),
           Effect("egoX", cov)]
    d = prepare_effects(eff, n, 1)
    for seed in (1,):
        a = _net(n, 4.0, 100 + seed)
        kw = dict(type_ids=d.type_ids, p1=d.p1, p2=d.p2, vmat=d.vmat, vmiss=d.vmiss,
                  dyn_flags=d.dyn_flags, vb=d.v_dyn, wtab=d.wtab, symflag=0)
        got = gpu_lib.network_stats(a, **kw)
        ref = oracle_lib.network_stats(a, d.type_ids, d.p1, d.p2, d.vmat, d.vmiss, d.dyn_flags,
                                       d.v_dyn, None, None, d.wtab, 0)
        np.testing.assert_array_equal(np.asarray(got).view(np.uint64),
                                      np.asarray(ref, np.float64).view(np.uint64))


def test_behavior_statistics_above_lds(gpu_lib, oracle_lib):
    n = 4200
    rng = np.random.default_rng(11)
    x = _net(n, 3.0, 12)
    z = (rng.binomial(4, 0.4, size=(3, n)) + 1).astype(np.int32)
    bty = np.array([0, 1, 2, 3, 4], np.int32)
    bk = (rng.random(n) < 0.95).astype(np.int32)
    got = gpu_lib.behavior_stats(z, bty, x, bk, 2.6, 4.0, 0.7)
    ref = oracle_lib.behavior_stats(z, bty, x, bk, 2.6, 4.0, 0.7)
    np.testing.assert_array_equal(np.asarray(got).view(np.uint64),
                                  np.asarray(ref, np.float64).view(np.uint64))
