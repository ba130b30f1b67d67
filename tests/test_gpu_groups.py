"""GPU: several groups in one launch, and the block-shape autotuner.

* cusna_saom_simulate_groups returns, for every group, what the one-group
  call returns (same kernel, same block shape: identical bits), for groups of
  different sizes, effect sets and behaviour, across launch shapes;
* the groups also match the oracle (the bar of the parity tests);
* the autotuner changes the block shape, not the trajectories: dist, steps,
  statistics and end networks identical; scores to the parity tolerance.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_period_equal

pytestmark = pytest.mark.gpu


def _groups():
    from cusna_amd.effects import Effect
    from cusna_amd.workloads import (bernoulli_network, n500_problem, coevo300_problem,
                                     effect_problem, poisson_capped_network)

    rng = np.random.default_rng(21)
    out = []
    x = poisson_capped_network(50, 2.3, 6, 0.4, rng)
    eff = [Effect("density"), Effect("recip"), Effect("transTrip")]
    out.append(effect_problem(x, eff, [-2.5, 2.0, 0.4], 6.0, 24, 101))
    out.append(effect_problem(x, eff, [-2.0, 1.5, 0.2], 4.0, 9, 102))
    x2 = bernoulli_network(130, 4.0, rng)
    cov = rng.integers(1, 6, 130).astype(float)
    eff2 = [Effect("density"), Effect("recip"), Effect("cycle3"), Effect("inPop"),
            Effect("outAct"), Effect("egoX", cov), Effect("simX", cov)]
    out.append(effect_problem(x2, eff2, [-2.6, 1.7, -0.1, 0.05, -0.05, 0.1, 0.5], 5.0, 17,
                              103))
    out.append(n500_problem(M=12, n=500))
    out.append(coevo300_problem(M=8, n=300))
    # conditional group of the first network
    out.append(effect_problem(x, eff, [-2.5, 2.0, 0.4], 1.0, 11, 104, conditional=True,
                              target_dist=40))
    return out


def test_groups_match_single_calls(gpu_lib, oracle_lib):
    probs = _groups()
    got = gpu_lib.simulate_groups(probs, want_end_nets=True)
    for g, (pr, r) in enumerate(zip(probs, got)):
        one = gpu_lib.simulate_period(pr, want_end_nets=True)
        for key in one:
            np.testing.assert_array_equal(np.asarray(r[key]), np.asarray(one[key]),
                                          err_msg=f"group {g} {key}")
        ref = oracle_lib.simulate(pr)
        assert_period_equal(r, ref, label=f"group {g}")


def test_groups_single_group_and_errors(gpu_lib):
    from cusna_amd.abi import CusnaError

    pr = _groups()[0]
    a = gpu_lib.simulate_groups([pr])[0]
    b = gpu_lib.simulate_period(pr)
    np.testing.assert_array_equal(a["steps"], b["steps"])
    bad = _groups()[1]
    bad.type_ids = np.array([0, 1, 99], np.int32)  # unknown effect id
    with pytest.raises(CusnaError):
        gpu_lib.simulate_groups([pr, bad])


def test_autotune_keeps_trajectories(gpu_lib):
    from cusna_amd.workloads import n500_problem

    pr = n500_problem(M=48, n=200)
    plan = gpu_lib.PeriodPlan(pr)
    try:
        plan.run(77)
        a = plan.fetch()
        threads = plan.autotune(probe_chains=24)
        assert threads in (64, 128, 256, 512, 1024)
        assert plan.block_threads == threads
        plan.run(77)
        b = plan.fetch()
    finally:
        plan.close()
    for key in ("dist", "dist_stat", "steps", "stats", "ntim"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    np.testing.assert_allclose(np.asarray(b["scores"], float), np.asarray(a["scores"], float),
                               rtol=1e-5, atol=1e-5)
