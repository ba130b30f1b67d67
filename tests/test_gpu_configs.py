"""GPU parity at the benchmark configurations' own shapes.

* the compiled reference's outputs at 500 actors (8 chains), 300-actor
  co-evolution, 1000 actors with twelve effects, and groups of 1045-1188
  actors (tests/golden/large_cases.json.gz): the generic and the
  runtime-specialised kernels reproduce them (end networks by SHA-256);
* replay at 500 actors: with the host's recorded noise the device reproduces
  the reference's free run exactly, ntim included;
* the uneven multi-group batch: 24 groups of 44..1188 actors (five of them
  above 1024 actors, rows in device memory) with the batch's effect list, 4
  chains each, through one cusna_saom_simulate_groups call, per group against
  the oracle;
* a multi-group launch mixing groups with and without rate effects under
  forced specialisation runs each group's own rate path.

Bar as the parity tests (tests/test_gpu_parity.py).
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from conftest import load_golden
from test_gpu_parity import assert_period_equal

pytestmark = pytest.mark.gpu


def _large_cases():
    return load_golden("large_cases.json.gz")["cases"]


def _host_workers() -> int:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(16, cores))


def _plan_run(gpu_lib, prob, specialize):
    plan = gpu_lib.PeriodPlan(prob, want_end_nets=True)
    try:
        plan.specialize(specialize)
        plan.run(prob.seed)
        got = plan.fetch(want_end_nets=True)
        used = plan.specialized
    finally:
        plan.close()
    return got, used


@pytest.mark.parametrize("specialize", [0, 1])
@pytest.mark.parametrize("idx", range(6))
def test_large_reference_case(gpu_lib, idx, specialize):
    from large_cases import end_hashes, large_problem

    case = _large_cases()[idx]
    prob = large_problem(case["spec"])
    print(f"case {case['spec']['name']} n={prob.n} k={prob.k} M={prob.M} "
          f"specialize={specialize}", flush=True)
    got, used = _plan_run(gpu_lib, prob, specialize)
    assert used == bool(specialize)
    assert_period_equal(got, case["out"], label=case["spec"]["name"])
    assert end_hashes(got["a_ends"]) == case["out"]["ends_sha256"]


def test_replay_n500_reproduces_reference(gpu_lib, oracle_lib):
    """Replay of the 500-actor case with the noise the host C library computes
    (log u0, log(-log u_j)): identical to the reference's free run, ntim
    bit for bit."""
    from large_cases import end_hashes, large_problem

    case = _large_cases()[0]
    assert case["spec"]["name"] == "c2_n500"
    prob = large_problem(case["spec"])
    smax = int(max(case["out"]["steps"])) + 64
    print(f"case replay c2_n500 M={prob.M} smax={smax}", flush=True)
    lu0, llu = oracle_lib.replay_noise(prob.seed, prob.M, prob.n, smax)
    got = gpu_lib.simulate_period_replay(prob, lu0, llu, want_end_nets=True)
    assert_period_equal(got, case["out"], exact_ntim=True, label="replay c2_n500")
    assert end_hashes(got["a_ends"]) == case["out"]["ends_sha256"]


def _oracle_groups(oracle_lib, probs):
    """Oracle outputs of every group, groups on parallel host threads (each
    call's own OpenMP team runs the group's chains)."""
    per = max(1, min(4, _host_workers()))

    def one(pr):
        oracle_lib.oracle().oracle_set_threads(per)
        return oracle_lib.simulate(pr, want_end_nets=True)

    with cf.ThreadPoolExecutor(max(1, _host_workers() // per)) as ex:
        return list(ex.map(one, probs))


def test_config3_mixed_batch(gpu_lib, oracle_lib):
    """24 groups of the uneven batch (44..1188 actors, 5 streamed) with its
    effects density/recip/transTrip/inPop/egoX/altX on a Bernoulli(0.5)
    covariate, 4 chains each, in one groups call."""
    from cusna_amd.workloads import mixed_group_problems

    probs = mixed_group_problems(groups=24, chains=4)
    ns = [p.n for p in probs]
    assert sum(n > 1024 for n in ns) >= 4 and min(ns) < 100
    print(f"case config3 batch G=24 n={min(ns)}..{max(ns)}", flush=True)
    got = gpu_lib.simulate_groups(probs, want_end_nets=True)
    print("device done; oracle", flush=True)
    refs = _oracle_groups(oracle_lib, probs)
    for g, (pr, r, ref) in enumerate(zip(probs, got, refs)):
        assert_period_equal(r, ref, label=f"group {g} n={pr.n}")
        np.testing.assert_array_equal(r["a_ends"], ref["a_ends"], err_msg=f"group {g}")


def test_groups_rate_and_plain_specialised(gpu_lib, oracle_lib, monkeypatch):
    """Two groups with the same effect list and launch shape, the first (the
    longer, so the bucket's lead) without rate effects, the second with
    outRate and a covariate rate effect: under forced specialisation each
    group runs its own rate path."""
    from rate_cases import _BASE, rate_problem

    a = dict(_BASE, name="plain", n=200, seed=701, net_seed=702, lam=6.0, M=4)
    b = dict(_BASE, name="rated", n=210, seed=703, net_seed=704, lam=1.5, M=4,
             rate=[[1, 0.05], [0, 0.2]])
    probs = [rate_problem(a), rate_problem(b)]
    assert probs[0].nrate == 0 and probs[1].nrate == 2
    monkeypatch.setenv("CUSNA_SPECIALIZE", "1")
    got = gpu_lib.simulate_groups(probs, want_end_nets=True)
    for g, (pr, r) in enumerate(zip(probs, got)):
        ref = oracle_lib.simulate(pr, want_end_nets=True)
        assert_period_equal(r, ref, label=f"group {g}")
        np.testing.assert_array_equal(r["a_ends"], ref["a_ends"])
        if pr.nrate:
            np.testing.assert_array_equal(r["actor_dist"], ref["actor_dist"])
            g_, r_ = np.asarray(r["rscores"], float), np.asarray(ref["rscores"], float)
            assert np.all(np.abs(g_ - r_) <= 1e-5 * np.maximum(1.0, np.abs(r_)))
