"""GPU: non-basic rate effects (nrate > 0) and composition change in the
period kernel, against the reference engine's outputs
(tests/golden/rate_cases.json.gz) and the oracle.

Bar as the parity tests: trajectories, end networks, distances, statistics
and per-actor distances bit for bit; ntim to 1e-12 relative; scores (the
rate score steps/lam - integral of the total rate included) and rate-effect
scores within 1e-5 * max(1, |ref|) (device vs glibc exp/log in the
rates and waiting times; the covariate-weighted rate sums behind the
rate-effect scores are summed in parallel on the device, sequentially in the
reference).
"""
import numpy as np
import pytest

from conftest import load_golden
from rate_cases import rate_problem, rate_specs
from test_gpu_parity import SCORE_TOL, assert_period_equal

pytestmark = pytest.mark.gpu


def _assert_rate_equal(got, ref, label):
    assert_period_equal(got, ref, label=label)
    if "rscores" in ref:
        g = np.asarray(got["rscores"], np.float64)
        r = np.asarray(ref["rscores"], np.float64).reshape(g.shape)
        assert np.all(np.abs(g - r) <= SCORE_TOL * np.maximum(1.0, np.abs(r))), label
        np.testing.assert_array_equal(got["actor_dist"],
                                      np.asarray(ref["actor_dist"]).reshape(got["actor_dist"].shape),
                                      err_msg=f"{label} actor_dist")


@pytest.mark.parametrize("q", range(len(rate_specs())))
def test_rate_and_composition_match_reference(q, gpu_lib):
    from cusna_amd.abi import unpack_rows

    case = load_golden("rate_cases.json.gz")["cases"][q]
    prob = rate_problem(case["spec"])
    print(f"case {case['spec']['name']}", flush=True)
    got = gpu_lib.simulate_period(prob, want_end_nets=True)
    _assert_rate_equal(got, case["out"], case["spec"]["name"])
    ends = np.stack([unpack_rows(np.asarray(p, np.uint32), prob.n)
                     for p in case["out"]["ends_packed"]])
    np.testing.assert_array_equal(got["a_ends"], ends)


@pytest.mark.parametrize("n,specialize", [(500, False), (500, True), (1100, False)])
def test_rate_large_matches_oracle(n, specialize, gpu_lib, oracle_lib):
    """n = 500 (LDS-resident, generic and runtime-specialised kernels) and
    n = 1100 (streamed rows) with all three rate kinds and composition."""
    spec = dict(rate_specs()[3], name=f"rate_n{n}", n=n, M=3, seed=900 + n, net_seed=901 + n,
                lam=1.5, comp=dict(nact=n - 40, inactive=20))
    prob = rate_problem(spec)
    plan = gpu_lib.PeriodPlan(prob, want_end_nets=True)
    plan.specialize(1 if specialize else 0)
    plan.set_params(prob.theta, prob.lam)
    plan.run(prob.seed)
    got = plan.fetch(want_end_nets=True)
    if specialize:
        assert plan.specialized
    plan.close()
    ref = oracle_lib.simulate(prob, want_end_nets=True)
    _assert_rate_equal(got, ref, spec["name"])
    np.testing.assert_array_equal(got["a_ends"], ref["a_ends"])
    assert np.all(got["steps"] > 100)


def test_cow_composition_with_rate_effect(gpu_lib, oracle_lib):
    """A period built by the host layer from the reference's COW data
    (symmetric, n = 116, composition change) with an outRate effect."""
    import gzip
    import json
    import os

    from conftest import GOLDEN
    from cusna_amd.data import build_saom_data
    from cusna_amd.effects import Effect, RateEffect
    from cusna_amd.period import period_problem

    with gzip.open(os.path.join(GOLDEN, "data", "cow.json.gz"), "rt") as f:
        g = json.load(f)
    mat = [np.array([[np.nan if v is None else v for v in r] for r in w], float)
           for w in g["waves_input"]]
    data = build_saom_data(mat, composition=g["composition_input"])
    for p in (1, 2):
        prob = period_problem(data, [Effect("density"), Effect("transTriads")], p, [-2.0, 0.3],
                              3.0, 6, 11 + p, rate_effects=[RateEffect("outRate")],
                              rate_theta=[0.05], composition=True)
        got = gpu_lib.simulate_period(prob, want_end_nets=True)
        ref = oracle_lib.simulate(prob, want_end_nets=True)
        _assert_rate_equal(got, ref, f"cow p{p}")
        np.testing.assert_array_equal(got["a_ends"], ref["a_ends"])
