"""GPU: the runtime-specialised period kernel (hipRTC) against the generic
kernel and the reference engine's outputs: identical trajectories, end
networks, statistics and scores (same arithmetic in the same order)."""
import numpy as np
import pytest

from conftest import load_golden
from test_gpu_parity import assert_period_equal

pytestmark = pytest.mark.gpu


def _both(gpu_lib, prob, seed=7):
    plan = gpu_lib.PeriodPlan(prob, want_end_nets=True)
    try:
        plan.specialize(0)
        plan.run(seed)
        a = plan.fetch(want_end_nets=True)
        assert not plan.specialized
        plan.specialize(1)
        plan.run(seed)
        b = plan.fetch(want_end_nets=True)
        assert plan.specialized, "specialised kernel did not run"
    finally:
        plan.close()
    for key in a:
        np.testing.assert_array_equal(np.asarray(b[key]), np.asarray(a[key]), err_msg=key)
    return b


REF_IDX = list(range(0, 38, 3)) + [33, 34, 35, 36, 37]


@pytest.mark.parametrize("idx", sorted(set(REF_IDX)))
def test_specialised_matches_generic_and_reference(gpu_lib, idx):
    from ref_cases_loader import spec_problem

    case = load_golden("ref_cases.json.gz")["cases"][idx]
    prob = spec_problem(case["spec"])
    print(f"case {case['spec']['name']} n={prob.n} k={prob.k}", flush=True)
    got = _both(gpu_lib, prob, seed=case["spec"]["seed"])
    assert_period_equal(got, case["out"], label=case["spec"]["name"])


def test_specialised_c2_and_streamed(gpu_lib):
    from cusna_amd.workloads import n500_problem
    from test_gpu_stream import CASES, _problem

    _both(gpu_lib, n500_problem(M=64, n=500), seed=11)
    n, names, theta, lam, M = CASES["n1100_simple"]
    _both(gpu_lib, _problem(n, names, theta, lam, M, 31), seed=31)


def test_specialised_large_chain_state(gpu_lib):
    """n = 1000 (16 words per row, 136 KB of chain state in LDS): the
    specialised kernel, or the generic one where the runtime refuses the
    module launch, gives the generic kernel's results."""
    from test_gpu_stream import _problem

    pr = _problem(1000, ["density", "recip", "transTrip", "egoX"], [-3.0, 1.8, 0.3, 0.1],
                  1.0, 6, 17)
    plan = gpu_lib.PeriodPlan(pr)
    try:
        plan.specialize(0)
        plan.run(5)
        a = plan.fetch()
        plan.specialize(1)
        plan.run(5)
        b = plan.fetch()
        print("specialised kernel used:", plan.specialized, flush=True)
    finally:
        plan.close()
    for key in a:
        np.testing.assert_array_equal(np.asarray(b[key]), np.asarray(a[key]), err_msg=key)


def test_lean_specialisation_matches_generic(gpu_lib, monkeypatch):
    """The lean specialisation (static covariates at the alternatives read
    per ministep rather than held in registers for the period, the build the
    launch falls back to when the other exceeds the four-wave register
    budget, as the rate-effect path at 500 actors does) gives the generic
    kernel's results, on the 500-actor problem with altX and simX."""
    from cusna_amd.workloads import n500_problem

    monkeypatch.setenv("CUSNA_RTC_LEAN", "1")
    _both(gpu_lib, n500_problem(M=64, n=500), seed=13)
