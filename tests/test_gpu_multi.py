"""GPU: the multi-network period simulator (cusna_saom_simulate_multi, the
multi plan) and the cross-network statistics (cusna_cross_stats).

Bar (as the single-network parity tests): trajectories bit for bit — end
networks, ministeps, rate scores (steps_d / lam_d - nact) — against the
reference's compiled engine (tests/golden/multi_cases.json.gz) and the
oracle; elapsed time within 1e-12 relative (device vs glibc log); effect
scores (float32 of float64 sums, tree vs sequential order) within
1e-5 * max(1, |ref|). Cross statistics bit for bit.
"""
import numpy as np
import pytest

from conftest import load_golden
from multi_cases import CROSS_SPECS, cross_inputs, multi_problem, multi_specs

pytestmark = pytest.mark.gpu


def _unpack(rows, n):
    bits = np.unpackbits(np.asarray(rows, np.uint32).view(np.uint8).reshape(
        *np.shape(rows)[:-1], -1), axis=-1, bitorder="little")
    return bits[..., :n]


def assert_multi_equal(got, ref, label=""):
    np.testing.assert_array_equal(got["steps"], ref["steps"], err_msg=f"{label} steps")
    np.testing.assert_array_equal(got["a_ends"], ref["a_ends"], err_msg=f"{label} a_ends")
    np.testing.assert_array_equal(got["rscores"], ref["rscores"], err_msg=f"{label} rscores")
    np.testing.assert_allclose(got["ntim"], ref["ntim"], rtol=1e-12, atol=0,
                               err_msg=f"{label} ntim")
    rs = np.asarray(ref["scores"], np.float64)
    tol = 1e-5 * np.maximum(1.0, np.abs(rs))
    assert np.all(np.abs(np.asarray(got["scores"], np.float64) - rs) <= tol), label


@pytest.fixture(scope="module")
def gold():
    return load_golden("multi_cases.json.gz")


@pytest.mark.parametrize("q", range(len(multi_specs())))
def test_multi_matches_reference_fixture(q, gold, gpu_lib):
    case = gold["multi"][q]
    pr = multi_problem(case["spec"])
    print(f"case {case['spec']['name']} n={pr.n} D={pr.D} M={pr.M}", flush=True)
    got = gpu_lib.simulate_multi(pr)
    ref = dict(steps=np.array(case["steps"]), ntim=np.array(case["ntim"]),
               rscores=np.array(case["rscores"], np.float32), scores=np.array(case["scores"]),
               a_ends=_unpack(case["ends"], pr.n))
    assert_multi_equal(got, ref, case["spec"]["name"])


LARGE = {
    # n, D, lam, effects per network (O(1) change statistics keep the oracle fast)
    "n500_lds": (500, 2, [3.0, 2.0], [[("density", -3.0), ("recip", 1.5), ("crprod", 1.0, 1)],
                                      [("density", -3.2), ("recip", 1.2),
                                       ("crprodRecip", 0.6, 0)]]),
    "n600_rows_global": (600, 4, [1.0, 1.0, 0.8, 0.6],
                         [[("density", -3.0), ("crprodMutual", 0.8, 1)],
                          [("density", -3.0), ("recip", 1.4)],
                          [("density", -3.1), ("crprod", 0.9, 0)],
                          [("density", -3.2), ("inPop", 0.02)]]),
    "n1100_streamed": (1100, 2, [1.0, 0.8], [[("density", -3.5), ("recip", 1.8),
                                             ("crprod", 0.7, 1)],
                                            [("density", -3.5), ("outAct", -0.02)]]),
}


@pytest.mark.parametrize("case", sorted(LARGE))
def test_multi_large_shapes_match_oracle(case, gpu_lib, oracle_lib):
    """LDS-resident (D x 64 KB rows), device-memory rows (four networks of
    600 actors exceed the LDS) and the streamed word bounds (n > 1024)."""
    n, D, lam, eff = LARGE[case]
    spec = dict(name=case, n=n, deg=[2.0] * D, lam=lam, M=3, seed=40 + n, eff=eff)
    pr = multi_problem(spec)
    print(f"case {case}", flush=True)
    got = gpu_lib.simulate_multi(pr)
    ref = oracle_lib.simulate_multi(pr)
    assert_multi_equal(got, ref, case)
    assert np.all(got["steps"] > 100)


def test_multi_plan_reruns_and_distances(gpu_lib, oracle_lib):
    spec = multi_specs()[1]
    pr = multi_problem(spec)
    rng = np.random.default_rng(5)
    kdist = (rng.random((pr.D, pr.n, pr.n)) < 0.8).astype(np.uint8)
    plan = gpu_lib.MultiPlan(pr, want_end_nets=True, kdist=kdist)
    plan.run(pr.seed)
    a = plan.fetch(want_end_nets=True)
    ref = oracle_lib.simulate_multi(pr)
    assert_multi_equal(a, ref, "plan")
    dist = plan.dist()
    x0 = np.asarray(pr.x0)
    for d in range(pr.D):
        want = ((a["a_ends"][:, d] != x0[d][None]) & (kdist[d][None] != 0)).sum((1, 2))
        np.testing.assert_array_equal(dist[d], want)
    plan.run(pr.seed)  # same key, same result
    np.testing.assert_array_equal(plan.fetch(want_end_nets=True)["a_ends"], a["a_ends"])
    plan.set_params(np.asarray(pr.theta) * 0.5, np.asarray(pr.lam) * 2.0)
    plan.run(pr.seed)
    b = plan.fetch()
    assert b["steps"].mean() > 1.5 * a["steps"].mean()
    plan.close()


@pytest.mark.parametrize("q", range(len(CROSS_SPECS)))
def test_cross_stats_match_reference_fixture(q, gold, gpu_lib, oracle_lib):
    case = gold["cross"][q]
    a, wp = cross_inputs(case["seed"])
    got = gpu_lib.cross_stats(a, wp, case["t"], case["parm"])
    np.testing.assert_array_equal(got, case["out"])
    a2, wp2 = cross_inputs(case["seed"] + 50, n=130, M=4)  # word boundary, self-loops kept
    a2[:, 3, 3] = 1
    np.testing.assert_array_equal(gpu_lib.cross_stats(a2, wp2, case["t"], case["parm"]),
                                  oracle_lib.cross_stats(a2, wp2, case["t"], case["parm"]))


@pytest.mark.parametrize("parm", [1.0, 2.0])
def test_cross_from_several_term_windows(parm, gpu_lib, oracle_lib):
    """`from` (t = 36) at n = 600: the square-root form (parm >= 2) adds its
    tie terms window by window (several 256-actor windows), against the
    oracle bit for bit; sparse and dense predictors."""
    for seed, dens in ((4242, 0.12), (4243, 0.01)):
        a, wp = cross_inputs(seed, n=600, M=3)
        if dens < 0.1:
            rng = np.random.default_rng(seed)
            a = (rng.random(a.shape) < dens).astype(np.uint8)
            for c in range(a.shape[0]):
                np.fill_diagonal(a[c], 0)
        print(f"case cross from n=600 parm={parm} density={dens}", flush=True)
        np.testing.assert_array_equal(gpu_lib.cross_stats(a, wp, 36, parm),
                                      oracle_lib.cross_stats(a, wp, 36, parm))


def test_multi_argument_checks(gpu_lib):
    from cusna_amd.abi import CusnaError

    pr = multi_problem(multi_specs()[0])
    bad = multi_problem(multi_specs()[0])
    bad.netof = np.array([0, 0, 2, 1], np.int32)  # network 2 of D = 2
    with pytest.raises(CusnaError, match="status 2"):
        gpu_lib.simulate_multi(bad)
    bad = multi_problem(multi_specs()[0])
    bad.x0 = np.asarray(pr.x0)[:1]  # D = 1
    bad.lam = np.asarray(pr.lam)[:1]
    with pytest.raises(CusnaError, match="status 2"):
        gpu_lib.simulate_multi(bad)
    a, wp = cross_inputs(1, n=6, M=1)
    with pytest.raises(CusnaError, match="status 5"):
        gpu_lib.cross_stats(a, wp, 38)


@pytest.mark.parametrize("q", range(len(multi_specs())))
def test_specialised_multi_matches_generic(q, gpu_lib, oracle_lib, monkeypatch):
    """The kernel specialised to the effect list (hipRTC) against the generic
    kernel and the oracle: trajectories bit for bit, scores to the bar (the
    specialised kernel sums the score terms in another tree order)."""
    pr = multi_problem(multi_specs()[q])
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("CUSNA_SPECIALIZE", mode)
        plan = gpu_lib.MultiPlan(pr, want_end_nets=True)
        plan.run(pr.seed)
        out[mode] = plan.fetch(want_end_nets=True)
        assert plan.specialized == (mode == "1")
        plan.close()
    assert_multi_equal(out["1"], out["0"], "specialised vs generic")
    assert_multi_equal(out["1"], oracle_lib.simulate_multi(pr), "specialised vs oracle")
