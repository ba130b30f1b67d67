"""CPU: the oracle on periods with non-basic rate effects and composition
change, against tests/golden/rate_cases.json.gz (outputs of the reference
engine compiled from source, libcusna_saom_cpu.cpp:170-193, 283-302,
356-361, 406-409): trajectories, end networks, distances, statistics, the
rate score steps/lam - integral of the total rate, the rate-effect scores and
the per-actor distances, all bit for bit; plus closed-form checks of the
rate-effect bookkeeping."""
import numpy as np
import pytest

from conftest import load_golden
from rate_cases import rate_problem, rate_specs


@pytest.mark.parametrize("q", range(len(rate_specs())))
def test_oracle_matches_reference_rate_cases(q, oracle_lib):
    from cusna_amd.abi import unpack_rows

    case = load_golden("rate_cases.json.gz")["cases"][q]
    assert case["spec"]["name"] == rate_specs()[q]["name"]
    prob = rate_problem(case["spec"])
    r = oracle_lib.simulate(prob, want_end_nets=True)
    o = case["out"]
    for key in ("dist", "dist_stat", "steps", "ntim", "stats"):
        np.testing.assert_array_equal(np.asarray(r[key]).ravel(), np.asarray(o[key]).ravel(),
                                      err_msg=f"{case['spec']['name']} {key}")
    np.testing.assert_array_equal(r["scores"], np.asarray(o["scores"], np.float32))
    if prob.nrate:
        np.testing.assert_array_equal(r["rscores"], np.asarray(o["rscores"], np.float32))
        np.testing.assert_array_equal(r["actor_dist"], np.asarray(o["actor_dist"]))
    ends = np.stack([unpack_rows(np.asarray(p, np.uint32), prob.n) for p in o["ends_packed"]])
    np.testing.assert_array_equal(r["a_ends"], ends)


def test_composition_restores_inactive_actors(oracle_lib):
    """Inactive actors' rows and columns return to the start network."""
    spec = [s for s in rate_specs() if s["name"] == "comp_restore"][0]
    prob = rate_problem(spec)
    r = oracle_lib.simulate(prob, want_end_nets=True)
    x0 = np.asarray(prob.x0)
    inact = np.asarray(prob.inact) != 0
    for c in range(prob.M):
        a = r["a_ends"][c]
        np.testing.assert_array_equal(a[inact], x0[inact])
        np.testing.assert_array_equal(a[:, inact], x0[:, inact])
    assert (r["actor_dist"][:, inact] == 0).all() if "actor_dist" in r else True


def test_actor_dist_sums_to_dist_stat(oracle_lib):
    spec = [s for s in rate_specs() if s["name"] == "rate_four"][0]
    r = oracle_lib.simulate(rate_problem(spec), want_end_nets=True)
    np.testing.assert_array_equal(r["actor_dist"].sum(1), r["dist_stat"])
