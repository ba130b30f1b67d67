"""CPU: the oracle's multi-network period simulator and cross-network
statistics, pinned to the reference.

* against tests/golden/multi_cases.json.gz, outputs of the reference's
  cusna_saom_simulate_multi_cpu / cusna_cross_stats compiled from source
  (libcusna_saom_cpu.cpp:427-604, 819-857): ten problems covering every cross
  effect (33-37, from with and without the square root), D = 2, 3, 4, every
  structural and covariate effect id, a dyadic covariate, creation/endowment,
  symmetric networks, permitted-toggle masks, composition (nact < n),
  per-chain parameter rows and a word boundary (n = 70, 130);
* against the compiled reference itself where oracle/_ref is built;
* closed forms: cross statistics as matrix products (numpy), and a
  multi-network period whose cross effects have theta = 0 and whose networks
  share no effect equals, network by network, the single-network model's
  trajectory statistics in distribution (steps ~ Poisson rate).
"""
import numpy as np
import pytest

from conftest import load_golden
from multi_cases import CROSS_SPECS, cross_inputs, multi_problem, multi_specs


def _unpack(rows, n):
    bits = np.unpackbits(np.asarray(rows, np.uint32).view(np.uint8).reshape(
        *np.shape(rows)[:-1], -1), axis=-1, bitorder="little")
    return bits[..., :n]


@pytest.fixture(scope="module")
def gold():
    return load_golden("multi_cases.json.gz")


@pytest.mark.parametrize("q", range(len(multi_specs())))
def test_oracle_multi_matches_reference_fixture(q, gold, oracle_lib):
    case = gold["multi"][q]
    spec = case["spec"]
    assert spec["name"] == multi_specs()[q]["name"]
    pr = multi_problem(spec)
    r = oracle_lib.simulate_multi(pr)
    np.testing.assert_array_equal(r["steps"], case["steps"])
    np.testing.assert_array_equal(r["ntim"], np.array(case["ntim"]))
    np.testing.assert_array_equal(r["rscores"], np.array(case["rscores"], np.float32))
    np.testing.assert_array_equal(r["scores"], np.array(case["scores"], np.float32))
    np.testing.assert_array_equal(r["a_ends"], _unpack(case["ends"], pr.n))


def test_oracle_multi_matches_compiled_reference(oracle_lib):
    if oracle_lib.reference() is None:
        pytest.skip("oracle/_ref not built")
    for spec in multi_specs()[:4]:
        pr = multi_problem(spec, M=3)
        pr.seed += 17
        a = oracle_lib.simulate_multi(pr)
        b = oracle_lib.simulate_multi_reference(pr)
        for key in ("steps", "ntim", "rscores", "scores", "a_ends"):
            np.testing.assert_array_equal(a[key], b[key], err_msg=f"{spec['name']} {key}")


@pytest.mark.parametrize("q", range(len(CROSS_SPECS)))
def test_oracle_cross_stats(q, gold, oracle_lib):
    case = gold["cross"][q]
    a, wp = cross_inputs(case["seed"])
    got = oracle_lib.cross_stats(a, wp, case["t"], case["parm"])
    np.testing.assert_array_equal(got, case["out"])
    # closed forms
    W = wp.astype(np.int64)
    star = W @ W.T
    for c in range(a.shape[0]):
        A = a[c].astype(np.int64)
        want = {33: (A * W).sum(), 34: (A * W.T).sum(), 35: (A * W * W.T).sum(),
                36: (A * (np.sqrt(star) if case["parm"] >= 2 else star)).sum(),
                37: (A * (W @ A)).sum()}[case["t"]]
        assert got[c] == pytest.approx(want, rel=1e-12)


def test_oracle_cross_stats_argument_checks(oracle_lib):
    a, wp = cross_inputs(1, n=5, M=1)
    with pytest.raises(RuntimeError, match="status 5"):
        oracle_lib.cross_stats(a, wp, 32)


def test_oracle_multi_uncoupled_rates(oracle_lib):
    """Networks with only density/recip and equal rates: each ministep picks a
    network with probability lam_d / sum(lam); over many chains the per-network
    step counts (n * lam_d * rscore identity) follow from the rate scores:
    rscore_d = steps_d / lam_d - nact."""
    spec = dict(multi_specs()[0])
    pr = multi_problem(spec, M=64)
    r = oracle_lib.simulate_multi(pr)
    lam = np.asarray(spec["lam"])
    steps_d = (r["rscores"].astype(np.float64) + pr.n) * lam[None, :]
    np.testing.assert_allclose(steps_d.sum(1), r["steps"], rtol=1e-5)
    share = steps_d.sum(0) / steps_d.sum()
    np.testing.assert_allclose(share, lam / lam.sum(), atol=0.03)
