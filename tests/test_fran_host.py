"""CPU: the FRAN closure's contract checks (R/fran.R:41-51, test-fran.R:83-88):
a theta whose length does not match the effects and a conditional flag that
does not match the algorithm are rejected before anything is simulated, and
the parameter layout is the estimator's."""
from dataclasses import replace

import numpy as np
import pytest

from mom_cases import estimate_fixture


def _waves():
    g = estimate_fixture()
    return [np.array(w, float) for w in g["inputs"]["s50_waves"]]


@pytest.fixture
def net(oracle_lib):
    """The oracle's network statistics as the targets' carrier (no device)."""
    def stats(a, **kw):
        return oracle_lib.network_stats(a, kw["type_ids"], kw["p1"], kw["p2"], kw["vmat"],
                                        kw["vmiss"], kw["dyn_flags"], kw["vb"], kw["od0"],
                                        kw["id0"], kw["wtab"], kw["symflag"])
    return stats


def test_fran_rejects_mismatched_theta_and_condition(net):
    from cusna_amd.fran import cusna_fran

    fran = cusna_fran(_waves(), ["density", "recip"], net_stats=net)
    with pytest.raises(ValueError, match="3 parameters"):
        fran({"theta": [1.0, 2.0, 3.0]})
    with pytest.raises(ValueError, match="conditional"):
        fran({"theta": [-2.0, 1.0], "cconditional": False})
    assert fran.counter == 0


def test_fran_parameter_layout(net):
    from cusna_amd.fran import cusna_fran

    waves = _waves()
    cond = cusna_fran(waves, ["density", "recip", "transTrip"], net_stats=net)
    assert cond.est.p == 3 and cond.est.param_names == ["density", "recip", "transTrip"]
    unc = cusna_fran(waves, ["density", "recip"], conditional=False, net_stats=net)
    assert unc.est.p == 4 and unc.est.param_names[:2] == ["rate (period 1)", "rate (period 2)"]
    cov = np.linspace(-1.0, 1.0, 50)
    withx = cusna_fran(waves, ["density", "egoX"], covariate=cov, net_stats=net)
    assert withx.est.effects[1].cov is not None and withx.est.effects[0].cov is None


@pytest.mark.parametrize("conditional", [True, False])
def test_simulate_layout_with_the_oracle_as_period_simulator(net, oracle_lib, monkeypatch,
                                                             conditional):
    """MomEstimator.simulate (the mirror of .mom_simulate, R/mom.R:133-183)
    with the oracle standing in for the device simulator: the rows are the
    period calls' statistics in the effect columns, the unconditional rate
    statistic (distance plus the rate constant) and the rate score in the
    period's own column, and zeros in the other period's rate column."""
    from cusna_amd import mom
    from cusna_amd.fran import cusna_fran

    monkeypatch.setattr(mom.lib, "simulate_period", lambda pr: oracle_lib.simulate(pr))
    fran = cusna_fran(_waves(), ["density", "recip"], conditional=conditional, net_stats=net)
    est = fran.est
    theta = [-2.5, 2.4] if conditional else [6.0, 5.0, -2.5, 2.4]
    np.random.seed(7)
    r = fran({"theta": theta})
    np.random.seed(7)
    seeds = [int(np.random.randint(1, 2**31)) for _ in range(est.P)]
    off = 0 if conditional else est.P
    assert r["fra"].shape == (est.P, est.p) and r["sc"].shape == (est.P, est.p)
    for q in range(est.P):
        ref = oracle_lib.simulate(replace(est.problems[q], theta=np.array(theta[off:]),
                                          lam=1.0 if conditional else theta[q], M=1,
                                          seed=seeds[q]))
        np.testing.assert_array_equal(r["fra"][q, off:], np.asarray(ref["stats"])[0])
        np.testing.assert_array_equal(r["sc"][q, off:],
                                      np.asarray(ref["scores"]).reshape(1, -1)[0, 1:])
        assert r["ntim0"][q] == np.asarray(ref["ntim"])[0]
        if not conditional:
            assert r["fra"][q, q] == np.asarray(ref["dist_stat"])[0] + est.rate_consts[q]
            assert r["fra"][q, 1 - q] == 0 and r["sc"][q, 1 - q] == 0
            assert r["sc"][q, q] == np.asarray(ref["scores"]).reshape(1, -1)[0, 0]
