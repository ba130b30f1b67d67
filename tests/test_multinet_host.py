"""CPU: the multi-network estimator's host layer (data container, parameter
layout, targets, start values, guarded step) against the reference's
multinet fixture (tests/testthat/fixtures/multinet, test-multinet.R) with
the oracle as the statistic carrier."""
import numpy as np
import pytest

from conftest import load_golden


def _fixture():
    g = load_golden("multinet.json.gz")
    netA = [np.array(w, float) for w in g["netA_waves"]]
    netB = [np.array(w, float) for w in g["netB_waves"]]
    return g, netA, netB


def _est(oracle_lib, cross=False):
    from cusna_amd.effects import Effect
    from cusna_amd.multinet import MultiMomEstimator, saom_multinet_data

    g, netA, netB = _fixture()
    mn = saom_multinet_data([netA, netB])
    ebn = [[Effect("density"), Effect("recip")], [Effect("density"), Effect("recip")]]
    if cross:
        ebn[0].append(Effect("crprod", net_ref=1))
        ebn[1] += [Effect("from", net_ref=0, parm=2.0), Effect("to", net_ref=0)]
    est = MultiMomEstimator(mn, ebn, net_stats=oracle_lib.network_stats,
                            cross_stats=oracle_lib.cross_stats)
    return g, est


def test_layout_and_targets_match_fixture(oracle_lib):
    g, est = _est(oracle_lib)
    assert est.param_names == g["param_names"]
    np.testing.assert_array_equal(est.targets, g["targets"])
    th0 = est.default_theta0()
    n = est.n
    np.testing.assert_allclose(th0, [1.5 * 115 / n + 1, -1, 0, 1.5 * 133 / n + 1, -1, 0])


def test_cross_targets_closed_form(oracle_lib):
    g, est = _est(oracle_lib, cross=True)
    a_end = [nd.target_network(1).astype(np.int64) for nd in est.mn.nets]
    wp = [nd.network_less_missing_start(1).astype(np.int64) for nd in est.mn.nets]
    t = dict(zip(est.param_names, est.targets))
    assert t["net0 crprod<-net1"] == (a_end[0] * wp[1]).sum()
    star = wp[0] @ wp[0].T
    assert t["net1 from<-net0"] == pytest.approx((a_end[1] * np.sqrt(star)).sum(), rel=1e-13)
    assert t["net1 to<-net0"] == (a_end[1] * (wp[0] @ a_end[1])).sum()


def test_guard_step(oracle_lib):
    _, est = _est(oracle_lib)
    theta = np.array([2.0, -1.0, 0.5, 3.0, -2.0, 4.0])
    step = np.array([1.9, 5.0, -5.0, 0.1, 0.2, -3.0])
    got = est.guard_step(theta, step.copy())
    # caps max(1, |theta|/2); a rate step leaving less than 10 % halves the rate
    np.testing.assert_allclose(got, [1.0, 1.0, -1.0, 0.1, 0.2, -2.0])


def test_data_validation():
    from cusna_amd.multinet import saom_multinet_data

    _, netA, netB = _fixture()
    with pytest.raises(ValueError, match=">= 2 networks"):
        saom_multinet_data([netA])
    with pytest.raises(ValueError, match="same number of waves"):
        saom_multinet_data([netA, netB + [netB[0]]])
    with pytest.raises(ValueError, match="actor set size"):
        saom_multinet_data([netA, [w[:40, :40] for w in netB]])
