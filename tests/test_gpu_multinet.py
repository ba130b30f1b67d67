"""GPU: the multi-network estimator (cusna_amd.multinet) over the device plan.

* the device statistics of the multi plan (per-network distances, network
  statistics and cross statistics of the masked end networks) equal the
  oracle's on the end networks of the same run, bit for bit;
* the estimate of the reference's multinet fixture (test-multinet.R: two
  networks of 50 actors, density + recip each, its schedule) lands within the
  reference's bar |theta - theta_ref| < se_ref + 0.2; the reference's R
  random stream cannot run here, so this is its statistical bar, not bit
  parity.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _fixture():
    g = load_golden("multinet.json.gz")
    return (g, [np.array(w, float) for w in g["netA_waves"]],
            [np.array(w, float) for w in g["netB_waves"]])


def _est(cross, control=None):
    from cusna_amd.effects import Effect
    from cusna_amd.multinet import MultiMomEstimator, saom_multinet_data

    g, netA, netB = _fixture()
    mn = saom_multinet_data([netA, netB])
    ebn = [[Effect("density"), Effect("recip")], [Effect("density"), Effect("recip")]]
    if cross:
        ebn[0] += [Effect("crprod", net_ref=1), Effect("transTrip"), Effect("inPop")]
        ebn[1] += [Effect("from", net_ref=0, parm=2.0), Effect("to", net_ref=0),
                   Effect("crprodMutual", net_ref=0)]
    return g, MultiMomEstimator(mn, ebn, control)


@pytest.mark.parametrize("cross", [False, True])
def test_plan_statistics_match_oracle(cross, gpu_lib, oracle_lib):
    g, est = _est(cross)
    theta = est.default_theta0()
    theta[est.eff_pos] += 0.3 * np.sin(np.arange(est.k))  # away from the start values
    M = 16
    plan = est._plan(1, M, want_end_nets=True)
    plan.set_params(theta[est.eff_pos], [theta[est.rate_pos[d, 0]] for d in range(est.D)])
    plan.run(12345)
    out = plan.fetch(want_end_nets=True)
    dist = plan.dist()
    st = plan.stats()
    for d in range(est.D):
        nd = est.mn.nets[d]
        kd = nd.kmask_dist(1)
        s = est._stat_inputs(d, 1)
        cols = np.nonzero(est.netof == d)[0]
        for c in range(M):
            a = out["a_ends"][c, d]
            assert dist[d, c] == ((a != nd.imputed[0]) & kd).sum()
            masked = ((a != 0) & s["keep"]) | s["force"]
            for q, e in enumerate(est.specs_by_net[d]):
                if e.type_id >= 33:
                    want = oracle_lib.cross_stats(masked[None], s["cross_wp"][q], e.type_id,
                                                  s["p1"][q])[0]
                else:
                    want = oracle_lib.network_stats(
                        masked, [s["type_ids"][q]], [s["p1"][q]], [s["p2"][q]],
                        s["vmat"][q:q + 1], vmiss=s["vmiss"][q:q + 1], od0=s["od0"],
                        id0=s["id0"], wtab=s["wtab"][q:q + 1], symflag=s["symflag"])[0]
                assert st[c, cols[q]] == want, (d, c, e.name)
    est.close()


def test_estimate_within_reference_bar(gpu_lib):
    """test-multinet.R:21-32: |theta - theta_py| < 1.0 * se + 0.2."""
    from cusna_amd.mom import MomControl

    g, _, _ = _fixture()
    S = g["settings"]
    ctl = MomControl(firstg=S["firstg"], n2=tuple(S["n2"]), batch2=S["batch2"], n1=S["n1"],
                     n3=S["n3"], nD=S["nD"], diagonalize=S["diagonalize"], seed=S["seed"])
    _, est = _est(False, ctl)
    np.testing.assert_array_equal(est.targets, g["targets"])
    fit = est.run()
    est.close()
    th_ref, se_ref = np.array(g["theta"]), np.array(g["se"])
    print("theta", fit.theta, "se", fit.se, "wall", fit.wall_time)
    assert np.all(np.isfinite(fit.theta)) and np.all(np.isfinite(fit.se))
    bad = np.abs(fit.theta - th_ref) >= se_ref + 0.2
    assert not bad.any(), (fit.theta, th_ref, se_ref)
    assert fit.n_sims == S["n1"] + S["n3"] + sum(S["nD"] + n2 * S["batch2"] for n2 in S["n2"])
