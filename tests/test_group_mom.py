"""CPU: the multi-group estimator's parameter layout and targets.

Groups are composed as the periods of one cusna_mom_estimate call: one rate
parameter per group period, the effect targets summed over every group's
periods (as R/mom.R:152-166 sums them over periods). Pinned by closed forms:
the density target is the number of ties of every end wave, recip
sum_ij x_ij x_ji over them, and the rate target of a period the Hamming
distance between its two waves (R/targets.R)."""
import numpy as np

from cusna_amd.data import build_saom_data
from cusna_amd.effects import Effect
from cusna_amd.mom import GroupMomEstimator


def _waves(rng, n, T):
    w = [(rng.random((n, n)) < 0.1).astype(np.uint8)]
    for _ in range(T - 1):
        flip = (rng.random((n, n)) < 0.05).astype(np.uint8)
        w.append(w[-1] ^ flip)
    for x in w:
        np.fill_diagonal(x, 0)
    return w


def test_group_layout_and_targets(oracle_lib):
    rng = np.random.default_rng(4)
    groups = [(30, 3), (40, 2)]  # (n, waves): 2 + 1 periods
    datas, effs, waves_all = [], [], []
    for n, T in groups:
        w = _waves(rng, n, T)
        waves_all.append(w)
        datas.append(build_saom_data(w))
        cov = rng.integers(1, 6, size=n).astype(float)
        effs.append([Effect("density"), Effect("recip"), Effect("egoX", cov)])

    def net(a, **kw):
        return oracle_lib.network_stats(a, kw["type_ids"], kw["p1"], kw["p2"], kw["vmat"],
                                        kw["vmiss"], kw["dyn_flags"], kw["vb"], kw["od0"],
                                        kw["id0"], kw["wtab"], kw["symflag"])

    est = GroupMomEstimator(datas, effs, net_stats=net)
    assert est.P == 3 and est.k == 3 and est.p == 6
    assert len(est.problems) == 3 and len(est.param_names) == 6
    ham = [int((w[q] != w[q + 1]).sum()) for w in waves_all for q in range(len(w) - 1)]
    np.testing.assert_allclose(est.targets[:3], ham)
    ends = [w[q + 1] for w in waves_all for q in range(len(w) - 1)]
    np.testing.assert_allclose(est.targets[3], sum(int(x.sum()) for x in ends))
    np.testing.assert_allclose(est.targets[4], sum(int((x & x.T).sum()) for x in ends))
    th0 = est.default_theta0()
    np.testing.assert_allclose(th0[:3], [1.5 * ham[0] / 30 + 1, 1.5 * ham[1] / 30 + 1,
                                         1.5 * ham[2] / 40 + 1])
    np.testing.assert_array_equal(th0[3:], [-1.0, 0.0, 0.0])
    # the groups' problems keep their own sizes and start waves
    assert [pr.n for pr in est.problems] == [30, 30, 40]
