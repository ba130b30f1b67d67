"""CPU: the MoM oracle pinned to closed forms and brute force, and the host
estimator context (targets, parameter layout) against estimate.json.gz."""
import numpy as np
import pytest

import mom_oracle as MO
from mom_cases import CONFIGS, control, estimate_fixture, setup


def test_derivative_brute_force():
    rng = np.random.default_rng(3)
    P, M, p = 2, 7, 3
    sp, scp = rng.normal(size=(P, M, p)), rng.normal(size=(P, M, p))
    D = np.zeros((p, p))
    for q in range(P):
        for a in range(p):
            for b in range(p):
                ma = sum(sp[q, c, a] for c in range(M)) / M
                mb = sum(scp[q, c, b] for c in range(M)) / M
                D[a, b] += sum((sp[q, c, a] - ma) * (scp[q, c, b] - mb) for c in range(M)) / M
    np.testing.assert_allclose(MO.derivative(sp, scp), D, rtol=1e-12)
    # one period with scores equal to the moments: the (biased) covariance
    np.testing.assert_allclose(MO.derivative(sp[:1], sp[:1]), np.cov(sp[0], rowvar=False,
                                                                    bias=True), rtol=1e-12)
    # a transposed operand would swap the roles: not symmetric in general
    assert not np.allclose(MO.derivative(sp, scp), MO.derivative(sp, scp).T)


def test_pinv_and_safe_solve():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(4, 4))
    np.testing.assert_allclose(MO.pinv(a), np.linalg.inv(a), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(MO.safe_solve(a), np.linalg.inv(a), rtol=1e-12)
    s = a.copy()
    s[3] = s[0] + s[1]  # rank 3
    np.testing.assert_allclose(MO.safe_solve(s), np.linalg.pinv(s), rtol=1e-8, atol=1e-10)
    z = np.zeros((3, 3))
    np.testing.assert_array_equal(MO.safe_solve(z), np.zeros((3, 3)))


def test_dinv_diagonalize_limits():
    rng = np.random.default_rng(5)
    D = rng.normal(size=(3, 3)) + 3 * np.eye(3)
    np.testing.assert_allclose(MO.dinv(D, 0.0), np.linalg.inv(D), rtol=1e-12)
    np.testing.assert_allclose(MO.dinv(D, 1.0), np.diag(1 / np.diag(D)), rtol=1e-12)


def test_guard_step_hand_cases():
    th = np.array([0.5, 10.0, -4.0, 2.0])
    st = np.array([0.48, -30.0, 5.0, 1.95])
    out = MO.guard_step(th, st.copy(), n_rates=2)
    # rate 0.5: 0.5 - 0.48 < 0.05 -> halve to 0.25; rate 10: capped at 10/2;
    # eval -4: capped at 2; eval 2: cap max(1, 1) = 1
    np.testing.assert_array_equal(out, [0.25, -5.0, 2.0, 1.0])
    out = MO.guard_step(th, st.copy(), n_rates=0)
    np.testing.assert_array_equal(out, [0.48, -5.0, 2.0, 1.0])


def test_phase3_summary_closed_forms():
    rng = np.random.default_rng(6)
    M, p = 4000, 3
    stats = rng.normal(size=(M, p)) * [1.0, 2.0, 0.5] + [1.0, 2.0, 3.0]
    tg = np.array([1.1, 1.8, 3.0])
    r = MO.phase3_summary(stats, stats[None], stats[None], tg)
    np.testing.assert_allclose(r["sim_means"], stats.mean(0), rtol=1e-12)
    np.testing.assert_allclose(r["tconv"], (stats.mean(0) - tg) / stats.std(0, ddof=1),
                               rtol=1e-12)
    # D = biased covariance, so se = sqrt(diag(D^-1 msf D^-1)) ~ 1/sd * M/(M-1)
    np.testing.assert_allclose(r["se"], np.sqrt(np.diag(np.linalg.inv(
        np.cov(stats, rowvar=False)))) * M / (M - 1), rtol=1e-9)
    # with a diagonal moment covariance tconv_max is the norm of the t-ratios
    d = rng.normal(size=(M, p))
    d -= d.mean(0)
    q, _ = np.linalg.qr(d)
    z = q * np.sqrt(M - 1) * [1.0, 2.0, 3.0] + [0.1, 0.2, 0.3]
    r = MO.phase3_summary(z, z[None], z[None], np.zeros(p))
    np.testing.assert_allclose(r["tconv_max"], np.linalg.norm(r["tconv"]), rtol=1e-9)


def test_rm_update_moves_towards_target():
    # linear model s = A theta: one full step with gain 1 and dinv = A^-1 lands on target
    A = np.array([[2.0, 0.3], [0.1, 1.5]])
    th = np.array([0.4, -0.2])
    tg = np.array([1.0, 0.5])
    stats = (A @ th)[None, :].repeat(3, 0)
    new = MO.rm_update(th, stats, tg, np.linalg.inv(A), 1.0, 0)
    np.testing.assert_allclose(A @ new, tg, rtol=1e-12)


@pytest.mark.parametrize("name", CONFIGS)
def test_estimator_context_matches_fixture(name, oracle_lib):
    """Targets and parameter names of the four test-estimate.R set-ups."""
    from cusna_amd.mom import MomEstimator

    g = estimate_fixture()
    data, eff, bef, cond = setup(name, g)

    def net(a, **kw):
        return oracle_lib.network_stats(a, kw["type_ids"], kw["p1"], kw["p2"], kw["vmat"],
                                        kw["vmiss"], kw["dyn_flags"], kw["vb"], kw["od0"],
                                        kw["id0"], kw["wtab"], kw["symflag"])

    est = MomEstimator(data, eff, bef, cond, control(g), net_stats=net,
                       beh_stats=oracle_lib.behavior_stats)
    gold = g["configs"][name]
    assert est.param_names == gold["param_names"]
    np.testing.assert_allclose(est.targets, gold["targets"], rtol=1e-9)
    assert est.p == len(gold["theta"])
    th0 = est.default_theta0()
    assert th0.shape == (est.p,) and np.all(np.isfinite(th0))
