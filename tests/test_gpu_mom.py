"""GPU: the method-of-moments engine (cusna_mom_estimate).

* the device reductions of phase 3 (sim means, D, msf, t-ratios, se) against
  the MoM oracle on the phase-3 rows the engine returns;
* phase-2 iterations replayed from the captured HIP graph give the same bits
  as the same iterations launched one by one;
* the estimates of the four set-ups of the reference's test-estimate.R land
  within its bar (|theta - theta_ref| < se_ref + 0.1) of estimate.json.gz;
  the reference's R estimator uses another random stream, so this is the
  reference's own statistical bar, not bit parity.
"""
import numpy as np
import pytest

import mom_oracle as MO
from mom_cases import CONFIGS, control, estimate_fixture, setup

pytestmark = pytest.mark.gpu


def _est(name, **over):
    from cusna_amd.mom import MomEstimator

    g = estimate_fixture()
    data, eff, bef, cond = setup(name, g)
    return MomEstimator(data, eff, bef, cond, control(g, **over)), g


SMALL = dict(n1=64, nD=64, batch2=32, n2=(4, 6), n3=128)


@pytest.mark.parametrize("name", ["baseline", "baseline_cond", "coevol"])
def test_phase3_reductions_match_oracle(name, gpu_lib):
    est, _ = _est(name, **SMALL)
    fit = est.run()
    stats, sp, scp = fit.phase3_stats, fit.phase3_sp, fit.phase3_scp
    assert stats.shape == (SMALL["n3"], est.p)
    np.testing.assert_allclose(sp.sum(0), stats, rtol=0, atol=1e-9)
    ref = MO.phase3_summary(stats, sp, scp, est.targets)
    for key in ("sim_means", "D", "msf", "tconv", "se"):
        np.testing.assert_allclose(getattr(fit, key), ref[key], rtol=1e-9, atol=1e-12,
                                   err_msg=key)
    np.testing.assert_allclose(fit.tconv_max, ref["tconv_max"], rtol=1e-9)
    assert fit.n_sims == (SMALL["n1"] + SMALL["n3"] + 2 * SMALL["nD"]
                          + sum(SMALL["n2"]) * SMALL["batch2"])
    if est.conditional:
        assert fit.rate_est.shape == (est.P,) and np.all(fit.rate_est > 0)
        assert np.all(fit.rate_sd > 0)


@pytest.mark.parametrize("name", ["baseline", "coevol"])
def test_graph_replay_matches_eager(name, gpu_lib):
    est, _ = _est(name, **SMALL)
    a = est.run(use_graph=True)
    b = est.run(use_graph=False)
    np.testing.assert_array_equal(a.theta, b.theta)
    np.testing.assert_array_equal(a.phase3_stats, b.phase3_stats)
    np.testing.assert_array_equal(a.D, b.D)
    c = est.run(use_graph=True)  # a fresh engine starts from the same state
    np.testing.assert_array_equal(a.theta, c.theta)


def test_seed_changes_the_stream(gpu_lib):
    est, _ = _est("baseline", **SMALL)
    a = est.run()
    est.control.seed += 1
    b = est.run()
    assert not np.array_equal(a.phase3_stats, b.phase3_stats)


@pytest.mark.parametrize("name", CONFIGS)
def test_estimate_within_reference_bar(name, gpu_lib):
    """test-estimate.R:21-27: |theta - theta_ref| < 1.0 * se_ref + 0.1."""
    est, g = _est(name)
    fit = est.run()
    gold = g["configs"][name]
    th_ref, se_ref = np.array(gold["theta"]), np.array(gold["se"])
    assert np.all(np.isfinite(fit.theta)) and np.all(np.isfinite(fit.se))
    bad = np.abs(fit.theta - th_ref) >= se_ref + 0.1
    assert not bad.any(), (name, fit.theta, th_ref, se_ref)


def test_until_converged_reaches_tolerance(gpu_lib):
    """Extra last-gain sub-phases until max|tconv| <= 0.1 (bounded repeats)."""
    est, _ = _est("baseline")
    fit = est.run(max_repeats=6, conv_tol=0.1)
    assert fit.max_abs_tconv <= 0.1 or fit.repeats == 6
    assert fit.max_abs_tconv == pytest.approx(np.abs(fit.tconv).max())


def test_rejects_bad_schedule(gpu_lib):
    from cusna_amd.abi import CusnaError

    est, _ = _est("baseline", n3=1)
    with pytest.raises(CusnaError):
        est.run()


def test_surrogate_ablation(gpu_lib):
    """Surrogate phase 2 (perceptron root + simulation-checked Newton
    corrections) against the plain Robbins-Monro phase 2 on the 50-actor fit
    (50 actors, conditional, reference default schedule, n3 = 1000): both
    reach max|t| <= 0.1 (repeats allowed), the estimates agree within one
    standard error of the plain fit, and the surrogate uses fewer simulated
    chains and less wall time; without a trained network the same path falls
    back to the plain sub-phases."""
    import time

    from cusna_amd.mom import MomControl, MomEstimator
    from cusna_amd.workloads import fit50_estimation

    data, eff = fit50_estimation(lambda p: gpu_lib.simulate_period(p, want_end_nets=True)["a_ends"])
    est = MomEstimator(data, eff, None, True, MomControl(n3=1000))
    est.run(max_repeats=0)  # code objects loaded
    t0 = time.perf_counter()
    plain = est.run(max_repeats=10, conv_tol=0.1)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    sur = est.run(max_repeats=10, conv_tol=0.1, surrogate=True)
    t_sur = time.perf_counter() - t0
    assert plain.max_abs_tconv <= 0.1 and sur.max_abs_tconv <= 0.1
    assert np.all(np.abs(sur.theta - plain.theta) <= plain.se), (sur.theta, plain.theta)
    assert sur.n_sims < plain.n_sims
    assert t_sur < t_plain, (t_sur, t_plain)
    # the same path without a trained network (64 pairs, one Adam step): its
    # corrections cannot finish, the plain sub-phases run as the fallback, and
    # it simulates more chains than plain MoM (chain counts are deterministic)
    none = est.run(max_repeats=10, conv_tol=0.1, surrogate=True, surrogate_pairs_log2=6,
                   surrogate_steps=1)
    assert none.max_abs_tconv <= 0.1
    assert none.n_sims > plain.n_sims > sur.n_sims
    print(f"plain {t_plain:.3f} s, {plain.n_sims} chains; surrogate {t_sur:.3f} s, "
          f"{sur.n_sims} chains; untrained network {none.n_sims} chains")


def test_group_estimator_plain_and_surrogate(gpu_lib):
    """Three groups of 60 actors fitted together (GroupMomEstimator: the
    groups are the periods of one call, one rate each): the plain estimator
    and the surrogate with one warm plain sub-phase both reach max|t| <= 0.1
    and agree within one standard error of the plain fit; the estimates lie
    within three standard errors of the theta the waves were simulated at."""
    from cusna_amd.data import build_saom_data
    from cusna_amd.effects import Effect
    from cusna_amd.mom import GroupMomEstimator, MomControl
    from cusna_amd.workloads import bernoulli_network, effect_problem

    rng = np.random.default_rng(21)
    theta_true = np.array([-2.2, 1.8, 0.3])
    datas, effs = [], []
    for g in range(3):
        w1 = bernoulli_network(60, 3.0, rng)
        eff = [Effect("density"), Effect("recip"), Effect("transTrip")]
        pr = effect_problem(w1, eff, theta_true, 4.0, 1, 900 + g)
        w2 = np.asarray(gpu_lib.simulate_period(pr, want_end_nets=True)["a_ends"][0], np.uint8)
        datas.append(build_saom_data([w1, w2]))
        effs.append(eff)
    est = GroupMomEstimator(datas, effs, MomControl(n3=1000))
    assert est.p == 6
    plain = est.run(max_repeats=10, conv_tol=0.1)
    sur = est.run(max_repeats=10, conv_tol=0.1, surrogate=True, surrogate_warmup=1)
    for fit in (plain, sur):
        assert fit.max_abs_tconv <= 0.1, fit.max_abs_tconv
    assert np.all(np.abs(sur.theta - plain.theta) <= plain.se), (sur.theta, plain.theta)
    assert np.all(np.abs(plain.theta[3:] - theta_true) <= 3 * plain.se[3:]), plain.theta
