"""GPU: sienaGOF auxiliary distributions (cusna_gof_distribution and its
device-resident form cusna_period_plan_gof) against the oracle, bit for bit.

The oracle follows libcusna_saom_cpu.cpp:716-765 and is itself checked
against the compiled reference and against breadth-first search in
test_gof_oracle.py. Cases cover the degree kinds and the geodesic kind at
word bounds 1..16 (n up to 1024), ragged n, self-loops, unsorted and repeated
levels, levels beyond n, level 0, and the reference's argument checks. The
geodesic form that reads rows from device memory (networks above the LDS) is
checked against the oracle where the oracle's boolean powers are affordable
(forced at small n, and n = 1100 at two levels) and against breadth-first
distances (tests/geodesic_check.py, itself checked against the oracle) up to
n = 4200 (two words per lane).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _nets(M, n, deg, seed, self_loops=True):
    rng = np.random.default_rng(seed)
    a = (rng.random((M, n, n)) < deg / max(n - 1, 1)).astype(np.uint8)
    if not self_loops:
        for c in range(M):
            np.fill_diagonal(a[c], 0)
    return a


DEG_LEVELS = [0, 1, 2, 3, 5, 8, 2, -1, 10000]
GEO_LEVELS = [3, 1, 0, 2, 5, 4, 4]


@pytest.mark.parametrize("n", [1, 2, 37, 64, 65, 200, 500, 1500])
@pytest.mark.parametrize("kind", [0, 1])
def test_degree_distribution_matches_oracle(n, kind, gpu_lib, oracle_lib):
    M = 6 if n <= 500 else 2
    a = _nets(M, n, 3.0, 100 + n + kind)
    got = gpu_lib.gof_distribution(a, kind, DEG_LEVELS)
    ref = oracle_lib.gof_distribution(a, kind, DEG_LEVELS)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("n,deg", [(1, 0.0), (5, 1.0), (37, 1.5), (64, 2.0), (65, 2.0),
                                   (129, 1.2), (300, 1.5), (600, 1.1), (1024, 1.05)])
def test_geodesic_distribution_matches_oracle(n, deg, gpu_lib, oracle_lib):
    """Sparse networks keep the reference's boolean powers cheap on the CPU
    and give long shortest paths (many breadth-first levels)."""
    M = 3 if n <= 300 else 1
    a = _nets(M, n, deg, 7 * n)
    lv = GEO_LEVELS if n < 300 else ([1, 2, 4, 8, 16] if n < 1000 else [1, 2, 3])
    got = gpu_lib.gof_distribution(a, 2, lv)
    ref = oracle_lib.gof_distribution(a, 2, lv)
    np.testing.assert_array_equal(got, ref)


def test_geodesic_levels_beyond_n(gpu_lib, oracle_lib):
    """Levels above n saturate at the transitive closure (the reference keeps
    multiplying; no shortest path is longer than n - 1)."""
    a = _nets(4, 12, 1.2, 5)
    lv = [1, 40, 11, 12, 13, 0]
    got = gpu_lib.gof_distribution(a, 2, lv)
    ref = oracle_lib.gof_distribution(a, 2, lv)
    np.testing.assert_array_equal(got, ref)


def test_geodesic_only_level_zero(gpu_lib, oracle_lib):
    a = _nets(2, 20, 2.0, 9)
    got = gpu_lib.gof_distribution(a, 2, [0, 0])
    ref = oracle_lib.gof_distribution(a, 2, [0, 0])
    np.testing.assert_array_equal(got, ref)
    assert np.all(got[:, 2] == 20 * 19)


def test_empty_and_complete_networks(gpu_lib, oracle_lib):
    n = 70
    a = np.zeros((2, n, n), np.uint8)
    a[1] = 1
    for kind, lv in ((0, [0, n - 1, n]), (1, [0, n - 1, n]), (2, [1, 2, 3])):
        got = gpu_lib.gof_distribution(a, kind, lv)
        np.testing.assert_array_equal(got, oracle_lib.gof_distribution(a, kind, lv))
    geo = gpu_lib.gof_distribution(a, 2, [1])
    assert geo[0].tolist() == [0, n * (n - 1)] and geo[1].tolist() == [n * (n - 1), 0]


def test_argument_checks(gpu_lib):
    """Status codes as libcusna_saom_cpu.cpp:719,764; a geodesic call above
    the LDS (n = 1100, empty network) succeeds: every pair unreached."""
    from cusna_amd.abi import CusnaError

    a = _nets(1, 6, 1.0, 1)
    with pytest.raises(CusnaError, match="UNSUPPORTED|status 5"):
        gpu_lib.gof_distribution(a, 3, [1])
    with pytest.raises(CusnaError, match="NULL_ARG|status 1"):
        gpu_lib.gof_distribution(a, 0, [])
    with pytest.raises(CusnaError, match="SHAPE|status 2"):
        gpu_lib.gof_distribution(a, 2, [1, -1])
    big = np.zeros((1, 1100, 1100), np.uint8)
    np.testing.assert_array_equal(gpu_lib.gof_distribution(big, 2, [1]), [[0, 1100 * 1099]])
    np.testing.assert_array_equal(gpu_lib.gof_distribution(big, 0, [0]), [[1100]])


@pytest.mark.parametrize("n,deg", [(1, 0.0), (5, 1.0), (37, 1.5), (65, 2.0), (300, 1.5),
                                   (600, 1.1)])
def test_geodesic_device_memory_form_matches_oracle(n, deg, gpu_lib, oracle_lib, monkeypatch):
    """The form that walks rows in device memory, forced where the LDS form
    would run (CUSNA_GOF_STREAM=1), against the oracle bit for bit."""
    monkeypatch.setenv("CUSNA_GOF_STREAM", "1")
    M = 3 if n <= 300 else 1
    a = _nets(M, n, deg, 11 * n + 1)
    lv = GEO_LEVELS if n < 300 else [1, 2, 4, 8, 16]
    print(f"case geodesic device-memory form n={n}", flush=True)
    np.testing.assert_array_equal(gpu_lib.gof_distribution(a, 2, lv),
                                  oracle_lib.gof_distribution(a, 2, lv))


def test_geodesic_above_lds_matches_oracle(gpu_lib, oracle_lib):
    """n = 1100 (18 words, rows beyond the LDS): against the oracle's boolean
    powers at levels 1 and 2."""
    a = _nets(1, 1100, 1.2, 1100)
    print("case geodesic n=1100 vs oracle", flush=True)
    np.testing.assert_array_equal(gpu_lib.gof_distribution(a, 2, [2, 1]),
                                  oracle_lib.gof_distribution(a, 2, [2, 1]))


@pytest.mark.parametrize("n,deg,M", [(1100, 1.3, 2), (2100, 1.6, 1), (4200, 2.0, 1)])
def test_geodesic_above_lds_matches_bfs(n, deg, M, gpu_lib):
    """Networks above the LDS (one and two words per lane) against
    breadth-first distances at levels up to the diameter of the giant
    component and beyond."""
    from geodesic_check import geodesic_counts

    a = _nets(M, n, deg, 3 * n + 7)
    lv = [1, 2, 3, 5, 8, 13, 21, 40]
    print(f"case geodesic n={n} vs breadth-first distances", flush=True)
    np.testing.assert_array_equal(gpu_lib.gof_distribution(a, 2, lv), geodesic_counts(a, lv))


def _plan_problem(n, M, seed, masks):
    from cusna_amd.effects import Effect
    from cusna_amd.workloads import bernoulli_network, effect_problem

    x0 = bernoulli_network(n, 2.0, np.random.default_rng(seed))
    pr = effect_problem(x0, [Effect("density"), Effect("recip"), Effect("transTrip")],
                        [-2.6, 1.8, 0.3], 1.5, M, seed)
    if masks:
        rng = np.random.default_rng(seed + 1)
        pr.keep = (rng.random((n, n)) > 0.1).astype(np.uint8)
        pr.force = ((rng.random((n, n)) < 0.01) & (pr.keep == 0)).astype(np.uint8)
    return pr


@pytest.mark.parametrize("n,masks", [(50, False), (50, True), (300, True), (1100, False)])
def test_plan_gof_matches_dense(n, masks, gpu_lib, oracle_lib):
    """GoF of the end networks where they lie in device memory equals the
    oracle's GoF of the dense end networks the same run returns; with the
    masks, of (a_end & keep) | force (the statistic network)."""
    M = 8 if n <= 300 else 2
    pr = _plan_problem(n, M, 11 + n, masks)
    plan = gpu_lib.PeriodPlan(pr, want_end_nets=True)
    plan.set_params(pr.theta, pr.lam)
    plan.run(pr.seed)
    ends = plan.fetch(want_end_nets=True)["a_ends"].reshape(M, n, n).astype(np.uint8)
    stat = ends
    if masks:
        stat = (ends & pr.keep[None]) | pr.force[None]
    from geodesic_check import geodesic_counts

    for kind, lv in ((0, DEG_LEVELS), (1, DEG_LEVELS), (2, [1, 2, 3, 6])):
        # above the LDS the geodesic counts come from breadth-first distances
        check = (lambda a: geodesic_counts(a, lv)) if (kind == 2 and n > 1024) else \
            (lambda a: oracle_lib.gof_distribution(a, kind, lv))
        raw = plan.gof(kind, lv, masked=False)
        np.testing.assert_array_equal(raw, check(ends))
        got = plan.gof(kind, lv, masked=True)
        np.testing.assert_array_equal(got, check(stat))
