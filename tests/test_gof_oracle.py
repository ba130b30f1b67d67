"""CPU: the oracle's GoF distributions against the reference compiled from
source (libcusna_saom_cpu.cpp:716-765) and against closed forms (numpy degree
counts; breadth-first search distances)."""
import numpy as np
import pytest


def _nets(M, n, deg, seed, self_loops=False):
    rng = np.random.default_rng(seed)
    a = (rng.random((M, n, n)) < deg / (n - 1)).astype(np.uint8)
    if not self_loops:
        for c in range(M):
            np.fill_diagonal(a[c], 0)
    return a


def _bfs_counts(a, maxk):
    """counts[k] = ordered pairs i != j with shortest-path length <= k."""
    n = a.shape[0]
    counts = np.zeros(maxk + 1, np.int64)
    for i in range(n):
        dist = np.full(n, -1)
        dist[i] = 0
        front = [i]
        d = 0
        while front and d < maxk:
            d += 1
            nxt = []
            for u in front:
                for v in np.nonzero(a[u])[0]:
                    if dist[v] < 0:
                        dist[v] = d
                        nxt.append(v)
            front = nxt
        for k in range(1, maxk + 1):
            counts[k] += int(np.sum((dist >= 1) & (dist <= k)))
    return counts


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_gof_matches_compiled_reference(oracle_lib, kind):
    R = oracle_lib.reference()
    if R is None:
        pytest.skip("oracle/_ref not built")
    a = _nets(5, 37, 3.0, kind, self_loops=True)
    lv = [3, 0, 1, 5, 2] if kind == 2 else [0, 1, 2, 3, 4, 5, 8]
    got = oracle_lib.gof_distribution(a, kind, lv)
    ref = oracle_lib.gof_distribution(a, kind, lv, lib=R, fn="cusna_gof_distribution")
    np.testing.assert_array_equal(got, ref)


def test_gof_closed_forms(oracle_lib):
    a = _nets(3, 40, 2.5, 7)
    lv = list(range(0, 9))
    ind = oracle_lib.gof_distribution(a, 0, lv)
    outd = oracle_lib.gof_distribution(a, 1, lv)
    for c in range(3):
        np.testing.assert_array_equal(ind[c], [(a[c].sum(0) <= L).sum() for L in lv])
        np.testing.assert_array_equal(outd[c], [(a[c].sum(1) <= L).sum() for L in lv])
    geo = oracle_lib.gof_distribution(a, 2, [1, 2, 3, 4, 6])
    for c in range(3):
        counts = _bfs_counts(a[c], 6)
        np.testing.assert_array_equal(geo[c, :5], counts[[1, 2, 3, 4, 6]])
        assert geo[c, 5] == 40 * 39 - counts[6]
        assert geo[c, 0] == a[c].sum()  # distance 1 = the ties


def test_gof_argument_checks(oracle_lib):
    a = _nets(1, 5, 1.0, 1)
    with pytest.raises(RuntimeError, match="status 5"):
        oracle_lib.gof_distribution(a, 3, [1])
    with pytest.raises(RuntimeError, match="status 1"):
        oracle_lib.gof_distribution(a, 0, [])


def test_shortest_path_checker_matches_oracle(oracle_lib):
    """The breadth-first checker that the device tests use above the oracle's
    sizes gives the oracle's counts (self-loops, unreached pairs, level 0)."""
    from geodesic_check import geodesic_counts

    a = _nets(3, 150, 1.3, 21, self_loops=True)
    for lv in ([1, 2, 3, 5, 8], [0, 2], [4, 1, 4]):
        np.testing.assert_array_equal(geodesic_counts(a, lv),
                                      oracle_lib.gof_distribution(a, 2, lv))
