"""Geodesic distribution from breadth-first distances (scipy's unweighted
shortest paths): the checker for networks too large for the oracle's
boolean matrix powers (libcusna_saom_cpu.cpp:716-765 counts ordered pairs
i != j within distance k, which is what shortest-path distances give)."""
import numpy as np


def geodesic_counts(a, levels):
    """a: M x n x n dense networks; levels: geodesic levels (>= 0). Returns
    M x (len(levels) + 1): pairs i != j at distance <= level, then the pairs
    not reached within the largest level."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path

    a = np.asarray(a)
    M, n = a.shape[0], a.shape[1]
    kmax = max(levels)
    out = np.zeros((M, len(levels) + 1))
    for c in range(M):
        d = shortest_path(csr_matrix(a[c].astype(np.int8)), method="D", directed=True,
                          unweighted=True)
        np.fill_diagonal(d, np.inf)
        for L, lev in enumerate(levels):
            out[c, L] = np.count_nonzero(d <= lev) if lev >= 1 else 0
        out[c, -1] = n * (n - 1) - (np.count_nonzero(d <= kmax) if kmax >= 1 else 0)
    return out
