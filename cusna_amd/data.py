"""SAOM panel data conventions.

Host-side restatement of the reference's R data engine (cran/cusna R/data.R):
wave parsing (:27-55), imputation and symmetry (:60-118), behaviour data
(:213-259), period masks (:264-362) and behaviour masks (:366-383).
Periods are 1-based, as in the R code: period p spans waves p and p+1.
Composition change (R/data.R:123-210) is not restated: the R glue never
passes composition inputs to the period call (cusna_saom_glue.cpp:106).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

NA_CODE, STRUCT_ZERO, STRUCT_ONE = 9, 10, 11


def parse_wave(w) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(vals 0/1, miss, struct) of one raw wave (R/data.R:27-55)."""
    w = np.asarray(w)
    n = w.shape[0]
    if w.ndim != 2 or w.shape[1] != n:
        raise ValueError("each wave must be a square matrix.")
    if np.issubdtype(w.dtype, np.floating) and np.isnan(w).any():
        miss = np.isnan(w)
        wi = np.trunc(np.where(miss, 0.0, w)).astype(np.int64)
    else:
        wi = np.rint(w).astype(np.int64)
        miss = wi == NA_CODE
    struct = (wi == STRUCT_ZERO) | (wi == STRUCT_ONE)
    vals = wi.copy()
    vals[wi == STRUCT_ZERO] = 0
    vals[wi == STRUCT_ONE] = 1
    vals[miss] = 0
    chk = vals[~miss & ~struct]
    if not np.isin(chk, (0, 1)).all():
        raise ValueError("tie values must be 0/1/9(NA)/10/11 or NaN")
    np.fill_diagonal(vals, 0)
    np.fill_diagonal(miss, False)
    np.fill_diagonal(struct, False)
    return vals.astype(np.int32), miss, struct


@dataclass
class BehaviorData:
    n: int
    n_waves: int
    values: np.ndarray
    miss: np.ndarray
    imputed: np.ndarray
    overall_mean: float
    zmin: int
    zmax: int
    range: float
    sim_mean: float


def build_behavior_data(values, n: int, nw: int) -> BehaviorData:
    """Behaviour scalars and imputation (R/data.R:213-259)."""
    v = np.asarray(values, dtype=float).reshape(n, nw)
    miss = np.isnan(v)
    obs = v[~miss]
    wave_means = [float(np.mean(v[~miss[:, m], m])) for m in range(nw)]
    overall_mean = float(np.mean(wave_means))
    zmin, zmax = int(obs.min()), int(obs.max())
    rng = float(zmax - zmin)
    tot, cnt = 0.0, 0
    for m in range(nw - 1):
        x = v[~np.isnan(v[:, m]), m]
        if len(x) > 1:
            d = 1.0 - np.abs(np.subtract.outer(x, x)) / rng
            s = 0.0
            for q in d.ravel(order="F").tolist():
                s += q
            tot = tot + (s - len(x))
            cnt += len(x) * (len(x) - 1)
    sim_mean = tot / cnt if cnt > 0 else 0.0
    imp = v.copy()
    allna = miss.all(axis=1)
    if allna.any():
        vals, counts = np.unique(obs, return_counts=True)
        imp[allna, :] = vals[np.argmax(counts)]
    first = imp[:, 0].copy()
    for m in range(1, nw):
        first = np.where(np.isnan(first), imp[:, m], first)
    imp[:, 0] = first
    for m in range(1, nw):
        imp[:, m] = np.where(np.isnan(imp[:, m]), imp[:, m - 1], imp[:, m])
    if np.isnan(imp).any():
        raise ValueError("behavior imputation failed.")
    return BehaviorData(n, nw, v, miss, np.rint(imp).astype(np.int32), overall_mean, zmin, zmax,
                        rng, sim_mean)


@dataclass
class SaomData:
    """Internal SAOM data object (R/data.R:60-118)."""

    n: int
    n_waves: int
    vals: list
    miss: list
    struct: list
    imputed: list
    symmetric: bool
    beh: BehaviorData | None = None

    @property
    def n_periods(self) -> int:
        return self.n_waves - 1

    # -------------------------------------------------- period masks (1-based p)
    def kmask_dist(self, p: int) -> np.ndarray:
        k = ~(self.miss[p - 1] | self.miss[p] | self.struct[p - 1] | self.struct[p])
        np.fill_diagonal(k, False)
        return k

    def observed_distance(self, p: int) -> int:
        excl = self.miss[p - 1] | self.miss[p] | self.struct[p - 1] | self.struct[p]
        diff = (self.vals[p - 1] != self.vals[p]) & ~excl
        np.fill_diagonal(diff, False)
        return int(diff.sum())

    def rate_target(self, p: int) -> int:
        a = (self.imputed[p - 1] != 0) & ~self.miss[p - 1]
        b = (self.imputed[p] != 0) & ~self.miss[p - 1] & ~self.miss[p]
        diff = (a ^ b) & ~self.struct[p - 1] & ~self.struct[p]
        np.fill_diagonal(diff, False)
        return int(diff.sum())

    def network_less_missing_start(self, p: int) -> np.ndarray:
        return ((self.imputed[p - 1] != 0) & ~self.miss[p - 1] & ~self.miss[p]).astype(np.uint8)

    def stat_keep_force(self, p: int) -> tuple[np.ndarray, np.ndarray]:
        """(keep, force): statistic network = (A & keep) | force (R/data.R:319-331)."""
        keep = ~(self.miss[p - 1] | self.miss[p])
        force = np.zeros((self.n, self.n), dtype=bool)
        s1 = self.struct[p]
        keep &= ~s1
        force[s1] = self.imputed[p][s1] != 0
        s0 = self.struct[p - 1]
        keep &= ~s0
        force[s0] = self.imputed[p - 1][s0] != 0
        np.fill_diagonal(keep, False)
        np.fill_diagonal(force, False)
        return keep, force

    def stat_network(self, a: np.ndarray, p: int) -> np.ndarray:
        keep, force = self.stat_keep_force(p)
        return (((np.asarray(a) != 0) & keep) | force).astype(np.uint8)

    def target_network(self, p: int) -> np.ndarray:
        return self.stat_network(self.imputed[p], p)

    def permitted(self, p: int) -> np.ndarray:
        pm = ~self.struct[p - 1]
        np.fill_diagonal(pm, False)
        return pm

    # ----------------------------------------------------- behaviour masks
    def beh_kmask_dist(self, p: int) -> np.ndarray:
        return ~(self.beh.miss[:, p - 1] | self.beh.miss[:, p])

    def beh_observed_distance(self, p: int) -> int:
        d = np.abs(self.beh.imputed[:, p] - self.beh.imputed[:, p - 1])
        return int(d[self.beh_kmask_dist(p)].sum())

    def beh_values_less_missing_starts(self, p: int) -> np.ndarray:
        return np.where(self.beh.miss[:, p - 1], 0.0, self.beh.imputed[:, p - 1].astype(float))


def build_saom_data(waves, behavior=None) -> SaomData:
    """Parse, impute (NA -> previous imputed wave) and flag symmetry (R/data.R:60-118)."""
    if len(waves) < 2:
        raise ValueError("`waves` must be a list of at least two adjacency matrices.")
    parsed = [parse_wave(w) for w in waves]
    n = parsed[0][0].shape[0]
    if any(p[0].shape[0] != n for p in parsed):
        raise ValueError("all waves must have the same node set.")
    vals = [p[0] for p in parsed]
    miss = [p[1] for p in parsed]
    struct = [p[2] for p in parsed]
    imputed = [vals[0].copy()]
    for m in range(1, len(waves)):
        w = vals[m].copy()
        mm = miss[m]
        w[mm] = imputed[m - 1][mm]
        imputed.append(w)
    symmetric = all((vals[m] == vals[m].T).all() and (miss[m] == miss[m].T).all()
                    and (struct[m] == struct[m].T).all() for m in range(len(waves)))
    beh = build_behavior_data(behavior, n, len(waves)) if behavior is not None else None
    return SaomData(n, len(waves), vals, miss, struct, imputed, bool(symmetric), beh)
