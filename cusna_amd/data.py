"""SAOM panel data conventions.

Host-side restatement of the reference's R data engine (cran/cusna R/data.R):
wave parsing (:27-55), imputation and symmetry (:60-118), behaviour data
(:213-259), period masks (:264-362) and behaviour masks (:366-383).
Periods are 1-based, as in the R code: period p spans waves p and p+1.
Composition change (R/data.R:123-210) is restated for the data treatment and
the targets; ``period_actors`` gives the active list and the inactive mask
of a period for the period call (the R glue itself never passes them,
cusna_saom_glue.cpp:106).
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
    rawvals: list | None = None   # vals after the composition treatment
    has_composition: bool = False
    comp_active: np.ndarray | None = None  # (P, n): active during period p
    cc_option: int = 1

    def __post_init__(self):
        if self.rawvals is None:
            self.rawvals = [v.copy() for v in self.vals]

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
        diff = (self.rawvals[p - 1] != self.rawvals[p]) & ~excl
        np.fill_diagonal(diff, False)
        return int(diff.sum())

    def rate_target(self, p: int) -> int:
        a = (self.imputed[p - 1] != 0) & ~self.miss[p - 1]
        b = (self.imputed[p] != 0) & ~self.miss[p - 1] & ~self.miss[p]
        diff = (a ^ b) & ~self.struct[p - 1] & ~self.struct[p]
        np.fill_diagonal(diff, False)
        return int(diff.sum())

    def rate_target_rows(self, p: int) -> np.ndarray:
        """Per-ego tie differences of period p (R/data.R:294-300)."""
        a = (self.imputed[p - 1] != 0) & ~self.miss[p - 1]
        b = (self.imputed[p] != 0) & ~self.miss[p - 1] & ~self.miss[p]
        diff = (a ^ b) & ~self.struct[p - 1] & ~self.struct[p]
        np.fill_diagonal(diff, False)
        return diff.sum(1).astype(np.int64)

    def network_less_missing(self, p: int) -> np.ndarray:
        """Wave p with its missing ties removed (R/data.R:309-311)."""
        return ((self.imputed[p - 1] != 0) & ~self.miss[p - 1]).astype(np.uint8)

    def period_actors(self, p: int):
        """(act, nact, inactive mask) of period p for the period call: active
        actors first in index order, then the inactive ones; None without
        composition change."""
        if not self.has_composition:
            return None
        active = np.asarray(self.comp_active[p - 1], bool)
        act = np.concatenate([np.nonzero(active)[0], np.nonzero(~active)[0]]).astype(np.int32)
        return act, int(active.sum()), (~active).astype(np.uint8)

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

    def stat_keep_force_struct(self, p: int) -> tuple[np.ndarray, np.ndarray]:
        """(keep, force) from the structural values only (R/data.R:335-347)."""
        keep = np.ones((self.n, self.n), dtype=bool)
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

    def rate_const_rows(self, p: int) -> np.ndarray:
        """Per-ego ties observed at the start and missing at the end (R/data.R:302-307)."""
        cc = ((self.imputed[p - 1] != 0) & ~self.miss[p - 1] & self.miss[p] & ~self.struct[p - 1]
              & ~self.struct[p])
        np.fill_diagonal(cc, False)
        return cc.sum(1).astype(np.int64)

    def rate_const(self, p: int) -> int:
        return int(self.rate_const_rows(p).sum())

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

    def beh_stat_values(self, z_end, p: int) -> np.ndarray:
        """Centred behaviour, 0 where missing at either end (R/data.R:379-383)."""
        u = np.asarray(z_end, float) - self.beh.overall_mean
        u[~self.beh_kmask_dist(p)] = 0.0
        return u

    def beh_values_less_missing_starts(self, p: int) -> np.ndarray:
        return np.where(self.beh.miss[:, p - 1], 0.0, self.beh.imputed[:, p - 1].astype(float))


def apply_composition(obj: SaomData, composition, cc_option: int = 1) -> SaomData:
    """Composition-change data treatment (R/data.R:123-210): actors inactive
    at the start of a wave are, by their action code (1 before joining,
    2 after leaving, 3 a gap) and ccOption, either set to zero where their row
    is missing (row and column, imputation from the previous wave for gaps and
    leavers) or made missing entirely; waves in ascending order, codes in the
    order 1, 3, 2 within a wave."""
    n, nw = obj.n, obj.n_waves
    cc = int(cc_option)
    if cc not in (1, 2, 3):
        raise ValueError("ccOption must be 1, 2 or 3.")
    if len(composition) != n:
        raise ValueError("composition needs one interval list per actor.")
    active_start = np.zeros((n, nw), bool)
    for j in range(n):
        flat = np.asarray(composition[j], float).ravel()
        if len(flat) < 2 or len(flat) % 2:
            raise ValueError(f"actor {j + 1}: interval list must have an even length.")
        if not np.all(flat == np.round(flat)):
            raise ValueError(f"actor {j + 1}: only integer event times are supported.")
        for q in range(len(flat) // 2):
            s, e = int(flat[2 * q]), int(flat[2 * q + 1])
            if not 1 <= s <= e <= nw:
                raise ValueError(f"actor {j + 1}: interval ({s},{e}) outside 1..{nw}")
            active_start[j, s - 1:e] = True
    action = np.zeros((n, nw), int)
    for j in range(n):
        for m in range(nw):
            if active_start[j, m]:
                continue
            prec = active_start[j, :m].any()
            succ = active_start[j, m + 1:].any()
            action[j, m] = 1 if not prec else (2 if not succ else 3)
    for m in range(nw):
        for code in (1, 3, 2):
            for j in np.nonzero(action[:, m] == code)[0]:
                row_na = obj.miss[m][j, :].copy()
                zero = ((code == 1 and cc in (1, 2)) or (code == 3 and cc in (1, 2))
                        or (code == 2 and cc == 1))
                if zero:
                    obj.miss[m][j, row_na] = False
                    obj.miss[m][row_na, j] = False
                    obj.rawvals[m][j, row_na] = 0
                    obj.rawvals[m][row_na, j] = 0
                    if code == 1 or m == 0:
                        obj.imputed[m][j, row_na] = 0
                        obj.imputed[m][row_na, j] = 0
                    else:
                        obj.imputed[m][j, row_na] = obj.imputed[m - 1][j, row_na]
                        obj.imputed[m][row_na, j] = obj.imputed[m - 1][row_na, j]
                else:
                    obj.miss[m][j, :] = True
                    obj.miss[m][:, j] = True
                    obj.miss[m][j, j] = False
    obj.has_composition = True
    obj.cc_option = cc
    obj.comp_active = active_start[:, :nw - 1].T.copy()
    return obj


def build_saom_data(waves, behavior=None, composition=None, cc_option: int = 1) -> SaomData:
    """Parse, impute (NA -> previous imputed wave) and flag symmetry
    (R/data.R:60-118); optional composition change (:106-108)."""
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
    obj = SaomData(n, len(waves), vals, miss, struct, imputed, bool(symmetric), None)
    if composition is not None:
        apply_composition(obj, composition, cc_option)
    if behavior is not None:
        if composition is not None:
            raise ValueError("composition change with a dependent behavior is not supported.")
        obj.beh = build_behavior_data(behavior, n, len(waves))
    return obj
