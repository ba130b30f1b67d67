"""Effect registry and descriptor preparation.

Host-side restatement of the reference's R effect layer
(cran/cusna R/effects.R): effect ids (:14-48), gwesp weight tables (:53-63),
effect specs (:66-134), covariate centring (:139-147), simX parameters
(:150-166), ``prepare_effects`` (:176-238), interaction indices (:241-262),
etypes (:265-274) and weight tables (:285-294). The descriptor bundle it
returns is what the period call consumes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# numeric effect ids shared with the kernels (R/effects.R:14-23)
EFF = dict(
    density=0, recip=1, transTrip=2, cycle3=3, inPop=4, transTies=5, transRecTrip=6,
    inPopSqrt=7, outPop=8, outPopSqrt=9, egoX=10, altX=11, simX=12, sameX=13, inAct=14,
    inActSqrt=15, outAct=16, outActSqrt=17, outTrunc=18, reciAct=19, higher=20,
    egoXaltX=21, X=22, transTriads=23, gwesp=24, egoPlusAltX=25, gwespFB=26, gwespBF=27,
    gwespBB=28, gwdspFF=29, gwdspFB=30, altRThresholdX=31, interaction=32)
STRUCTURAL = {k: v for k, v in EFF.items() if k in (
    "density", "recip", "transTrip", "cycle3", "inPop", "transTies", "transRecTrip",
    "inPopSqrt", "outPop", "outPopSqrt", "inAct", "inActSqrt", "outAct", "outActSqrt",
    "outTrunc", "reciAct", "transTriads", "gwesp", "gwespFB", "gwespBF", "gwespBB",
    "gwdspFF", "gwdspFB")}
STRUCTURAL["gwespFF"] = 24
COVARIATE = {k: EFF[k] for k in ("egoX", "altX", "simX", "sameX", "higher", "egoXaltX",
                                 "egoPlusAltX", "altRThresholdX")}
DYADIC = {"X": 22}
# cross-network effects of the multi-network model (R/effects.R:33-37)
CROSSNET = dict(crprod=33, crprodRecip=34, crprodMutual=35, **{"from": 36, "to": 37})
CROSS_IDS = (33, 34, 35, 36, 37)
GW_FAMILY = (24, 26, 27, 28, 29, 30)
BEH = dict(linear=0, quad=1, avAlt=2, avSim=3, totSim=4)
ETYPES = dict(eval=0, creation=1, endow=2)
DEG_P1 = (18, 8, 9, 14, 15)
DEFAULT_PARM = dict(outTrunc=1.0, outPop=1.0, inAct=1.0, outPopSqrt=0.0, inActSqrt=0.0,
                    gwesp=69.0, gwespFF=69.0, gwespFB=69.0, gwespBF=69.0, gwespBB=69.0,
                    gwdspFF=69.0, gwdspFB=69.0, altRThresholdX=0.0)
INT_COERCE = ("gwesp", "gwespFF", "gwespFB", "gwespBF", "gwespBB", "gwdspFF", "gwdspFB",
              "altRThresholdX")


def gwesp_weights(parm: float, n: int) -> np.ndarray:
    """Cumulative geometric weights w[s] = e^a (1 - (1 - e^-a)^s), w[0] = 0,
    a = parm/100, built by the running product of R/effects.R:53-63."""
    alpha = 0.01 * float(parm)
    w = np.zeros(n + 1)
    factor = 1.0 - math.exp(-alpha)
    pw = 1.0
    for s in range(1, n + 1):
        pw = pw * factor
        w[s] = math.exp(alpha) * (1.0 - pw)
    return w


@dataclass
class Effect:
    """A network effect spec (R/effects.R:66-111)."""

    name: str
    cov: np.ndarray | None = None
    parm: float | None = None
    dyn: bool = False
    etype: str = "eval"
    comps: tuple | None = None  # interaction components (0-based, earlier effects)
    net_ref: int | None = None  # cross-network effects: source network (0-based)
    type_id: int = field(init=False)
    needs_cov: bool = field(init=False)

    def __post_init__(self):
        if self.etype not in ETYPES:
            raise ValueError(f"unknown effect type {self.etype}")
        if self.name == "interaction" or self.comps is not None:
            c = tuple(int(x) for x in self.comps or ())
            if len(c) not in (2, 3):
                raise ValueError("interaction needs 2 or 3 components")
            self.comps, self.type_id, self.needs_cov, self.parm = c, 32, False, 1.0
            return
        if self.name in CROSSNET:
            if self.net_ref is None:
                raise ValueError(f"{self.name} requires net_ref (the other dependent network "
                                 "index)")
            self.type_id, self.needs_cov = CROSSNET[self.name], False
        elif self.name in STRUCTURAL:
            self.type_id, self.needs_cov = STRUCTURAL[self.name], False
        elif self.name in COVARIATE:
            self.type_id, self.needs_cov = COVARIATE[self.name], not self.dyn
            if self.cov is None and self.needs_cov:
                raise ValueError(f"{self.name} requires a covariate")
        elif self.name in DYADIC:
            self.type_id, self.needs_cov = 22, True
            w = None if self.cov is None else np.asarray(self.cov)
            if w is None or w.ndim not in (2, 3) or w.shape[0] != w.shape[1]:
                raise ValueError(f"{self.name} requires an (n,n) or (n,n,periods) dyadic covariate")
        else:
            raise ValueError(f"unknown effect {self.name}")
        if self.parm is None:
            self.parm = DEFAULT_PARM.get(self.name, 1.0)
        if self.name in INT_COERCE:
            self.parm = float(math.trunc(self.parm))


@dataclass
class BehEffect:
    name: str
    type_id: int = field(init=False)

    def __post_init__(self):
        if self.name not in BEH:
            raise ValueError(f"unknown behavior effect {self.name}")
        self.type_id = BEH[self.name]


RATE_KINDS = dict(RateX=0, outRate=1, inRate=2)


@dataclass
class RateEffect:
    """A non-basic rate effect (R/effects.R:120-127): RateX (covariate),
    outRate (out-degree), inRate (in-degree)."""

    name: str
    cov: np.ndarray | None = None
    kind: int = field(init=False)

    def __post_init__(self):
        if self.name not in RATE_KINDS:
            raise ValueError(f"unknown rate effect {self.name}")
        self.kind = RATE_KINDS[self.name]
        if self.kind == 0 and self.cov is None:
            raise ValueError("RateX requires a covariate")


def rate_cov_values(e: RateEffect, p: int = 1) -> np.ndarray:
    """Centred rate covariate for period p, missing values 0 (R/effects.R:168-173)."""
    v = np.asarray(e.cov, dtype=float)
    gmean = np.nanmean(v)
    vm = v - gmean if v.ndim == 1 else v[:, min(p, v.shape[1]) - 1] - gmean
    return np.where(np.isnan(vm), 0.0, vm)


def effect_cov_values(e: Effect, p: int = 1) -> np.ndarray:
    """Centred covariate for 1-based period p; grand mean over all columns
    (R/effects.R:139-147)."""
    v = np.asarray(e.cov, dtype=float)
    gmean = np.nanmean(v)
    if v.ndim == 1:
        return v - gmean
    return v[:, min(p, v.shape[1]) - 1] - gmean


def effect_sim_params(e: Effect) -> tuple[float, float]:
    """(1/range, simMean) pooled over covariate columns (R/effects.R:150-166)."""
    v = np.asarray(e.cov, dtype=float)
    rng = np.nanmax(v) - np.nanmin(v)
    cols = v.reshape(v.shape[0], -1, order="F") if v.ndim > 1 else v.reshape(-1, 1)
    tot, cnt = 0.0, 0
    for m in range(cols.shape[1]):
        x = cols[~np.isnan(cols[:, m]), m]
        if len(x) > 1:
            d = 1.0 - np.abs(np.subtract.outer(x, x)) / rng
            tot = tot + (_rsum(d) - len(x))
            cnt += len(x) * (len(x) - 1)
    return 1.0 / rng, (tot / cnt if cnt > 0 else 0.0)


def _rsum(m: np.ndarray) -> float:
    """Sum in column-major order (the order R's sum() walks a matrix)."""
    return _seq_sum(np.asarray(m, dtype=float).ravel(order="F"))


def _seq_sum(x: np.ndarray) -> float:
    # R's sum() accumulates in long double; a sequential float64 sum differs
    # from it only in the last bits, far below every tolerance that uses it.
    s = 0.0
    for v in x.tolist():
        s += v
    return s


@dataclass
class Descriptors:
    type_ids: np.ndarray
    p1: np.ndarray
    p2: np.ndarray
    vmat: np.ndarray
    vmiss: np.ndarray
    dyn_flags: np.ndarray
    v_dyn: np.ndarray
    dymats: list
    dyc_idx: np.ndarray
    wtab: np.ndarray
    etypes: np.ndarray
    int_comp: np.ndarray


def effect_netrefs(effects: list) -> np.ndarray:
    """Source network of each cross-network effect, -1 otherwise
    (R/effects.R:277-283)."""
    return np.array([-1 if e.net_ref is None else int(e.net_ref) for e in effects], np.int32)


def prepare_effects(effects: list, n: int, p: int = 1, beh=None, z_start=None) -> Descriptors:
    """Float64 descriptor bundle for period p (R/effects.R:176-238, 241-294)."""
    k = len(effects)
    type_ids = np.array([e.type_id for e in effects], dtype=np.int32)
    p1 = np.zeros(k)
    p2 = np.zeros(k)
    vmat = np.zeros((k, n))
    vmiss = np.zeros((k, n), dtype=np.int32)
    dyn = np.array([1 if e.dyn else 0 for e in effects], dtype=np.int32)
    dymats, dyc_idx = [], -np.ones(k, dtype=np.int32)
    off = ~np.eye(n, dtype=bool)
    for idx, e in enumerate(effects):
        t = e.type_id
        if t == 12:
            if e.dyn:
                if beh is not None:
                    p1[idx] = 1.0 / beh.range
                    p2[idx] = beh.sim_mean
            else:
                p1[idx], p2[idx] = effect_sim_params(e)
        elif t in DEG_P1:
            p1[idx] = float(e.parm)
        elif t == 31:
            p1[idx] = float(math.trunc(e.parm))
        elif t in CROSS_IDS:
            p1[idx] = float(e.parm)
        if t == 22:
            w = np.asarray(e.cov, dtype=float)
            if w.ndim == 3:
                flat = w.reshape(n * n, -1, order="F")[off.ravel(order="F"), :]
                wm = w[:, :, min(p, w.shape[2]) - 1] - np.nanmean(flat)
            else:
                wm = w - np.mean(w[off])
            wm = np.array(wm, dtype=float)
            np.fill_diagonal(wm, 0.0)
            dymats.append(wm)
            dyc_idx[idx] = len(dymats) - 1
        elif e.needs_cov and not e.dyn:
            vv = effect_cov_values(e, p)
            if len(vv) != n:
                raise ValueError(f"{e.name}: covariate length != n")
            miss = np.isnan(vv)
            vmiss[idx] = miss.astype(np.int32)
            vmat[idx] = np.where(miss, 0.0, vv)
        elif e.dyn and beh is not None:
            vmiss[idx] = beh.miss[:, min(p, beh.miss.shape[1]) - 1].astype(np.int32)
    if (dyn == 1).any() and beh is not None and z_start is not None:
        v_dyn = np.asarray(z_start, dtype=float) - beh.overall_mean
    else:
        v_dyn = np.zeros(n)
    wtab = np.zeros((k, n + 1))
    for i, e in enumerate(effects):
        if e.type_id in GW_FAMILY:
            wtab[i] = gwesp_weights(e.parm, n)
    etypes = np.zeros(k, dtype=np.int32)
    for i, e in enumerate(effects):
        et = ETYPES[e.etype]
        if et != 0 and e.type_id != 25:
            raise ValueError("creation/endowment types are supported for egoPlusAltX only")
        etypes[i] = et
    ic = -np.ones((k, 3), dtype=np.int32)
    for ei, e in enumerate(effects):
        if e.type_id != 32:
            continue
        for slot, ci in enumerate(e.comps):
            if not (0 <= ci < ei):
                raise ValueError(f"interaction component {ci} must precede position {ei}")
            if effects[ci].type_id == 32:
                raise ValueError("nested interactions not supported")
            if effects[ci].etype != "eval":
                raise ValueError("interaction components must be eval effects")
            ic[ei, slot] = ci
    return Descriptors(type_ids, p1, p2, vmat, vmiss, dyn, v_dyn, dymats, dyc_idx, wtab,
                       etypes, ic)
