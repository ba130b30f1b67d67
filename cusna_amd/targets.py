"""Moment targets over the statistic carriers.

Restates the reference's R target assembly (cran/cusna R/targets.R): per-period
evaluation targets on the transformed observed end wave (:19-40), behaviour
targets (:43-59), basic-rate targets with the composition-change gate
(:66-71), non-basic rate-effect targets (:73-89), behaviour rate targets
(:128-132) and the theta-ordered target vector (:99-126). ``net_stats`` /
``beh_stats`` are the carriers: by default the GPU library's
cusna_network_stats / cusna_behavior_stats.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .data import SaomData
from .effects import prepare_effects, rate_cov_values


def _gpu_net_stats(a, **kw):
    from . import lib

    return lib.network_stats(a, **kw)


def _gpu_beh_stats(z, bty, xnet, bkdist, zbar, brange, bsimmean):
    from . import lib

    return lib.behavior_stats(z, bty, xnet, bkdist, zbar, brange, bsimmean)


def eval_targets(data: SaomData, effects: list, net_stats=_gpu_net_stats) -> np.ndarray:
    """(P, k) evaluation targets (R/targets.R:19-40)."""
    P, n, k = data.n_periods, data.n, len(effects)
    beh = data.beh
    symflag = 1 if data.symmetric else 0
    out = np.zeros((P, k))
    for p in range(1, P + 1):
        z_start = data.beh_values_less_missing_starts(p) if beh is not None else None
        d = prepare_effects(effects, n, p, beh, z_start)
        a_start = data.imputed[p - 1]
        a_tgt = data.target_network(p)
        out[p - 1] = net_stats(a_tgt, type_ids=d.type_ids, p1=d.p1, p2=d.p2, vmat=d.vmat,
                               vmiss=d.vmiss, dyn_flags=d.dyn_flags, vb=d.v_dyn,
                               od0=(a_start != 0).sum(1).astype(np.int32),
                               id0=(a_start != 0).sum(0).astype(np.int32), wtab=d.wtab,
                               symflag=symflag)
    return out


def beh_targets(data: SaomData, beh_effects: list, beh_stats=_gpu_beh_stats) -> np.ndarray:
    """(P, kb) behaviour targets on the observed end behaviour (R/targets.R:43-59)."""
    beh = data.beh
    bty = np.array([e.type_id for e in beh_effects], np.int32)
    out = np.zeros((data.n_periods, len(beh_effects)))
    for p in range(1, data.n_periods + 1):
        z_end = beh.imputed[:, p]
        out[p - 1] = beh_stats(z_end[None, :], bty, data.network_less_missing_start(p),
                               data.beh_kmask_dist(p).astype(np.int32), beh.overall_mean,
                               beh.range, beh.sim_mean)[0]
    return out


def basic_rate_targets(data: SaomData) -> np.ndarray:
    """Per period: the rate target (missing-aware tie changes) with
    composition change, else the observed Hamming distance (:66-71)."""
    fn = data.rate_target if data.has_composition else data.observed_distance
    return np.array([float(fn(p)) for p in range(1, data.n_periods + 1)])


def rate_effect_targets(data: SaomData, rate_effects: list) -> np.ndarray:
    """(P, nr): per-ego tie differences weighted by the centred covariate
    (RateX) or the out-/in-degrees of the period-start wave less its missing
    ties (outRate/inRate) (:73-89)."""
    P = data.n_periods
    out = np.zeros((P, len(rate_effects)))
    for p in range(1, P + 1):
        rows = data.rate_target_rows(p).astype(float)
        for r, e in enumerate(rate_effects):
            if e.kind == 0:
                w = rate_cov_values(e, p)
            elif e.kind == 1:
                w = data.network_less_missing(p).sum(1).astype(float)
            else:
                w = data.network_less_missing(p).sum(0).astype(float)
            out[p - 1, r] = float(np.sum(w * rows))
    return out


def beh_rate_targets(data: SaomData) -> np.ndarray:
    return np.array([float(data.beh_observed_distance(p))
                     for p in range(1, data.n_periods + 1)])


@dataclass
class MomTargets:
    eval: np.ndarray
    beh: np.ndarray | None
    basic_rate: np.ndarray
    beh_rate: np.ndarray | None
    target_vector: np.ndarray
    layout: list
    rate_eff: np.ndarray | None = None


def mom_targets(data: SaomData, effects: list, beh_effects: list | None = None,
                conditional: bool = True, net_stats=_gpu_net_stats,
                beh_stats=_gpu_beh_stats, rate_effects: list | None = None) -> MomTargets:
    """Theta-ordered target vector (R/targets.R:99-126):
    unconditional [rate x P, rate effects x nr, eval x k]; conditional
    [rate effects x nr, eval x k]; co-evolution [net rate x P, beh rate x P,
    eval x k, beh x kb]."""
    ev = eval_targets(data, effects, net_stats)
    br = basic_rate_targets(data)
    re_t = rate_effect_targets(data, rate_effects) if rate_effects else None
    bt = beh_targets(data, beh_effects, beh_stats) if beh_effects else None
    bh = beh_rate_targets(data) if bt is not None else None
    parts, layout = [], []
    if bt is not None:
        parts += [br, bh, ev.sum(0), bt.sum(0)]
        layout = ["net_rate", "beh_rate", "eval", "beh"]
    else:
        if not conditional:
            parts.append(br)
            layout.append("rate")
        if re_t is not None:
            parts.append(re_t.sum(0))
            layout.append("rate_eff")
        parts.append(ev.sum(0))
        layout.append("eval")
    return MomTargets(ev, bt, br, bh, np.concatenate(parts), layout, re_t)
