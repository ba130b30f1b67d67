"""Marshalling of one period call from SAOM data + effects.

Restates what the reference's R layer and cpp11 glue hand to the period
simulator: ``.run_period_sim`` (cran/cusna R/mom.R:91-125) and
``cusna_simulate_period_cpu_`` (src/cusna_saom_glue.cpp:57-189): the
imputed start wave, float64 descriptors, float32 theta/lambda broadcast over
chains, behaviour scalars rounded to float32, and NULL masks (the default
off-diagonal case).
"""
from __future__ import annotations

import numpy as np

from .abi import SaomProblem
from .data import SaomData
from .effects import prepare_effects


def period_problem(data: SaomData, effects: list, p: int, theta, lam, M: int, seed: int,
                   beh_effects: list | None = None, thetab=None, lamb=0.0,
                   conditional: bool = False, maxdegree: int = 0,
                   rate_effects: list | None = None, rate_theta=None,
                   composition: bool = False) -> SaomProblem:
    """SaomProblem for 1-based period p (R/mom.R:91-106, cusna_saom_glue.cpp:67-147).

    Beyond what the R glue passes (cusna_saom_glue.cpp:106 leaves them unset),
    ``composition=True`` with composition change in ``data`` sets the period's
    active list and inactive mask (restored at the period end, libcusna_saom_cpu.cpp:290-291, 356-361),
    and ``rate_effects`` with their parameters ``rate_theta`` set the
    non-basic rate effects (:170-193): RateX uses the centred covariate,
    outRate / inRate the current degrees."""
    n = data.n
    beh = data.beh
    kb = len(beh_effects) if beh_effects else 0
    z_start = data.beh_values_less_missing_starts(p) if (beh is not None and kb) else None
    if beh is not None and z_start is None and any(e.dyn for e in effects):
        z_start = data.beh_values_less_missing_starts(p)
    d = prepare_effects(effects, n, p, beh, z_start)
    prob = SaomProblem(
        x0=data.imputed[p - 1].astype(np.uint8), theta=np.asarray(theta, dtype=float),
        lam=float(lam), type_ids=d.type_ids, M=M, seed=int(seed), etypes=np.zeros(len(effects)),
        p1=d.p1, p2=d.p2, vmat=d.vmat, vmiss=d.vmiss, dyn_flags=d.dyn_flags, dyc_idx=d.dyc_idx,
        wtab=d.wtab, int_comp=d.int_comp, vb=d.v_dyn, maxdegree=maxdegree,
        conditional=1 if conditional else 0,
        target_dist=data.observed_distance(p) if conditional else 0,
        symflag=1 if data.symmetric else 0)
    comp = data.period_actors(p) if composition else None
    if comp is not None:
        prob.act, prob.nact, prob.inact = comp
        prob.has_comp = 1
    if rate_effects:
        from .effects import rate_cov_values

        prob.rate_kind = np.array([e.kind for e in rate_effects], np.int32)
        prob.rate_alpha = np.asarray(rate_theta, dtype=float)
        prob.rate_u = np.stack([rate_cov_values(e, p) if e.kind == 0 else np.zeros(n)
                                for e in rate_effects])
    if kb:
        prob.kb = kb
        prob.lamb = float(lamb)
        prob.thetab = np.asarray(thetab, dtype=float)
        prob.bty = np.array([e.type_id for e in beh_effects], dtype=np.int32)
        prob.z0 = beh.imputed[:, p - 1].astype(np.int32)
        prob.bkdist = data.beh_kmask_dist(p).astype(np.int32)
        prob.zmin, prob.zmax = beh.zmin, beh.zmax
        prob.zbar, prob.brange, prob.bsimmean = beh.overall_mean, beh.range, beh.sim_mean
    return prob
