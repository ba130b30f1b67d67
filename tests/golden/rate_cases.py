"""Synthetic period cases with non-basic rate effects and composition change,
for golden fixtures.

``rate_specs()`` lists JSON-serialisable specifications on top of the
ref_cases format (``rate``: list of [kind, alpha] with kind 0 a covariate,
1 out-degree, 2 in-degree; ``comp``: inactive actors and active list);
``rate_problem(spec)`` builds the SaomProblem; ``build_rate_cases(O)`` runs
the reference engine compiled from its source (oracle/_ref) on every spec.
"""
from __future__ import annotations

import numpy as np

from ref_cases import _pack_list, spec_problem

_BASE = dict(M=4, lam=3.0, theta=[-1.8, 1.0, 0.2],
             effects=[{"name": "density"}, {"name": "recip"}, {"name": "transTrip"}], deg=3.0)


def rate_specs() -> list:
    S = []

    def add(name, n, seed, **kw):
        d = dict(_BASE, name=name, n=n, seed=seed, net_seed=seed + 9)
        d.update(kw)
        S.append(d)

    add("rate_outdeg", 45, 301, rate=[[1, 0.08]])
    add("rate_indeg", 45, 302, rate=[[2, -0.06]])
    add("rate_cov", 50, 303, rate=[[0, 0.4]])
    add("rate_four", 70, 304, rate=[[0, 0.2], [1, 0.03], [2, -0.02], [0, -0.15]])
    add("rate_clamp", 40, 305, rate=[[1, 9.0]], lam=0.02)
    add("rate_conditional", 50, 306, rate=[[1, 0.05], [0, 0.2]], lam=1.0, conditional=True,
        target_dist=60)
    add("rate_symmetric", 50, 307, rate=[[1, 0.05]], sym=True,
        theta=[-1.5, 0.2], effects=[{"name": "density"}, {"name": "transTrip"}])
    add("rate_active", 60, 308, rate=[[2, 0.05], [0, -0.3]], comp=dict(nact=45, inactive=0))
    add("rate_wide", 130, 309, rate=[[1, 0.04], [0, 0.3]], lam=2.0, M=3)
    add("comp_restore", 60, 310, comp=dict(nact=48, inactive=12))
    add("comp_restore_rate", 70, 311, rate=[[1, 0.06]], comp=dict(nact=55, inactive=15))
    add("comp_conditional", 50, 312, lam=1.0, conditional=True, target_dist=50,
        comp=dict(nact=40, inactive=10))
    return S


def rate_problem(spec: dict, M: int | None = None):
    prob = spec_problem(spec, M)
    n = spec["n"]
    rng = np.random.default_rng(spec["seed"] + 500)
    if spec.get("rate"):
        kinds = [int(r[0]) for r in spec["rate"]]
        prob.rate_kind = np.array(kinds, np.int32)
        prob.rate_alpha = np.array([float(r[1]) for r in spec["rate"]])
        u = np.zeros((len(kinds), n))
        for q, kd in enumerate(kinds):
            if kd == 0:
                u[q] = rng.integers(1, 6, n) - 3.0
        prob.rate_u = u
    if spec.get("comp"):
        c = spec["comp"]
        perm = rng.permutation(n).astype(np.int32)
        prob.act = perm
        prob.nact = int(c["nact"])
        if c["inactive"]:
            prob.has_comp = 1
            inact = np.zeros(n, np.uint8)
            inact[perm[n - int(c["inactive"]):]] = 1
            prob.inact = inact
    return prob


def build_rate_cases(O) -> dict:
    cases = []
    for spec in rate_specs():
        prob = rate_problem(spec)
        r = O.simulate_reference(prob, want_end_nets=True)
        out = {k: np.asarray(v).tolist() for k, v in r.items() if k != "a_ends"}
        out["ends_packed"] = _pack_list(r["a_ends"])
        cases.append(dict(spec=spec, out=out))
    return {"cases": cases}
