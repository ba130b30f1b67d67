"""Synthetic period-simulator and statistics cases for golden fixtures.

``case_specs()`` lists JSON-serialisable problem specifications;
``spec_problem(spec)`` turns one into a SaomProblem. ``build_cases(O)`` runs
the reference engine compiled from its source (oracle/_ref) on every spec and
records its outputs; tests replay the same specs on the oracle and on the GPU.
"""
from __future__ import annotations

import numpy as np

from cusna_amd.abi import SaomProblem, pack_rows
from cusna_amd.data import build_behavior_data
from cusna_amd.effects import BehEffect, Effect
from cusna_amd.workloads import bernoulli_network, effect_problem

# every network effect id with a change statistic, with a moderate parameter
EFFECT_SWEEP = [
    ("recip", 0.8), ("transTrip", 0.2), ("cycle3", -0.3), ("inPop", 0.05), ("transTies", 0.3),
    ("transRecTrip", 0.2), ("inPopSqrt", 0.3), ("outPop", -0.05), ("outPopSqrt", -0.2),
    ("egoX", 0.2), ("altX", -0.2), ("simX", 0.7), ("sameX", 0.4), ("inAct", 0.05),
    ("inActSqrt", 0.2), ("outAct", -0.05), ("outActSqrt", -0.2), ("outTrunc", -0.3),
    ("reciAct", 0.1), ("higher", 0.3), ("egoXaltX", 0.2), ("X", 0.5), ("transTriads", 0.2),
    ("gwesp", 0.4), ("egoPlusAltX", 0.1), ("gwespFB", 0.3), ("gwespBF", 0.3), ("gwespBB", 0.3),
    ("gwdspFF", -0.1), ("gwdspFB", -0.1), ("altRThresholdX", 0.2),
]
COV_EFFECTS = {"egoX", "altX", "simX", "sameX", "higher", "egoXaltX", "egoPlusAltX",
               "altRThresholdX"}


def _net(n, deg, seed, sym=False):
    rng = np.random.default_rng(seed)
    a = bernoulli_network(n, deg, rng)
    if sym:
        a = np.triu(a, 1)
        a = a | a.T
    return a


def case_specs() -> list:
    specs = []
    rng = np.random.default_rng(20261203)
    # one case per effect id: [density, effect], n crosses a word boundary
    for q, (name, th) in enumerate(EFFECT_SWEEP):
        n = 70 if q % 2 == 0 else 45
        cov = rng.integers(1, 6, size=n).astype(float).tolist() if name in COV_EFFECTS else None
        eff = [{"name": "density"}, {"name": name, "cov": cov}]
        if name == "X":
            eff[1]["dyadic"] = rng.normal(size=(n, n)).round(3).tolist()
        if name == "altRThresholdX":
            eff[1]["parm"] = 3.0
        specs.append(dict(name=f"eff_{name}", n=n, M=4, seed=1000 + q, lam=4.0,
                          theta=[-1.6, th], effects=eff, deg=4.0, net_seed=q))
    # several effects at once, two-word rows
    n = 130
    cov = rng.integers(1, 6, size=n).astype(float).tolist()
    specs.append(dict(name="mix_c2_like", n=n, M=6, seed=77, lam=6.0,
                      theta=[-2.8, 1.8, 0.3, -0.1, 0.1, 0.1, 0.6, 0.02, -0.03],
                      effects=[{"name": "density"}, {"name": "recip"}, {"name": "transTrip"},
                               {"name": "cycle3"}, {"name": "egoX", "cov": cov},
                               {"name": "altX", "cov": cov}, {"name": "simX", "cov": cov},
                               {"name": "inPop"}, {"name": "outAct"}],
                      deg=5.0, net_seed=500))
    # interaction (3 components) and creation / endowment
    n = 60
    cov = rng.integers(1, 6, size=n).astype(float).tolist()
    specs.append(dict(name="interaction", n=n, M=4, seed=78, lam=5.0,
                      theta=[-2.0, 0.8, 0.2, 0.3, -0.1, 0.05],
                      effects=[{"name": "density"}, {"name": "recip"}, {"name": "transTrip"},
                               {"name": "egoX", "cov": cov},
                               {"name": "interaction", "comps": [1, 3]},
                               {"name": "interaction", "comps": [1, 2, 3]}],
                      deg=4.0, net_seed=501))
    specs.append(dict(name="creation_endow", n=n, M=4, seed=79, lam=5.0,
                      theta=[-2.0, 0.8, 0.3, -0.2],
                      effects=[{"name": "density"}, {"name": "recip"},
                               {"name": "egoPlusAltX", "cov": cov, "etype": "creation"},
                               {"name": "egoPlusAltX", "cov": cov, "etype": "endow"}],
                      deg=4.0, net_seed=502))
    # symmetric network, conditional stopping, maxdegree
    specs.append(dict(name="symmetric", n=50, M=4, seed=80, lam=4.0, theta=[-1.5, 0.2],
                      effects=[{"name": "density"}, {"name": "transTrip"}], deg=3.0,
                      net_seed=503, sym=True))
    specs.append(dict(name="conditional", n=50, M=5, seed=81, lam=1.0,
                      theta=[-2.2, 1.5, 0.3],
                      effects=[{"name": "density"}, {"name": "recip"}, {"name": "transTrip"}],
                      deg=3.0, net_seed=504, conditional=True, target_dist=90))
    specs.append(dict(name="maxdegree", n=50, M=4, seed=82, lam=6.0, theta=[-0.5, 0.4],
                      effects=[{"name": "density"}, {"name": "recip"}], deg=3.0,
                      net_seed=505, maxdegree=4))
    # co-evolution with every behaviour effect
    n = 60
    z = rng.binomial(4, 0.4, size=(n, 3)) + 1
    specs.append(dict(name="coevol_all_beh", n=n, M=4, seed=83, lam=4.0,
                      theta=[-2.0, 1.0, 0.2, 0.5],
                      effects=[{"name": "density"}, {"name": "recip"}, {"name": "transTrip"},
                               {"name": "simX", "dyn": True}],
                      deg=4.0, net_seed=506, z=z.tolist(),
                      beh_effects=["linear", "quad", "avAlt", "avSim", "totSim"],
                      thetab=[-0.2, -0.3, 0.8, 0.4, 0.1], lamb=3.0))
    return specs


def spec_problem(spec: dict, M: int | None = None) -> SaomProblem:
    n = spec["n"]
    x0 = _net(n, spec["deg"], spec["net_seed"], spec.get("sym", False))
    effects = []
    dyadics = []
    for e in spec["effects"]:
        cov = e.get("cov")
        if e["name"] == "X":
            cov = np.asarray(e["dyadic"])
            dyadics.append(cov)
        effects.append(Effect(e["name"], None if cov is None else np.asarray(cov),
                              parm=e.get("parm"), dyn=e.get("dyn", False),
                              etype=e.get("etype", "eval"), comps=e.get("comps")))
    beh = beh_effects = z0 = None
    if spec.get("z") is not None:
        beh = build_behavior_data(np.asarray(spec["z"], float), n, 3)
        z0 = beh.imputed[:, 0]
        beh_effects = [BehEffect(b) for b in spec["beh_effects"]]
    prob = effect_problem(x0, effects, spec["theta"], spec["lam"], M or spec["M"], spec["seed"],
                          beh=beh, z0=z0, beh_effects=beh_effects, thetab=spec.get("thetab"),
                          lamb=spec.get("lamb", 0.0), conditional=spec.get("conditional", False),
                          target_dist=spec.get("target_dist", 0))
    prob.symflag = 1 if spec.get("sym") else 0
    prob.maxdegree = spec.get("maxdegree", 0)
    return prob


def stats_specs() -> list:
    """Random networks for the statistic carriers (all statistic ids)."""
    out = []
    ids = [t for t in range(32) if t != 22]
    for q, (n, deg) in enumerate([(1, 0.0), (3, 1.0), (17, 3.0), (64, 5.0), (65, 6.0),
                                  (100, 8.0), (129, 4.0)]):
        out.append(dict(n=n, deg=deg, seed=900 + q, ids=ids))
    return out


def stats_inputs(spec: dict):
    n = spec["n"]
    rng = np.random.default_rng(spec["seed"])
    a = bernoulli_network(max(n, 2), spec["deg"], rng)[:n, :n] if n > 1 else np.zeros((1, 1),
                                                                                     np.uint8)
    k = len(spec["ids"])
    p1 = np.zeros(k)
    p2 = np.zeros(k)
    for q, t in enumerate(spec["ids"]):
        if t in (8, 9, 14, 15):
            p1[q] = [1.0, 0.0, -1.0][q % 3]
        elif t == 18:
            p1[q] = 3.0
        elif t == 12:
            p1[q], p2[q] = 0.25, 0.3
        elif t == 31:
            p1[q] = 0.5
    vmat = rng.normal(size=(k, n)).round(2)
    vmiss = (rng.random((k, n)) < 0.1).astype(np.uint8)
    dyn = np.zeros(k, np.int32)
    dyn[spec["ids"].index(11)] = 1
    vb = rng.normal(size=n).round(3)
    od0 = rng.integers(0, 5, n).astype(np.int32)
    id0 = rng.integers(0, 5, n).astype(np.int32)
    from cusna_amd.effects import gwesp_weights
    wtab = np.array([gwesp_weights(69.0, n) for _ in range(k)])
    return a, dict(type_ids=np.array(spec["ids"], np.int32), p1=p1, p2=p2, vmat=vmat,
                   vmiss=vmiss, dyn_flags=dyn, vb=vb, od0=od0, id0=id0, wtab=wtab, symflag=0)


def beh_stats_specs() -> list:
    return [dict(n=n, M=3, seed=950 + q) for q, n in enumerate([1, 5, 40, 70])]


def beh_stats_inputs(spec: dict):
    n, M = spec["n"], spec["M"]
    rng = np.random.default_rng(spec["seed"])
    z = rng.integers(1, 6, size=(M, n)).astype(np.int32)
    xnet = (rng.random((n, n)) < 0.1).astype(np.uint8)
    np.fill_diagonal(xnet, 0)
    bk = (rng.random(n) > 0.15).astype(np.int32)
    return z, np.arange(5, dtype=np.int32), xnet, bk, 2.8666666, 4.0, 0.61234


def _pack_list(a):
    return [pack_rows(m).tolist() for m in a]


def build_cases(O) -> dict:
    """Run the compiled reference engine on every spec."""
    R = O.reference()
    cases = []
    for spec in case_specs():
        prob = spec_problem(spec)
        r = O.simulate_reference(prob, want_end_nets=True)
        out = {k: np.asarray(v).tolist() for k, v in r.items() if k != "a_ends"}
        out["ends_packed"] = _pack_list(r["a_ends"])
        cases.append(dict(spec=spec, out=out))
    stats = []
    for spec in stats_specs():
        a, kw = stats_inputs(spec)
        ids = kw["type_ids"]
        res = []
        for lo in range(0, len(ids), 16):  # at most 16 effects per call
            sl = slice(lo, lo + 16)
            res += O.network_stats(a, ids[sl], kw["p1"][sl], kw["p2"][sl], kw["vmat"][sl],
                                   kw["vmiss"][sl], kw["dyn_flags"][sl], kw["vb"], kw["od0"],
                                   kw["id0"], kw["wtab"][sl], 0, lib=R,
                                   fn="cusna_network_stats").tolist()
        stats.append(dict(spec=spec, out=res))
    beh = []
    for spec in beh_stats_specs():
        z, bty, xnet, bk, zbar, br, bs = beh_stats_inputs(spec)
        beh.append(dict(spec=spec, out=O.behavior_stats(z, bty, xnet, bk, zbar, br, bs, lib=R,
                                                        fn="cusna_behavior_stats").tolist()))
    return dict(cases=cases, stats=stats, beh=beh)
