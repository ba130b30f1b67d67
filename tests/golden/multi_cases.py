"""Synthetic multi-network period cases for golden fixtures.

``multi_specs()`` lists JSON-serialisable specifications; ``multi_problem(spec)``
turns one into a MultiProblem (D networks over one actor set, per-network
effect blocks, cross-network effects). ``build_multi_cases(O)`` runs the
reference's multi-network engine compiled from its source (oracle/_ref) on
every spec and records its outputs; tests replay the same specs on the oracle
and on the GPU.
"""
from __future__ import annotations

import numpy as np

from cusna_amd.abi import MultiProblem, pack_rows
from cusna_amd.effects import Effect, effect_netrefs, prepare_effects
from cusna_amd.workloads import bernoulli_network

COV = {"egoX", "altX", "simX", "sameX", "higher", "egoXaltX", "egoPlusAltX", "altRThresholdX"}


def _net(n, deg, seed, sym=False):
    a = bernoulli_network(n, deg, np.random.default_rng(seed))
    if sym:
        a = np.triu(a, 1)
        a = a | a.T
    return a


def multi_specs() -> list:
    S = []
    S.append(dict(name="two_basic", n=30, deg=[2.0, 2.5], lam=[4.0, 6.0], M=8, seed=11,
                  eff=[[("density", -1.5), ("recip", 1.2)], [("density", -1.7), ("recip", 1.5)]]))
    S.append(dict(name="cross_all", n=40, deg=[2.0, 3.0], lam=[5.0, 4.0], M=8, seed=12,
                  eff=[[("density", -1.6), ("recip", 1.0), ("crprod", 1.2, 1),
                        ("crprodRecip", 0.5, 1), ("crprodMutual", 0.4, 1)],
                       [("density", -1.8), ("transTrip", 0.2), ("from", 0.3, 0, 1.0),
                        ("to", 0.2, 0)]]))
    S.append(dict(name="from_sqrt", n=36, deg=[3.0, 3.0], lam=[4.0, 4.0], M=6, seed=13,
                  eff=[[("density", -1.5), ("from", 0.4, 1, 2.0)],
                       [("density", -1.5), ("crprod", 0.8, 0), ("to", 0.1, 0)]]))
    S.append(dict(name="four_nets", n=25, deg=[2.0, 1.5, 2.5, 1.0], lam=[3.0, 2.0, 4.0, 1.5],
                  M=8, seed=14,
                  eff=[[("density", -1.5), ("recip", 1.0), ("crprod", 0.6, 3)],
                       [("density", -1.2), ("inPop", 0.05), ("outAct", -0.05)],
                       [("density", -1.4), ("cycle3", -0.2), ("crprodRecip", 0.4, 0)],
                       [("density", -2.0), ("to", 0.3, 2)]]))
    S.append(dict(name="structural_sweep", n=45, deg=[2.5, 2.0], lam=[5.0, 5.0], M=6, seed=15,
                  eff=[[("density", -1.6), ("recip", 0.9), ("transTies", 0.3),
                        ("transRecTrip", 0.2), ("inPopSqrt", 0.2), ("outPop", -0.05),
                        ("outPopSqrt", -0.1), ("inAct", 0.04), ("inActSqrt", 0.1),
                        ("outActSqrt", -0.2), ("outTrunc", -0.2), ("reciAct", 0.1),
                        ("transTriads", 0.2)],
                       [("density", -1.4), ("gwesp", 0.4), ("gwespFB", 0.2), ("gwespBF", 0.2),
                        ("gwespBB", 0.2), ("gwdspFF", -0.1), ("gwdspFB", -0.1),
                        ("crprodMutual", 0.3, 0)]]))
    S.append(dict(name="covariates", n=33, deg=[2.0, 2.0, 2.0], lam=[4.0, 3.0, 5.0], M=6,
                  seed=16,
                  eff=[[("density", -1.5), ("egoX", 0.2), ("altX", -0.2), ("simX", 0.6)],
                       [("density", -1.5), ("sameX", 0.4), ("higher", 0.2), ("egoXaltX", 0.1)],
                       [("density", -1.5), ("egoPlusAltX", 0.1), ("altRThresholdX", 0.2),
                        ("X", 0.5), ("crprod", 0.5, 0)]]))
    S.append(dict(name="symmetric_and_masks", n=30, deg=[2.0, 2.0], lam=[4.0, 4.0], M=8, seed=17,
                  sym=[1, 0], perm=True, act=24,
                  eff=[[("density", -1.2), ("transTrip", 0.1), ("crprod", 0.8, 1)],
                       [("density", -1.6), ("recip", 1.2), ("crprodRecip", 0.5, 0)]]))
    S.append(dict(name="creation_endowment", n=28, deg=[2.0, 2.5], lam=[4.0, 4.0], M=8, seed=18,
                  eff=[[("density", -1.5), ("recip", 1.0), ("egoPlusAltX", 0.3, None, None,
                                                            "creation")],
                       [("density", -1.5), ("crprod", 0.7, 0, None, "endow"),
                        ("to", 0.2, 0, None, "creation")]]))
    S.append(dict(name="word_boundary", n=70, deg=[2.5, 2.0, 3.0], lam=[3.0, 5.0, 2.0], M=4,
                  seed=19, per_chain=True,
                  eff=[[("density", -1.8), ("recip", 1.2), ("from", 0.2, 2, 1.0)],
                       [("density", -1.6), ("transTrip", 0.15), ("to", 0.1, 0)],
                       [("density", -2.0), ("crprod", 1.0, 1), ("cycle3", -0.1)]]))
    S.append(dict(name="big_rate", n=130, deg=[3.0, 3.0], lam=[10.0, 8.0], M=2, seed=20,
                  eff=[[("density", -2.2), ("recip", 1.5), ("transTrip", 0.2), ("crprod", 0.8, 1)],
                       [("density", -2.0), ("recip", 1.3), ("to", 0.1, 0)]]))
    return S


def _effect(spec_e, rng, n):
    name, theta = spec_e[0], spec_e[1]
    net_ref = spec_e[2] if len(spec_e) > 2 else None
    parm = spec_e[3] if len(spec_e) > 3 else None
    etype = spec_e[4] if len(spec_e) > 4 else "eval"
    cov = None
    if name in COV:
        cov = rng.integers(1, 6, n).astype(float)
    elif name == "X":
        cov = (rng.random((n, n)) < 0.3).astype(float)
    kw = {} if parm is None else {"parm": parm}
    e = Effect(name, cov, net_ref=net_ref, **kw)
    e.etype = etype
    return e, theta


def multi_problem(spec: dict, M: int | None = None) -> MultiProblem:
    n, D = spec["n"], len(spec["deg"])
    M = int(M or spec["M"])
    rng = np.random.default_rng(spec["seed"] + 1000)
    sym = spec.get("sym", [0] * D)
    x0 = np.stack([_net(n, spec["deg"][d], spec["seed"] * 10 + d, bool(sym[d]))
                   for d in range(D)]).astype(np.uint8)
    effs, th, netof = [], [], []
    for d in range(D):
        for se in spec["eff"][d]:
            e, t = _effect(se, rng, n)
            effs.append(e)
            th.append(t)
            netof.append(d)
    # descriptors as prepare_effects builds them for one network (etypes set
    # directly: the multi-network engine applies creation/endowment to any id)
    etypes = np.array([{"eval": 0, "creation": 1, "endow": 2}[e.etype] for e in effs], np.int32)
    for e in effs:
        e.etype = "eval"
    desc = prepare_effects(effs, n, 1)
    dyc = np.stack(desc.dymats) if desc.dymats else None
    theta = np.array(th, float)
    lam = np.array(spec["lam"], float)
    if spec.get("per_chain"):
        theta = theta[None, :] * (1.0 + 0.05 * rng.standard_normal((M, len(th))))
        lam = lam[None, :] * (1.0 + 0.1 * rng.random((M, D)))
    perm = None
    if spec.get("perm"):
        perm = (rng.random((D, n, n)) > 0.15).astype(np.uint8)
        for d in range(D):
            np.fill_diagonal(perm[d], 0)
            if sym[d]:
                perm[d] = np.triu(perm[d], 1)
                perm[d] = perm[d] | perm[d].T
    act, nact = None, None
    if spec.get("act"):
        act = rng.permutation(n).astype(np.int32)
        nact = int(spec["act"])
    return MultiProblem(
        x0=x0, lam=lam, theta=theta, type_ids=desc.type_ids, netof=np.array(netof, np.int32),
        M=M, seed=int(spec["seed"]) * 7919 + 5, netref=effect_netrefs(effs),
        etypes=etypes, p1=desc.p1, p2=desc.p2, vmat=desc.vmat, wtab=desc.wtab, dyc=dyc,
        dyc_idx=desc.dyc_idx, perm=perm, symf=np.array(sym, np.int32), act=act, nact=nact)


def build_multi_cases(O) -> dict:
    """Reference outputs (oracle/_ref) for every multi spec; end networks
    packed to bit rows (M, D, n, W) uint32."""
    out = []
    for spec in multi_specs():
        pr = multi_problem(spec)
        r = O.simulate_multi_reference(pr)
        packed = np.stack([[pack_rows(r["a_ends"][c, d]) for d in range(pr.D)]
                           for c in range(pr.M)])
        out.append(dict(spec=spec, steps=r["steps"].tolist(), ntim=r["ntim"].tolist(),
                        rscores=r["rscores"].astype(float).tolist(),
                        scores=r["scores"].astype(float).tolist(),
                        ends=packed.astype(np.int64).tolist()))
    return {"cases": out}


def cross_inputs(seed: int, n: int = 37, M: int = 3):
    """Random end networks (M, n, n) and a predictor network (n, n)."""
    rng = np.random.default_rng(seed)
    a = (rng.random((M, n, n)) < 0.12).astype(np.uint8)
    wp = (rng.random((n, n)) < 0.15).astype(np.uint8)
    for c in range(M):
        np.fill_diagonal(a[c], 0)
    np.fill_diagonal(wp, 0)
    return a, wp


CROSS_SPECS = [(33, 1.0), (34, 1.0), (35, 1.0), (36, 1.0), (36, 2.0), (37, 1.0)]


def build_cross_cases(O) -> dict:
    """Reference cusna_cross_stats outputs on random inputs."""
    R = O.reference()
    out = []
    for q, (t, parm) in enumerate(CROSS_SPECS):
        a, wp = cross_inputs(100 + q)
        v = O.cross_stats(a, wp, t, parm, lib=R, fn="cusna_cross_stats")
        out.append(dict(seed=100 + q, t=t, parm=parm, out=v.tolist()))
    return {"cases": out}
