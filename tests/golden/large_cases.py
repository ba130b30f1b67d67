"""Period cases at the benchmark configurations' own sizes, for golden fixtures.

``large_specs()`` names each case by the workload builder that makes its
inputs (cusna_amd/workloads.py, fixed seeds); ``large_problem(spec)`` rebuilds
the SaomProblem; ``build_large_cases(O)`` runs the reference engine compiled
from its source (oracle/_ref) on every case and records its outputs. End
networks are recorded as the SHA-256 of the dense uint8 matrix of each chain
(the packed bits of n = 500..1200 networks would make the fixture large).

  c2_n500       500 actors, seven effects, rate 8, 8 chains (the n = 500
                workload, period 1 from wave 1)
  c4_n300       network-behaviour co-evolution, 300 actors, avAlt, 4 chains
  c5_n1000      1000 actors, the twelve effects of the 16-group panel, rate 3,
                4 chains
  c3_g3_n1045   group 3 of the uneven multi-group batch (17 words per row)
  c3_g6_n1118   group 6 (18 words per row)
  c3_g12_n1188  group 12 (19 words per row)
"""
from __future__ import annotations

import hashlib

import numpy as np


def large_specs() -> list:
    return [
        dict(name="c2_n500", builder="n500_problem", kw=dict(M=8)),
        dict(name="c4_n300", builder="coevo300_problem", kw=dict(M=4)),
        dict(name="c5_n1000", builder="groups1000_problem", kw=dict(M=4)),
        dict(name="c3_g3_n1045", builder="mixed_group", kw=dict(group=3, M=3)),
        dict(name="c3_g6_n1118", builder="mixed_group", kw=dict(group=6, M=3)),
        dict(name="c3_g12_n1188", builder="mixed_group", kw=dict(group=12, M=2)),
    ]


def large_problem(spec: dict):
    from cusna_amd import workloads as W

    kw = dict(spec["kw"])
    if spec["builder"] == "mixed_group":
        g = kw.pop("group")
        return W.mixed_group_problems(groups=g + 1, chains=kw["M"])[g]
    return getattr(W, spec["builder"])(**kw)


def end_hashes(a_ends) -> list:
    """SHA-256 of each chain's dense uint8 end network (C order)."""
    a = np.ascontiguousarray(np.asarray(a_ends) != 0, np.uint8)
    return [hashlib.sha256(a[c].tobytes()).hexdigest() for c in range(a.shape[0])]


def build_large_cases(O) -> dict:
    cases = []
    for spec in large_specs():
        prob = large_problem(spec)
        r = O.simulate_reference(prob, want_end_nets=True)
        out = {k: np.asarray(v).tolist() for k, v in r.items() if k != "a_ends"}
        out["ends_sha256"] = end_hashes(r["a_ends"])
        cases.append(dict(spec=spec, n=prob.n, k=prob.k, out=out))
    return dict(cases=cases)
