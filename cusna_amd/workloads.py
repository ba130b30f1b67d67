"""Synthetic SAOM workloads of the benchmark configurations.

Each builder returns the inputs of one period call (a ``SaomProblem``) from a
fixed seed, using the same marshalling as the R glue (float64 descriptors,
NULL masks, float32 theta/lambda):

  fit50_waves         50 actors, 3 waves: wave 1 Poisson(2.3) out-ties capped at
                   6, reciprocated with p = 0.4; waves 2-3 simulated.
  n500_problem       500 actors, directed, rate 8 per period, mean outdegree 5;
                   density -2.8, recip 1.8, transTrip 0.3, cycle3 -0.1,
                   egoX/altX/simX 0.1/0.1/0.6 on a covariate uniform on 1..5.
  mixed_group_problems      256 networks of 30..1200 actors, 500 chains each (groups).
  coevo300_problem       300 actors with an ordinal behaviour 1..5 = Binomial(4, 0.4)+1;
                   network density/recip/transTrip/simX(behaviour), behaviour
                   rate 4, linear -0.2, quad -0.3, avAlt 0.8.
  groups1000_estimation  16 groups of 1000 actors, two waves each, 12 effects
                   (density, recip, transTrip, cycle3, inPop, outAct, egoX/altX/
                   simX on an N(0, 1) and on a 1..5 covariate), true theta and
                   rates drawn from the seed.
"""
from __future__ import annotations

import numpy as np

from .abi import SaomProblem
from .data import build_saom_data
from .effects import BehEffect, Effect, prepare_effects

N500_THETA = [-2.8, 1.8, 0.3, -0.1, 0.1, 0.1, 0.6]
N500_EFFECTS = ["density", "recip", "transTrip", "cycle3", "egoX", "altX", "simX"]


def bernoulli_network(n: int, mean_outdeg: float, rng: np.random.Generator) -> np.ndarray:
    """Directed Bernoulli graph with P(tie) = mean_outdeg / (n - 1), no loops."""
    a = (rng.random((n, n)) < mean_outdeg / (n - 1)).astype(np.uint8)
    np.fill_diagonal(a, 0)
    return a


def poisson_capped_network(n: int, lam: float, cap: int, p_recip: float,
                           rng: np.random.Generator) -> np.ndarray:
    """Each actor sends min(Poisson(lam), cap) ties to uniform alters; each tie
    is reciprocated with probability p_recip."""
    a = np.zeros((n, n), dtype=np.uint8)
    for i in range(n):
        d = min(int(rng.poisson(lam)), cap)
        alters = rng.choice(np.delete(np.arange(n), i), size=d, replace=False)
        a[i, alters] = 1
    ii, jj = np.nonzero(a)
    back = rng.random(len(ii)) < p_recip
    a[jj[back], ii[back]] = 1
    np.fill_diagonal(a, 0)
    return a


def effect_problem(x0: np.ndarray, effects: list, theta, lam: float, M: int, seed: int,
                   beh=None, z0=None, beh_effects=None, thetab=None, lamb: float = 0.0,
                   conditional: bool = False, target_dist: int = 0) -> SaomProblem:
    """SaomProblem for a start network and an effect list (descriptors as the
    R layer prepares them for period 1)."""
    n = x0.shape[0]
    z_start = None if z0 is None else np.asarray(z0, float)
    d = prepare_effects(effects, n, 1, beh, z_start)
    prob = SaomProblem(
        x0=x0, theta=np.asarray(theta, float), lam=lam, type_ids=d.type_ids, M=M, seed=seed,
        etypes=d.etypes, p1=d.p1, p2=d.p2, vmat=d.vmat, vmiss=d.vmiss, dyn_flags=d.dyn_flags,
        dyc_idx=d.dyc_idx, wtab=d.wtab, int_comp=d.int_comp, vb=d.v_dyn,
        conditional=1 if conditional else 0, target_dist=target_dist)
    if beh_effects:
        prob.kb = len(beh_effects)
        prob.lamb = lamb
        prob.thetab = np.asarray(thetab, float)
        prob.bty = np.array([e.type_id for e in beh_effects], np.int32)
        prob.z0 = np.asarray(z0, np.int32)
        prob.bkdist = np.ones(n, np.int32)
        prob.zmin, prob.zmax = beh.zmin, beh.zmax
        prob.zbar, prob.brange, prob.bsimmean = beh.overall_mean, beh.range, beh.sim_mean
    return prob


FIT50_THETA = [-2.5, 2.0, 0.4]
FIT50_EFFECTS = ["density", "recip", "transTrip"]


def fit50_waves(simulate, n: int = 50, seed: int = 1) -> list:
    """Three waves of 50 actors: wave 1 with min(Poisson(2.3), 6) out-ties per
    actor, each reciprocated with p = 0.4; waves 2 and 3 are seeded simulated
    periods at rate 6 with density -2.5, recip 2.0, transTrip 0.4
    (``simulate`` runs a SaomProblem and returns a_ends)."""
    rng = np.random.default_rng(seed)
    waves = [poisson_capped_network(n, 2.3, 6, 0.4, rng)]
    for q in range(2):
        prob = effect_problem(waves[-1], [Effect(e) for e in FIT50_EFFECTS], FIT50_THETA, 6.0, 1,
                              seed * 1000 + q)
        waves.append(np.asarray(simulate(prob)[0], np.uint8))
    return waves


def fit50_estimation(simulate, n: int = 50, seed: int = 1):
    """(data, effects) of the 50-actor fit: conditional, density/recip/transTrip."""
    return build_saom_data(fit50_waves(simulate, n, seed)), [Effect(e) for e in FIT50_EFFECTS]


def n500_covariate(n: int = 500, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed + 1000)
    return rng.integers(1, 6, size=n).astype(float)


def n500_effects(x: np.ndarray) -> list:
    return [Effect("density"), Effect("recip"), Effect("transTrip"), Effect("cycle3"),
            Effect("egoX", x), Effect("altX", x), Effect("simX", x)]


def n500_problem(M: int = 4096, n: int = 500, seed: int = 2, x0: np.ndarray | None = None,
               period_seed: int = 11) -> SaomProblem:
    """The n = 500 period problem (seven effects, rate 8) from wave 1 (or from the given start network)."""
    rng = np.random.default_rng(seed)
    w1 = bernoulli_network(n, 5.0, rng)
    x = n500_covariate(n, seed)
    start = w1 if x0 is None else x0
    return effect_problem(start, n500_effects(x), N500_THETA, 8.0, M, period_seed)


def n500_waves(simulate, n: int = 500, seed: int = 2) -> list:
    """Waves 1 and 2 of the n = 500 workload; wave 2 is one seeded simulated period from
    wave 1 (``simulate`` runs a SaomProblem and returns a_ends)."""
    rng = np.random.default_rng(seed)
    w1 = bernoulli_network(n, 5.0, rng)
    prob = n500_problem(M=1, n=n, seed=seed, x0=w1, period_seed=seed + 7)
    w2 = simulate(prob)[0]
    return [w1, np.asarray(w2, np.uint8)]


def mixed_group_problems(groups: int = 256, chains: int = 500, seed: int = 3) -> list:
    """`groups` independent networks, n uniform on 30..1200, mean
    outdegree 4, density/recip/transTrip/inPop/egoX/altX on a Bernoulli(0.5)
    covariate, rate 4, `chains` chains each (one period)."""
    rng = np.random.default_rng(seed)
    out = []
    for g in range(groups):
        n = int(rng.integers(30, 1201))
        x0 = bernoulli_network(n, 4.0, rng)
        cov = (rng.random(n) < 0.5).astype(float)
        eff = [Effect("density"), Effect("recip"), Effect("transTrip"), Effect("inPop"),
               Effect("egoX", cov), Effect("altX", cov)]
        out.append(effect_problem(x0, eff, [-2.8, 1.8, 0.3, 0.02, 0.1, 0.1], 4.0, chains,
                                  seed * 100000 + g))
    return out


def coevo300_problem(M: int = 2048, n: int = 300, seed: int = 4, period_seed: int = 13) -> SaomProblem:
    """Network-behaviour co-evolution period problem at n = 300 (period 1 from wave 1)."""
    rng = np.random.default_rng(seed)
    w1 = bernoulli_network(n, 4.0, rng)
    z = rng.binomial(4, 0.4, size=(n, 3)) + 1
    w2 = bernoulli_network(n, 4.0, rng)
    d = build_saom_data([w1, w2, w2], behavior=z.astype(float))
    effects = [Effect("density"), Effect("recip"), Effect("transTrip"),
               Effect("simX", dyn=True)]
    beh_effects = [BehEffect("linear"), BehEffect("quad"), BehEffect("avAlt")]
    return effect_problem(w1, effects, [-2.6, 1.6, 0.3, 0.8], 6.0, M, period_seed, beh=d.beh,
                          z0=d.beh.imputed[:, 0], beh_effects=beh_effects,
                          thetab=[-0.2, -0.3, 0.8], lamb=4.0)


GROUPS1000_THETA = [-4.5, 2.0, 0.3, -0.2, 0.03, -0.03, 0.1, 0.1, 0.5, 0.05, 0.1, 0.4]


def groups1000_effects(z: np.ndarray, c: np.ndarray) -> list:
    return [Effect("density"), Effect("recip"), Effect("transTrip"), Effect("cycle3"),
            Effect("inPop"), Effect("outAct"), Effect("egoX", z), Effect("altX", z),
            Effect("simX", z), Effect("egoX", c), Effect("altX", c), Effect("simX", c)]


def groups1000_estimation(simulate, groups: int = 16, n: int = 1000, seed: int = 5):
    """(datas, effects per group, theta, rates): ``groups`` groups of n actors;
    wave 1 Bernoulli with mean outdegree 5, covariates N(0, 1) and uniform on
    1..5 per group; the true theta is GROUPS1000_THETA perturbed by 10 % of
    each magnitude and the rates uniform on 2.5..3.5, both from the seed;
    wave 2 of each group is one seeded simulated period at them (``simulate``
    runs a SaomProblem and returns a_ends)."""
    rng = np.random.default_rng(seed)
    base = np.asarray(GROUPS1000_THETA)
    theta = base + 0.1 * np.abs(base) * rng.standard_normal(base.size)
    rates = rng.uniform(2.5, 3.5, groups)
    datas, effs = [], []
    for g in range(groups):
        w1 = bernoulli_network(n, 5.0, rng)
        z = rng.standard_normal(n)
        c = rng.integers(1, 6, size=n).astype(float)
        eff = groups1000_effects(z, c)
        prob = effect_problem(w1, eff, theta, float(rates[g]), 1, seed * 100000 + g)
        w2 = np.asarray(simulate(prob)[0], np.uint8)
        datas.append(build_saom_data([w1, w2]))
        effs.append(eff)
    return datas, effs, theta, rates


def groups1000_problem(M: int = 4, n: int = 1000, seed: int = 5, rate: float = 3.0,
                       period_seed: int = 51) -> SaomProblem:
    """One period of one group of the 16 x 1000-actor panel's shape: wave 1
    Bernoulli with mean outdegree 5, the 12 effects of ``groups1000_effects``
    on an N(0, 1) and a 1..5 covariate, GROUPS1000_THETA, rate ``rate``."""
    rng = np.random.default_rng(seed)
    w1 = bernoulli_network(n, 5.0, rng)
    z = rng.standard_normal(n)
    c = rng.integers(1, 6, size=n).astype(float)
    return effect_problem(w1, groups1000_effects(z, c), GROUPS1000_THETA, rate, M, period_seed)
