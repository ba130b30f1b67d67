"""Method-of-moments (Robbins-Monro) estimator over the device library.

Host side of the reference's R estimator: ``.mom_estimator`` (cran/cusna
R/mom.R:24-88) builds the context here -- targets, parameter layout and names,
per-period problems, default start (R/mom.R:199-212) -- and the three phases of
``.mom_run`` (R/mom.R:218-269) run inside ``cusna_mom_estimate`` on the device.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from . import lib
from .data import SaomData
from .period import period_problem
from .targets import mom_targets


@dataclass
class MomControl:
    """Schedule (defaults of R/mom.R:26-28)."""

    firstg: float = 0.4
    n2: tuple = (40, 40, 60, 80)
    batch2: int = 256
    n1: int = 1000
    n3: int = 4000
    nD: int = 1000
    diagonalize: float = 0.2
    seed: int = 1234
    maxdegree: int = 0

    @property
    def nsub(self) -> int:
        return len(self.n2)


@dataclass
class MomFit:
    theta: np.ndarray
    se: np.ndarray
    tconv: np.ndarray
    tconv_max: float
    max_abs_tconv: float
    param_names: list
    targets: np.ndarray
    sim_means: np.ndarray
    D: np.ndarray
    msf: np.ndarray
    conditional: bool
    rate_est: np.ndarray | None
    rate_sd: np.ndarray | None
    n_sims: int
    repeats: int
    phase_seconds: list
    wall_time: float
    phase3_stats: np.ndarray = field(repr=False, default=None)
    phase3_sp: np.ndarray = field(repr=False, default=None)
    phase3_scp: np.ndarray = field(repr=False, default=None)


class MomEstimator:
    """Estimator context (R/mom.R:24-88); ``run`` performs the estimation."""

    def __init__(self, data: SaomData, effects: list, beh_effects: list | None = None,
                 conditional: bool = False, control: MomControl | None = None,
                 **target_carriers):
        self.data = data
        self.effects = list(effects)
        self.beh_effects = list(beh_effects) if beh_effects else []
        self.control = control or MomControl()
        self.coevol = len(self.beh_effects) > 0
        if self.coevol and conditional:
            # the co-evolution layout carries the network rates as parameters
            # (R/mom.R:39); a conditional run would pin their moments
            raise ValueError("co-evolution estimation is unconditional")
        self.conditional = bool(conditional)
        self.n, self.P, self.k = data.n, data.n_periods, len(self.effects)
        self.kb = len(self.beh_effects)
        P, k, kb = self.P, self.k, self.kb
        # target_carriers: optional net_stats / beh_stats callables of mom_targets
        mt = mom_targets(data, self.effects, self.beh_effects or None, self.conditional,
                         **target_carriers)
        self.targets = mt.target_vector
        eff_names = [e.name for e in self.effects]
        if self.coevol:
            self.mode = 2
            self.p = 2 * P + k + kb
            self.param_names = ([f"rate net (p{q})" for q in range(1, P + 1)]
                                + [f"rate beh (p{q})" for q in range(1, P + 1)] + eff_names
                                + [f"beh {b.name}" for b in self.beh_effects])
        elif self.conditional:
            self.mode = 1
            self.p = k
            self.param_names = eff_names
        else:
            self.mode = 0
            self.p = P + k
            self.param_names = [f"rate (period {q})" for q in range(1, P + 1)] + eff_names
        # no composition change on this path: the rate constants are zero (mom.R:59-61)
        self.rate_consts = np.zeros(P)
        self.problems = []
        self.xnets = [] if self.coevol else None
        for q in range(1, P + 1):
            pr = period_problem(data, self.effects, q, np.zeros(k), 1.0, 1, 0,
                                beh_effects=self.beh_effects or None,
                                thetab=np.zeros(kb) if kb else None, lamb=1.0,
                                conditional=self.conditional,
                                maxdegree=self.control.maxdegree)
            self.problems.append(pr)
            if self.coevol:
                self.xnets.append(data.network_less_missing_start(q))

    def default_theta0(self) -> np.ndarray:
        """Start values (R/mom.R:199-212)."""
        ev0 = np.array([-1.0 if e.name == "density" else 0.0 for e in self.effects])
        if self.conditional:
            return ev0
        lam0 = 1.5 * self.targets[:self.P] / self.n + 1.0
        if self.coevol:
            lamb0 = 1.5 * self.targets[self.P:2 * self.P] / self.n + 1.0
            return np.concatenate([lam0, lamb0, ev0, np.zeros(self.kb)])
        return np.concatenate([lam0, ev0])

    def simulate(self, theta, M: int, rng: np.random.Generator | None = None) -> dict:
        """All periods for M chains at theta (``.mom_simulate``, R/mom.R:133-183),
        through the device period simulator: per period one
        cusna_saom_simulate_period call, its seed drawn from ``rng`` (numpy's
        global stream when None, as R's ``sample.int`` draws from R's stream),
        and the moment rows of the estimator's layout -- ``sp``, the period's
        statistics with its rate statistic (distance plus the rate constant) in
        the period's rate column when unconditional, and ``scp``, the scores.
        Returns stats (M x p, summed over periods), pieces [{sp, scp}] per
        period and ntim (M x P)."""
        if self.coevol:
            raise NotImplementedError("simulate: network-only estimators (the co-evolution "
                                      "layout runs inside cusna_mom_estimate)")
        theta = np.asarray(theta, dtype=float)
        if theta.size != self.p:
            raise ValueError(f"theta has {theta.size} values, the estimator {self.p}")
        P, p, k = self.P, self.p, self.k
        if self.conditional:
            lam, th, off = np.ones(P), theta, 0
        else:
            lam, th, off = theta[:P], theta[P:], P
        stats = np.zeros((M, p))
        ntim = np.zeros((M, P))
        pieces = []
        for q, pr in enumerate(self.problems):
            # sample.int(.Machine$integer.max, 1): uniform on 1 .. 2^31 - 1
            seed = int(rng.integers(1, 2**31)) if rng is not None else \
                int(np.random.randint(1, 2**31))
            r = lib.simulate_period(replace(pr, theta=th.copy(), lam=float(lam[q]), M=int(M),
                                            seed=seed))
            sp = np.zeros((M, p))
            scp = np.zeros((M, p))
            if not self.conditional:
                sp[:, q] = r["dist_stat"] + self.rate_consts[q]
                scp[:, q] = r["scores"][:, 0]
            sp[:, off:off + k] = r["stats"]
            scp[:, off:off + k] = r["scores"][:, 1:k + 1]
            ntim[:, q] = r["ntim"]
            stats += sp
            pieces.append({"sp": sp, "scp": scp, "seed": seed})
        return {"stats": stats, "pieces": pieces, "ntim": ntim}

    def run(self, theta0=None, use_graph: bool = True, max_repeats: int = 0,
            conv_tol: float = 0.1, surrogate: bool = False, surrogate_pairs_log2: int = 0,
            surrogate_steps: int = 0, surrogate_rounds: int = 0,
            surrogate_corrections: int = -1, specialize: int = 0,
            verbose: bool = False, surrogate_warmup: int = 0) -> MomFit:
        """Estimate. ``surrogate`` replaces the Robbins-Monro sub-phases by the
        perceptron surrogate (see cusna_mom_options in include/libcusna.h)."""
        c = self.control
        th0 = self.default_theta0() if theta0 is None else np.asarray(theta0, float)
        r = lib.mom_estimate(self.problems, self.targets, th0, self.mode, c.n1, c.nD, c.batch2,
                             c.n3, list(c.n2), c.firstg, c.diagonalize, c.seed,
                             xnets=self.xnets, rate_consts=self.rate_consts,
                             use_graph=use_graph, max_repeats=max_repeats, conv_tol=conv_tol,
                             surrogate=surrogate, surrogate_pairs_log2=surrogate_pairs_log2,
                             surrogate_steps=surrogate_steps, surrogate_rounds=surrogate_rounds,
                             surrogate_corrections=surrogate_corrections,
                             specialize=specialize, verbose=verbose,
                             surrogate_warmup=surrogate_warmup)
        return MomFit(theta=r["theta"], se=r["se"], tconv=r["tconv"],
                      tconv_max=r["tconv_max"], max_abs_tconv=r["max_abs_tconv"],
                      param_names=self.param_names, targets=self.targets,
                      sim_means=r["sim_means"], D=r["D"], msf=r["msf"],
                      conditional=self.conditional, rate_est=r.get("rate_est"),
                      rate_sd=r.get("rate_sd"), n_sims=r["n_sims"], repeats=r["repeats"],
                      phase_seconds=r["phase_seconds"], wall_time=r["wall_seconds"],
                      phase3_stats=r["phase3_stats"], phase3_sp=r["phase3_sp"],
                      phase3_scp=r["phase3_scp"])


def mom_estimate(data: SaomData, effects: list, beh_effects: list | None = None,
                 conditional: bool = False, control: MomControl | None = None,
                 theta0=None, **kw) -> MomFit:
    """One-call estimate (the reference's mom_estimate, R/estimate.R)."""
    return MomEstimator(data, effects, beh_effects, conditional, control).run(theta0, **kw)


class GroupMomEstimator:
    """Several independent groups (networks over different actor sets, each
    with its own waves) fitted with one parameter vector: the groups' periods
    are the periods of one cusna_mom_estimate call, each with its own start
    network and rate parameter, and the effect targets are the sums over the
    groups, as for the periods of one group (R/mom.R:152-166 sums the
    statistics over periods). The reference's R layer fits one group; this
    composes that per-period algebra over groups. Unconditional only."""

    def __init__(self, datas: list, effects_per_group: list, control: MomControl | None = None,
                 **target_carriers):
        if len(datas) != len(effects_per_group) or not datas:
            raise ValueError("one effect list per group")
        self.control = control or MomControl()
        self.groups = [MomEstimator(d, e, None, False, self.control, **target_carriers)
                       for d, e in zip(datas, effects_per_group)]
        k = self.groups[0].k
        if any(g.k != k for g in self.groups):
            raise ValueError("every group needs the same number of effects")
        self.k = k
        self.P = sum(g.P for g in self.groups)
        if self.P > 16:
            raise ValueError("at most 16 periods over all groups")
        self.p = self.P + k
        self.problems = [pr for g in self.groups for pr in g.problems]
        rates = np.concatenate([g.targets[:g.P] for g in self.groups])
        eff = np.sum([g.targets[g.P:] for g in self.groups], axis=0)
        self.targets = np.concatenate([rates, eff])
        g0 = self.groups[0]
        self.param_names = ([f"rate (group {gi + 1}, period {q})"
                             for gi, g in enumerate(self.groups) for q in range(1, g.P + 1)]
                            + g0.param_names[g0.P:])

    def default_theta0(self) -> np.ndarray:
        """R/mom.R:199-212 per group: rates 1.5 x target / n + 1, density -1."""
        lam0 = np.concatenate([1.5 * g.targets[:g.P] / g.n + 1.0 for g in self.groups])
        return np.concatenate([lam0, self.groups[0].default_theta0()[self.groups[0].P:]])

    def run(self, theta0=None, use_graph: bool = True, max_repeats: int = 0,
            conv_tol: float = 0.1, verbose: bool = False, **surrogate_kw) -> MomFit:
        c = self.control
        th0 = self.default_theta0() if theta0 is None else np.asarray(theta0, float)
        r = lib.mom_estimate(self.problems, self.targets, th0, 0, c.n1, c.nD, c.batch2, c.n3,
                             list(c.n2), c.firstg, c.diagonalize, c.seed,
                             rate_consts=np.zeros(self.P), use_graph=use_graph,
                             max_repeats=max_repeats, conv_tol=conv_tol, verbose=verbose,
                             **surrogate_kw)
        return MomFit(theta=r["theta"], se=r["se"], tconv=r["tconv"],
                      tconv_max=r["tconv_max"], max_abs_tconv=r["max_abs_tconv"],
                      param_names=self.param_names, targets=self.targets,
                      sim_means=r["sim_means"], D=r["D"], msf=r["msf"], conditional=False,
                      rate_est=None, rate_sd=None, n_sims=r["n_sims"], repeats=r["repeats"],
                      phase_seconds=r["phase_seconds"], wall_time=r["wall_seconds"],
                      phase3_stats=r["phase3_stats"], phase3_sp=r["phase3_sp"],
                      phase3_scp=r["phase3_scp"])
