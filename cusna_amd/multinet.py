"""Multi-network co-evolution: data container and method-of-moments estimator.

Restates the reference's R layer for several dependent networks over one actor
set: ``saom_multinet_data`` (cran/cusna R/multinet.R:8-58),
``mom_estimate_multinet`` (R/multinet.R:60-98) and the estimator of
R/mom.R:271-480 (``.batch_net_stats``, ``.multi_mom_estimator``,
``.multi_default_theta0``, ``.multi_guard_step``, ``.multi_mom_simulate``,
``.multi_mom_run``). Per-variable parameter blocks [rate of network d in
period m x P, effects of network d ...] for d = 0..D-1; unconditional only.

The simulations and every statistic run on the device: one multi-network plan
per period and batch size (cusna_multi_plan_*) simulates the D networks and
returns, per chain, the period distance of each network over its distance
dyads, the rate scores, the effect scores and the statistics of each masked
end network (network statistics and cross-network statistics against the
period-start predictors). The host keeps the Robbins-Monro algebra on p x p
matrices (R/mom.R:15-20, 173-187).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import lib
from .abi import MultiProblem
from .data import SaomData, build_saom_data
from .effects import CROSS_IDS, Effect, effect_netrefs, prepare_effects
from .mom import MomControl, MomFit


@dataclass
class MultiNetData:
    """Several networks over one actor set (R/multinet.R:8-22)."""

    nets: list
    D: int
    n: int
    n_periods: int


def saom_multinet_data(waves_list) -> MultiNetData:
    """R/multinet.R:35-50: >= 2 networks, each a list of >= 2 waves, all with
    the same number of waves and the same actor set."""
    if not isinstance(waves_list, (list, tuple)) or len(waves_list) < 2:
        raise ValueError("`waves_list` must be a list of >= 2 networks, each a list of wave "
                         "matrices.")
    n_waves = len(waves_list[0])
    for net in waves_list:
        if not isinstance(net, (list, tuple)) or len(net) < 2:
            raise ValueError("each network must be a list of >= 2 wave matrices.")
        if len(net) != n_waves:
            raise ValueError("all networks must have the same number of waves.")
    nets = [build_saom_data(list(w)) for w in waves_list]
    n0, p0 = nets[0].n, nets[0].n_periods
    for d, nd in enumerate(nets):
        if nd.n != n0:
            raise ValueError(f"network {d + 1}: actor set size differs")
        if nd.n_periods != p0:
            raise ValueError(f"network {d + 1}: wave count differs")
    return MultiNetData(nets, len(nets), n0, p0)


# ---------------------------------------------------------------- algebra
def _pinv(m: np.ndarray) -> np.ndarray:
    """SVD pseudo-inverse with R's tolerance (R/mom.R:15-19)."""
    u, s, vt = np.linalg.svd(m)
    keep = s > max(np.sqrt(np.finfo(float).eps) * s[0], 0.0)
    return (vt[keep].T / s[keep]) @ u[:, keep].T


def _safe_solve(m: np.ndarray) -> np.ndarray:
    """solve(), else the pseudo-inverse (R/mom.R:20)."""
    try:
        if np.linalg.cond(m) > 1.0 / np.finfo(float).eps:
            raise np.linalg.LinAlgError
        return np.linalg.inv(m)
    except np.linalg.LinAlgError:
        return _pinv(m)


def _derivative(pieces) -> np.ndarray:
    """Sum over periods of cov(sp, scp) with divisor M (R/mom.R:173-182)."""
    p = pieces[0][0].shape[1]
    D = np.zeros((p, p))
    for sp, scp in pieces:
        M = sp.shape[0]
        D += (sp - sp.mean(0)).T @ (scp - scp.mean(0)) / M
    return D


def _dinv(D: np.ndarray, diagonalize: float) -> np.ndarray:
    mix = (1 - diagonalize) * D + diagonalize * np.diag(np.diag(D))
    return _safe_solve(mix)


def _splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


class MultiMomEstimator:
    """Estimator context (R/mom.R:310-361); ``run`` estimates (:411-480)."""

    def __init__(self, mn: MultiNetData, effects_by_net: list, control: MomControl | None = None,
                 net_stats=None, cross_stats=None):
        # statistic carriers for the targets: the device library by default
        self._net_stats_fn = net_stats or lib.network_stats
        self._cross_stats_fn = cross_stats or lib.cross_stats
        if len(effects_by_net) != mn.D:
            raise ValueError(f"`effects_by_net` must be a list with one effects list per network "
                             f"({mn.D}).")
        self.mn = mn
        self.control = control or MomControl()
        self.D, self.P, self.n = mn.D, mn.n_periods, mn.n
        self.specs_by_net = [list(e) for e in effects_by_net]
        self.effects = [e for effs in self.specs_by_net for e in effs]
        for e in self.effects:
            if e.type_id == 32:
                raise ValueError("interactions are not supported in the multi-network model")
            if e.type_id in CROSS_IDS and not 0 <= int(e.net_ref) < self.D:
                raise ValueError(f"{e.name}: net_ref must name one of the {self.D} networks")
        self.netof = np.array([d for d, effs in enumerate(self.specs_by_net) for _ in effs],
                              np.int32)
        self.k = len(self.effects)
        # per-variable layout (:318-329)
        self.rate_pos = np.zeros((self.D, self.P), int)
        self.eff_pos = np.zeros(self.k, int)
        names, pos, ei = [], 0, 0
        for d in range(self.D):
            for m in range(self.P):
                self.rate_pos[d, m] = pos
                pos += 1
                names.append(f"rate net{d} (p{m + 1})")
            for e in self.specs_by_net[d]:
                self.eff_pos[ei] = pos
                ei += 1
                pos += 1
                nr = "" if e.net_ref is None else f"<-net{e.net_ref}"
                names.append(f"net{d} {e.name}{nr}")
        self.p = pos
        self.param_names = names
        nets = mn.nets
        self.symf = np.array([1 if nd.symmetric else 0 for nd in nets], np.int32)
        self.rate_targets = np.array([[float(nd.rate_target(m)) for m in range(1, self.P + 1)]
                                      for nd in nets])
        self.rate_consts = np.array([[float(self._rate_const(nd, m))
                                      for m in range(1, self.P + 1)] for nd in nets])
        # period-start cross predictors, less missing at both ends (:337-338)
        self.cross_pred = [[nd.network_less_missing_start(m) for nd in nets]
                           for m in range(1, self.P + 1)]
        self.targets = self._targets()
        self._plans: dict = {}
        self._calls = 0

    @staticmethod
    def _rate_const(nd: SaomData, p: int) -> int:
        """Ties observed at the start and missing at the end (R/data.R:287-292)."""
        cc = ((nd.imputed[p - 1] != 0) & ~nd.miss[p - 1] & nd.miss[p] & ~nd.struct[p - 1]
              & ~nd.struct[p])
        np.fill_diagonal(cc, False)
        return int(cc.sum())

    # ------------------------------------------------------------- targets
    def _stat_inputs(self, d: int, m: int) -> dict:
        """Descriptors of network d's statistics in period m (:280-308)."""
        nd = self.mn.nets[d]
        eff = self.specs_by_net[d]
        desc = prepare_effects(eff, self.n, m)
        a_start = nd.imputed[m - 1]
        keep, force = nd.stat_keep_force(m)
        cross_wp = np.zeros((len(eff), self.n, self.n), np.uint8)
        for q, e in enumerate(eff):
            if e.type_id in CROSS_IDS:
                cross_wp[q] = self.cross_pred[m - 1][int(e.net_ref)]
        return dict(type_ids=desc.type_ids, p1=desc.p1, p2=desc.p2, vmat=desc.vmat,
                    vmiss=desc.vmiss, wtab=desc.wtab,
                    od0=(a_start != 0).sum(1).astype(np.int32),
                    id0=(a_start != 0).sum(0).astype(np.int32), symflag=int(self.symf[d]),
                    keep=keep, force=force, cross_wp=cross_wp)

    def _net_stats(self, a_masked: np.ndarray, d: int, m: int) -> np.ndarray:
        """Statistics of one masked network through the device carriers
        (cusna_network_stats, cusna_cross_stats), as .batch_net_stats."""
        s = self._stat_inputs(d, m)
        eff = self.specs_by_net[d]
        out = np.zeros(len(eff))
        nc = [q for q, e in enumerate(eff) if e.type_id not in CROSS_IDS]
        if nc:
            out[nc] = self._net_stats_fn(
                a_masked, type_ids=s["type_ids"][nc], p1=s["p1"][nc], p2=s["p2"][nc],
                vmat=s["vmat"][nc], vmiss=s["vmiss"][nc], dyn_flags=np.zeros(len(nc), np.int32),
                vb=np.zeros(self.n), od0=s["od0"], id0=s["id0"], wtab=s["wtab"][nc],
                symflag=s["symflag"])
        for q, e in enumerate(eff):
            if e.type_id in CROSS_IDS:
                out[q] = self._cross_stats_fn(a_masked[None], s["cross_wp"][q], e.type_id,
                                              s["p1"][q])[0]
        return out

    def _targets(self) -> np.ndarray:
        t = np.zeros(self.p)
        for d in range(self.D):
            for m in range(self.P):
                t[self.rate_pos[d, m]] = self.rate_targets[d, m]
        ei = 0
        for d in range(self.D):
            eff = self.specs_by_net[d]
            for m in range(1, self.P + 1):
                tg = self._net_stats(self.mn.nets[d].target_network(m), d, m)
                for q in range(len(eff)):
                    t[self.eff_pos[ei + q]] += tg[q]
            ei += len(eff)
        return t

    # ----------------------------------------------------------- schedule
    def default_theta0(self) -> np.ndarray:
        """R/mom.R:363-373."""
        th = np.zeros(self.p)
        for d in range(self.D):
            for m in range(self.P):
                th[self.rate_pos[d, m]] = 1.5 * self.rate_targets[d, m] / self.n + 1.0
        for ei, e in enumerate(self.effects):
            if e.name == "density":
                th[self.eff_pos[ei]] = -1.0
        return th

    def guard_step(self, theta: np.ndarray, step: np.ndarray) -> np.ndarray:
        """R/mom.R:375-383."""
        cap = np.maximum(1.0, 0.5 * np.abs(theta))
        step = np.minimum(np.maximum(step, -cap), cap)
        for d in range(self.D):
            for m in range(self.P):
                q = self.rate_pos[d, m]
                if theta[q] - step[q] < 0.1 * theta[q]:
                    step[q] = 0.5 * theta[q]
        return step

    def _plan(self, m: int, M: int, want_end_nets: bool = False):
        key = (m, M, want_end_nets)
        if key not in self._plans:
            nets = self.mn.nets
            desc = prepare_effects(self.effects, self.n, m)
            etypes = np.array([{"eval": 0, "creation": 1, "endow": 2}[e.etype]
                               for e in self.effects], np.int32)
            prob = MultiProblem(
                x0=np.stack([nd.imputed[m - 1] for nd in nets]).astype(np.uint8),
                lam=np.ones(self.D), theta=np.zeros(self.k), type_ids=desc.type_ids,
                netof=self.netof, M=M, seed=0, netref=effect_netrefs(self.effects),
                etypes=etypes, p1=desc.p1, p2=desc.p2, vmat=desc.vmat, wtab=desc.wtab,
                perm=np.stack([nd.permitted(m) for nd in nets]).astype(np.uint8),
                symf=self.symf)
            kd = np.stack([nd.kmask_dist(m) for nd in nets]).astype(np.uint8)
            plan = lib.MultiPlan(prob, want_end_nets=want_end_nets, kdist=kd)
            for d in range(self.D):
                s = self._stat_inputs(d, m)
                plan.set_stats(d, **s)
            self._plans[key] = plan
        return self._plans[key]

    def simulate(self, theta: np.ndarray, M: int):
        """M chains of every period at theta (R/mom.R:385-409): summed
        statistics (M, p) and the per-period moment rows (sp, scp)."""
        th_flat = theta[self.eff_pos]
        stats_acc = np.zeros((M, self.p))
        pieces = []
        for m in range(1, self.P + 1):
            plan = self._plan(m, M)
            lam = np.array([theta[self.rate_pos[d, m - 1]] for d in range(self.D)])
            plan.set_params(th_flat, lam)
            self._calls += 1
            plan.run(_splitmix64((int(self.control.seed) << 20) ^ (self._calls << 4) ^ m))
            out = plan.fetch()
            dist = plan.dist()
            st = plan.stats()
            sp = np.zeros((M, self.p))
            scp = np.zeros((M, self.p))
            for d in range(self.D):
                q = self.rate_pos[d, m - 1]
                sp[:, q] = dist[d] + self.rate_consts[d, m - 1]
                scp[:, q] = out["rscores"][:, d].astype(np.float64)
            sp[:, self.eff_pos] = st
            scp[:, self.eff_pos] = out["scores"].astype(np.float64)
            stats_acc += sp
            pieces.append((sp, scp))
        return stats_acc, pieces

    def run(self, theta0=None) -> MomFit:
        """The three phases (R/mom.R:411-480)."""
        ctl = self.control
        t0 = time.perf_counter()
        self._calls = 0
        n_sims = ctl.n1 + ctl.n3 + sum(ctl.nD + ctl.n2[s] * ctl.batch2 for s in range(ctl.nsub))
        theta = np.array(theta0, float) if theta0 is not None else self.default_theta0()
        stats, pieces = self.simulate(theta, ctl.n1)
        dinv = _dinv(_derivative(pieces), ctl.diagonalize)
        dev = stats.mean(0) - self.targets
        theta = theta - self.guard_step(theta, 0.5 * ctl.firstg * (dinv @ dev))
        t1 = time.perf_counter()
        for sub in range(ctl.nsub):
            gain = ctl.firstg * (0.5 ** sub)
            _, pieces = self.simulate(theta, ctl.nD)
            dinv = _dinv(_derivative(pieces), ctl.diagonalize)
            thav = np.zeros(self.p)
            for _ in range(int(ctl.n2[sub])):
                stats, _ = self.simulate(theta, ctl.batch2)
                dev = stats.mean(0) - self.targets
                theta = theta - self.guard_step(theta, gain * (dinv @ dev))
                thav += theta
            theta = thav / ctl.n2[sub]
        t2 = time.perf_counter()
        stats, pieces = self.simulate(theta, ctl.n3)
        D = _derivative(pieces)
        dinv = _safe_solve(D)
        msf = np.cov(stats, rowvar=False)
        dev_mean = stats.mean(0) - self.targets
        tconv = dev_mean / np.sqrt(np.diag(msf))
        tconv_max = float(np.sqrt(dev_mean @ _safe_solve(msf) @ dev_mean))
        se = np.sqrt(np.abs(np.diag(dinv @ msf @ dinv.T)))
        t3 = time.perf_counter()
        return MomFit(theta=theta, se=se, tconv=tconv, tconv_max=tconv_max,
                      max_abs_tconv=float(np.abs(tconv).max()), param_names=self.param_names,
                      targets=self.targets, sim_means=stats.mean(0), D=D, msf=msf,
                      conditional=False, rate_est=None, rate_sd=None, n_sims=int(n_sims),
                      repeats=0, phase_seconds=[t1 - t0, t2 - t1, t3 - t2], wall_time=t3 - t0,
                      phase3_stats=stats)

    def close(self) -> None:
        for plan in self._plans.values():
            plan.close()
        self._plans.clear()


def mom_estimate_multinet(data: MultiNetData, effects_by_net: list,
                          control: MomControl | None = None) -> MomFit:
    """R/multinet.R:71-98."""
    if not isinstance(data, MultiNetData):
        raise TypeError("`data` must be created with saom_multinet_data().")
    est = MultiMomEstimator(data, effects_by_net, control)
    try:
        return est.run()
    finally:
        est.close()


__all__ = ["MultiNetData", "saom_multinet_data", "MultiMomEstimator", "mom_estimate_multinet",
           "Effect"]
