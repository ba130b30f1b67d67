"""Benchmark: SAOM chain-periods per second at n = 500 on one MI355X.

Workload: 500 directed actors, 4096 replicate chains, 2 periods,
rate 8 per period; wave 1 Bernoulli with mean outdegree 5; effects density
-2.8, recip 1.8, transTrip 0.3, cycle3 -0.1, egoX/altX/simX 0.1/0.1/0.6 on a
covariate uniform on 1..5; wave 2 is one seeded simulated period. Choice in
float64, statistics exact.

A step is one pass over the batch: both periods for all 4096 chains (8192
chain-periods), each period one launch of the period kernel plus the
statistics kernel, from inputs already resident in device memory (period
plans). The timed region is bracketed by device synchronisation.

Also reported: the period kernel against its VALU roof (HIP events on the
plan's stream), the reference CPU engine (compiled from the reference's own
source under oracle/_ref) timed on this host's cores on a bounded sample, and
("mom") the wall time of a method-of-moments fit to max|t| <= 0.1 on the
device next to one pass of the same schedule through the reference engine,
and ("groups") a batch of 256 networks of 30..1200 actors.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SAOM chain-periods/sec at n=500 (GPU vs host CPU); MoM wall time to "
          "max|t|≤0.1")
# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.3e12 VALU lane-ops/s
# (the FP64 vector peak of 78.6 TFLOP/s counts an FMA as two flops).
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9


def algorithmic_ops_per_alternative(n: int, k: int) -> dict:
    """Operations one alternative evaluation (i, j) needs, counted from the
    algorithm (not from the compiled code):
      integer: Philox4x32-10, about 100 (ten rounds of two 32x32 -> 64-bit
               products, xors and key additions), plus 6 per 64-bit word of a
               row (AND + popcount for the two-path, shared-out-tie and
               3-cycle counts): 6 W64 + 100, 148 at n = 500
      float64: 4 per effect (multiply and add of the linear predictor and of
               the softmax-weighted change-statistic sum): 28 at k = 7
      transcendental: 3 (the two logarithms of the Gumbel noise and the
               exponential of the softmax weight), counted as one each
    179 at n = 500, k = 7."""
    w64 = (n + 63) // 64
    ints, f64, trans = 6 * w64 + 100, 4 * k, 3
    return {"int": ints, "f64": f64, "transcendental": trans, "total": ints + f64 + trans}


def host_threads() -> dict:
    """Host threads for the CPU legs: every core this process may run on
    (its affinity mask), capped by OMP_NUM_THREADS where the environment
    sets it."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    return {"threads": threads, "affinity_cpus": aff, "cpu_count": os.cpu_count(),
            "omp_num_threads": omp or None}


def measured_traffic(n: int, chains: int):
    """HBM bytes per period launch of the n = 500 workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy); None for other shapes."""
    path = os.path.join(ROOT, "profiles", "period_kernel_rtc_c2_n500_traffic.md")
    if n != 500 or chains != 4096 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def cpu_baseline(prob_factories, threads: int, chains: int, host: dict) -> dict:
    """The reference CPU engine (compiled from the reference's source, else
    the oracle) on `chains` chains of each period, OpenMP over chains."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    lib = O.reference()
    kind = "reference"
    if lib is None:
        lib, kind = O.oracle(), "port"
    setter = getattr(O.oracle(), "oracle_set_threads")
    setter(threads)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    walls, steps = [], []
    for factory in prob_factories:
        prob = factory(chains)
        t0 = time.perf_counter()
        r = O.simulate_reference(prob) if kind == "reference" else O.simulate(prob)
        walls.append(time.perf_counter() - t0)
        steps.append(float(r["steps"].mean()))
    wall = sum(walls)
    P = len(prob_factories)
    return {"value": round(P * chains / wall, 4), "unit": "chain-periods/s", "cores": threads,
            "kind": kind, "affinity_cpus": host["affinity_cpus"], "cpu_count": host["cpu_count"],
            "omp_num_threads": host["omp_num_threads"],
            "sample": (f"{chains} chains of each of the {P} periods of the same workload "
                       f"(n=500, rate 8, 7 effects), OpenMP static schedule over chains, "
                       f"{threads} threads, {wall:.1f} s wall "
                       f"({' + '.join(f'{w:.1f}' for w in walls)}), mean ministeps "
                       f"{' / '.join(f'{x:.0f}' for x in steps)}")}


def groups_measure(lib) -> dict:
    """Group batch: 256 independent networks (n uniform on 30..1200, 500
    chains each) through one cusna_saom_simulate_groups call per run: one
    period launch per launch shape, items longest-first. Timed: the whole
    call (plan set-up and uploads, launches, statistics, copies back), after
    a 4-chain warm-up call that compiles and loads the same specialised
    kernels (specialisation forced for it: the 4-chain call alone is below
    the library's work threshold)."""
    import os

    import numpy as np

    from cusna_amd.workloads import mixed_group_problems

    small = mixed_group_problems(chains=4)
    prev = os.environ.get("CUSNA_SPECIALIZE")
    os.environ["CUSNA_SPECIALIZE"] = "1"
    try:
        lib.simulate_groups(small)
    finally:
        if prev is None:
            del os.environ["CUSNA_SPECIALIZE"]
        else:
            os.environ["CUSNA_SPECIALIZE"] = prev
    probs = mixed_group_problems()
    t0 = time.perf_counter()
    outs = lib.simulate_groups(probs)
    wall = time.perf_counter() - t0
    cp = sum(p.M for p in probs)
    alts = sum(float(o["steps"].astype(np.int64).sum()) * p.n for p, o in zip(probs, outs))
    ns = [p.n for p in probs]
    return {"workload": "groups: 256 networks, n uniform 30..1200 (%d..%d here), "
                        "500 chains each, rate 4, density/recip/transTrip/inPop/egoX/altX"
                        % (min(ns), max(ns)),
            "chain_periods": cp, "wall_s": round(wall, 3),
            "chain_periods_per_s": round(cp / wall, 1),
            "alternatives_per_s": float("%.4g" % (alts / wall)),
            "streamed_groups": sum(1 for v in ns if v > 1024)}


def gof_measure(lib, plan, x0, cpu: bool) -> dict:
    """sienaGOF distributions of the 4096 end networks of the last n = 500
    period, where they lie in device memory (cusna_period_plan_gof): in- and
    out-degree counts at levels 0..10 and geodesic counts at levels 1..5.
    CPU leg: the reference's cusna_gof_distribution (single-threaded, as the
    reference runs it) on 4 end networks of the same problem."""
    import numpy as np

    from cusna_amd.workloads import n500_problem

    res = {"workload": "GoF of 4096 end networks of the n=500 period: in/out-degree levels "
                       "0..10, geodesic levels 1..5"}
    plan.gof(0, range(11))  # code objects loaded
    for name, kind, lv in (("indegree", 0, range(11)), ("outdegree", 1, range(11)),
                           ("geodesic", 2, range(1, 6))):
        t0 = time.perf_counter()
        plan.gof(kind, lv, masked=True)
        w = time.perf_counter() - t0
        res[name] = {"wall_ms": round(1e3 * w, 2), "networks_per_s": round(4096 / w, 1)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        R = O.reference()
        if R is not None:
            ends = lib.simulate_period(n500_problem(M=4, x0=x0, period_seed=11),
                                       want_end_nets=True)["a_ends"]
            t0 = time.perf_counter()
            O.gof_distribution(ends, 2, list(range(1, 6)), lib=R, fn="cusna_gof_distribution")
            cw = time.perf_counter() - t0
            t1 = time.perf_counter()
            O.gof_distribution(ends, 0, list(range(11)), lib=R, fn="cusna_gof_distribution")
            cd = time.perf_counter() - t1
            res["cpu"] = {"kind": "reference", "cores": 1,
                          "geodesic_networks_per_s": round(4 / cw, 3),
                          "indegree_networks_per_s": round(4 / cd, 1),
                          "sample": f"4 end networks of the same problem, {cw:.1f} s geodesic, "
                                    f"{cd:.3f} s in-degree"}
    return res


def rate_measure(lib, x0, cpu: bool, threads: int) -> dict:
    """The n = 500 period with two non-basic rate effects (outRate and a
    covariate RateX), 4096 chains: plan run (launch + completion) after one
    warm-up; CPU leg: the reference engine on a bounded sample."""
    import numpy as np

    from cusna_amd.workloads import n500_problem

    def prob(M):
        pr = n500_problem(M=M, x0=x0, period_seed=11)
        pr.rate_kind = np.array([1, 0], np.int32)
        pr.rate_alpha = np.array([0.05, 0.2])
        pr.rate_u = np.stack([np.zeros(pr.n), np.asarray(pr.vmat)[4]])
        pr.lam = 0.6
        return pr

    plan = lib.PeriodPlan(prob(4096))
    plan.run(1)
    plan.sync()
    t0 = time.perf_counter()
    for s in range(2):
        plan.run(100 + s)
        plan.sync()
    wall = (time.perf_counter() - t0) / 2
    out = plan.fetch()
    res = {"workload": "n=500 period (4096 chains, 7 effects) with outRate and RateX rate "
                       "effects, basic rate 0.6", "ms_per_period": round(1e3 * wall, 2),
           "chain_periods_per_s": round(4096 / wall, 1),
           "mean_ministeps": round(float(out["steps"].mean()), 1),
           "specialised": plan.specialized}
    plan.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        if O.reference() is not None:
            os.environ["OMP_NUM_THREADS"] = str(threads)
            chains = 2 * threads
            t0 = time.perf_counter()
            O.simulate_reference(prob(chains))
            cw = time.perf_counter() - t0
            res["cpu"] = {"kind": "reference", "cores": threads,
                          "chain_periods_per_s": round(chains / cw, 3),
                          "sample": f"{chains} chains of the same problem, {threads} threads, "
                                    f"{cw:.1f} s wall"}
    return res


def coevo_measure(lib, cpu: bool, threads: int) -> dict:
    """Network-behaviour co-evolution period: 300 actors, 2048 chains,
    behaviour 1..5 (linear, quad, avAlt; rate 4), network density/recip/
    transTrip/simX(behaviour, dynamic), rate 6: plan run (launch +
    completion) after one warm-up; CPU leg: the reference engine on a
    bounded sample of the same problem."""
    from cusna_amd.workloads import coevo300_problem

    plan = lib.PeriodPlan(coevo300_problem(M=2048))
    plan.run(1)
    plan.sync()
    t0 = time.perf_counter()
    for s in range(2):
        plan.run(100 + s)
        plan.sync()
    wall = (time.perf_counter() - t0) / 2
    out = plan.fetch()
    res = {"workload": "co-evolution period: n=300, 2048 chains, network rate 6 "
                       "(density/recip/transTrip/simX dyn), behaviour rate 4 "
                       "(linear/quad/avAlt)", "ms_per_period": round(1e3 * wall, 2),
           "chain_periods_per_s": round(2048 / wall, 1),
           "kernel_ms": round(plan.kernel_ms(), 2),
           "mean_ministeps": round(float(out["steps"].mean()), 1),
           "specialised": plan.specialized}
    plan.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        if O.reference() is not None:
            os.environ["OMP_NUM_THREADS"] = str(threads)
            chains = 2 * threads
            t0 = time.perf_counter()
            O.simulate_reference(coevo300_problem(M=chains))
            cw = time.perf_counter() - t0
            res["cpu"] = {"kind": "reference", "cores": threads,
                          "chain_periods_per_s": round(chains / cw, 3),
                          "sample": f"{chains} chains of the same problem, {threads} threads, "
                                    f"{cw:.1f} s wall"}
    return res


def multi_problem_n500(M: int, x0s, seed: int = 77):
    """Two co-evolving networks of 500 actors (the waves of the n = 500 workload as starts),
    rate 4 each, density/recip/transTrip/crprod on network 0 and
    density/recip/to/crprodRecip on network 1."""
    import numpy as np

    from cusna_amd.abi import MultiProblem
    from cusna_amd.effects import Effect, effect_netrefs, prepare_effects

    effs = [Effect("density"), Effect("recip"), Effect("transTrip"), Effect("crprod", net_ref=1),
            Effect("density"), Effect("recip"), Effect("to", net_ref=0),
            Effect("crprodRecip", net_ref=0)]
    n = x0s[0].shape[0]
    d = prepare_effects(effs, n, 1)
    return MultiProblem(x0=np.stack(x0s).astype(np.uint8), lam=np.array([4.0, 4.0]),
                        theta=np.array([-2.6, 1.8, 0.3, 0.8, -2.7, 1.6, 0.1, 0.5]),
                        type_ids=d.type_ids, netof=np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int32),
                        M=M, seed=seed, netref=effect_netrefs(effs), p1=d.p1, p2=d.p2,
                        vmat=d.vmat, wtab=d.wtab, symf=np.zeros(2, np.int32))


def multi_measure(lib, x0s, cpu: bool, threads: int) -> dict:
    """Multi-network periods (two networks of 500 actors, 4096 chains):
    one cusna_multi_plan_run per step (launch + completion), after one
    warm-up run; the CPU leg runs the reference's multi-network engine
    (oracle/_ref) on a bounded sample of the same problem. Then the
    multi-network MoM fit of the reference's fixture (test-multinet.R)."""
    import numpy as np

    M = 4096
    prob = multi_problem_n500(M, x0s)
    plan = lib.MultiPlan(prob)
    plan.run(1)
    steps = 3
    t0 = time.perf_counter()
    for s in range(steps):
        plan.run(100 + s)
    wall = (time.perf_counter() - t0) / steps
    out = plan.fetch()
    plan.close()
    res = {"workload": "multi-network periods: D=2 networks of n=500 (the n=500 waves as "
                       "starts), %d chains, rate 4+4, 4+4 effects incl. crprod/to/crprodRecip"
                       % M,
           "ms_per_period": round(1e3 * wall, 2), "chain_periods_per_s": round(M / wall, 1),
           "alternatives_per_s": float("%.4g" % (float(out["steps"].astype(np.int64).sum())
                                                 * prob.n / wall)),
           "mean_ministeps": round(float(out["steps"].mean()), 1)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        if O.reference() is not None:
            os.environ["OMP_NUM_THREADS"] = str(threads)
            chains = 2 * threads
            t0 = time.perf_counter()
            O.simulate_multi_reference(multi_problem_n500(chains, x0s), want_end_nets=False)
            cw = time.perf_counter() - t0
            res["cpu"] = {"kind": "reference", "cores": threads,
                          "chain_periods_per_s": round(chains / cw, 3),
                          "sample": f"{chains} chains of the same problem through "
                                    f"cusna_saom_simulate_multi_cpu, {threads} threads, "
                                    f"{cw:.1f} s wall"}
    # the multinet fixture fit
    import gzip

    from cusna_amd.effects import Effect
    from cusna_amd.mom import MomControl
    from cusna_amd.multinet import MultiMomEstimator, saom_multinet_data

    with gzip.open(os.path.join(ROOT, "tests", "golden", "multinet.json.gz"), "rt") as f:
        g = json.load(f)
    S = g["settings"]
    mn = saom_multinet_data([[np.array(w, float) for w in g["netA_waves"]],
                             [np.array(w, float) for w in g["netB_waves"]]])
    ctl = MomControl(firstg=S["firstg"], n2=tuple(S["n2"]), batch2=S["batch2"], n1=S["n1"],
                     n3=S["n3"], nD=S["nD"], diagonalize=S["diagonalize"], seed=S["seed"])
    est = MultiMomEstimator(mn, [[Effect("density"), Effect("recip")]] * 2, ctl)
    est.run()  # plans created, code objects loaded
    fit = est.run()
    est.close()
    th_ref, se_ref = np.array(g["theta"]), np.array(g["se"])
    res["mom"] = {"workload": "multinet fixture fit: 2 networks of 50 actors, density+recip "
                              "each, schedule n1=nD=300, n2=15/20 x 128, n3=600",
                  "wall_s": round(fit.wall_time, 3), "chains_simulated": fit.n_sims,
                  "max_abs_tconv": round(fit.max_abs_tconv, 4),
                  "theta": [round(float(x), 4) for x in fit.theta],
                  "within_reference_bar": bool(np.all(np.abs(fit.theta - th_ref)
                                                      < se_ref + 0.2))}
    return res


def mlp_measure(lib) -> dict:
    """Training of the surrogate perceptron at the shape the surrogate phase
    uses (3 parameters in, 3 moments out, width 256, batch 1024, 2^13
    samples, 200 Adam steps): wall time of the train call after one warm-up
    call. Per step two launches: the batch in 16-row slices through the
    forward and backward passes, then the weight gradients and Adam per
    64 x 64 weight tile (BF16 MFMA products)."""
    import numpy as np

    rng = np.random.default_rng(3)
    p, n, steps, B, width = 3, 1 << 13, 200, 1024, 256
    x = rng.uniform(-1, 1, (n, p)).astype(np.float32)
    y = (np.tanh(x @ rng.normal(size=(p, p))) + 0.1 * rng.normal(size=(n, p))).astype(np.float32)
    m = lib.Mlp(p, p, width=width, batch=B, seed=5)
    try:
        m.train(x, y, steps=10, lr=2e-3, seed=1)
        t0 = time.perf_counter()
        loss = m.train(x, y, steps=steps, lr=2e-3, seed=2)
        wall = time.perf_counter() - t0
    finally:
        m.close()
    # forward 2 B (64 x 256 + 2 x 256^2 + 256 x 64) FLOP (padded shapes), backward
    # twice that: dW for all four layers, dH for the upper three
    fwd = 2.0 * B * (64 * width + 2 * width * width + width * 64)
    flops = fwd + 2.0 * fwd - 2.0 * B * 64 * width
    return {"workload": "perceptron training: p=3 -> 256 -> 256 -> 256 -> 2p, batch 1024, "
                        "2^13 samples, 200 Adam steps (the surrogate phase's shape)",
            "ms_per_step": round(1e3 * wall / steps, 4), "final_loss": round(float(loss), 4),
            "gflop_per_step": round(flops / 1e9, 4),
            "tflops": round(flops * steps / wall / 1e12, 3)}


GROUPS_MOM_RUNS = {
    "plain": {},
    # the surrogate with the automatic warm start (one plain sub-phase first
    # when phase 1 lies far from the targets), trained on 2^12 (theta,
    # statistics) pairs: 2^14 and 2^16 pairs gave the same estimates to four
    # digits at 1.4 and 3.2 x the wall time, and simulating 2^18 pairs alone
    # (about 1.8 ms per pair at 16 x 1000 actors) would outlast the plain fit
    "surrogate": dict(surrogate=True, surrogate_pairs_log2=12),
    # the surrogate straight from phase 1 (its plain fallback engages when
    # the corrections leave a large residual)
    "surrogate_from_phase1": dict(surrogate=True, surrogate_pairs_log2=12, surrogate_warmup=-1),
    # larger training sets: 2^14 and 2^16 (theta, statistics) pairs
    "surrogate_p14": dict(surrogate=True, surrogate_pairs_log2=14),
    "surrogate_p16": dict(surrogate=True, surrogate_pairs_log2=16),
}


def groups_mom_measure(lib, runs: list) -> dict:
    """Surrogate-assisted against plain MoM on 16 groups of 1000 actors, 12
    effects, one period each (28 parameters: 16 rates + 12 effects),
    phase 3 with 2000 chains, the reference's default schedule. Minutes of
    device time: run with --groups-mom only. Progress lines on stderr."""
    import numpy as np

    from cusna_amd.mom import GroupMomEstimator, MomControl
    from cusna_amd.workloads import groups1000_estimation

    t0 = time.perf_counter()
    datas, effs, theta, rates = groups1000_estimation(
        lambda p: lib.simulate_period(p, want_end_nets=True)["a_ends"])
    print(f"[groups-mom] data built in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    est = GroupMomEstimator(datas, effs, MomControl(n3=2000))
    out = {"workload": "16 groups x 1000 actors, one period each, 12 effects (density, recip, "
                       "transTrip, cycle3, inPop, outAct, egoX/altX/simX on N(0,1) and on a "
                       "1..5 covariate) + 16 rates; unconditional; n1=nD=1000, n2=40/40/60/80 "
                       "x 256, n3=2000; then repeats until max|t|<=0.1",
           "true_theta": [round(float(x), 4) for x in theta]}
    fits = {}
    for key in runs:
        kw = GROUPS_MOM_RUNS[key]
        print(f"[groups-mom] {key}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        fit = est.run(max_repeats=10, conv_tol=0.1, verbose=True, **kw)
        wall = time.perf_counter() - t0
        fits[key] = fit
        out[key] = {"wall_s": round(wall, 2), "max_abs_tconv": round(float(fit.max_abs_tconv), 4),
                    "repeats": fit.repeats, "chains_simulated": fit.n_sims,
                    "theta": [round(float(x), 4) for x in fit.theta[est.P:]]}
    if "plain" in fits:
        plain = fits["plain"]
        out["plain"]["se"] = [round(float(x), 4) for x in plain.se[est.P:]]
        for key, fit in fits.items():
            if key != "plain":
                out[key]["within_one_se_of_plain"] = bool(np.all(np.abs(fit.theta - plain.theta)
                                                                 <= plain.se))
                out[key]["wall_fraction_of_plain"] = round(out[key]["wall_s"]
                                                           / out["plain"]["wall_s"], 3)
    return out


def v1_call_measure(lib) -> dict:
    """Per-call wall time of the one-shot entry points at the R glue's
    phase-2 shape (R/mom.R:238-244): one period of the 50-actor fit's
    problem, 256 chains, conditional, host inputs to host outputs, one call
    per Robbins-Monro step. The ctypes structs are built once; each call is
    cusna_saom_simulate_period (or its no-allocation variant) as the glue
    makes it. Next to it the period kernel's own time on a device-resident
    plan of the same problem."""
    import ctypes as C

    import numpy as np

    from cusna_amd.abi import SaomResult
    from cusna_amd.effects import Effect
    from cusna_amd.workloads import effect_problem, poisson_capped_network

    rng = np.random.default_rng(8)
    x = poisson_capped_network(50, 2.3, 6, 0.4, rng)
    prob = effect_problem(x, [Effect("density"), Effect("recip"), Effect("transTrip")],
                          [-2.5, 2.0, 0.4], 1.0, 256, 44, conditional=True, target_dist=40)
    L = lib.load()
    p = prob.build()
    res = SaomResult(prob.n, prob.k, prob.M, prob.kb, False)
    out = {"workload": "one period call: 50 actors, density/recip/transTrip, conditional "
                       "(target distance 40), 256 chains, host buffers in and out"}
    for key, fn in (("simulate_period", L.cusna_saom_simulate_period),
                    ("simulate_period_noalloc", L.cusna_saom_simulate_period_noalloc)):
        for _ in range(10):
            assert fn(C.byref(p), C.byref(res.c)) == 0
        calls = 300
        t0 = time.perf_counter()
        for _ in range(calls):
            fn(C.byref(p), C.byref(res.c))
        out[key + "_ms_per_call"] = round(1e3 * (time.perf_counter() - t0) / calls, 4)
    plan = lib.PeriodPlan(prob)
    try:
        ms = []
        for s in range(20):
            plan.run(100 + s)
            plan.sync()
            ms.append(plan.kernel_ms())
        out["kernel_ms"] = round(float(np.median(ms[5:])), 4)
    finally:
        plan.close()
    return out


def mom_measure(lib, cpu: bool, threads: int) -> dict:
    """Fit of 50 actors, 3 waves, conditional MoM on density/recip/
    transTrip with the reference's default schedule (n3 = 1000), repeated
    last-gain sub-phases until max|tconv| <= 0.1. Wall time of the whole
    cusna_mom_estimate call (set-up, three phases, host solves)."""
    import numpy as np

    from cusna_amd.mom import MomControl, MomEstimator
    from cusna_amd.workloads import fit50_estimation

    data, eff = fit50_estimation(lambda p: lib.simulate_period(p, want_end_nets=True)["a_ends"])
    est = MomEstimator(data, eff, None, True, MomControl(n3=1000))
    est.run(max_repeats=0)  # warm-up: code objects and first-launch set-up
    t0 = time.perf_counter()
    fit = est.run(max_repeats=20, conv_tol=0.1)
    wall = time.perf_counter() - t0
    out = {"workload": "MoM fit: n=50, 3 waves, conditional, density/recip/transTrip, "
                       "schedule n1=nD=1000, n2=40/40/60/80 x 256, n3=1000, then last-gain "
                       "sub-phase + phase 3 until max|t|<=0.1",
           "wall_s": round(wall, 3), "max_abs_tconv": round(float(fit.max_abs_tconv), 4),
           "repeats": fit.repeats, "chains_simulated": fit.n_sims,
           "phase_s": [round(x, 3) for x in fit.phase_seconds],
           "theta": [round(float(x), 4) for x in fit.theta]}
    # same fit with the surrogate phase 2 (perceptron on 2^13 simulated pairs,
    # 200 Adam steps, then simulation-checked Newton corrections), and the
    # ablation without a trained network (64 pairs, one step: the corrections
    # alone, from the phase-1 Newton point)
    for key, kw in (("surrogate", {}),
                    ("newton_only", dict(surrogate_pairs_log2=6, surrogate_steps=1))):
        t0 = time.perf_counter()
        sf = est.run(max_repeats=20, conv_tol=0.1, surrogate=True, **kw)
        out[key] = {"wall_s": round(time.perf_counter() - t0, 3),
                    "max_abs_tconv": round(float(sf.max_abs_tconv), 4),
                    "repeats": sf.repeats, "chains_simulated": sf.n_sims,
                    "theta": [round(float(x), 4) for x in sf.theta],
                    "within_one_se_of_plain": bool(np.all(np.abs(sf.theta - fit.theta)
                                                          <= fit.se)),
                    "wall_fraction_of_plain": round((time.perf_counter() - t0) / wall, 3)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import mom_oracle as MO
        import oracle as O

        if O.reference() is not None:
            sim, kind = O.simulate_reference, "reference"
        else:
            sim, kind = O.simulate, "port"
        O.oracle().oracle_set_threads(threads)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        c = est.control
        rows = MO.reference_rows(est, sim)
        t0 = time.perf_counter()
        r = MO.mom_run(rows, est.P, est.p, est.targets, est.default_theta0(), est.mode, c.n1,
                       c.nD, c.batch2, c.n3, c.n2, c.firstg, c.diagonalize, c.seed)
        cw = time.perf_counter() - t0
        out["cpu"] = {"kind": kind, "cores": threads, "one_pass_wall_s": round(cw, 3),
                      "one_pass_chains": r["n_sims"],
                      "one_pass_max_abs_tconv": round(float(np.abs(r["tconv"]).max()), 4),
                      "wall_s_at_gpu_chains": round(cw * fit.n_sims / r["n_sims"], 3),
                      "sample": "one pass of the same schedule through the reference "
                                "period engine (Python driver of R/mom.R's algebra); "
                                "scaled by chains to the GPU run's chain count"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chains", type=int, default=4096)
    ap.add_argument("--n", type=int, default=500)
    ap.add_argument("--threads", type=int, default=0, help="block size (0 = library default)")
    ap.add_argument("--cpu-chains", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mom", action="store_true")
    ap.add_argument("--no-groups", action="store_true")
    ap.add_argument("--no-multi", action="store_true")
    ap.add_argument("--no-gof", action="store_true")
    ap.add_argument("--no-rate", action="store_true")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--no-coevo", action="store_true")
    ap.add_argument("--no-groups-mom", action="store_true",
                    help="skip the 16 x 1000-actor plain vs surrogate MoM comparison "
                         "(about five minutes)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="only the perceptron training section (for counter runs)")
    ap.add_argument("--no-v1", action="store_true",
                    help="skip the per-call time of the one-shot entry points")
    ap.add_argument("--v1-only", action="store_true",
                    help="only the per-call time of the one-shot period entry points")
    ap.add_argument("--mom-only", action="store_true",
                    help="only the 50-actor MoM fit, without its CPU leg (for profiler runs)")
    ap.add_argument("--groups-mom", default="", metavar="RUNS",
                    help="only the 16 x 1000-actor MoM comparison, comma-separated runs of "
                         + "/".join(GROUPS_MOM_RUNS) + " (minutes each)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one GPU, and the "
              "path shards over chains without cross-GPU traffic.",
              file=sys.stderr)
        return 2

    import numpy as np

    from cusna_amd import lib
    from cusna_amd.workloads import n500_problem, n500_waves

    if not lib.has_device():
        print("no HIP device", file=sys.stderr)
        return 3
    if args.mlp_only:
        print(json.dumps({"mlp": mlp_measure(lib)}), flush=True)
        return 0
    if args.mom_only:
        print(json.dumps({"mom": mom_measure(lib, False, args.threads)}), flush=True)
        return 0
    if args.v1_only:
        print(json.dumps({"v1_call": v1_call_measure(lib)}), flush=True)
        return 0
    if args.groups_mom:
        runs = args.groups_mom.split(",")
        if any(r not in GROUPS_MOM_RUNS for r in runs):
            print(f"--groups-mom runs: {sorted(GROUPS_MOM_RUNS)}", file=sys.stderr)
            return 2
        print(json.dumps({"groups_mom": groups_mom_measure(lib, runs)}), flush=True)
        return 0
    n, M = args.n, args.chains
    w1, w2 = n500_waves(lambda p: lib.simulate_period(p, want_end_nets=True)["a_ends"], n=n)
    probs = [n500_problem(M=M, n=n, x0=w, period_seed=11 + q) for q, w in enumerate((w1, w2))]
    for p in probs:
        p.threads = args.threads
    plans = [lib.PeriodPlan(p) for p in probs]

    def step(s: int) -> list:
        ms = []
        for q, pl in enumerate(plans):
            pl.run(1_000_003 * s + q)
            pl.sync()
            ms.append(pl.kernel_ms())
        return ms

    for s in range(args.warmup):
        step(10_000 + s)
    for pl in plans:
        pl.sync()
    kern = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        kern += step(s)
    for pl in plans:
        pl.sync()
    elapsed = time.perf_counter() - t0
    outs = [pl.fetch() for pl in plans]
    gof = None if args.no_gof else gof_measure(lib, plans[1], w1, not args.no_cpu)
    chain_periods = 2 * M * args.steps
    value = chain_periods / elapsed
    # algorithmic work of the last launches: alternatives = ministeps x n
    units = [float(o["steps"].astype(np.int64).sum()) * n for o in outs]
    ops = algorithmic_ops_per_alternative(n, probs[0].k)
    kms = float(np.mean(kern))
    achieved = float(np.mean(units)) * ops["total"] / (kms * 1e-3)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "chain-periods/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"period simulation: n={n}, {M} chains x 2 periods, "
                               "rate 8, density/recip/transTrip/cycle3/egoX/altX/simX, "
                               "unconditional", "n": n, "chains": M, "periods": 2,
                   "block_threads": plans[0].block_threads},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": round(VALU_PEAK_OPS / 1e12, 3), "unit": "Tops/s",
                     "frac": round(achieved / VALU_PEAK_OPS, 4),
                     "traffic": measured_traffic(n, M) if plans[0].specialized else None,
                     "kernel": "saom_period_rtc" if plans[0].specialized
                               else "saom_period_kernel", "kernel_ms": round(kms, 3),
                     "alternatives_per_launch": float(np.mean(units)),
                     "ops_per_alternative": ops["total"],
                     "ops_split": {k: v for k, v in ops.items() if k != "total"}},
    }
    host = host_threads()
    threads = host["threads"]
    if not args.no_cpu:
        chains = args.cpu_chains or 8 * threads  # at least 8 chains per core
        line["cpu_baseline"] = cpu_baseline(
            [lambda m, w=w, q=q: n500_problem(M=m, n=n, x0=w, period_seed=11 + q)
             for q, w in enumerate((w1, w2))], threads, chains, host)
    for pl in plans:
        pl.close()
    if gof is not None:
        line["gof"] = gof
    if not args.no_rate:
        line["rate"] = rate_measure(lib, w1, not args.no_cpu, threads)
    if not args.no_coevo:
        line["coevo"] = coevo_measure(lib, not args.no_cpu, threads)
    if not args.no_multi:
        line["multi"] = multi_measure(lib, (w1, w2), not args.no_cpu, threads)
    # the one-shot call's cost is measured before the groups batch: after a
    # full 256-group call the same calls took 0.58 instead of 0.36 ms for the
    # rest of the process (not reproduced by streams, events, code objects or
    # large plans alone)
    if not args.no_v1:
        line["v1_call"] = v1_call_measure(lib)
    if not args.no_groups:
        line["groups"] = groups_measure(lib)
    if not args.no_mom:
        line["mom"] = mom_measure(lib, not args.no_cpu, threads)
    if not args.no_mlp:
        line["mlp"] = mlp_measure(lib)
    if not args.no_groups_mom:
        line["groups_mom"] = groups_mom_measure(lib, ["plain", "surrogate"])
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
