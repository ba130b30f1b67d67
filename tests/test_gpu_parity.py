"""GPU parity: the HIP library against the reference fixtures and the oracle.

Bar (integer / index work bit-exact; floating point as stated per test):
  dist, dist_stat, steps, bdist, z_end, end networks: identical
  stats: identical bits (the statistic kernels sum in the reference's order)
  ntim: identical bits in replay mode; |rel| <= 1e-12 free running (device vs
        host C library log() may differ by an ulp)
  scores (float32 outputs of float64 sums): |diff| <= 1e-5 * max(1, |ref|);
        the device sums the softmax weights as a tree, the reference
        sequentially, so the float64 sums agree to ~1e-15 relative.
"""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

SCORE_TOL = 1e-5


def _dense_from_packed(packed, n):
    from cusna_amd.abi import unpack_rows

    return unpack_rows(np.asarray(packed, dtype=np.uint32), n)


def assert_period_equal(got, ref, *, exact_ntim=False, label=""):
    for key in ("dist", "dist_stat", "steps"):
        np.testing.assert_array_equal(np.asarray(got[key]), np.asarray(ref[key]).reshape(
            np.asarray(got[key]).shape), err_msg=f"{label} {key}")
    gs, rs = np.asarray(got["stats"]), np.asarray(ref["stats"]).reshape(np.asarray(got["stats"]).shape)
    np.testing.assert_array_equal(gs.view(np.uint64), rs.astype(np.float64).view(np.uint64),
                                  err_msg=f"{label} stats bits")
    gn, rn = np.asarray(got["ntim"]), np.asarray(ref["ntim"], dtype=np.float64)
    if exact_ntim:
        np.testing.assert_array_equal(gn, rn, err_msg=f"{label} ntim")
    else:
        np.testing.assert_allclose(gn, rn, rtol=1e-12, atol=0, err_msg=f"{label} ntim")
    gsc = np.asarray(got["scores"], np.float64)
    rsc = np.asarray(ref["scores"], np.float64).reshape(gsc.shape)
    assert np.all(np.abs(gsc - rsc) <= SCORE_TOL * np.maximum(1.0, np.abs(rsc))), (
        f"{label} scores max diff {np.max(np.abs(gsc - rsc))}")
    if "bdist" in ref:
        np.testing.assert_array_equal(np.asarray(got["bdist"]), np.asarray(ref["bdist"]),
                                      err_msg=f"{label} bdist")
        np.testing.assert_array_equal(np.asarray(got["z_end"]),
                                      np.asarray(ref["z_end"]).reshape(np.asarray(got["z_end"]).shape),
                                      err_msg=f"{label} z_end")
        gb = np.asarray(got["bscores"], np.float64)
        rb = np.asarray(ref["bscores"], np.float64).reshape(gb.shape)
        assert np.all(np.abs(gb - rb) <= SCORE_TOL * np.maximum(1.0, np.abs(rb))), f"{label} bscores"


def test_philox_known_answers(gpu_lib):
    # Random123 philox4x32-10 known-answer vectors (Salmon et al. 2011)
    kat = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
           ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
           ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]
    for ctr, key, want in kat:
        np.testing.assert_array_equal(gpu_lib.philox4x32(ctr, key), np.array(want, np.uint32))


def _sim_fixture_problems():
    from cusna_amd.data import build_saom_data
    from cusna_amd.effects import BehEffect, Effect as E
    from cusna_amd.period import period_problem

    g = load_golden("sim.json.gz")
    waves = [np.array(w, float) for w in g["inputs"]["s50_waves"]]
    alc = np.array(g["inputs"]["alcohol"], float)
    cfgs = {
        "baseline": (build_saom_data(waves), [E("density"), E("recip"), E("transTrip")], None),
        "covariate": (build_saom_data(waves), [E("density"), E("recip"), E("transTrip"),
                                               E("egoX", alc[:, 0]), E("altX", alc[:, 0]),
                                               E("simX", alc[:, 0])], None),
        "coevol": (build_saom_data(waves, behavior=alc),
                   [E("density"), E("recip"), E("transTrip"), E("egoX", dyn=True),
                    E("altX", dyn=True), E("simX", dyn=True)],
                   [BehEffect("linear"), BehEffect("quad"), BehEffect("avSim")]),
    }
    out = []
    for name, (d, eff, beff) in cfgs.items():
        c = g[name]
        prob = period_problem(d, eff, c["m"] + 1, c["theta"], c["lam"], c["M"], c["seed"], beff,
                              c.get("thetab"), c.get("lamb", 0.0))
        out.append((name, prob, c))
    return out


@pytest.mark.parametrize("cfg", ["baseline", "covariate", "coevol"])
def test_reference_sim_fixture(gpu_lib, cfg):
    """The reference's own frozen period-simulator outputs (tests/testthat/
    fixtures/sim/sim.json.gz; test-sim.R:39-72)."""
    name, prob, c = [x for x in _sim_fixture_problems() if x[0] == cfg][0]
    print(f"case sim.json.gz/{name} n={prob.n} M={prob.M}", flush=True)
    got = gpu_lib.simulate_period(prob)
    assert_period_equal(got, c, label=name)


def _ref_cases():
    return load_golden("ref_cases.json.gz")["cases"]


@pytest.mark.parametrize("idx", range(38))
def test_reference_engine_cases(gpu_lib, idx):
    """Outputs of the reference engine compiled from source, every effect id."""
    from ref_cases_loader import spec_problem

    case = _ref_cases()[idx]
    spec = case["spec"]
    prob = spec_problem(spec)
    print(f"case {spec['name']} n={prob.n} M={prob.M} k={prob.k}", flush=True)
    got = gpu_lib.simulate_period(prob, want_end_nets=True)
    assert_period_equal(got, case["out"], label=spec["name"])
    ends = np.stack([_dense_from_packed(p, prob.n) for p in case["out"]["ends_packed"]])
    np.testing.assert_array_equal(got["a_ends"], ends)


def test_replay_n500_bit_exact(gpu_lib, oracle_lib):
    """The n = 500, seven-effect workload on 8 chains in replay mode: the
    device reads the host's recorded log(u0) and log(-log u_j), so the whole
    trajectory, ntim and statistics must be identical to the oracle's."""
    from cusna_amd.workloads import n500_problem

    prob = n500_problem(M=8)
    print("case replay n500 M=8", flush=True)
    free = oracle_lib.simulate(prob)
    smax = int(free["steps"].max()) + 64
    lu0, llu = oracle_lib.replay_noise(prob.seed, prob.M, prob.n, smax)
    ref = oracle_lib.simulate(prob, want_end_nets=True, replay=(lu0, llu, smax))
    got = gpu_lib.simulate_period_replay(prob, lu0, llu, want_end_nets=True)
    assert_period_equal(got, ref, exact_ntim=True, label="replay n500")
    np.testing.assert_array_equal(got["a_ends"], ref["a_ends"])
    # replay reproduces the free-running oracle (same recorded noise)
    np.testing.assert_array_equal(ref["steps"], free["steps"])


def test_free_n500_against_oracle(gpu_lib, oracle_lib):
    """Free running at n = 500: device log/exp, trajectories still identical."""
    from cusna_amd.workloads import n500_problem

    prob = n500_problem(M=16)
    print("case free n500 M=16", flush=True)
    ref = oracle_lib.simulate(prob, want_end_nets=True)
    got = gpu_lib.simulate_period(prob, want_end_nets=True)
    assert_period_equal(got, ref, label="free n500")
    np.testing.assert_array_equal(got["a_ends"], ref["a_ends"])


def test_coevolution_avalt_n300(gpu_lib, oracle_lib):
    """Co-evolution at n = 300: behaviour avAlt dynamics at n = 300."""
    from cusna_amd.workloads import coevo300_problem

    prob = coevo300_problem(M=8)
    print("case coevo n=300 M=8", flush=True)
    ref = oracle_lib.simulate(prob)
    got = gpu_lib.simulate_period(prob)
    assert_period_equal(got, ref, label="coevo n300")


def test_per_chain_theta(gpu_lib, oracle_lib):
    """Chain c uses theta row c: equal to a run where every row is row c."""
    from cusna_amd.workloads import n500_problem

    prob = n500_problem(M=6, n=200)
    rng = np.random.default_rng(5)
    th = np.array(prob.theta, float)[None, :] + rng.normal(scale=0.1, size=(6, prob.k))
    prob.theta = th
    prob.lam = np.linspace(6.0, 9.0, 6)
    got = gpu_lib.simulate_period(prob)
    ref = oracle_lib.simulate(prob, per_chain=True)
    assert_period_equal(got, ref, label="per-chain")


@pytest.mark.parametrize("idx", range(7))
def test_network_stats_bit_exact(gpu_lib, idx):
    """cusna_network_stats on random networks, every statistic id, against the
    reference engine's outputs."""
    from ref_cases_loader import stats_inputs

    case = load_golden("ref_cases.json.gz")["stats"][idx]
    a, kw = stats_inputs(case["spec"])
    ids = kw["type_ids"]
    got = []
    for lo in range(0, len(ids), 16):
        sl = slice(lo, lo + 16)
        got += gpu_lib.network_stats(a, ids[sl], kw["p1"][sl], kw["p2"][sl], kw["vmat"][sl],
                                     kw["vmiss"][sl], kw["dyn_flags"][sl], kw["vb"], kw["od0"],
                                     kw["id0"], kw["wtab"][sl], 0).tolist()
    np.testing.assert_array_equal(np.array(got).view(np.uint64),
                                  np.array(case["out"], np.float64).view(np.uint64))


@pytest.mark.parametrize("idx", range(4))
def test_behavior_stats_bit_exact(gpu_lib, idx):
    from ref_cases_loader import beh_stats_inputs

    case = load_golden("ref_cases.json.gz")["beh"][idx]
    z, bty, xnet, bk, zbar, br, bs = beh_stats_inputs(case["spec"])
    got = gpu_lib.behavior_stats(z, bty, xnet, bk, zbar, br, bs)
    np.testing.assert_array_equal(got.view(np.uint64),
                                  np.array(case["out"], np.float64).view(np.uint64))


def test_packed_stats_match_dense(gpu_lib, oracle_lib):
    """The batched packed-row entry equals the oracle on every network."""
    rng = np.random.default_rng(3)
    n, M = 90, 5
    nets = (rng.random((M, n, n)) < 0.06).astype(np.uint8)
    for c in range(M):
        np.fill_diagonal(nets[c], 0)
    ids = np.array([0, 1, 2, 3, 4, 10, 11, 12, 16, 24], np.int32)
    k = len(ids)
    vmat = rng.integers(1, 6, size=(k, n)).astype(float) - 3.0
    p1 = np.where(ids == 12, 0.25, 0.0)
    from cusna_amd.effects import gwesp_weights

    wtab = np.array([gwesp_weights(69, n)] * k)
    got = gpu_lib.network_stats_packed(nets, ids, p1=p1, vmat=vmat, wtab=wtab)
    for c in range(M):
        ref = oracle_lib.network_stats(nets[c], ids, p1, np.zeros(k), vmat, wtab=wtab)
        np.testing.assert_array_equal(got[c].view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("n", [600, 1000])
def test_gwesp_family_stats_several_windows(n, gpu_lib, oracle_lib):
    """The gwesp/gwdsp statistics (ids 24, 26-30), whose tie terms the
    statistics kernel adds window by window, on networks of several
    windows (LDS-resident rows), against the oracle bit for bit."""
    from cusna_amd.effects import gwesp_weights

    rng = np.random.default_rng(n)
    M = 3 if n < 1000 else 1  # the oracle's gwdsp sums take ~5 s per network at n = 1000
    nets = (rng.random((M, n, n)) < 8.0 / n).astype(np.uint8)
    for c in range(M):
        np.fill_diagonal(nets[c], 0)
    ids = np.array([24, 26, 27, 28, 29, 30, 0, 2], np.int32)
    k = len(ids)
    wtab = np.array([gwesp_weights(69, n)] * k)
    print(f"case gwesp family stats n={n}", flush=True)
    got = gpu_lib.network_stats_packed(nets, ids, wtab=wtab)
    for c in range(M):
        ref = oracle_lib.network_stats(nets[c], ids, np.zeros(k), np.zeros(k), np.zeros((k, n)),
                                       wtab=wtab)
        np.testing.assert_array_equal(got[c].view(np.uint64), ref.view(np.uint64))


def test_period_plan_reuse(gpu_lib, oracle_lib):
    """A device-resident plan run twice with new parameters equals one-shot calls."""
    from cusna_amd.workloads import n500_problem

    prob = n500_problem(M=8, n=150)
    plan = gpu_lib.PeriodPlan(prob)
    plan.run(prob.seed)
    a = plan.fetch()
    ref = oracle_lib.simulate(prob)
    assert_period_equal(a, ref, label="plan run 1")
    th2 = np.array(prob.theta) + 0.1
    plan.set_params(th2, 7.0)
    plan.run(99)
    b = plan.fetch()
    prob.theta, prob.lam, prob.seed = th2, 7.0, 99
    ref2 = oracle_lib.simulate(prob)
    assert_period_equal(b, ref2, label="plan run 2")
    assert plan.kernel_ms() > 0.0
    plan.close()


def test_gumbel_float_bound(gpu_lib):
    """The float32 Gumbel bound that prunes alternatives stays far inside its
    2^-8 slack over the whole 32-bit uniform range: every x in steps of 97
    (44 million values) plus both ends and the u = 1/2 switch-over."""
    worst = gpu_lib.gumbel_bound_error(0, 97, (1 << 32) // 97)
    for x0 in (0, 0x7FFFF000, 0x80000000 - 4096, 0xFFFFF000, 0xFEFFF000):
        worst = max(worst, gpu_lib.gumbel_bound_error(x0, 1, 4096))
    print("max |approx - exact| =", worst)
    assert worst < 2.0 ** -12


def test_softmax_weight_exponential(gpu_lib):
    """The period kernel's own exp (table of 2^(j/32), degree-6 polynomial)
    against the host's exp: within 4 ulps (relative 1e-15) where the weights
    are normal numbers, on random arguments, on the boundaries of its range
    reduction (multiples of ln2/32 and half-steps) and at the ends of its range;
    below the normal range it may differ by a denormal rounding only."""
    rng = np.random.default_rng(17)
    grid = np.arange(-32 * 40, 32 * 4 + 1) * (np.log(2.0) / 32)
    xs = np.concatenate([rng.uniform(-700.0, 4.0, 200_000), rng.uniform(-1.0, 1.0, 100_000),
                         grid, grid + np.log(2.0) / 64, grid - np.log(2.0) / 64,
                         np.nextafter(grid, 1.0), np.nextafter(grid, -1.0),
                         [0.0, -0.0, 4.0, 3.15, -1e-300, 1e-300, -708.0, -700.0]])
    got = gpu_lib.exp_weight(xs)
    ref = np.exp(xs)
    rel = np.abs(got - ref) / ref
    print("max relative error", rel.max())
    assert rel.max() <= 1e-15, (xs[np.argmax(rel)], rel.max())
    low = np.array([-720.0, -740.0, -745.0, -800.0, -1000.0])
    assert np.all(np.abs(gpu_lib.exp_weight(low) - np.exp(low)) <= 1e-310)
