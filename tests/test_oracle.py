"""CPU tests of the oracle: pinned to the reference's fixtures, to the
reference engine compiled from source, and to known answers."""
import numpy as np
import pytest

from conftest import load_golden


def test_philox_known_answers(oracle_lib):
    # Random123 philox4x32-10 known-answer vectors (Salmon et al. 2011)
    kat = [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
           ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
           ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])]
    for ctr, key, want in kat:
        np.testing.assert_array_equal(oracle_lib.philox(ctr, key), np.array(want, np.uint32))


def test_reference_sim_fixture(oracle_lib):
    """The reference's frozen simulator outputs, bit for bit (test-sim.R:39-72)."""
    from test_gpu_parity import _sim_fixture_problems

    for name, prob, c in _sim_fixture_problems():
        r = oracle_lib.simulate(prob)
        for key in ("dist", "dist_stat", "steps"):
            np.testing.assert_array_equal(r[key], np.asarray(c[key]), err_msg=name)
        np.testing.assert_array_equal(r["stats"], np.asarray(c["stats"]), err_msg=name)
        np.testing.assert_array_equal(r["ntim"], np.asarray(c["ntim"]), err_msg=name)
        np.testing.assert_array_equal(r["scores"], np.asarray(c["scores"], np.float32),
                                      err_msg=name)
        if "bdist" in c:
            np.testing.assert_array_equal(r["bdist"], np.asarray(c["bdist"]))
            np.testing.assert_array_equal(r["z_end"], np.asarray(c["z_end"]))
            np.testing.assert_array_equal(r["bscores"], np.asarray(c["bscores"], np.float32))


def test_reference_engine_cases(oracle_lib):
    """Every synthetic case: oracle outputs equal the compiled reference's."""
    from cusna_amd.abi import unpack_rows
    from ref_cases_loader import spec_problem

    for case in load_golden("ref_cases.json.gz")["cases"]:
        prob = spec_problem(case["spec"])
        r = oracle_lib.simulate(prob, want_end_nets=True)
        o = case["out"]
        for key in ("dist", "dist_stat", "steps", "ntim", "stats"):
            np.testing.assert_array_equal(np.asarray(r[key]).ravel(), np.asarray(o[key]).ravel(),
                                          err_msg=f"{case['spec']['name']} {key}")
        np.testing.assert_array_equal(r["scores"], np.asarray(o["scores"], np.float32))
        ends = np.stack([unpack_rows(np.asarray(p, np.uint32), prob.n) for p in o["ends_packed"]])
        np.testing.assert_array_equal(r["a_ends"], ends)
        if prob.kb:
            np.testing.assert_array_equal(r["z_end"], np.asarray(o["z_end"]))
            np.testing.assert_array_equal(r["bdist"], np.asarray(o["bdist"]))


def test_reference_stats_cases(oracle_lib):
    from ref_cases_loader import beh_stats_inputs, stats_inputs

    g = load_golden("ref_cases.json.gz")
    for case in g["stats"]:
        a, kw = stats_inputs(case["spec"])
        ids = kw["type_ids"]
        got = []
        for lo in range(0, len(ids), 16):
            sl = slice(lo, lo + 16)
            got += oracle_lib.network_stats(a, ids[sl], kw["p1"][sl], kw["p2"][sl], kw["vmat"][sl],
                                            kw["vmiss"][sl], kw["dyn_flags"][sl], kw["vb"],
                                            kw["od0"], kw["id0"], kw["wtab"][sl], 0).tolist()
        np.testing.assert_array_equal(np.array(got), np.array(case["out"]))
    for case in g["beh"]:
        z, bty, xnet, bk, zbar, br, bs = beh_stats_inputs(case["spec"])
        np.testing.assert_array_equal(oracle_lib.behavior_stats(z, bty, xnet, bk, zbar, br, bs),
                                      np.array(case["out"]))


def test_against_compiled_reference_live(oracle_lib):
    """Where oracle/_ref is built (it needs the reference sources), compare on
    a fresh n = 500, seven-effect problem (3 chains) and a conditional run of
    the same effects at n = 200."""
    if oracle_lib.reference() is None:
        pytest.skip("oracle/_ref not built here")
    from cusna_amd.workloads import n500_problem

    cond = n500_problem(M=3, n=200, period_seed=29)
    cond.conditional, cond.target_dist, cond.lam = 1, 150, 1.0
    for prob in (n500_problem(M=3, n=500, period_seed=23), cond):
        a = oracle_lib.simulate(prob, want_end_nets=True)
        b = oracle_lib.simulate_reference(prob, want_end_nets=True)
        for key in a:
            np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert np.all(a["dist"] == 150)


def _large_cases():
    return load_golden("large_cases.json.gz")["cases"]


@pytest.mark.parametrize("idx", range(6))
def test_large_reference_cases(oracle_lib, idx):
    """The compiled reference's outputs at the configurations' own sizes
    (500 actors x 8 chains, 300-actor co-evolution, 1000 actors with twelve
    effects, groups of 1045-1188 actors): the oracle equals them bit for bit,
    end networks by SHA-256."""
    from large_cases import end_hashes, large_problem

    case = _large_cases()[idx]
    prob = large_problem(case["spec"])
    assert (prob.n, prob.k) == (case["n"], case["k"])
    r = oracle_lib.simulate(prob, want_end_nets=True)
    o = case["out"]
    for key in ("dist", "dist_stat", "steps", "ntim", "stats"):
        np.testing.assert_array_equal(np.asarray(r[key]).ravel(), np.asarray(o[key]).ravel(),
                                      err_msg=f"{case['spec']['name']} {key}")
    np.testing.assert_array_equal(r["scores"].ravel(), np.asarray(o["scores"], np.float32).ravel())
    assert end_hashes(r["a_ends"]) == o["ends_sha256"]
    if prob.kb:
        for key in ("z_end", "bdist"):
            np.testing.assert_array_equal(np.asarray(r[key]).ravel(), np.asarray(o[key]).ravel())
        np.testing.assert_array_equal(r["bscores"].ravel(),
                                      np.asarray(o["bscores"], np.float32).ravel())


def test_stats_closed_forms(oracle_lib):
    """Statistics against matrix algebra on random graphs: a dropped term, a
    transposed operand or a wrong divisor fails one of these."""
    rng = np.random.default_rng(7)
    for n in (5, 23, 60):
        a = (rng.random((n, n)) < 0.15).astype(np.uint8)
        np.fill_diagonal(a, 0)
        A = a.astype(np.int64)
        ids = np.array([0, 1, 2, 3, 4, 16, 23, 19], np.int32)
        k = len(ids)
        s = oracle_lib.network_stats(a, ids, np.zeros(k), np.zeros(k), np.zeros((k, n)))
        AA = A @ A
        ind, outd = A.sum(0), A.sum(1)
        want = [A.sum(), (A * A.T).sum(), (A * AA).sum(), np.trace(AA @ A) / 3.0,
                (ind ** 2).sum(), (outd ** 2).sum(), (A * AA).sum() / 6.0,
                (outd * (A * A.T).sum(1)).sum()]
        np.testing.assert_array_equal(s, np.array(want, float))


def test_hand_known_answers(oracle_lib):
    """test-stats.R:19-49: 3-node network and 4-node behaviour known answers."""
    a = np.zeros((3, 3), np.uint8)
    a[0, 1] = a[1, 0] = a[1, 2] = 1
    s = oracle_lib.network_stats(a, [0, 1, 2], np.zeros(3), np.zeros(3), np.zeros((3, 3)))
    np.testing.assert_array_equal(s, [3.0, 2.0, 0.0])
    z = np.array([[1, 2, 3, 2]])
    xnet = np.zeros((4, 4), np.uint8)
    xnet[0, 1:] = 1
    b = oracle_lib.behavior_stats(z, [0, 1], xnet, np.ones(4), 2.0, 2.0, 0.0)
    u = np.array([1, 2, 3, 2]) - 2.0
    np.testing.assert_array_equal(b[0], [u.sum(), (u ** 2).sum()])


def test_conditional_stops_at_target(oracle_lib):
    """A conditional run ends exactly at the observed distance (test-nets.R:42)."""
    from ref_cases_loader import case_specs, spec_problem

    spec = [s for s in case_specs() if s["name"] == "conditional"][0]
    r = oracle_lib.simulate(spec_problem(spec), want_end_nets=True)
    assert np.all(r["dist"] == spec["target_dist"])
    assert np.all(r["dist_stat"] == spec["target_dist"])


def test_thread_count_invariance(oracle_lib):
    """Results do not depend on the OpenMP thread count (test-nets.R:108-114)."""
    from cusna_amd.workloads import n500_problem

    prob = n500_problem(M=6, n=120)
    prev = oracle_lib.oracle().oracle_set_threads(1)
    a = oracle_lib.simulate(prob, want_end_nets=True)
    oracle_lib.oracle().oracle_set_threads(max(prev, 2))
    b = oracle_lib.simulate(prob, want_end_nets=True)
    for key in a:
        np.testing.assert_array_equal(a[key], b[key])


def test_pow_three_halves_table_matches_libm():
    """The device's x^1.5 (correctly rounded, plus the nine-entry table of
    degrees where the C library rounds the other way) equals the host libm's
    pow on every degree 0..8192."""
    import ctypes
    import os
    import subprocess
    import tempfile

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = r'''
#include <math.h>
#include <stdio.h>
#include "saom_common.h"
int main(void) { int bad = 0;
  for (int x = 0; x <= 8192; ++x)
    if (pow((double)x, 1.5) != saom::pow_three_halves_libm((double)x)) { printf("%d\n", x); ++bad; }
  return bad != 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "t.cpp")
        open(cpp, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I",
                        os.path.join(root, "cusna_amd", "csrc"), cpp, "-o", exe, "-lm"], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0, f"degrees that differ: {r.stdout}"
