"""CPU tests of the drop-in boundary: layout, exports, argument checks."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "libcusna.h")).read()
    return sorted(set(re.findall(r"CUSNA_API\s+[\w\s\*]+?\b(cusna_\w+)\s*\(", text)))


def test_header_and_binding_list_agree():
    from cusna_amd.lib import EXPORTS

    assert declared_symbols() == sorted(EXPORTS)


def test_library_exports_every_declared_symbol():
    from cusna_amd import lib

    L = lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (cusna_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    for s in declared_symbols():
        assert hasattr(L, s)


def test_struct_layout_matches_reference_header():
    """Offsets of every field equal those of the reference libcusna.h
    (recorded in tests/golden/abi_offsets.txt from the reference header)."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    ours = subprocess.run([os.path.join(ROOT, "oracle", "_build", "abi_offsets")],
                          capture_output=True, text=True, check=True).stdout
    ref = open(os.path.join(ROOT, "tests", "golden", "abi_offsets.txt")).read()
    assert ours == ref


def test_ctypes_mirror_matches_c_layout():
    from cusna_amd.abi import cusna_saom_problem, cusna_saom_result

    ref = dict(line.split() for line in open(os.path.join(ROOT, "tests", "golden",
                                                          "abi_offsets.txt")))
    assert ctypes.sizeof(cusna_saom_problem) == int(ref["sizeof.problem"])
    assert ctypes.sizeof(cusna_saom_result) == int(ref["sizeof.result"])
    for name, _ in cusna_saom_problem._fields_:
        assert getattr(cusna_saom_problem, name).offset == int(ref[f"cusna_saom_problem.{name}"])
    for name, _ in cusna_saom_result._fields_:
        assert getattr(cusna_saom_result, name).offset == int(ref[f"cusna_saom_result.{name}"])


def test_introspection():
    from cusna_amd import lib

    L = lib.load()
    assert L.cusna_abi_version() == 1
    assert b"gfx950" in L.cusna_version_string()


def test_argument_errors_without_device_work():
    """Status codes of the reference's checks; none of these reaches a kernel."""
    from cusna_amd import lib
    from cusna_amd.abi import SaomResult, cusna_saom_problem
    from cusna_amd.workloads import n500_problem

    L = lib.load()
    assert L.cusna_saom_simulate_period(None, None) == 1
    prob = n500_problem(M=2, n=40)
    p = prob.build()
    res = SaomResult(40, prob.k, 2)
    p.struct_version = 2
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 2
    p = prob.build()
    p.W = 1
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 2
    p = prob.build()
    p.k = 25
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 2
    p = prob.build()
    p.nrate = 5  # above CUSNA_MAXRATE
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 2
    p = prob.build()
    p.nrate = 1  # rate effects without kinds / parameters
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 1
    kinds, alpha = (ctypes.c_int32 * 1)(3), (ctypes.c_float * 1)(0.1)
    p.rkind, p.ralpha = ctypes.addressof(kinds), ctypes.addressof(alpha)
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 5
    # network stats: the reference returns NULL_ARG for shape errors too
    assert L.cusna_network_stats(None, 3, 1, None, None, None, None, None, None, None, None,
                                 None, None, 0, None) == 1
    assert L.cusna_behavior_stats(None, 3, 1, 1, None, None, None, 0.0, 1.0, 0.0, None) == 1


def test_no_device_status_on_cpu():
    from cusna_amd import lib
    from cusna_amd.abi import SaomResult
    from cusna_amd.workloads import n500_problem

    L = lib.load()
    if L.cusna_has_cuda():
        pytest.skip("a device is present")
    prob = n500_problem(M=2, n=40)
    p = prob.build()
    res = SaomResult(40, prob.k, 2)
    assert L.cusna_saom_simulate_period(ctypes.byref(p), ctypes.byref(res.c)) == 3
    with pytest.raises(RuntimeError):
        lib.simulate_period(prob)


def test_pack_bits_layout():
    from cusna_amd import lib
    from cusna_amd.abi import pack_rows

    rng = np.random.default_rng(1)
    n = 77
    a = (rng.random((n, n)) < 0.2).astype(np.uint8)
    W = (n + 31) // 32
    out = np.zeros((n, W), np.uint32)
    inn = np.zeros((n, W), np.uint32)
    w = ctypes.c_int32()
    st = lib.load().cusna_pack_bits(a.ctypes.data, n, out.ctypes.data, inn.ctypes.data,
                                    ctypes.byref(w))
    assert st == 0 and w.value == W
    np.testing.assert_array_equal(out, pack_rows(a))
    np.testing.assert_array_equal(inn, pack_rows(a.T))
    # bit (j & 31) of word (j >> 5) is the tie to j
    i, j = np.argwhere(a)[0]
    assert (out[i, j >> 5] >> (j & 31)) & 1 == 1


def test_mom_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of cusna_mom_options / cusna_mom_result against the C
    layout of include/libcusna.h (no reference counterpart: the reference
    leaves these types opaque, libcusna.h:541-542)."""
    from cusna_amd.abi import cusna_mom_options, cusna_mom_result

    lines = []
    for T in (cusna_mom_options, cusna_mom_result):
        lines.append(f'printf("sizeof.{T.__name__} %zu\\n", sizeof({T.__name__}));')
        for name, _ in T._fields_:
            lines.append(f'printf("{T.__name__}.{name} %zu\\n", offsetof({T.__name__}, {name}));')
    src = tmp_path / "mom_offsets.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "libcusna.h"\n'
                   "int main(void) {\n" + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "mom_offsets"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                       text=True, check=True).stdout.split("\n")
               if line)
    for T in (cusna_mom_options, cusna_mom_result):
        assert ctypes.sizeof(T) == int(got[f"sizeof.{T.__name__}"])
        for name, _ in T._fields_:
            assert getattr(T, name).offset == int(got[f"{T.__name__}.{name}"]), name
