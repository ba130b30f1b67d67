"""ctypes access to the CPU oracle and to the reference engine built from source.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of bench.py, as the checker. The product package
never imports this module.

  oracle()    -> oracle/_build/libsaom_oracle.so (the C restatement)
  reference() -> oracle/_ref/libcusna_ref.so (cran/cusna
                 src/libcusna_saom_cpu.cpp compiled by oracle/Makefile), or
                 None where it was not built
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libsaom_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcusna_ref.so")

_vp = C.c_void_p


class OracleOpts(C.Structure):
    _fields_ = [("per_chain", C.c_int32), ("log_u0", _vp), ("loglog_u", _vp),
                ("smax", C.c_int32)]


def build(force: bool = False) -> None:
    """Run oracle/Makefile (the reference part only where its sources exist)."""
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


_cache: dict = {}


def oracle() -> C.CDLL:
    if "o" not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_saom_simulate_period.restype = C.c_int
        lib.oracle_saom_simulate_period.argtypes = [_vp, _vp, _vp]
        lib.oracle_network_stats.restype = C.c_int
        lib.oracle_network_stats.argtypes = [_vp, C.c_int32, C.c_int32] + [_vp] * 10 + [
            C.c_int32, _vp]
        lib.oracle_behavior_stats.restype = C.c_int
        lib.oracle_behavior_stats.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp,
                                              _vp, C.c_double, C.c_double, C.c_double, _vp]
        lib.oracle_philox4x32.restype = C.c_int
        lib.oracle_philox4x32.argtypes = [_vp, _vp, _vp]
        lib.oracle_replay_noise.restype = C.c_int
        lib.oracle_replay_noise.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, _vp,
                                            _vp]
        lib.oracle_set_threads.restype = C.c_int
        lib.oracle_set_threads.argtypes = [C.c_int]
        lib.oracle_saom_simulate_multi.restype = C.c_int
        lib.oracle_saom_simulate_multi.argtypes = [_vp, _vp, C.c_int32]
        lib.oracle_cross_stats.restype = C.c_int
        lib.oracle_cross_stats.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_double, _vp]
        lib.oracle_gof_distribution.restype = C.c_int
        lib.oracle_gof_distribution.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp,
                                                C.c_int32, _vp]
        _cache["o"] = lib
    return _cache["o"]


def reference():
    """The compiled reference engine, or None where oracle/_ref was not built."""
    if "r" not in _cache:
        lib = None
        if os.path.exists(REF_SO):
            lib = C.CDLL(REF_SO)
            lib.cusna_saom_simulate_period_cpu.restype = C.c_int
            lib.cusna_saom_simulate_period_cpu.argtypes = [_vp, _vp]
            lib.cusna_network_stats.restype = C.c_int
            lib.cusna_network_stats.argtypes = [_vp, C.c_int32, C.c_int32] + [_vp] * 10 + [
                C.c_int32, _vp]
            lib.cusna_behavior_stats.restype = C.c_int
            lib.cusna_behavior_stats.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp,
                                                 _vp, C.c_double, C.c_double, C.c_double, _vp]
            lib.cusna_saom_simulate_multi_cpu.restype = C.c_int
            lib.cusna_saom_simulate_multi_cpu.argtypes = [_vp, _vp]
            lib.cusna_cross_stats.restype = C.c_int
            lib.cusna_cross_stats.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32,
                                              C.c_double, _vp]
            lib.cusna_gof_distribution.restype = C.c_int
            lib.cusna_gof_distribution.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp,
                                                   C.c_int32, _vp]
        _cache["r"] = lib
    return _cache["r"]


def _p(a):
    return None if a is None else a.ctypes.data


def simulate(prob, want_end_nets=False, per_chain=False, replay=None, lib=None) -> dict:
    """Run the oracle period simulator. replay = (log_u0, loglog_u, smax)."""
    from cusna_amd.abi import run_period

    L = lib or oracle()
    opts = OracleOpts()
    opts.per_chain = 1 if per_chain else 0
    if replay is not None:
        lu0, llu, smax = replay
        opts.log_u0, opts.loglog_u, opts.smax = _p(lu0), _p(llu), int(smax)
    return run_period(L.oracle_saom_simulate_period, prob, want_end_nets, "oracle",
                      extra_args=(C.byref(opts),))


def simulate_reference(prob, want_end_nets=False) -> dict:
    from cusna_amd.abi import run_period

    R = reference()
    if R is None:
        raise FileNotFoundError("oracle/_ref not built")
    return run_period(R.cusna_saom_simulate_period_cpu, prob, want_end_nets, "reference")


def replay_noise(seed: int, M: int, n: int, smax: int):
    lu0 = np.zeros((M, smax), np.float64)
    llu = np.zeros((M, smax, n), np.float64)
    st = oracle().oracle_replay_noise(int(seed) & 0xFFFFFFFFFFFFFFFF, M, n, smax, _p(lu0),
                                      _p(llu))
    if st != 0:
        raise RuntimeError(f"oracle_replay_noise status {st}")
    return lu0, llu


def _stats_args(a, type_ids, p1, p2, vmat, vmiss, dyn_flags, vb, od0, id0, wtab, symflag):
    n = a.shape[0]
    k = len(type_ids)
    arrs = dict(
        a=np.ascontiguousarray(a, np.uint8), t=np.ascontiguousarray(type_ids, np.int32),
        p1=np.ascontiguousarray(p1, np.float64), p2=np.ascontiguousarray(p2, np.float64),
        vmat=np.ascontiguousarray(vmat, np.float64).reshape(k, n),
        vmiss=None if vmiss is None else np.ascontiguousarray(vmiss, np.uint8),
        dyn=None if dyn_flags is None else np.ascontiguousarray(dyn_flags, np.int32),
        vb=None if vb is None else np.ascontiguousarray(vb, np.float64),
        od0=None if od0 is None else np.ascontiguousarray(od0, np.int32),
        id0=None if id0 is None else np.ascontiguousarray(id0, np.int32),
        wtab=np.ascontiguousarray(wtab, np.float64))
    return n, k, arrs


def network_stats(a, type_ids, p1, p2, vmat, vmiss=None, dyn_flags=None, vb=None, od0=None,
                  id0=None, wtab=None, symflag=0, lib=None, fn="oracle_network_stats"):
    n = np.asarray(a).shape[0]
    if wtab is None:
        wtab = np.zeros((len(type_ids), n + 1))
    n, k, A = _stats_args(a, type_ids, p1, p2, vmat, vmiss, dyn_flags, vb, od0, id0, wtab,
                          symflag)
    out = np.zeros(k, np.float64)
    f = getattr(lib or oracle(), fn)
    st = f(_p(A["a"]), n, k, _p(A["t"]), _p(A["p1"]), _p(A["p2"]), _p(A["vmat"]),
           _p(A["vmiss"]), _p(A["dyn"]), _p(A["vb"]), _p(A["od0"]), _p(A["id0"]),
           _p(A["wtab"]), int(symflag), _p(out))
    if st != 0:
        raise RuntimeError(f"{fn} status {st}")
    return out


def behavior_stats(z, bty, xnet, bkdist, zbar, brange, bsimmean, lib=None,
                   fn="oracle_behavior_stats"):
    z = np.ascontiguousarray(np.atleast_2d(z), np.int32)
    M, n = z.shape
    bty = np.ascontiguousarray(bty, np.int32)
    xnet = np.ascontiguousarray(xnet, np.uint8)
    bk = None if bkdist is None else np.ascontiguousarray(bkdist, np.int32)
    out = np.zeros((M, len(bty)), np.float64)
    f = getattr(lib or oracle(), fn)
    st = f(_p(z), n, M, len(bty), _p(bty), _p(xnet), _p(bk), float(zbar), float(brange),
           float(bsimmean), _p(out))
    if st != 0:
        raise RuntimeError(f"{fn} status {st}")
    return out


def gof_distribution(a_batch, kind, levls, lib=None, fn="oracle_gof_distribution"):
    """Degree (kind 0 in, 1 out) or geodesic (kind 2) distributions of M dense
    networks (M, n, n); the reference's own function with lib=reference(),
    fn="cusna_gof_distribution"."""
    a = np.ascontiguousarray(np.asarray(a_batch) != 0, np.uint8)
    M, n, _ = a.shape
    lv = np.ascontiguousarray(levls, np.int32)
    nlev = len(lv)
    out = np.zeros((M, nlev + (1 if kind == 2 else 0)), np.float64)
    f = getattr(lib or oracle(), fn)
    st = f(_p(a), n, M, int(kind), _p(lv), nlev, _p(out))
    if st != 0:
        raise RuntimeError(f"{fn} status {st}")
    return out


def simulate_multi(prob, want_end_nets=True) -> dict:
    """The oracle's multi-network period (MultiProblem)."""
    from cusna_amd.abi import run_multi

    L = oracle()
    return run_multi(L.oracle_saom_simulate_multi, prob, want_end_nets, "oracle multi",
                     extra_args=(C.c_int32(prob.nd),))


def simulate_multi_reference(prob, want_end_nets=True) -> dict:
    from cusna_amd.abi import run_multi

    R = reference()
    if R is None:
        raise FileNotFoundError("oracle/_ref not built")
    return run_multi(R.cusna_saom_simulate_multi_cpu, prob, want_end_nets, "reference multi")


def cross_stats(a_batch, wp, t, parm=0.0, lib=None, fn="oracle_cross_stats"):
    """Cross-network statistic of M dense networks (M, n, n) against wp (n, n);
    the reference's own function with lib=reference(), fn="cusna_cross_stats"."""
    a = np.ascontiguousarray(np.asarray(a_batch) != 0, np.uint8)
    M, n, _ = a.shape
    w = np.ascontiguousarray(np.asarray(wp) != 0, np.uint8)
    out = np.zeros(M, np.float64)
    st = getattr(lib or oracle(), fn)(_p(a), _p(w), n, M, int(t), float(parm), _p(out))
    if st != 0:
        raise RuntimeError(f"{fn} status {st}")
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    oracle().oracle_philox4x32(_p(c), _p(k), _p(out))
    return out
