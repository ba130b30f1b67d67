"""Loader and thin Python bindings of libcusna_amd.so (include/libcusna.h).

The library is the product: every function here calls into it. There is no
CPU fallback; if the library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (CusnaError, SaomProblem, SaomResult, cusna_mom_options, cusna_mom_result,
                  pack_rows, run_period)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libcusna_amd.so")
_vp = C.c_void_p

# Every symbol include/libcusna.h declares.
EXPORTS = (
    "cusna_saom_simulate_period", "cusna_saom_simulate_period_replay", "cusna_pack_bits",
    "cusna_philox4x32", "cusna_network_stats", "cusna_network_stats_packed",
    "cusna_behavior_stats", "cusna_abi_version", "cusna_version_string", "cusna_has_cuda",
    "cusna_last_error", "cusna_period_plan_create", "cusna_period_plan_set_params",
    "cusna_period_plan_run", "cusna_period_plan_sync", "cusna_period_plan_fetch",
    "cusna_period_plan_kernel_ms", "cusna_period_plan_shape", "cusna_period_plan_destroy",
    "cusna_gumbel_bound_check", "cusna_exp_weight_eval", "cusna_mom_estimate", "cusna_period_plan_autotune",
    "cusna_saom_simulate_groups", "cusna_mlp_create", "cusna_mlp_param_count",
    "cusna_mlp_layer_shape", "cusna_mlp_get_params", "cusna_mlp_set_params",
    "cusna_mlp_forward_backward", "cusna_mlp_train", "cusna_mlp_predict", "cusna_mlp_destroy",
    "cusna_period_plan_specialize", "cusna_period_plan_specialized", "cusna_gof_distribution",
    "cusna_period_plan_gof", "cusna_saom_simulate_multi", "cusna_cross_stats",
    "cusna_multi_plan_create", "cusna_multi_plan_set_params", "cusna_multi_plan_run",
    "cusna_multi_plan_fetch", "cusna_multi_plan_dist", "cusna_multi_plan_destroy",
    "cusna_multi_plan_set_stats", "cusna_multi_plan_stats", "cusna_multi_plan_specialized",
    "cusna_saom_simulate_period_noalloc",
)

_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Load the gfx950 library (built by ``cusna_amd.build``); raise if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m cusna_amd.build` "
                           "(the GPU library has no CPU fallback)")
    L = C.CDLL(p)
    L.cusna_saom_simulate_period.restype = C.c_int
    L.cusna_saom_simulate_period.argtypes = [_vp, _vp]
    L.cusna_saom_simulate_period_noalloc.restype = C.c_int
    L.cusna_saom_simulate_period_noalloc.argtypes = [_vp, _vp]
    L.cusna_saom_simulate_period_replay.restype = C.c_int
    L.cusna_saom_simulate_period_replay.argtypes = [_vp, _vp, _vp, _vp, C.c_int32]
    L.cusna_pack_bits.restype = C.c_int
    L.cusna_pack_bits.argtypes = [_vp, C.c_int32, _vp, _vp, _vp]
    L.cusna_philox4x32.restype = C.c_int
    L.cusna_philox4x32.argtypes = [_vp, _vp, _vp]
    L.cusna_network_stats.restype = C.c_int
    L.cusna_network_stats.argtypes = [_vp, C.c_int32, C.c_int32] + [_vp] * 10 + [C.c_int32, _vp]
    L.cusna_network_stats_packed.restype = C.c_int
    L.cusna_network_stats_packed.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32] + [_vp] * 10 + [
        C.c_int32, _vp, _vp, _vp]
    L.cusna_behavior_stats.restype = C.c_int
    L.cusna_behavior_stats.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp,
                                       C.c_double, C.c_double, C.c_double, _vp]
    L.cusna_abi_version.restype = C.c_int32
    L.cusna_version_string.restype = C.c_char_p
    L.cusna_has_cuda.restype = C.c_int32
    L.cusna_last_error.restype = C.c_char_p
    L.cusna_period_plan_create.restype = C.c_int
    L.cusna_period_plan_create.argtypes = [_vp, C.c_int32, C.POINTER(_vp)]
    L.cusna_period_plan_set_params.restype = C.c_int
    L.cusna_period_plan_set_params.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.cusna_period_plan_run.restype = C.c_int
    L.cusna_period_plan_run.argtypes = [_vp, C.c_uint64, C.c_int32]
    L.cusna_period_plan_sync.restype = C.c_int
    L.cusna_period_plan_sync.argtypes = [_vp]
    L.cusna_period_plan_fetch.restype = C.c_int
    L.cusna_period_plan_fetch.argtypes = [_vp, _vp]
    L.cusna_period_plan_kernel_ms.restype = C.c_double
    L.cusna_period_plan_kernel_ms.argtypes = [_vp]
    L.cusna_period_plan_shape.restype = C.c_int32
    L.cusna_period_plan_shape.argtypes = [_vp]
    L.cusna_exp_weight_eval.restype = C.c_int
    L.cusna_exp_weight_eval.argtypes = [_vp, C.c_int32, _vp]
    L.cusna_gumbel_bound_check.restype = C.c_int
    L.cusna_gumbel_bound_check.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.POINTER(C.c_double)]
    L.cusna_period_plan_destroy.restype = None
    L.cusna_period_plan_destroy.argtypes = [_vp]
    L.cusna_period_plan_autotune.restype = C.c_int
    L.cusna_period_plan_autotune.argtypes = [_vp, C.c_int32, C.POINTER(C.c_int32)]
    L.cusna_saom_simulate_groups.restype = C.c_int
    L.cusna_saom_simulate_groups.argtypes = [_vp, C.c_int32, _vp]
    L.cusna_saom_simulate_multi.restype = C.c_int
    L.cusna_saom_simulate_multi.argtypes = [_vp, _vp]
    L.cusna_cross_stats.restype = C.c_int
    L.cusna_cross_stats.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, _vp]
    L.cusna_multi_plan_create.restype = C.c_int
    L.cusna_multi_plan_create.argtypes = [_vp, C.c_int32, _vp, _vp]
    L.cusna_multi_plan_set_params.restype = C.c_int
    L.cusna_multi_plan_set_params.argtypes = [_vp, _vp, _vp]
    L.cusna_multi_plan_run.restype = C.c_int
    L.cusna_multi_plan_run.argtypes = [_vp, C.c_uint64]
    L.cusna_multi_plan_fetch.restype = C.c_int
    L.cusna_multi_plan_fetch.argtypes = [_vp, _vp]
    L.cusna_multi_plan_dist.restype = C.c_int
    L.cusna_multi_plan_dist.argtypes = [_vp, _vp]
    L.cusna_multi_plan_set_stats.restype = C.c_int
    L.cusna_multi_plan_set_stats.argtypes = [_vp, C.c_int32, _vp]
    L.cusna_multi_plan_stats.restype = C.c_int
    L.cusna_multi_plan_stats.argtypes = [_vp, _vp]
    L.cusna_multi_plan_specialized.restype = C.c_int32
    L.cusna_multi_plan_specialized.argtypes = [_vp]
    L.cusna_multi_plan_destroy.restype = C.c_int
    L.cusna_multi_plan_destroy.argtypes = [_vp]
    L.cusna_gof_distribution.restype = C.c_int
    L.cusna_gof_distribution.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_int32,
                                         _vp]
    L.cusna_period_plan_gof.restype = C.c_int
    L.cusna_period_plan_gof.argtypes = [_vp, C.c_int32, C.c_int32, _vp, C.c_int32, _vp]
    L.cusna_period_plan_specialize.restype = C.c_int
    L.cusna_period_plan_specialize.argtypes = [_vp, C.c_int32]
    L.cusna_period_plan_specialized.restype = C.c_int32
    L.cusna_period_plan_specialized.argtypes = [_vp]
    L.cusna_mlp_create.restype = C.c_int
    L.cusna_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                   C.POINTER(_vp)]
    L.cusna_mlp_param_count.restype = C.c_int64
    L.cusna_mlp_param_count.argtypes = [_vp]
    L.cusna_mlp_layer_shape.restype = C.c_int
    L.cusna_mlp_layer_shape.argtypes = [_vp, C.c_int32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]
    for f in ("cusna_mlp_get_params", "cusna_mlp_set_params"):
        getattr(L, f).restype = C.c_int
        getattr(L, f).argtypes = [_vp, _vp]
    L.cusna_mlp_forward_backward.restype = C.c_int
    L.cusna_mlp_forward_backward.argtypes = [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double)]
    L.cusna_mlp_train.restype = C.c_int
    L.cusna_mlp_train.argtypes = [_vp, _vp, _vp, C.c_int64, C.c_int32, C.c_float, C.c_uint64,
                                  C.POINTER(C.c_double)]
    L.cusna_mlp_predict.restype = C.c_int
    L.cusna_mlp_predict.argtypes = [_vp, _vp, C.c_int64, _vp]
    L.cusna_mlp_destroy.restype = None
    L.cusna_mlp_destroy.argtypes = [_vp]
    L.cusna_mom_estimate.restype = C.c_int
    L.cusna_mom_estimate.argtypes = [_vp, _vp, _vp, _vp, _vp, C.POINTER(cusna_mom_options),
                                     C.POINTER(cusna_mom_result)]
    if path is None:
        _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def _check(name: str, st: int) -> None:
    if st != 0:
        detail = load().cusna_last_error().decode(errors="replace")
        raise CusnaError(name, st, detail)


def has_device() -> bool:
    return bool(load().cusna_has_cuda())


def simulate_period(prob: SaomProblem, want_end_nets: bool = False) -> dict:
    """One period for prob.M chains on the GPU (cusna_saom_simulate_period)."""
    L = load()
    out = run_period(L.cusna_saom_simulate_period, prob, want_end_nets, "cusna_saom_simulate_period")
    return out


def simulate_period_noalloc(prob: SaomProblem, want_end_nets: bool = False) -> dict:
    """One period without device allocation (cusna_saom_simulate_period_noalloc):
    the call's device buffers sit in the library's static per-GPU array of
    CUSNA_NOALLOC_SCRATCH_BYTES (64 MiB); larger calls raise CusnaError (status 2)."""
    name = "cusna_saom_simulate_period_noalloc"
    L = load()
    try:
        return run_period(L.cusna_saom_simulate_period_noalloc, prob, want_end_nets, name)
    except CusnaError as e:
        raise CusnaError(name, e.status, L.cusna_last_error().decode(errors="replace")) from None


def simulate_multi(prob, want_end_nets: bool = True) -> dict:
    """One multi-network period for prob.M chains on the GPU
    (cusna_saom_simulate_multi); prob is a MultiProblem."""
    from .abi import run_multi

    return run_multi(load().cusna_saom_simulate_multi, prob, want_end_nets,
                     "cusna_saom_simulate_multi")


def cross_stats(a_batch, wp, t: int, parm: float = 1.0) -> np.ndarray:
    """Cross-network statistic of M dense networks against wp (cusna_cross_stats)."""
    a = np.ascontiguousarray(np.asarray(a_batch) != 0, np.uint8)
    M, n, _ = a.shape
    w = np.ascontiguousarray(np.asarray(wp) != 0, np.uint8)
    out = np.zeros(M, np.float64)
    _check("cusna_cross_stats",
           load().cusna_cross_stats(_p(a), _p(w), n, M, int(t), float(parm), _p(out)))
    return out


class MultiPlan:
    """Device-resident multi-network problem (cusna_multi_plan_*). kdist
    (D, n, n), when given, enables the per-network period distances."""

    def __init__(self, prob, want_end_nets: bool = False, kdist=None):
        from .abi import pack_rows

        self.L = load()
        self.prob = prob
        self._c = prob.build()
        self._kd = None
        if kdist is not None:
            kd = np.asarray(kdist)
            self._kd = np.ascontiguousarray(np.concatenate([pack_rows(kd[d])
                                                            for d in range(prob.D)]))
        self.h = _vp()
        _check("cusna_multi_plan_create",
               self.L.cusna_multi_plan_create(C.byref(self._c), int(want_end_nets),
                                              _p(self._kd), C.byref(self.h)))
        self.want_end_nets = want_end_nets

    def set_params(self, theta, lam) -> None:
        M, k, D = self.prob.M, self.prob.k, self.prob.D
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float32), (M, k)))
        lm = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float32), (M, D)))
        _check("cusna_multi_plan_set_params",
               self.L.cusna_multi_plan_set_params(self.h, _p(th), _p(lm)))

    def run(self, seed: int) -> None:
        _check("cusna_multi_plan_run",
               self.L.cusna_multi_plan_run(self.h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def fetch(self, want_end_nets: bool = False) -> dict:
        from .abi import MultiResult

        res = MultiResult(self.prob.n, self.prob.k, self.prob.M, self.prob.D, want_end_nets)
        _check("cusna_multi_plan_fetch", self.L.cusna_multi_plan_fetch(self.h, C.byref(res.c)))
        return res.as_dict()

    def set_stats(self, d: int, type_ids, p1, p2, vmat, vmiss=None, wtab=None, od0=None,
                  id0=None, symflag: int = 0, keep=None, force=None, cross_wp=None) -> None:
        """Statistics of network d's end states (cusna_multi_plan_set_stats):
        its effects in list order; cross_wp (k, n, n) predictor networks of
        the cross effects (other rows unused); keep/force (n, n) masks."""
        from .abi import ABI_VERSION, cusna_multi_stat_spec, pack_rows

        n = self.prob.n
        k = len(type_ids)
        hold = []

        def arr(x, dt):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, dt))
            hold.append(a)
            return a.ctypes.data

        spec = cusna_multi_stat_spec()
        spec.struct_version = ABI_VERSION
        spec.k = k
        spec.type_ids = arr(type_ids, np.int32)
        spec.p1, spec.p2 = arr(p1, np.float64), arr(p2, np.float64)
        spec.vmat = arr(np.reshape(vmat, (k, n)), np.float64)
        spec.vmiss = arr(None if vmiss is None else np.asarray(vmiss) != 0, np.uint8)
        spec.wtab = arr(wtab, np.float64)
        spec.od0, spec.id0 = arr(od0, np.int32), arr(id0, np.int32)
        spec.symflag = int(symflag)
        spec.keep = None if keep is None else arr(pack_rows(np.asarray(keep)), np.uint32)
        spec.force = None if force is None else arr(pack_rows(np.asarray(force)), np.uint32)
        if cross_wp is not None:
            cw = np.asarray(cross_wp)
            spec.cross_wp = arr(np.concatenate([pack_rows(cw[q]) for q in range(k)]), np.uint32)
        _check("cusna_multi_plan_set_stats",
               self.L.cusna_multi_plan_set_stats(self.h, int(d), C.byref(spec)))

    def stats(self) -> np.ndarray:
        out = np.zeros((self.prob.M, self.prob.k), np.float64)
        _check("cusna_multi_plan_stats", self.L.cusna_multi_plan_stats(self.h, _p(out)))
        return out

    @property
    def specialized(self) -> bool:
        """Whether the last run launched the specialised kernel."""
        return bool(self.L.cusna_multi_plan_specialized(self.h))

    def dist(self) -> np.ndarray:
        out = np.zeros((self.prob.D, self.prob.M), np.int32)
        _check("cusna_multi_plan_dist", self.L.cusna_multi_plan_dist(self.h, _p(out)))
        return out

    def close(self) -> None:
        if self.h:
            self.L.cusna_multi_plan_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate_period_replay(prob: SaomProblem, log_u0: np.ndarray, loglog_u: np.ndarray,
                           want_end_nets: bool = False) -> dict:
    L = load()
    lu0 = np.ascontiguousarray(log_u0, np.float64)
    llu = np.ascontiguousarray(loglog_u, np.float64)
    M, smax = lu0.shape
    if llu.shape != (M, smax, prob.n) or M != prob.M:
        raise ValueError("replay arrays must be (M, smax) and (M, smax, n)")
    return run_period(L.cusna_saom_simulate_period_replay, prob, want_end_nets,
                      "cusna_saom_simulate_period_replay",
                      extra_args=(_p(lu0), _p(llu), int(smax)))


def network_stats(a, type_ids, p1=None, p2=None, vmat=None, vmiss=None, dyn_flags=None, vb=None,
                  od0=None, id0=None, wtab=None, symflag=0) -> np.ndarray:
    """Statistics of one dense network (cusna_network_stats)."""
    a = np.ascontiguousarray(np.asarray(a) != 0, np.uint8)
    n = a.shape[0]
    t = np.ascontiguousarray(type_ids, np.int32)
    k = len(t)
    arr = lambda x, dt, shape: None if x is None else np.ascontiguousarray(
        np.asarray(x, dt).reshape(shape))
    p1 = arr(p1 if p1 is not None else np.zeros(k), np.float64, (k,))
    p2 = arr(p2 if p2 is not None else np.zeros(k), np.float64, (k,))
    vmat = arr(vmat if vmat is not None else np.zeros((k, n)), np.float64, (k, n))
    wtab = arr(wtab if wtab is not None else np.zeros((k, n + 1)), np.float64, (k, n + 1))
    vmiss = arr(None if vmiss is None else (np.asarray(vmiss) != 0), np.uint8, (k, n))
    dyn = arr(dyn_flags, np.int32, (k,))
    vb = arr(vb, np.float64, (n,))
    od0 = arr(od0, np.int32, (n,))
    id0 = arr(id0, np.int32, (n,))
    out = np.zeros(k, np.float64)
    st = load().cusna_network_stats(_p(a), n, k, _p(t), _p(p1), _p(p2), _p(vmat), _p(vmiss),
                                    _p(dyn), _p(vb), _p(od0), _p(id0), _p(wtab), int(symflag),
                                    _p(out))
    _check("cusna_network_stats", st)
    return out


def network_stats_packed(nets, type_ids, p1=None, p2=None, vmat=None, vmiss=None,
                         dyn_flags=None, vb=None, od0=None, id0=None, wtab=None, symflag=0,
                         keep=None, force=None) -> np.ndarray:
    """Statistics of M networks (M, n, n) in one launch (cusna_network_stats_packed)."""
    nets = np.asarray(nets)
    M, n, _ = nets.shape
    rows = np.ascontiguousarray(np.stack([pack_rows(nets[c]) for c in range(M)]), np.uint32)
    t = np.ascontiguousarray(type_ids, np.int32)
    k = len(t)
    f64 = lambda x, shape: np.ascontiguousarray(np.asarray(x, np.float64).reshape(shape))
    p1 = f64(p1 if p1 is not None else np.zeros(k), (k,))
    p2 = f64(p2 if p2 is not None else np.zeros(k), (k,))
    vmat = f64(vmat if vmat is not None else np.zeros((k, n)), (k, n))
    wtab = f64(wtab if wtab is not None else np.zeros((k, n + 1)), (k, n + 1))
    vm = None if vmiss is None else np.ascontiguousarray(np.asarray(vmiss) != 0, np.uint8)
    dyn = None if dyn_flags is None else np.ascontiguousarray(dyn_flags, np.int32)
    vbb = None if vb is None else f64(vb, (n,))
    o0 = None if od0 is None else np.ascontiguousarray(od0, np.int32)
    i0 = None if id0 is None else np.ascontiguousarray(id0, np.int32)
    kp = None if keep is None else pack_rows(keep)
    fc = None if force is None else pack_rows(force)
    out = np.zeros((M, k), np.float64)
    st = load().cusna_network_stats_packed(_p(rows), n, M, k, _p(t), _p(p1), _p(p2), _p(vmat),
                                           _p(vm), _p(dyn), _p(vbb), _p(o0), _p(i0), _p(wtab),
                                           int(symflag), _p(kp), _p(fc), _p(out))
    _check("cusna_network_stats_packed", st)
    return out


def behavior_stats(z, bty, xnet, bkdist, zbar, brange, bsimmean) -> np.ndarray:
    """Behaviour statistics of M behaviour vectors (cusna_behavior_stats)."""
    z = np.ascontiguousarray(np.atleast_2d(z), np.int32)
    M, n = z.shape
    bty = np.ascontiguousarray(bty, np.int32)
    xnet = np.ascontiguousarray(np.asarray(xnet) != 0, np.uint8)
    bk = None if bkdist is None else np.ascontiguousarray(bkdist, np.int32)
    out = np.zeros((M, len(bty)), np.float64)
    st = load().cusna_behavior_stats(_p(z), n, M, len(bty), _p(bty), _p(xnet), _p(bk),
                                     float(zbar), float(brange), float(bsimmean), _p(out))
    _check("cusna_behavior_stats", st)
    return out


class Mlp:
    """Surrogate perceptron on the BF16 matrix cores (cusna_mlp_*)."""

    def __init__(self, p_in: int, p_out: int, width: int = 256, batch: int = 1024,
                 seed: int = 1):
        self.L = load()
        self.h = _vp()
        _check("cusna_mlp_create", self.L.cusna_mlp_create(p_in, p_out, width, batch,
                                                            int(seed), C.byref(self.h)))
        self.p_in, self.p_out, self.width, self.batch = p_in, p_out, width, batch
        self.n_params = int(self.L.cusna_mlp_param_count(self.h))
        K, N = C.c_int32(), C.c_int32()
        self.layer_shapes = []
        for layer in range(4):
            _check("cusna_mlp_layer_shape",
                   self.L.cusna_mlp_layer_shape(self.h, layer, C.byref(K), C.byref(N)))
            self.layer_shapes.append((K.value, N.value))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        _check("cusna_mlp_get_params", self.L.cusna_mlp_get_params(self.h, _p(out)))
        return out

    def set_params(self, flat) -> None:
        f = np.ascontiguousarray(flat, np.float32)
        assert f.shape == (self.n_params,)
        _check("cusna_mlp_set_params", self.L.cusna_mlp_set_params(self.h, _p(f)))

    def forward_backward(self, x, y) -> dict:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        assert x.shape == (self.batch, self.p_in) and y.shape == (self.batch, self.p_out)
        out = np.zeros((self.batch, self.layer_shapes[3][1]), np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check("cusna_mlp_forward_backward",
               self.L.cusna_mlp_forward_backward(self.h, _p(x), _p(y), _p(out), _p(g),
                                                 C.byref(loss)))
        return {"out": out, "grads": g, "loss": loss.value}

    def train(self, x_all, y_all, steps: int, lr: float = 1e-3, seed: int = 7) -> float:
        x = np.ascontiguousarray(x_all, np.float32)
        y = np.ascontiguousarray(y_all, np.float32)
        loss = C.c_double()
        _check("cusna_mlp_train", self.L.cusna_mlp_train(self.h, _p(x), _p(y), x.shape[0],
                                                         int(steps), float(lr), int(seed),
                                                         C.byref(loss)))
        return loss.value

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((x.shape[0], 2 * self.p_out), np.float32)
        _check("cusna_mlp_predict", self.L.cusna_mlp_predict(self.h, _p(x), x.shape[0], _p(out)))
        return out

    def close(self) -> None:
        if self.h:
            self.L.cusna_mlp_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gof_distribution(a_batch, kind: int, levls) -> np.ndarray:
    """sienaGOF auxiliary distributions of M dense networks (cusna_gof_distribution)."""
    a = np.ascontiguousarray(np.asarray(a_batch) != 0, np.uint8)
    M, n, _ = a.shape
    lv = np.ascontiguousarray(levls, np.int32)
    out = np.zeros((M, len(lv) + (1 if kind == 2 else 0)), np.float64)
    _check("cusna_gof_distribution",
           load().cusna_gof_distribution(_p(a), n, M, int(kind), _p(lv), len(lv), _p(out)))
    return out


def simulate_groups(problems: list, want_end_nets: bool = False) -> list:
    """Several period problems in one call (cusna_saom_simulate_groups)."""
    L = load()
    G = len(problems)
    built = [pr.build() for pr in problems]
    res = [SaomResult(pr.n, pr.k, pr.M, pr.kb, want_end_nets, nrate=pr.nrate) for pr in problems]
    parr = (_vp * G)(*[C.addressof(b) for b in built])
    rarr = (_vp * G)(*[C.addressof(r.c) for r in res])
    _check("cusna_saom_simulate_groups", L.cusna_saom_simulate_groups(parr, G, rarr))
    return [r.as_dict() for r in res]


def mom_estimate(problems: list, targets, theta0, mode: int, n1: int, nD: int, batch2: int,
                 n3: int, n2: list, firstg: float = 0.4, diagonalize: float = 0.2,
                 seed: int = 1234, xnets: list | None = None, rate_consts=None,
                 use_graph: bool = True, max_repeats: int = 0, conv_tol: float = 0.1,
                 surrogate: bool = False, surrogate_pairs_log2: int = 0,
                 surrogate_steps: int = 0, surrogate_rounds: int = 0,
                 surrogate_corrections: int = -1, specialize: int = 0,
                 verbose: bool = False, surrogate_warmup: int = 0) -> dict:
    """Method-of-moments estimate on the device (cusna_mom_estimate).

    ``problems``: one SaomProblem per period (inputs only); ``xnets``: the
    co-evolution networks of the behaviour statistics, one dense (n, n) per
    period (mode 2)."""
    L = load()
    P = len(problems)
    targets = np.ascontiguousarray(targets, dtype=np.float64)
    p = targets.shape[0]
    built = [pr.build() for pr in problems]
    parr = (_vp * P)(*[C.addressof(b) for b in built])
    xarr = None
    xs = []
    if xnets is not None:
        xs = [np.ascontiguousarray(np.asarray(x) != 0, dtype=np.uint8) for x in xnets]
        xarr = (_vp * P)(*[x.ctypes.data for x in xs])
    rc = None if rate_consts is None else np.ascontiguousarray(rate_consts, dtype=np.float64)
    th0 = np.ascontiguousarray(theta0, dtype=np.float64)
    if th0.shape != (p,):
        raise ValueError(f"theta0 has shape {th0.shape}, expected ({p},)")
    if len(n2) > 8:
        raise ValueError("at most 8 phase-2 sub-phases")
    o = cusna_mom_options()
    o.struct_version = 1
    o.mode, o.P, o.nsub = int(mode), P, len(n2)
    for i, v in enumerate(n2):
        o.n2[i] = int(v)
    o.n1, o.nD, o.batch2, o.n3 = int(n1), int(nD), int(batch2), int(n3)
    o.firstg, o.diagonalize = float(firstg), float(diagonalize)
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.use_graph, o.max_repeats, o.conv_tol = int(use_graph), int(max_repeats), float(conv_tol)
    o.surrogate, o.surrogate_pairs_log2 = int(surrogate), int(surrogate_pairs_log2)
    o.surrogate_steps, o.surrogate_rounds = int(surrogate_steps), int(surrogate_rounds)
    o.surrogate_corrections = int(surrogate_corrections)
    o.specialize = int(specialize)
    o.verbose = int(bool(verbose))
    o.surrogate_warmup = int(surrogate_warmup)
    if mode == 2:
        pr0 = problems[0]
        o.beh_zbar, o.beh_brange, o.beh_bsimmean = (float(pr0.zbar), float(pr0.brange),
                                                     float(pr0.bsimmean))
    out = {name: np.zeros(p) for name in ("theta", "se", "tconv", "sim_means")}
    out["D"] = np.zeros((p, p))
    out["msf"] = np.zeros((p, p))
    out["phase3_stats"] = np.zeros((int(n3), p))
    out["phase3_sp"] = np.zeros((P, int(n3), p))
    out["phase3_scp"] = np.zeros((P, int(n3), p))
    if mode == 1:
        out["rate_est"] = np.zeros(P)
        out["rate_sd"] = np.zeros(P)
    r = cusna_mom_result()
    r.struct_version = 1
    for name, a in out.items():
        setattr(r, name, a.ctypes.data)
    _check("cusna_mom_estimate",
           L.cusna_mom_estimate(parr, xarr, targets.ctypes.data, _p(rc), th0.ctypes.data,
                                C.byref(o), C.byref(r)))
    del built, xs
    out.update(tconv_max=r.tconv_max, max_abs_tconv=r.max_abs_tconv, n_sims=int(r.n_sims),
               repeats=int(r.repeats), phase_seconds=list(r.phase_seconds),
               wall_seconds=r.wall_seconds)
    return out


def gumbel_bound_error(x0: int, stride: int, count: int) -> float:
    """Largest |float32 bound - float64 log(-log u)| on the device over a range
    of 32-bit uniforms (cusna_gumbel_bound_check)."""
    out = C.c_double()
    _check("cusna_gumbel_bound_check",
           load().cusna_gumbel_bound_check(x0 & 0xFFFFFFFF, stride & 0xFFFFFFFF, count,
                                           C.byref(out)))
    return out.value


def exp_weight(x) -> np.ndarray:
    """The period kernel's softmax-weight exponential on the device
    (cusna_exp_weight_eval)."""
    xs = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(xs)
    _check("cusna_exp_weight_eval", load().cusna_exp_weight_eval(_p(xs), len(xs), _p(out)))
    return out


def philox4x32(counter, key) -> np.ndarray:
    c = np.ascontiguousarray(counter, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    _check("cusna_philox4x32", load().cusna_philox4x32(_p(c), _p(k), _p(out)))
    return out


class PeriodPlan:
    """Device-resident period problem (cusna_period_plan_*)."""

    def __init__(self, prob: SaomProblem, want_end_nets: bool = False):
        self.L = load()
        self.prob = prob
        self._c = prob.build()
        self.h = _vp()
        _check("cusna_period_plan_create",
               self.L.cusna_period_plan_create(C.byref(self._c), int(want_end_nets),
                                               C.byref(self.h)))
        self.want_end_nets = want_end_nets

    def set_params(self, theta, lam, thetab=None, lamb=None) -> None:
        M, k = self.prob.M, self.prob.k
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float32), (M, k)))
        lm = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float32), (M,)))
        thb = lmb = None
        if self.prob.kb:
            thb = np.ascontiguousarray(np.broadcast_to(np.asarray(thetab, np.float32),
                                                       (M, self.prob.kb)))
            lmb = np.ascontiguousarray(np.broadcast_to(np.asarray(lamb, np.float32), (M,)))
        _check("cusna_period_plan_set_params",
               self.L.cusna_period_plan_set_params(self.h, _p(th), _p(lm), _p(thb), _p(lmb)))

    def run(self, seed: int, want_stats: bool = True) -> None:
        _check("cusna_period_plan_run",
               self.L.cusna_period_plan_run(self.h, int(seed) & 0xFFFFFFFFFFFFFFFF,
                                            int(want_stats)))

    def sync(self) -> None:
        _check("cusna_period_plan_sync", self.L.cusna_period_plan_sync(self.h))

    def fetch(self, want_end_nets: bool = False) -> dict:
        res = SaomResult(self.prob.n, self.prob.k, self.prob.M, self.prob.kb, want_end_nets,
                         nrate=self.prob.nrate)
        _check("cusna_period_plan_fetch", self.L.cusna_period_plan_fetch(self.h, C.byref(res.c)))
        return res.as_dict()

    def gof(self, kind: int, levls, masked: bool = True) -> np.ndarray:
        """GoF distributions of the end networks of the last run, computed in
        device memory (cusna_period_plan_gof)."""
        lv = np.ascontiguousarray(levls, np.int32)
        out = np.zeros((self.prob.M, len(lv) + (1 if kind == 2 else 0)), np.float64)
        _check("cusna_period_plan_gof",
               self.L.cusna_period_plan_gof(self.h, int(masked), int(kind), _p(lv), len(lv),
                                            _p(out)))
        return out

    def specialize(self, mode: int = 1) -> None:
        """Runtime-specialised period kernel: 1 on, 0 off, -1 by launch size."""
        _check("cusna_period_plan_specialize",
               self.L.cusna_period_plan_specialize(self.h, int(mode)))

    @property
    def specialized(self) -> bool:
        """Whether the last run launched the specialised kernel."""
        return bool(self.L.cusna_period_plan_specialized(self.h))

    def autotune(self, probe_chains: int = 0) -> int:
        """Pick the fastest block shape (cusna_period_plan_autotune); returns
        threads per chain."""
        out = C.c_int32()
        _check("cusna_period_plan_autotune",
               self.L.cusna_period_plan_autotune(self.h, int(probe_chains), C.byref(out)))
        return int(out.value)

    def kernel_ms(self) -> float:
        return float(self.L.cusna_period_plan_kernel_ms(self.h))

    @property
    def block_threads(self) -> int:
        return int(self.L.cusna_period_plan_shape(self.h))

    def close(self) -> None:
        if self.h:
            self.L.cusna_period_plan_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
