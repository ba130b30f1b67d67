"""ctypes mirror of the libcusna C ABI (include/libcusna.h).

The structs follow the reference ABI field for field (cran/cusna
src/libcusna.h:90-208). ``SaomProblem`` owns the numpy arrays a call needs and
builds the C struct over them; ``SaomResult`` owns the output buffers.
``run_period`` drives any library symbol with the reference's signature
``status fn(const cusna_saom_problem*, cusna_saom_result*)``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

ABI_VERSION = 1
MAXEFF = 24
MAXBEFF = 6
MAX_N = 8192

STATUS = {
    0: "CUSNA_OK",
    1: "CUSNA_ERR_NULL_ARG",
    2: "CUSNA_ERR_SHAPE",
    3: "CUSNA_ERR_NO_CUDA",
    4: "CUSNA_ERR_CUDA",
    5: "CUSNA_ERR_UNSUPPORTED",
    6: "CUSNA_ERR_INTERNAL",
}


class CusnaError(RuntimeError):
    """A libcusna entry point returned a non-zero status."""

    def __init__(self, fn: str, status: int, detail: str = ""):
        self.status = status
        msg = f"{fn} failed (status {status} {STATUS.get(status, '?')})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


_p = C.c_void_p


class cusna_saom_problem(C.Structure):
    _fields_ = [
        ("struct_version", C.c_int32),
        ("n", C.c_int32), ("W", C.c_int32), ("k", C.c_int32), ("M", C.c_int32),
        ("threads", C.c_int32),
        ("x0_out", _p), ("x0_in", _p), ("sim_out", _p), ("sim_in", _p),
        ("lam", _p), ("theta", _p),
        ("type_ids", _p), ("etypes", _p), ("wtab", _p), ("p1", _p), ("p2", _p),
        ("vmat", _p), ("vmiss", _p), ("dyc_idx", _p), ("dyc", _p),
        ("int0", _p), ("int1", _p), ("int2", _p), ("dyn_flags", _p),
        ("kdist", _p), ("keep_out", _p), ("keep_in", _p), ("force_out", _p),
        ("force_in", _p), ("perm", _p),
        ("od0", _p), ("id0", _p),
        ("maxdegree", C.c_int32), ("need_mask", C.c_int32), ("conditional", C.c_int32),
        ("target_dist", C.c_int32), ("symflag", C.c_int32), ("seed", C.c_uint64),
        ("has_comp", C.c_int32), ("nact", C.c_int32), ("act", _p), ("inactw", _p),
        ("kb", C.c_int32), ("lamb", _p), ("thetab", _p), ("bty", _p), ("z0", _p),
        ("vb", _p), ("bkdist", _p), ("zmin", C.c_int32), ("zmax", C.c_int32),
        ("zbar", C.c_float), ("brange", C.c_float), ("bsimmean", C.c_float),
        ("nrate", C.c_int32), ("rkind", _p), ("ralpha", _p), ("ru", _p),
        ("vmat64", _p), ("wtab64", _p), ("vb64", _p), ("p1_64", _p), ("p2_64", _p),
    ]


class cusna_saom_result(C.Structure):
    _fields_ = [
        ("struct_version", C.c_int32),
        ("stats", _p), ("ntim", _p), ("scores", _p), ("dist", _p), ("dist_stat", _p),
        ("steps", _p),
        ("bdist", _p), ("bscores", _p), ("z_end", _p),
        ("rscores", _p), ("actor_dist", _p),
        ("want_end_nets", C.c_int32), ("a_ends", _p),
    ]


class cusna_mom_options(C.Structure):
    """Mirror of struct cusna_mom_options (include/libcusna.h)."""
    _fields_ = [
        ("struct_version", C.c_int32), ("mode", C.c_int32), ("P", C.c_int32),
        ("nsub", C.c_int32), ("n2", C.c_int32 * 8),
        ("n1", C.c_int32), ("nD", C.c_int32), ("batch2", C.c_int32), ("n3", C.c_int32),
        ("firstg", C.c_double), ("diagonalize", C.c_double), ("seed", C.c_uint64),
        ("use_graph", C.c_int32), ("max_repeats", C.c_int32), ("conv_tol", C.c_double),
        ("beh_zbar", C.c_double), ("beh_brange", C.c_double), ("beh_bsimmean", C.c_double),
        ("surrogate", C.c_int32), ("surrogate_pairs_log2", C.c_int32),
        ("surrogate_steps", C.c_int32), ("surrogate_rounds", C.c_int32),
        ("surrogate_corrections", C.c_int32), ("specialize", C.c_int32), ("verbose", C.c_int32),
        ("surrogate_warmup", C.c_int32),
    ]


class cusna_mom_result(C.Structure):
    """Mirror of struct cusna_mom_result (include/libcusna.h)."""
    _fields_ = [
        ("struct_version", C.c_int32), ("p", C.c_int32),
        ("theta", _p), ("se", _p), ("tconv", _p), ("sim_means", _p),
        ("D", _p), ("msf", _p), ("rate_est", _p), ("rate_sd", _p), ("phase3_stats", _p),
        ("phase3_sp", _p), ("phase3_scp", _p),
        ("tconv_max", C.c_double), ("max_abs_tconv", C.c_double), ("n_sims", C.c_int64),
        ("repeats", C.c_int32), ("reserved", C.c_int32),
        ("phase_seconds", C.c_double * 3), ("wall_seconds", C.c_double),
    ]


def pack_rows(dense: np.ndarray) -> np.ndarray:
    """Dense (n, n) 0/1 -> packed out-rows (n, W) uint32, bit j&31 of word j>>5."""
    d = np.ascontiguousarray(dense != 0)
    n = d.shape[0]
    W = (n + 31) // 32
    pad = np.zeros((n, W * 32), dtype=np.uint8)
    pad[:, :n] = d
    return np.packbits(pad.reshape(n, W, 32), axis=2, bitorder="little").view("<u4").reshape(n, W)


def unpack_rows(rows: np.ndarray, n: int) -> np.ndarray:
    """Packed (..., n, W) uint32 rows -> dense (..., n, n) uint8."""
    r = np.ascontiguousarray(rows, dtype="<u4")
    b = np.unpackbits(r.view(np.uint8).reshape(*r.shape[:-1], -1), axis=-1, bitorder="little")
    return b[..., :n]


def _ptr(a):
    return None if a is None else a.ctypes.data


def _arr(x, dtype):
    if x is None:
        return None
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


@dataclass
class SaomProblem:
    """Inputs of one period call. Arrays use the ABI dtypes and shapes.

    ``x0`` is the dense (n, n) start network; masks are dense (n, n) 0/1 or
    None (None = the reference default: off-diagonal, force = 0). ``theta`` is
    (M, k) or (k,) (broadcast, as the R glue does at cusna_saom_glue.cpp:75-78);
    ``lam`` is (M,) or a scalar. Descriptors are float64 (the *64 fields)."""

    x0: np.ndarray
    theta: np.ndarray
    lam: np.ndarray | float
    type_ids: np.ndarray
    M: int
    seed: int
    etypes: np.ndarray | None = None
    p1: np.ndarray | None = None
    p2: np.ndarray | None = None
    vmat: np.ndarray | None = None
    vmiss: np.ndarray | None = None
    dyn_flags: np.ndarray | None = None
    dyc_idx: np.ndarray | None = None
    dyc: np.ndarray | None = None
    wtab: np.ndarray | None = None
    int_comp: np.ndarray | None = None  # (k, 3) component indices or -1
    vb: np.ndarray | None = None
    od0: np.ndarray | None = None
    id0: np.ndarray | None = None
    sim_start: np.ndarray | None = None
    kdist: np.ndarray | None = None
    keep: np.ndarray | None = None
    force: np.ndarray | None = None
    perm: np.ndarray | None = None
    maxdegree: int = 0
    conditional: int = 0
    target_dist: int = 0
    symflag: int = 0
    threads: int = 0
    # behaviour
    kb: int = 0
    lamb: np.ndarray | float | None = None
    thetab: np.ndarray | None = None
    bty: np.ndarray | None = None
    z0: np.ndarray | None = None
    bkdist: np.ndarray | None = None
    zmin: int = 0
    zmax: int = 0
    zbar: float = 0.0
    brange: float = 1.0
    bsimmean: float = 0.0
    # composition change
    has_comp: int = 0
    act: np.ndarray | None = None
    nact: int | None = None
    inact: np.ndarray | None = None  # (n,) 0/1
    # non-basic rate effects: kinds 0 covariate (rate_u row), 1 out-degree,
    # 2 in-degree; alpha the rate parameters
    rate_kind: np.ndarray | None = None
    rate_alpha: np.ndarray | None = None
    rate_u: np.ndarray | None = None  # (nrate, n)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nrate(self) -> int:
        return 0 if self.rate_kind is None else int(np.asarray(self.rate_kind).shape[0])

    @property
    def n(self) -> int:
        return int(np.asarray(self.x0).shape[0])

    @property
    def k(self) -> int:
        return int(np.asarray(self.type_ids).shape[0])

    def build(self) -> cusna_saom_problem:
        n, k, M = self.n, self.k, int(self.M)
        W = (n + 31) // 32
        keep = self._keep = []

        def hold(a):
            keep.append(a)
            return _ptr(a)

        def rows(d):
            return None if d is None else hold(pack_rows(np.asarray(d)))

        th = np.asarray(self.theta, dtype=np.float32)
        if th.ndim == 1:
            th = np.broadcast_to(th, (M, k))
        th = _arr(th, np.float32)
        lam = np.asarray(self.lam, dtype=np.float32)
        if lam.ndim == 0:
            lam = np.full(M, lam, dtype=np.float32)
        lam = _arr(lam, np.float32)
        if th.shape != (M, k) or lam.shape != (M,):
            raise ValueError("theta must be (M, k) or (k,), lam (M,) or scalar")
        p = cusna_saom_problem()
        p.struct_version = ABI_VERSION
        p.n, p.W, p.k, p.M, p.threads = n, W, k, M, int(self.threads)
        x0_out = pack_rows(self.x0)
        x0_in = pack_rows(np.asarray(self.x0).T)
        p.x0_out, p.x0_in = hold(x0_out), hold(x0_in)
        if self.sim_start is not None:
            p.sim_out = rows(self.sim_start)
            p.sim_in = rows(np.asarray(self.sim_start).T)
        p.lam, p.theta = hold(lam), hold(th)
        p.type_ids = hold(_arr(self.type_ids, np.int32))
        p.etypes = hold(_arr(self.etypes if self.etypes is not None else np.zeros(k), np.int32))
        dyn = self.dyn_flags if self.dyn_flags is not None else np.zeros(k)
        p.dyn_flags = hold(_arr(dyn, np.int32))
        dci = self.dyc_idx if self.dyc_idx is not None else -np.ones(k)
        p.dyc_idx = hold(_arr(dci, np.int32))
        if self.dyc is not None:
            p.dyc = hold(_arr(self.dyc, np.float32))
        ic = self.int_comp if self.int_comp is not None else -np.ones((k, 3))
        ic = np.asarray(ic, dtype=np.int32).reshape(k, 3)
        p.int0 = hold(_arr(ic[:, 0], np.int32))
        p.int1 = hold(_arr(ic[:, 1], np.int32))
        p.int2 = hold(_arr(ic[:, 2], np.int32))
        p1 = _arr(self.p1 if self.p1 is not None else np.zeros(k), np.float64)
        p2 = _arr(self.p2 if self.p2 is not None else np.zeros(k), np.float64)
        vmat = _arr(self.vmat if self.vmat is not None else np.zeros((k, n)), np.float64)
        wtab = _arr(self.wtab if self.wtab is not None else np.zeros((k, n + 1)), np.float64)
        vb = _arr(self.vb if self.vb is not None else np.zeros(n), np.float64)
        p.p1_64, p.p2_64, p.vmat64, p.wtab64, p.vb64 = (hold(p1), hold(p2), hold(vmat),
                                                        hold(wtab), hold(vb))
        # float32 copies for callers that only read the kernel descriptors
        p.p1, p.p2 = hold(p1.astype(np.float32)), hold(p2.astype(np.float32))
        p.vmat, p.wtab = hold(vmat.astype(np.float32)), hold(wtab.astype(np.float32))
        p.vb = hold(vb.astype(np.float32))
        vmiss = self.vmiss if self.vmiss is not None else np.zeros((k, n))
        p.vmiss = hold(_arr(np.asarray(vmiss) != 0, np.uint8))
        # start degrees stay NULL unless given: the R glue never sets them
        # (cusna_saom_glue.cpp:93-106), so degree statistics then use 0.
        p.od0 = hold(_arr(self.od0, np.int32)) if self.od0 is not None else None
        p.id0 = hold(_arr(self.id0, np.int32)) if self.id0 is not None else None
        p.kdist = rows(self.kdist)
        p.perm = rows(self.perm)
        p.keep_out = rows(self.keep)
        p.keep_in = rows(None if self.keep is None else np.asarray(self.keep).T)
        p.force_out = rows(self.force)
        p.force_in = rows(None if self.force is None else np.asarray(self.force).T)
        p.maxdegree, p.conditional = int(self.maxdegree), int(self.conditional)
        p.target_dist, p.symflag = int(self.target_dist), int(self.symflag)
        p.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        if self.nrate > 0:
            p.nrate = self.nrate
            p.rkind = hold(_arr(self.rate_kind, np.int32))
            p.ralpha = hold(_arr(self.rate_alpha, np.float32))
            if self.rate_u is not None:
                p.ru = hold(_arr(np.reshape(self.rate_u, (self.nrate, n)), np.float32))
        p.has_comp = int(self.has_comp)
        p.nact = int(self.nact if self.nact is not None else n)
        if self.act is not None:
            p.act = hold(_arr(self.act, np.int32))
        if self.inact is not None:
            bits = np.zeros(((n + 31) // 32) * 32, np.uint8)
            bits[:n] = np.asarray(self.inact).reshape(n) != 0
            p.inactw = hold(np.packbits(bits, bitorder="little").view("<u4").copy())
        p.kb = int(self.kb)
        p.brange = 1.0
        if self.kb > 0:
            kb = int(self.kb)
            thb = np.asarray(self.thetab, dtype=np.float32)
            if thb.ndim == 1:
                thb = np.broadcast_to(thb, (M, kb))
            lamb = np.asarray(self.lamb, dtype=np.float32)
            if lamb.ndim == 0:
                lamb = np.full(M, lamb, dtype=np.float32)
            p.lamb, p.thetab = hold(_arr(lamb, np.float32)), hold(_arr(thb, np.float32))
            p.bty = hold(_arr(self.bty, np.int32))
            p.z0 = hold(_arr(self.z0, np.int32))
            bk = self.bkdist if self.bkdist is not None else np.ones(n)
            p.bkdist = hold(_arr(bk, np.int32))
            p.zmin, p.zmax = int(self.zmin), int(self.zmax)
            p.zbar, p.brange, p.bsimmean = (float(self.zbar), float(self.brange),
                                            float(self.bsimmean))
        return p


class SaomResult:
    """Output buffers for one period call (all optional outputs allocated)."""

    def __init__(self, n: int, k: int, M: int, kb: int = 0, want_end_nets: bool = False,
                 rate_outputs: bool = True, nrate: int = 0):
        self.n, self.k, self.M, self.kb, self.nrate = n, k, M, kb, nrate
        self.rscores = np.zeros((M, max(nrate, 1)), np.float32)
        self.stats = np.zeros((M, k), np.float64)
        self.ntim = np.zeros(M, np.float64)
        self.scores = np.zeros((M, k + 1), np.float32)
        self.dist = np.zeros(M, np.int32)
        self.dist_stat = np.zeros(M, np.int32)
        self.steps = np.zeros(M, np.int32)
        self.bdist = np.zeros(M, np.int32)
        self.bscores = np.zeros((M, kb + 1), np.float32)
        self.z_end = np.zeros((M, n), np.int32)
        self.actor_dist = np.zeros((M, n), np.int32) if rate_outputs else None
        self.a_ends = np.zeros((M, n, n), np.uint8) if want_end_nets else None
        r = cusna_saom_result()
        r.struct_version = ABI_VERSION
        r.stats, r.ntim, r.scores = _ptr(self.stats), _ptr(self.ntim), _ptr(self.scores)
        r.dist, r.dist_stat, r.steps = _ptr(self.dist), _ptr(self.dist_stat), _ptr(self.steps)
        if kb > 0:
            r.bdist, r.bscores, r.z_end = _ptr(self.bdist), _ptr(self.bscores), _ptr(self.z_end)
        r.actor_dist = _ptr(self.actor_dist)
        if nrate > 0:
            r.rscores = _ptr(self.rscores)
        r.want_end_nets = 1 if want_end_nets else 0
        r.a_ends = _ptr(self.a_ends)
        self.c = r

    def as_dict(self) -> dict:
        d = {
            "stats": self.stats, "ntim": self.ntim, "scores": self.scores, "dist": self.dist,
            "dist_stat": self.dist_stat, "steps": self.steps,
        }
        if self.kb > 0:
            d.update(bdist=self.bdist, bscores=self.bscores, z_end=self.z_end)
        if self.nrate > 0:
            d.update(rscores=self.rscores, actor_dist=self.actor_dist)
        if self.a_ends is not None:
            d["a_ends"] = self.a_ends
        return d


def run_period(fn, prob: SaomProblem, want_end_nets: bool = False, name: str = "period",
               extra_args=()) -> dict:
    """Call ``fn(problem*, result*, *extra_args)`` and return the outputs."""
    p = prob.build()
    res = SaomResult(prob.n, prob.k, int(prob.M), int(prob.kb), want_end_nets,
                     nrate=prob.nrate)
    st = fn(C.byref(p), C.byref(res.c), *extra_args)
    if st != 0:
        raise CusnaError(name, st)
    return res.as_dict()


# ------------------------------------------------------- multi-network ABI
MAXNET = 4
CROSS_IDS = (33, 34, 35, 36, 37)


class cusna_saom_multi_problem(C.Structure):
    """Reference libcusna.h:238-267."""
    _fields_ = [
        ("struct_version", C.c_int32), ("n", C.c_int32), ("W", C.c_int32), ("k", C.c_int32),
        ("M", C.c_int32), ("D", C.c_int32), ("threads", C.c_int32),
        ("x0_out", _p), ("x0_in", _p), ("lam", _p), ("theta", _p),
        ("type_ids", _p), ("etypes", _p), ("netof", _p), ("netref", _p),
        ("wtab", _p), ("p1", _p), ("p2", _p), ("vmat", _p), ("dyc", _p), ("dyc_idx", _p),
        ("perm", _p), ("symf", _p), ("seed", C.c_uint64), ("act", _p), ("nact", C.c_int32),
    ]


class cusna_saom_multi_result(C.Structure):
    """Reference libcusna.h:271-278."""
    _fields_ = [
        ("struct_version", C.c_int32), ("rscores", _p), ("scores", _p), ("ntim", _p),
        ("steps", _p), ("a_ends", _p),
    ]


class cusna_multi_stat_spec(C.Structure):
    """Statistics of one network of a multi-network plan (include/libcusna.h)."""
    _fields_ = [
        ("struct_version", C.c_int32), ("k", C.c_int32), ("type_ids", _p), ("p1", _p),
        ("p2", _p), ("vmat", _p), ("vmiss", _p), ("wtab", _p), ("od0", _p), ("id0", _p),
        ("symflag", C.c_int32), ("keep", _p), ("force", _p), ("cross_wp", _p),
    ]


@dataclass
class MultiProblem:
    """One multi-network period (D networks over one actor set) as the R
    estimator hands it to the native simulator (R/mom.R:352-364,
    cusna_saom_glue.cpp:229-279). x0 and perm are (D, n, n); lam (D,) or
    (M, D); theta (k,) or (M, k); descriptors float32 on the wire."""
    x0: np.ndarray
    lam: np.ndarray
    theta: np.ndarray
    type_ids: np.ndarray
    netof: np.ndarray
    M: int
    seed: int
    netref: np.ndarray | None = None
    etypes: np.ndarray | None = None
    p1: np.ndarray | None = None
    p2: np.ndarray | None = None
    vmat: np.ndarray | None = None
    wtab: np.ndarray | None = None
    dyc: np.ndarray | None = None       # (nd, n, n)
    dyc_idx: np.ndarray | None = None
    perm: np.ndarray | None = None      # (D, n, n)
    symf: np.ndarray | None = None      # (D,)
    act: np.ndarray | None = None
    nact: int | None = None
    threads: int = 0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def D(self) -> int:
        return int(np.asarray(self.x0).shape[0])

    @property
    def n(self) -> int:
        return int(np.asarray(self.x0).shape[1])

    @property
    def k(self) -> int:
        return int(np.asarray(self.type_ids).shape[0])

    @property
    def nd(self) -> int:
        return 0 if self.dyc is None else int(np.asarray(self.dyc).shape[0])

    def build(self) -> cusna_saom_multi_problem:
        n, k, M, D = self.n, self.k, int(self.M), self.D
        W = (n + 31) // 32
        keep = self._keep = []

        def hold(a):
            keep.append(a)
            return _ptr(a)

        th = np.asarray(self.theta, np.float32)
        th = np.broadcast_to(th, (M, k)) if th.ndim == 1 else th
        lam = np.asarray(self.lam, np.float32)
        lam = np.broadcast_to(lam, (M, D)) if lam.ndim == 1 else lam
        if th.shape != (M, k) or lam.shape != (M, D):
            raise ValueError("theta must be (M, k) or (k,), lam (M, D) or (D,)")
        x0 = np.asarray(self.x0)
        p = cusna_saom_multi_problem()
        p.struct_version = ABI_VERSION
        p.n, p.W, p.k, p.M, p.D, p.threads = n, W, k, M, D, int(self.threads)
        p.x0_out = hold(np.concatenate([pack_rows(x0[d]) for d in range(D)]))
        p.x0_in = hold(np.concatenate([pack_rows(x0[d].T) for d in range(D)]))
        p.lam, p.theta = hold(_arr(lam, np.float32)), hold(_arr(th, np.float32))
        p.type_ids = hold(_arr(self.type_ids, np.int32))
        z = np.zeros(k)
        p.etypes = hold(_arr(self.etypes if self.etypes is not None else z, np.int32))
        p.netof = hold(_arr(self.netof, np.int32))
        p.netref = hold(_arr(self.netref if self.netref is not None else z, np.int32))
        p.wtab = hold(_arr(self.wtab if self.wtab is not None else np.zeros((k, n + 1)),
                           np.float32))
        p.p1 = hold(_arr(self.p1 if self.p1 is not None else z, np.float32))
        p.p2 = hold(_arr(self.p2 if self.p2 is not None else z, np.float32))
        p.vmat = hold(_arr(self.vmat if self.vmat is not None else np.zeros((k, n)),
                           np.float32))
        if self.dyc is not None:
            p.dyc = hold(_arr(self.dyc, np.float32))
        p.dyc_idx = hold(_arr(self.dyc_idx if self.dyc_idx is not None else -np.ones(k),
                              np.int32))
        if self.perm is not None:
            pm = np.asarray(self.perm)
            p.perm = hold(np.concatenate([pack_rows(pm[d]) for d in range(D)]))
        p.symf = hold(_arr(self.symf if self.symf is not None else np.zeros(D), np.int32))
        p.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        if self.act is not None:
            p.act = hold(_arr(self.act, np.int32))
        p.nact = int(self.nact if self.nact is not None else n)
        return p


class MultiResult:
    """Output buffers of one multi-network period call."""

    def __init__(self, n: int, k: int, M: int, D: int, want_end_nets: bool = True):
        self.rscores = np.zeros((M, D), np.float32)
        self.scores = np.zeros((M, k), np.float32)
        self.ntim = np.zeros(M, np.float64)
        self.steps = np.zeros(M, np.int32)
        self.a_ends = np.zeros((M, D, n, n), np.uint8) if want_end_nets else None
        r = cusna_saom_multi_result()
        r.struct_version = ABI_VERSION
        r.rscores, r.scores = _ptr(self.rscores), _ptr(self.scores)
        r.ntim, r.steps, r.a_ends = _ptr(self.ntim), _ptr(self.steps), _ptr(self.a_ends)
        self.c = r

    def as_dict(self) -> dict:
        d = {"rscores": self.rscores, "scores": self.scores, "ntim": self.ntim,
             "steps": self.steps}
        if self.a_ends is not None:
            d["a_ends"] = self.a_ends
        return d


def run_multi(fn, prob: MultiProblem, want_end_nets: bool = True, name: str = "multi",
              extra_args=()) -> dict:
    """Call ``fn(multi_problem*, multi_result*, *extra_args)``; return the outputs."""
    p = prob.build()
    res = MultiResult(prob.n, prob.k, int(prob.M), prob.D, want_end_nets)
    st = fn(C.byref(p), C.byref(res.c), *extra_args)
    if st != 0:
        raise CusnaError(name, st)
    return res.as_dict()
