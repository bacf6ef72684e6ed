"""brisk_amd - Python binding of the MI355X SDP interior-point solver (libbrisk_amd.so).

Mirrors the reference's Python interface for this path (python/brisk/_core.py of
MilanKorda/brisk: ``solve_sdpa``, ``solve_file`` and the ``Result`` fields) over the same flat
C interface, plus ``solve_batch`` and ``Server`` for batched and continuous-batching solves.
The library is the product path: there is no CPU fallback, and a missing library is an error.
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
I, D = ctypes.c_int, ctypes.c_double
PI, PD = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
P = ctypes.c_void_p


class BriskProblem(ctypes.Structure):
    _fields_ = [("m", I), ("nblk", I), ("bs", PI), ("c", PD), ("nnz", ctypes.c_int64),
                ("mat", PI), ("blk", PI), ("i", PI), ("j", PI), ("v", PD)]


def library_path():
    return os.environ.get("BRISK_AMD_LIBRARY", os.path.join(_HERE, "libbrisk_amd.so"))


def load_library(path=None):
    """Load libbrisk_amd.so (built by __graft_entry__.build()); raises if it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or library_path()
    if not os.path.exists(path):
        raise RuntimeError("libbrisk_amd.so not found at %s: build it first (python -c 'import __graft_entry__ as g; g.build()')" % path)
    L = ctypes.CDLL(path)
    L.brisk_version.restype = ctypes.c_char_p
    L.brisk_result_new.restype = P
    L.brisk_result_delete.argtypes = [P]
    for name, rt in (("status", I), ("exit_code", I), ("iterations", I), ("m", I), ("nblk", I), ("have_x", I),
                     ("pobj", D), ("dobj", D), ("time", D)):
        f = getattr(L, "brisk_result_" + name)
        f.argtypes, f.restype = [P], rt
    L.brisk_result_status_str.argtypes, L.brisk_result_status_str.restype = [P], ctypes.c_char_p
    L.brisk_result_cause.argtypes, L.brisk_result_cause.restype = [P], ctypes.c_char_p
    L.brisk_result_blocksize.argtypes, L.brisk_result_blocksize.restype = [P, I], I
    L.brisk_result_dimacs.argtypes, L.brisk_result_dimacs.restype = [P, PD], None
    L.brisk_result_phases.argtypes, L.brisk_result_phases.restype = [P, PD], None
    L.brisk_result_counts.argtypes, L.brisk_result_counts.restype = [P, PI], None
    L.brisk_result_trace.argtypes, L.brisk_result_trace.restype = [P, PD, PD, PD, I], I
    L.brisk_result_y.argtypes, L.brisk_result_y.restype = [P], PD
    L.brisk_result_X.argtypes, L.brisk_result_X.restype = [P, I], PD
    L.brisk_result_Z.argtypes, L.brisk_result_Z.restype = [P, I], PD
    ppc = ctypes.POINTER(ctypes.c_char_p)
    L.brisk_solve_data.argtypes = [I, I, PI, PD, ctypes.c_int64, PI, PI, PI, PI, PD, I, ppc, P]
    L.brisk_solve_data.restype = I
    L.brisk_solve_file.argtypes = [ctypes.c_char_p, I, ppc, P]
    L.brisk_solve_file.restype = I
    L.brisk_solve_batch.argtypes = [I, ctypes.POINTER(BriskProblem), I, ppc, ctypes.POINTER(P)]
    L.brisk_solve_batch.restype = I
    L.brisk_set_output.argtypes, L.brisk_set_output.restype = [I, P], None
    L.brisk_interrupt.argtypes, L.brisk_interrupt.restype = [], None
    L.brisk_serve_new.argtypes, L.brisk_serve_new.restype = [I, ppc], P
    L.brisk_serve_submit.argtypes, L.brisk_serve_submit.restype = [P, ctypes.POINTER(BriskProblem)], I
    L.brisk_serve_step.argtypes, L.brisk_serve_step.restype = [P], I
    L.brisk_serve_take.argtypes, L.brisk_serve_take.restype = [P, I, P, PD], I
    L.brisk_serve_inflight.argtypes, L.brisk_serve_inflight.restype = [P], I
    L.brisk_serve_done.argtypes, L.brisk_serve_done.restype = [P, PI, I], I
    L.brisk_serve_delete.argtypes, L.brisk_serve_delete.restype = [P], None
    L.bk_prof_enable.argtypes, L.bk_prof_enable.restype = [I], None
    L.bk_prof_reset.argtypes, L.bk_prof_reset.restype = [], None
    L.bk_prof_read.argtypes, L.bk_prof_read.restype = [I, PD], None
    L.bk_device_name.restype = ctypes.c_char_p
    _LIB = L
    return L


def version():
    return load_library().brisk_version().decode()


@dataclass
class Result:
    """The solution of the problem as given (fields as the reference's Result): y is BRISK's y
    (the SDPA x is -y); X, Z one array per block, n x n (SDP) or n (LP); dimacs: errors 1..6."""
    status: int
    status_str: str
    exit_code: int
    iterations: int
    pobj: float
    dobj: float
    time: float
    dimacs: np.ndarray
    blocksizes: List[int] = field(default_factory=list)
    y: Optional[np.ndarray] = None
    X: Optional[List[np.ndarray]] = None
    Z: Optional[List[np.ndarray]] = None
    cause: str = ""
    phases: Optional[np.ndarray] = None      # host setup + upload, iterations, DIMACS, download (s)
    counts: Optional[List[int]] = None       # iterations with an f32 factor, with an f64 factor, PCG steps
    trace: Optional[dict] = None             # pobj, dobj, mu per iteration

    @property
    def x(self):
        return None if self.y is None else -self.y


def _options(options):
    if options is None:
        return []
    if isinstance(options, dict):
        out = []
        for k, v in options.items():
            out.append("-" + str(k))
            if v is not True:
                out.append(str(v))
        return out
    return [str(o) for o in options]


def _argv(options):
    opts = _options(options)
    arr = (ctypes.c_char_p * max(len(opts), 1))(*[o.encode() for o in opts])
    return len(opts), arr


def _collect(L, r):
    nb = L.brisk_result_nblk(r)
    bs = [L.brisk_result_blocksize(r, k) for k in range(nb)]
    err = np.zeros(6)
    L.brisk_result_dimacs(r, err.ctypes.data_as(PD))
    ph = np.zeros(4)
    L.brisk_result_phases(r, ph.ctypes.data_as(PD))
    cn = (I * 3)()
    L.brisk_result_counts(r, cn)
    m = L.brisk_result_m(r)
    yp = L.brisk_result_y(r)
    y = np.ctypeslib.as_array(yp, (m,)).copy() if yp and m > 0 else (np.zeros(0) if yp else None)
    X = Z = None
    if L.brisk_result_have_x(r):
        X, Z = [], []
        for k, n in enumerate(bs):
            shp = (n, n) if n > 0 else (-n,)
            ln = n * n if n > 0 else -n
            xp, zp = L.brisk_result_X(r, k), L.brisk_result_Z(r, k)
            X.append(np.ctypeslib.as_array(xp, (ln,)).copy().reshape(shp, order="F"))
            Z.append(np.ctypeslib.as_array(zp, (ln,)).copy().reshape(shp, order="F"))
    po, do, mu = np.zeros(128), np.zeros(128), np.zeros(128)
    nt = L.brisk_result_trace(r, po.ctypes.data_as(PD), do.ctypes.data_as(PD), mu.ctypes.data_as(PD), 128)
    return Result(status=L.brisk_result_status(r), status_str=L.brisk_result_status_str(r).decode(),
                  exit_code=L.brisk_result_exit_code(r), iterations=L.brisk_result_iterations(r),
                  pobj=L.brisk_result_pobj(r), dobj=L.brisk_result_dobj(r), time=L.brisk_result_time(r),
                  dimacs=err, blocksizes=bs, y=y, X=X, Z=Z, cause=L.brisk_result_cause(r).decode(),
                  phases=ph, counts=list(cn), trace=dict(pobj=po[:nt], dobj=do[:nt], mu=mu[:nt]))


class _Arrays:
    """Keeps contiguous int32/float64 copies alive for the duration of a call."""

    def __init__(self, blocksizes, c, mat, blk, i, j, v):
        self.bs = np.ascontiguousarray(blocksizes, dtype=np.int32)
        self.c = np.ascontiguousarray(c, dtype=np.float64)
        self.mat = np.ascontiguousarray(mat, dtype=np.int32)
        self.blk = np.ascontiguousarray(blk, dtype=np.int32)
        self.i = np.ascontiguousarray(i, dtype=np.int32)
        self.j = np.ascontiguousarray(j, dtype=np.int32)
        self.v = np.ascontiguousarray(v, dtype=np.float64)
        n = len(self.mat)
        if not (len(self.blk) == len(self.i) == len(self.j) == len(self.v) == n):
            raise ValueError("mat, blk, i, j, v must have the same length")

    def problem(self):
        return BriskProblem(len(self.c), len(self.bs), self.bs.ctypes.data_as(PI), self.c.ctypes.data_as(PD),
                            len(self.mat), self.mat.ctypes.data_as(PI), self.blk.ctypes.data_as(PI),
                            self.i.ctypes.data_as(PI), self.j.ctypes.data_as(PI), self.v.ctypes.data_as(PD))


def solve_sdpa(blocksizes, c, mat, blk, i, j, v, options=None, verbose=False):
    """Solve the problem given by the numbers of an SDPA sparse file (reference
    python/brisk/_core.py:1366 solve_sdpa): 1-based indices, mat 0 = F0 (C = -F0)."""
    L = load_library()
    a = _Arrays(blocksizes, c, mat, blk, i, j, v)
    n, argv = _argv(options)
    L.brisk_set_output(0 if verbose else 1, None)
    r = L.brisk_result_new()
    try:
        rc = L.brisk_solve_data(len(a.c), len(a.bs), a.bs.ctypes.data_as(PI), a.c.ctypes.data_as(PD), len(a.mat),
                                a.mat.ctypes.data_as(PI), a.blk.ctypes.data_as(PI), a.i.ctypes.data_as(PI),
                                a.j.ctypes.data_as(PI), a.v.ctypes.data_as(PD), n, argv, r)
        if rc < 0:
            raise RuntimeError("brisk_solve_data aborted (%d): no usable GPU or out of device memory" % rc)
        if rc in (1, 2):
            raise ValueError("brisk_solve_data: %s" % ("bad option" if rc == 1 else "bad problem data"))
        return _collect(L, r)
    finally:
        L.brisk_result_delete(r)


def solve_file(path, options=None, verbose=False):
    """Solve an SDPA sparse file (reference python/brisk/_core.py:1013 solve_file)."""
    L = load_library()
    n, argv = _argv(options)
    L.brisk_set_output(0 if verbose else 1, None)
    r = L.brisk_result_new()
    try:
        rc = L.brisk_solve_file(str(path).encode(), n, argv, r)
        if rc < 0:
            raise RuntimeError("brisk_solve_file aborted (%d)" % rc)
        if rc in (1, 2):
            raise ValueError("brisk_solve_file: %s" % ("bad option" if rc == 1 else "cannot read the file"))
        return _collect(L, r)
    finally:
        L.brisk_result_delete(r)


def solve_batch(problems, options=None, verbose=False):
    """Solve several problems together, one batched iteration for all. problems: list of
    (blocksizes, c, mat, blk, i, j, v) tuples. Returns a list of Result."""
    L = load_library()
    arrs = [_Arrays(*p) for p in problems]
    probs = (BriskProblem * max(len(arrs), 1))(*[a.problem() for a in arrs])
    res = (P * max(len(arrs), 1))(*[L.brisk_result_new() for _ in arrs])
    n, argv = _argv(options)
    L.brisk_set_output(0 if verbose else 1, None)
    try:
        rc = L.brisk_solve_batch(len(arrs), probs, n, argv, res)
        if rc < 0:
            raise RuntimeError("brisk_solve_batch aborted (%d)" % rc)
        if rc in (1, 2):
            raise ValueError("brisk_solve_batch: %s" % ("bad option" if rc == 1 else "bad problem data"))
        return [_collect(L, res[k]) for k in range(len(arrs))]
    finally:
        for k in range(len(arrs)):
            L.brisk_result_delete(res[k])


def profile(enable=True):
    """Per-kernel-class event timing on the library stream (for measurement)."""
    load_library().bk_prof_enable(1 if enable else 0)


def profile_reset():
    load_library().bk_prof_reset()


def profile_read(cls):
    """{ms, launches, flops, bytes} of a kernel class (see dev.h: 0-3 f64 GEMM by transposes,
    4-7 f32 GEMM, 8 Cholesky diagonal blocks, 9 GEMV, 10 Schur assembly, 11 A(.), 12 Lanczos,
    13 dot products, 14 linear combinations)."""
    out = np.zeros(4)
    load_library().bk_prof_read(cls, out.ctypes.data_as(PD))
    return dict(ms=out[0], launches=out[1], flops=out[2], bytes=out[3])


class Server:
    """The continuous-batching request loop (brisk_serve_*): submit problems as they arrive,
    step() advances every solve in flight by one iteration and admits waiting requests."""

    def __init__(self, options=None):
        self.L = load_library()
        n, argv = _argv(options)
        self.h = self.L.brisk_serve_new(n, argv)
        if not self.h:
            raise RuntimeError("brisk_serve_new failed (bad option or no GPU)")
        self._keep = {}

    def submit(self, blocksizes, c, mat, blk, i, j, v):
        a = _Arrays(blocksizes, c, mat, blk, i, j, v)
        p = a.problem()
        rid = self.L.brisk_serve_submit(self.h, ctypes.byref(p))
        if rid < 0:
            raise ValueError("brisk_serve_submit: bad problem data")
        return rid

    def step(self):
        rc = self.L.brisk_serve_step(self.h)
        if rc < 0:
            raise RuntimeError("brisk_serve_step failed")
        return rc

    def take(self, rid):
        r = self.L.brisk_result_new()
        lat = D()
        try:
            st = self.L.brisk_serve_take(self.h, rid, r, ctypes.byref(lat))
            if st < 0:
                raise KeyError(rid)
            if st == 0:
                return None
            res = _collect(self.L, r)
            return res, lat.value
        finally:
            self.L.brisk_result_delete(r)

    def inflight(self):
        return self.L.brisk_serve_inflight(self.h)

    def done(self, cap=4096):
        """Ids of the requests finished since the last call, in finishing order."""
        buf = (I * cap)()
        k = self.L.brisk_serve_done(self.h, buf, cap)
        return list(buf[:k])

    def close(self):
        if self.h:
            self.L.brisk_serve_delete(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
