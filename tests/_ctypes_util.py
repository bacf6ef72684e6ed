"""ctypes plumbing shared by the tests: the instance generators (product side), the oracle
(test infrastructure) and, where it was built in this container, the reference library."""
import ctypes, os, sys
import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from brisk_amd.generators import Inst, Instance, gen_lib  # noqa: E402,F401  (the product's generators)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
I, D = ctypes.c_int, ctypes.c_double
PI, PD = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)


class OracleData(ctypes.Structure):
    _fields_ = [("m", I), ("nblk", I), ("bs", PI), ("c", PD), ("nnz", ctypes.c_int64),
                ("mat", PI), ("blk", PI), ("i", PI), ("j", PI), ("v", PD)]


TR = 128


class OracleOptions(ctypes.Structure):
    _fields_ = [("tol", D), ("maxit", I), ("method", I)]


class OracleResult(ctypes.Structure):
    _fields_ = [("status", I), ("iters", I), ("m", I), ("nblk", I), ("pobj", D), ("dobj", D), ("acc", D),
                ("err", D * 6), ("y", PD), ("X", PD), ("xlen", ctypes.c_size_t),
                ("trace_pobj", D * TR), ("trace_dobj", D * TR), ("trace_mu", D * TR), ("trace_score", D * TR),
                ("trace_ap", D * TR), ("trace_ad", D * TR)]


_orc = None


def oracle_lib():
    global _orc
    if _orc is None:
        _orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        pO = ctypes.POINTER(OracleData)
        _orc.oracle_solve.argtypes = [pO, ctypes.POINTER(OracleOptions), ctypes.POINTER(OracleResult)]
        _orc.oracle_result_free.argtypes = [ctypes.POINTER(OracleResult)]
        _orc.oracle_dimacs.argtypes = [pO, PD, PD, PD, PD, PD]
        _orc.oracle_schur.argtypes = [pO, PD, PD, PD]
        _orc.oracle_lammin.argtypes = [I, PD]
        _orc.oracle_lammin.restype = D
        _orc.oracle_lanczos.argtypes = [I, PD, PD, I, D, PD]
        _orc.oracle_lanczos.restype = D
        _orc.oracle_step_lambda.argtypes = [I, PD, PD]
        _orc.oracle_step_lambda.restype = D
        _orc.oracle_correct.argtypes = [pO, PD, PD, I, I, PD, PD]
    return _orc


def oracle_data(inst):
    s = inst.s
    return OracleData(s.m, s.nblk, s.bs, s.c, s.nnz, s.mat, s.blk, s.i, s.j, s.v)


def oracle_solve(inst, tol=1e-8, maxit=100, method=0):
    lib = oracle_lib()
    d = oracle_data(inst)
    r = OracleResult()
    rc = lib.oracle_solve(ctypes.byref(d), ctypes.byref(OracleOptions(tol, maxit, method)), ctypes.byref(r))
    if rc != 0:
        raise ValueError("oracle_solve: bad data (%d)" % rc)
    out = dict(status=r.status, iters=r.iters, pobj=r.pobj, dobj=r.dobj, err=list(r.err), acc=r.acc,
               y=np.ctypeslib.as_array(r.y, (r.m,)).copy(), X=np.ctypeslib.as_array(r.X, (r.xlen,)).copy(),
               trace_pobj=list(r.trace_pobj[:r.iters + 1]), trace_dobj=list(r.trace_dobj[:r.iters + 1]),
               trace_mu=list(r.trace_mu[:r.iters + 1]))
    lib.oracle_result_free(ctypes.byref(r))
    return out


def oracle_correct(inst, X, y, half, maxpass):
    """The oracle's end-phase metric correction of (X, y): (return code, corrected X, steps)."""
    lib = oracle_lib()
    d = oracle_data(inst)
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros_like(X)
    steps = np.zeros(maxpass)
    rc = lib.oracle_correct(ctypes.byref(d), X.ctypes.data_as(PD), y.ctypes.data_as(PD), half, maxpass,
                            out.ctypes.data_as(PD), steps.ctypes.data_as(PD))
    return rc, out, steps


def oracle_dimacs(inst, X, y):
    lib = oracle_lib()
    d = oracle_data(inst)
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    err = np.zeros(6)
    po, do = D(), D()
    lib.oracle_dimacs(ctypes.byref(d), X.ctypes.data_as(PD), y.ctypes.data_as(PD), err.ctypes.data_as(PD),
                      ctypes.byref(po), ctypes.byref(do))
    return err, po.value, do.value


def read_sdpa(path):
    """The numbers of an SDPA sparse file (.dat-s): comment lines starting with '"' or '*',
    then m, nblk, the block sizes (separators ",(){}" allowed), the m entries of c and the
    lines "mat blk i j value" (the format as the reference's read_sdpa, problem.c:426)."""
    toks = []
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines) and lines[k].lstrip()[:1] in ('"', '*'):
        k += 1
    body = " ".join(lines[k:])
    for ch in ",(){}=":
        body = body.replace(ch, " ")
    toks = body.split()

    def num(t):
        try:
            return float(t)
        except ValueError:
            return None
    head, q = [], 0
    while len(head) < 2 or len(head) < 2 + int(head[1]) + int(head[0]):
        x = num(toks[q])
        q += 1
        if x is not None:          # words in the header ("mDIM") are skipped
            head.append(x)
    m, nblk = int(head[0]), int(head[1])
    bs = [int(t) for t in head[2:2 + nblk]]
    c = head[2 + nblk:2 + nblk + m]
    rest = toks[q:]
    rest = rest[:len(rest) // 5 * 5]
    e = np.array(rest, dtype=np.float64).reshape(-1, 5)
    return dict(m=m, nblk=nblk, bs=np.array(bs, np.int32), c=np.array(c), mat=e[:, 0].astype(np.int32),
                blk=e[:, 1].astype(np.int32), i=e[:, 2].astype(np.int32), j=e[:, 3].astype(np.int32),
                v=np.ascontiguousarray(e[:, 4]))


class ArrInstance:
    """An instance held as numpy arrays (read from a file or built in a test); the same
    interface as Instance."""

    def __init__(self, d):
        self.d = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in d.items()}
        a = self.d
        a["bs"] = np.ascontiguousarray(a["bs"], np.int32)
        a["c"] = np.ascontiguousarray(a["c"], np.float64)
        for k in ("mat", "blk", "i", "j"):
            a[k] = np.ascontiguousarray(a[k], np.int32)
        a["v"] = np.ascontiguousarray(a["v"], np.float64)
        self.s = Inst(a["m"], a["nblk"], a["bs"].ctypes.data_as(PI), a["c"].ctypes.data_as(PD), len(a["v"]),
                      a["mat"].ctypes.data_as(PI), a["blk"].ctypes.data_as(PI), a["i"].ctypes.data_as(PI),
                      a["j"].ctypes.data_as(PI), a["v"].ctypes.data_as(PD))

    @property
    def m(self):
        return self.d["m"]

    @property
    def bs(self):
        return list(self.d["bs"])

    def arrays(self):
        return dict(self.d)

    def data_args(self):
        s = self.s
        return (s.m, s.nblk, s.bs, s.c, s.nnz, s.mat, s.blk, s.i, s.j, s.v)


def golden(name="small"):
    import json
    with open(os.path.join(ROOT, "tests", "golden", name + ".json")) as f:
        return json.load(f)


def golden_instance(rec):
    """The instance a golden record was computed on."""
    if "file" in rec:
        return ArrInstance(read_sdpa(os.path.join(ROOT, "tests", "golden", rec["file"])))
    return Instance(rec["generator"], *rec["args"])


def oracle_lanczos(D_, L=None, K=30, tol=0.02):
    """The oracle's restatement of the reference's Lanczos bound on L^{-1} D L^{-T} (or D):
    (bound, steps, theta, res)."""
    n = D_.shape[0]
    Df = np.asfortranarray(D_, dtype=np.float64)
    Lf = None if L is None else np.asfortranarray(L, dtype=np.float64)
    info = np.zeros(3)
    b = oracle_lib().oracle_lanczos(n, None if Lf is None else Lf.ctypes.data_as(PD), Df.ctypes.data_as(PD), K, tol,
                                    info.ctypes.data_as(PD))
    return b, int(info[0]), info[1], info[2]


def oracle_step_lambda(D_, L=None):
    """The reference's step-length bound of one block: exact for order <= 32, else the Lanczos bound."""
    n = D_.shape[0]
    Df = np.asfortranarray(D_, dtype=np.float64)
    Lf = None if L is None else np.asfortranarray(L, dtype=np.float64)
    return oracle_lib().oracle_step_lambda(n, None if Lf is None else Lf.ctypes.data_as(PD), Df.ctypes.data_as(PD))
