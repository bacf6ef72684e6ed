"""Seeded instance generators (libbrisk_gen.so, include/brisk_gen.h) and the serving stream.

The workloads the solver is measured on: the Lovasz theta SDP of G(n, p), the max-cut SDP of a
+-1-weighted G(n, p) or of a cycle, a dense random SDP feasible by construction, a multi-block
sparse SDP with an LP block, and the serving stream of theta / max-cut requests of uneven sizes
with Poisson arrivals. Every random number of an instance is a pure function of its seed."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
I, D = ctypes.c_int, ctypes.c_double
PI, PD = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)


class Inst(ctypes.Structure):
    """BriskInstance of brisk_gen.h: the numbers of an SDPA sparse file (owned by the library)."""
    _fields_ = [("m", I), ("nblk", I), ("bs", PI), ("c", PD), ("nnz", ctypes.c_int64),
                ("mat", PI), ("blk", PI), ("i", PI), ("j", PI), ("v", PD)]


_gen = None


def gen_lib():
    global _gen
    if _gen is None:
        path = os.path.join(_HERE, "libbrisk_gen.so")
        if not os.path.exists(path):
            raise RuntimeError("libbrisk_gen.so is not built: run __graft_entry__.build()")
        g = ctypes.CDLL(path)
        pI = ctypes.POINTER(Inst)
        u64 = ctypes.c_uint64
        g.brisk_gen_theta.argtypes = [I, D, u64, pI]
        g.brisk_gen_theta_edges.argtypes = [I, ctypes.c_int64, PI, pI]
        g.brisk_gen_maxcut.argtypes = [I, D, u64, pI]
        g.brisk_gen_maxcut_cycle.argtypes = [I, pI]
        g.brisk_gen_dense.argtypes = [I, I, u64, pI]
        g.brisk_gen_multi.argtypes = [I, I, I, I, u64, pI]
        g.brisk_gen_free.argtypes = [pI]
        g.brisk_gen_u01.argtypes = [u64, u64, u64]
        g.brisk_gen_u01.restype = D
        _gen = g
    return _gen


class Instance:
    """A generated instance (kind: theta, theta_edges, maxcut, maxcut_cycle, dense, multi, with
    the generator's arguments); frees its arrays when collected."""

    def __init__(self, kind, *args):
        self.s = Inst()
        g = gen_lib()
        if kind == "theta_edges":
            n, edges = args
            e = np.ascontiguousarray(np.asarray(edges, dtype=np.int32).reshape(-1))
            rc = g.brisk_gen_theta_edges(n, len(e) // 2, e.ctypes.data_as(PI), ctypes.byref(self.s))
        else:
            rc = getattr(g, "brisk_gen_" + kind)(*args, ctypes.byref(self.s))
        if rc != 0:
            raise ValueError("generator %s%r failed" % (kind, args))
        self.kind = kind
        self.args = args

    def __del__(self):
        try:
            gen_lib().brisk_gen_free(ctypes.byref(self.s))
        except Exception:
            pass

    @property
    def m(self):
        return self.s.m

    @property
    def bs(self):
        return [self.s.bs[k] for k in range(self.s.nblk)]

    def arrays(self):
        s = self.s
        nnz = s.nnz
        return dict(m=s.m, nblk=s.nblk, bs=np.ctypeslib.as_array(s.bs, (s.nblk,)).copy(),
                    c=np.ctypeslib.as_array(s.c, (s.m,)).copy(),
                    mat=np.ctypeslib.as_array(s.mat, (nnz,)).copy() if nnz else np.zeros(0, np.int32),
                    blk=np.ctypeslib.as_array(s.blk, (nnz,)).copy() if nnz else np.zeros(0, np.int32),
                    i=np.ctypeslib.as_array(s.i, (nnz,)).copy() if nnz else np.zeros(0, np.int32),
                    j=np.ctypeslib.as_array(s.j, (nnz,)).copy() if nnz else np.zeros(0, np.int32),
                    v=np.ctypeslib.as_array(s.v, (nnz,)).copy() if nnz else np.zeros(0))

    def sdpa(self):
        """(blocksizes, c, mat, blk, i, j, v) as solve_sdpa and Server.submit take them."""
        a = self.arrays()
        return (a["bs"], a["c"], a["mat"], a["blk"], a["i"], a["j"], a["v"])

    def data_args(self):
        s = self.s
        return (s.m, s.nblk, s.bs, s.c, s.nnz, s.mat, s.blk, s.i, s.j, s.v)


# the serving stream: theta / max-cut requests of uneven sizes with Poisson arrivals
STREAM = dict(count=4096, seed=0, nmin=20, nmax=200, pmin=0.05, pmax=0.3,
              desc="serving stream: 4096 theta/max-cut SDPs (50/50), n ~ U{20..200}, p ~ U[0.05,0.3], Poisson arrivals, seed 0")


def stream(count=STREAM["count"], seed=STREAM["seed"]):
    """Request k: (kind, n, p, instance seed 1000 + k) and a unit-rate exponential inter-arrival
    gap (scaled by 1 / lambda at run time)."""
    rng = np.random.default_rng(seed)
    theta = rng.random(count) < 0.5
    n = rng.integers(STREAM["nmin"], STREAM["nmax"] + 1, count)
    p = rng.uniform(STREAM["pmin"], STREAM["pmax"], count)
    gaps = rng.exponential(1.0, count)
    reqs = [("theta" if theta[k] else "maxcut", int(n[k]), float(p[k]), 1000 + k) for k in range(count)]
    return reqs, gaps
