"""Kernel-level parity of the hot path on the GPU, through the probe entry points of
libbrisk_amd.so (include/brisk_probe.h), on random points of real problems: the Schur
complement assembly of every route against the oracle's (pinned by brute force in
test_oracle.py), the constraint operators A(D) and A*(y) against dense numpy, the Lanczos bound of
the step-length test against eigenvalues, the DIMACS errors (including the cone errors of
non-PSD points) against the oracle's, whose cone errors are eigenvalues, and the end phase's
metric corrections against the oracle's (pinned by their definitions in test_oracle.py)."""
import ctypes
import numpy as np
import pytest
import brisk_amd
from _ctypes_util import (Instance, ArrInstance, oracle_lib, oracle_data, oracle_dimacs, oracle_solve, oracle_lanczos,
                          oracle_step_lambda, oracle_correct, PD)
from _dense import dense_constraints, random_spd, random_sym, pack, unpack

pytestmark = pytest.mark.gpu


class _BP(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("nblk", ctypes.c_int), ("bs", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("nnz", ctypes.c_int64), ("mat", ctypes.c_void_p), ("blk", ctypes.c_void_p), ("i", ctypes.c_void_p),
                ("j", ctypes.c_void_p), ("v", ctypes.c_void_p)]


def _lib():
    L = brisk_amd.load_library()
    P = ctypes.c_void_p
    L.brisk_probe_schur.argtypes = [P, PD, PD, ctypes.c_int, PD]
    L.brisk_probe_aop.argtypes = [P, PD, PD]
    L.brisk_probe_aty.argtypes = [P, PD, PD]
    L.brisk_probe_lammin.argtypes = [P, PD, PD, ctypes.c_int, PD]
    L.brisk_probe_dimacs.argtypes = [P, PD, PD, PD, PD]
    L.brisk_probe_correct.argtypes = [P, PD, PD, ctypes.c_int, PD, PD]
    return L


def bp(inst):
    s = inst.s
    return _BP(s.m, s.nblk, ctypes.cast(s.bs, ctypes.c_void_p), ctypes.cast(s.c, ctypes.c_void_p), s.nnz,
               ctypes.cast(s.mat, ctypes.c_void_p), ctypes.cast(s.blk, ctypes.c_void_p), ctypes.cast(s.i, ctypes.c_void_p),
               ctypes.cast(s.j, ctypes.c_void_p), ctypes.cast(s.v, ctypes.c_void_p))


def arr(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return x, x.ctypes.data_as(PD)


# every assembly route: single-entry gathers with a heavy identity row (theta), the Hadamard form
# (max-cut), blocks + LP with heavy groups (multi), the dense route (dense)
# (many blocks with m >= 64: only the pairs of constraints sharing a block are assembled)
SCHUR_CASES = [("theta", (30, 0.3, 3)), ("theta", (60, 0.15, 1)), ("maxcut", (80, 0.1, 0)),
               ("multi", (4, 10, 20, 60, 0)), ("multi", (6, 30, 40, 300, 2)), ("multi", (3, 12, 0, 40, 5)),
               ("multi", (24, 12, 60, 500, 7)), ("dense", (12, 30, 0)), ("dense", (40, 120, 1))]


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("case", SCHUR_CASES, ids=["%s%r" % c for c in SCHUR_CASES])
def test_schur_vs_oracle(case, prec):
    """M_ij = <A_i, X A_j Z^-1> from the GPU's assembly route against the oracle's assembly at
    the same random SPD X, Z (Z^-1 by numpy): FP64 to 1e-11 of max |M|; the dense route's FP32
    products (prec 0, the mixed path) to 2e-5."""
    if prec == 0 and case[0] != "dense":
        pytest.skip("the FP32 assembly is the dense route's only")
    inst = Instance(case[0], *case[1])
    print("case", case, "prec", prec, "m", inst.m, flush=True)
    A, C, bs, lp, norms = dense_constraints(inst)
    rng = np.random.default_rng(11)
    X = [random_spd(n, rng, l) for n, l in zip(bs, lp)]
    Z = [random_spd(n, rng, l) for n, l in zip(bs, lp)]
    Zi = [np.linalg.inv(B) for B in Z]
    m = inst.m
    xp, xpp = arr(pack(X, lp))
    zp, zpp = arr(pack(Z, lp))
    zip_, zipp = arr(pack(Zi, lp))
    Mo = np.zeros(m * m)
    assert oracle_lib().oracle_schur(ctypes.byref(oracle_data(inst)), xpp, zipp, Mo.ctypes.data_as(PD)) == 0
    Mo = Mo.reshape(m, m, order="F") * np.outer(norms, norms)     # the oracle's M is of the equilibrated A_i
    Mg = np.zeros(m * m)
    rc = _lib().brisk_probe_schur(ctypes.byref(bp(inst)), xpp, zpp, prec, Mg.ctypes.data_as(PD))
    assert rc == 0
    Mg = Mg.reshape(m, m, order="F")
    err = np.abs(Mg - Mo).max() / np.abs(Mo).max()
    print("max |M_gpu - M_oracle| / max |M| = %.2e" % err, flush=True)
    assert err <= (1e-11 if prec else 2e-5)
    assert np.allclose(Mg, Mg.T, rtol=0, atol=(1e-12 if prec else 1e-5) * np.abs(Mo).max())


OP_CASES = SCHUR_CASES + [("maxcut_cycle", (7,))]


@pytest.mark.parametrize("case", OP_CASES, ids=["%s%r" % c for c in OP_CASES])
def test_operators_vs_dense(case):
    """A(D)_i = <A_i, D> and A*(y) = sum y_i A_i at random symmetric D and y, against dense
    numpy on the problem as given (1e-13 relative)."""
    inst = Instance(case[0], *case[1])
    A, C, bs, lp, _ = dense_constraints(inst, scaled=False)
    rng = np.random.default_rng(5)
    D = [random_sym(n, rng, l) for n, l in zip(bs, lp)]
    y = rng.standard_normal(inst.m)
    want_aop = np.array([sum(float(np.sum(Ai[k] * D[k])) for k in range(len(bs))) for Ai in A])
    want_aty = [sum(y[i] * A[i][k] for i in range(inst.m)) for k in range(len(bs))]
    L = _lib()
    dp, dpp = arr(pack(D, lp))
    got = np.zeros(inst.m)
    assert L.brisk_probe_aop(ctypes.byref(bp(inst)), dpp, got.ctypes.data_as(PD)) == 0
    sc = np.abs(want_aop).max() + 1e-300
    print("A(D): max err %.2e of %.2e" % (np.abs(got - want_aop).max(), sc), flush=True)
    assert np.abs(got - want_aop).max() <= 1e-13 * sc * max(1, max(bs))
    yp, ypp = arr(y)
    out = np.zeros(len(dp))
    assert L.brisk_probe_aty(ctypes.byref(bp(inst)), ypp, out.ctypes.data_as(PD)) == 0
    got_b = unpack(out, bs, lp)
    sc = max(np.abs(B).max() for B in want_aty) + 1e-300
    for g, w in zip(got_b, want_aty):
        assert np.abs(g - w).max() <= 1e-13 * sc * max(1, inst.m ** 0.5)


def one_block(n):
    """A problem with one SDP block of order n and one constraint (tr X = 1): the step-length
    probe takes its block structure from it."""
    return ArrInstance(dict(m=1, nblk=1, bs=np.array([n]), c=np.array([1.0]), mat=np.ones(n, np.int32),
                            blk=np.ones(n, np.int32), i=np.arange(1, n + 1, dtype=np.int32),
                            j=np.arange(1, n + 1, dtype=np.int32), v=np.ones(n)))


@pytest.mark.parametrize("n", [5, 17, 30, 32, 64, 200, 256, 300, 700])
def test_lanczos_bound_vs_oracle(n):
    """The Lanczos bound of the step-length test (blocks of order <= 256: the fused one-workgroup
    kernel; above: the per-step launches) against the oracle's restatement of the reference's
    bound (pinned by the Rayleigh-Ritz identity in test_oracle.py): for order <= 30 the exact
    lambda_min as the reference's small_lammin (1e-9), for 31 and 32 the 30-step bound run to
    the end, above the 30-step bound with Ritz accuracy 0.02 (1e-8 relative). That bound,
    theta - res, lies within res of an eigenvalue but not always below the smallest (before
    the Krylov space reaches its eigenvector: -0.924 for a lambda_min of -0.934 here, in the
    reference's algorithm as in ours); the step's Cholesky test keeps the iterate in the cone.
    Mode 0 tests lambda_min(X), mode 1 lambda_min(L^-1 dX L^-T)."""
    rng = np.random.default_rng(n)
    inst = one_block(n)
    L = _lib()
    for trial in range(3):
        G = rng.standard_normal((n, n))
        X = G @ G.T / n + (0.05 + trial) * np.eye(n)
        dX = random_sym(n, rng, False) / np.sqrt(n)
        Lc = np.linalg.cholesky(X)
        for mode in (0, 1):
            D_, Lf = (X, None) if mode == 0 else (dX, Lc)
            if mode == 0:
                lam = np.linalg.eigvalsh(X)[0]
            else:
                Li = np.linalg.inv(Lc)
                lam = np.linalg.eigvalsh(Li @ dX @ Li.T)[0]
            if n <= 30:
                want = oracle_step_lambda(D_, Lf)
            elif n <= 32:
                want = oracle_lanczos(D_, Lf, K=30, tol=0.0)[0]
            else:
                want = oracle_lanczos(D_, Lf, K=30, tol=0.02)[0]
            xp, xpp = arr(X.reshape(-1, order="F"))
            dp, dpp = arr(dX.reshape(-1, order="F"))
            out = ctypes.c_double()
            assert L.brisk_probe_lammin(ctypes.byref(bp(inst)), xpp, dpp, mode, ctypes.byref(out)) == 0
            got = out.value
            print("n %d trial %d mode %d: gpu %.12g oracle %.12g lambda_min %.12g" % (n, trial, mode, got, want, lam), flush=True)
            assert abs(got - want) <= (1e-9 if n <= 30 else 1e-8) * (1 + abs(want))


DIMACS_CASES = [("theta", (30, 0.3, 3)), ("multi", (4, 10, 20, 60, 0)), ("maxcut", (60, 0.1, 2)), ("dense", (12, 30, 0))]


@pytest.mark.parametrize("shift", [0.0, 1e-9, 1e-7, 1e-5])
@pytest.mark.parametrize("case", DIMACS_CASES, ids=["%s%r" % c for c in DIMACS_CASES])
def test_dimacs_vs_oracle(case, shift):
    """The six DIMACS errors of a point (X, y) as the GPU measures a returned point, against the
    oracle's (eigenvalues for the cone errors): the oracle's optimal point with X shifted by
    -shift I and y moved by shift times a random vector, so that X and Z = C - A'y leave the
    cone by about shift (inside it, or on its boundary, for shift 0). Errors 1, 3, 5 and 6 and the
    objectives to 1e-10 relative; the cone errors, from the factorization ladder, no smaller than
    the eigenvalue's and within 1.1e-4 relative above it."""
    inst = Instance(case[0], *case[1])
    A, C, bs, lp, _ = dense_constraints(inst, scaled=False)
    o = oracle_solve(inst)
    assert o["status"] == 0
    X = unpack(o["X"], bs, lp)
    X = [B - shift * np.eye(B.shape[0]) for B in X]
    rng = np.random.default_rng(3)
    y = o["y"] + shift * rng.standard_normal(inst.m)
    xv = pack(X, lp)
    want, po, do = oracle_dimacs(inst, xv, y)
    xp, xpp = arr(xv)
    yp, ypp = arr(y)
    err = np.zeros(6)
    obj = np.zeros(2)
    assert _lib().brisk_probe_dimacs(ctypes.byref(bp(inst)), xpp, ypp, err.ctypes.data_as(PD), obj.ctypes.data_as(PD)) == 0
    print("gpu", err, obj, flush=True)
    print("oracle", want, po, do, flush=True)
    assert obj[0] == pytest.approx(po, rel=1e-10, abs=1e-13) and obj[1] == pytest.approx(do, rel=1e-10, abs=1e-13)
    for e in (0, 2, 4, 5):
        assert err[e] == pytest.approx(want[e], rel=1e-8, abs=1e-14)
    for e in (1, 3):
        # below the ladder's first rung (1e-12 of the norms) a cone error is resolution, not a value
        floor = 2e-12
        assert err[e] >= want[e] * (1 - 1e-9) - floor
        assert err[e] <= want[e] * (1 + 1.1e-4) + floor


# every assembly route at the correction's scaling pairs: gathers (theta), Hadamard (max-cut),
# blocks + LP with heavy groups (multi), the dense route with full factors (dense)
CORRECT_CASES = [("theta", (9, 0.4, 1)), ("theta", (30, 0.3, 3)), ("maxcut", (60, 0.1, 2)), ("multi", (2, 6, 5, 12, 4)),
                 ("multi", (4, 10, 20, 60, 0)), ("dense", (6, 10, 2)), ("dense", (12, 30, 0))]


@pytest.mark.parametrize("point", [(1e-3, 1.0), (1.0, 30.0), (0.0, 1.001)], ids=["near", "far", "singular"])
@pytest.mark.parametrize("case", CORRECT_CASES, ids=["%s%r" % c for c in CORRECT_CASES])
def test_sqrt_restore_vs_oracle(case, point):
    """The square-root-metric restoration (three passes) of X = scale X* + shift I at the oracle's
    optimum X* (1.001 X*: the numerically singular X the restoration is for) through the GPU's
    eigendecomposition, scalings, assembly, FP64 factor and PCG, against the oracle's: the same
    steps, X within 1e-9 of |X| plus 1e-6 of the correction (near-zero eigenvalues of a
    singular X enter D through their square roots)."""
    inst = Instance(case[0], *case[1])
    A, C, bs, lp, _ = dense_constraints(inst, scaled=False)
    o = oracle_solve(inst)
    assert o["status"] == 0
    shift, scale = point
    X = [scale * B + shift * np.eye(B.shape[0]) for B in unpack(o["X"], bs, lp)]
    xv = pack(X, lp)
    rc, want, wst = oracle_correct(inst, xv, o["y"], 1, 3)
    assert rc == 0
    xp, xpp = arr(xv)
    yp, ypp = arr(o["y"])
    got = np.zeros_like(xv)
    st = np.zeros(3)
    print("case", case, point, flush=True)
    assert _lib().brisk_probe_correct(ctypes.byref(bp(inst)), xpp, ypp, 1, got.ctypes.data_as(PD), st.ctypes.data_as(PD)) == 0
    print("steps gpu", st, "oracle", wst, "diff", np.abs(got - want).max(), "corr", np.abs(want - xv).max(), flush=True)
    assert list(st) == list(wst)
    assert np.abs(got - want).max() <= 1e-9 * np.abs(xv).max() + 1e-6 * np.abs(want - xv).max()


@pytest.mark.parametrize("point", [(1e-3, 1.0), (0.1, 3.0)], ids=["near", "far"])
@pytest.mark.parametrize("case", [("theta", (30, 0.3, 3)), ("maxcut", (60, 0.1, 2))], ids=["theta", "maxcut"])
def test_xmetric_polish_vs_oracle(case, point):
    """The X-metric polish (two passes) at y = y* - c with A*(c) = 0.05 I (Z = C - A*(y)
    positive definite), against the oracle's: the same steps, X within 1e-9 of |X| plus 1e-7 of
    the correction."""
    inst = Instance(case[0], *case[1])
    A, C, bs, lp, _ = dense_constraints(inst, scaled=False)
    o = oracle_solve(inst)
    shift, scale = point
    X = [scale * B + shift * np.eye(B.shape[0]) for B in unpack(o["X"], bs, lp)]
    basis = np.array([np.concatenate([Ai[k].ravel() for k in range(len(X))]) for Ai in A]).T
    eye = np.concatenate([np.eye(n).ravel() for n in bs])
    y = o["y"] - np.linalg.lstsq(basis, 0.05 * eye, rcond=None)[0]
    xv = pack(X, lp)
    rc, want, wst = oracle_correct(inst, xv, y, 0, 2)
    assert rc == 0
    xp, xpp = arr(xv)
    yp, ypp = arr(y)
    got = np.zeros_like(xv)
    st = np.zeros(3)
    print("case", case, point, flush=True)
    assert _lib().brisk_probe_correct(ctypes.byref(bp(inst)), xpp, ypp, 0, got.ctypes.data_as(PD), st.ctypes.data_as(PD)) == 0
    print("steps gpu", st, "oracle", wst, "diff", np.abs(got - want).max(), "corr", np.abs(want - xv).max(), flush=True)
    assert list(st[:2]) == list(wst)
    assert np.abs(got - want).max() <= 1e-9 * np.abs(xv).max() + 1e-7 * np.abs(want - xv).max()
