"""The oracle pinned on the CPU: against the reference's own outputs (golden fixtures written
by the reference library compiled from its sources), against closed-form optima, and against
brute-force restatements of its building blocks (Schur complement, smallest eigenvalue, DIMACS
errors, the end phase's metric corrections) from their definitions."""
import math
import numpy as np
import pytest

from _dense import dense_constraints, random_spd, pack, unpack  # noqa: F401
from _ctypes_util import oracle_lanczos, oracle_step_lambda  # noqa: E402
from _ctypes_util import (Instance, ArrInstance, golden, golden_instance, oracle_solve, oracle_dimacs,
                          oracle_lib, oracle_data, oracle_correct, PD, D)
import ctypes

GOLD = golden()
TOL_OBJ = 1e-7       # objectives: relative to 1 + |value|
TOL_DIMACS = 1e-8


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def objective_ok(ours, rec, key):
    """Within 1e-7 of the reference's FP64 solve, or at least as close to the optimum (the
    reference's double-double solve) as the reference's FP64 solve is."""
    ref = rec[key]
    if rel(ours, ref) <= TOL_OBJ:
        return True
    dd = rec["dd"][key]
    return rel(ours, dd) <= max(rel(ref, dd), 1e-9)


@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_vs_reference(name):
    rec = GOLD[name]
    inst = golden_instance(rec)
    o = oracle_solve(inst)
    assert rec["status"] == 0 and o["status"] == 0, (rec["status"], o["status"])
    assert objective_ok(o["pobj"], rec, "pobj"), (o["pobj"], rec["pobj"], rec["dd"]["pobj"])
    assert objective_ok(o["dobj"], rec, "dobj"), (o["dobj"], rec["dobj"], rec["dd"]["dobj"])
    assert max(abs(e) for e in o["err"]) <= TOL_DIMACS, o["err"]
    # the errors the oracle reports are those of the point it returns
    err, po, do = oracle_dimacs(inst, o["X"], o["y"])
    assert np.allclose(err, o["err"], rtol=1e-6, atol=1e-15)
    assert po == pytest.approx(o["pobj"], rel=1e-13) and do == pytest.approx(o["dobj"], rel=1e-13)


@pytest.mark.parametrize("name", sorted(k for k in GOLD if "X" in GOLD[k]))
def test_oracle_dimacs_on_reference_point(name):
    """The DIMACS errors and objectives the oracle computes at the reference's own returned
    point equal those the reference reported for it."""
    rec = GOLD[name]
    inst = golden_instance(rec)
    err, po, do = oracle_dimacs(inst, np.array(rec["X"]), np.array(rec["y"]))
    assert po == pytest.approx(rec["pobj"], rel=1e-12, abs=1e-13)
    assert do == pytest.approx(rec["dobj"], rel=1e-12, abs=1e-13)
    ref = np.array(rec["dimacs"])
    # the cone errors (2 and 4) are eigenvalues resolved to about eps |X| (eigen solvers differ)
    tol = 1e-13 + 1e-6 * np.abs(ref) + np.array([0, 1e-11, 0, 1e-11, 0, 0])
    assert np.all(np.abs(err - ref) <= tol), (err, ref)


# ---- closed forms -----------------------------------------------------------------------------

def cycle_edges(n):
    return [(i, (i + 1) % n) for i in range(n)]


def petersen():
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    return outer + spokes + inner


THETA = [
    ("C5", 5, cycle_edges(5), math.sqrt(5.0)),
    ("C7", 7, cycle_edges(7), 7 * math.cos(math.pi / 7) / (1 + math.cos(math.pi / 7))),
    ("C9", 9, cycle_edges(9), 9 * math.cos(math.pi / 9) / (1 + math.cos(math.pi / 9))),
    ("C41", 41, cycle_edges(41), 41 * math.cos(math.pi / 41) / (1 + math.cos(math.pi / 41))),
    ("C6", 6, cycle_edges(6), 3.0),
    ("K6", 6, [(i, j) for i in range(6) for j in range(i + 1, 6)], 1.0),
    ("empty5", 5, [], 5.0),
    ("Petersen", 10, petersen(), 4.0),
]


@pytest.mark.parametrize("case", THETA, ids=[c[0] for c in THETA])
def test_oracle_theta_closed_form(case):
    """Lovasz theta: odd cycle n cos(pi/n)/(1+cos(pi/n)), even cycle n/2, complete graph 1,
    empty graph n, Petersen graph 4; the solve's primal value is -theta."""
    _, n, edges, theta = case
    inst = Instance("theta_edges", n, edges)
    assert inst.m == 1 + len(edges)
    o = oracle_solve(inst)
    assert o["status"] == 0
    assert o["pobj"] == pytest.approx(-theta, rel=1e-7)
    assert o["dobj"] == pytest.approx(-theta, rel=1e-7)


@pytest.mark.parametrize("n", [3, 5, 7, 8, 12, 15])
def test_oracle_maxcut_cycle_closed_form(n):
    """Max-cut relaxation of the unit-weight cycle: n for even n, n (1 + cos(pi/n)) / 2 for odd n
    (the Goemans-Williamson bound of C_n); the solve minimises -<L/4, X>."""
    o = oracle_solve(Instance("maxcut_cycle", n))
    want = n if n % 2 == 0 else n * (1 + math.cos(math.pi / n)) / 2
    assert o["status"] == 0
    assert o["pobj"] == pytest.approx(-want, rel=1e-7)


def test_oracle_example1_value():
    """The SDPA manual's example 1: optimal value 41.9 (BRISK's sign convention)."""
    rec = GOLD["file_example1"]
    o = oracle_solve(golden_instance(rec))
    assert o["status"] == 0 and o["dobj"] == pytest.approx(41.9, abs=1e-6)


# SDPLIB's published optimal values (7 digits), in BRISK's sign convention
SDPLIB = {"file_truss1": 8.999996, "file_control1": -17.78463, "file_theta1": -23.0, "file_theta3": -42.16698,
          "file_arch0": -0.5665172, "file_mcp250-1": -317.2643}


@pytest.mark.parametrize("name", sorted(SDPLIB))
def test_oracle_sdplib_values(name):
    o = oracle_solve(golden_instance(GOLD[name]))
    assert o["status"] == 0 and abs(o["pobj"] - SDPLIB[name]) <= 6e-7 * (1 + abs(SDPLIB[name]))


# ---- building blocks from their definitions ---------------------------------------------------

SCHUR_CASES = [("multi", (2, 6, 5, 12, 4)), ("multi", (3, 5, 0, 9, 1)), ("dense", (6, 10, 2)),
               ("theta", (9, 0.4, 1)), ("maxcut", (10, 0.3, 2))]


@pytest.mark.parametrize("case", SCHUR_CASES, ids=["%s%r" % c for c in SCHUR_CASES])
def test_oracle_schur_brute_force(case):
    """M_ab = <A_a, X A_b Zi> summed over blocks, from dense matrices (the HKM Schur complement),
    against the oracle's assembly (route-0 expansion, Hadamard form or dense formula)."""
    inst = Instance(case[0], *case[1])
    A, C, bs, lp, _ = dense_constraints(inst)
    rng = np.random.default_rng(7)
    X = [random_spd(n, rng, l) for n, l in zip(bs, lp)]
    Zi = [random_spd(n, rng, l) for n, l in zip(bs, lp)]
    m = inst.m
    M = np.zeros((m, m))
    for a in range(m):
        for b in range(m):
            M[a, b] = sum(float(np.sum(A[a][k] * (X[k] @ A[b][k] @ Zi[k]))) for k in range(len(bs)))
    Mo = np.zeros(m * m)
    xp, zp = pack(X, lp), pack(Zi, lp)
    rc = oracle_lib().oracle_schur(ctypes.byref(oracle_data(inst)), xp.ctypes.data_as(PD), zp.ctypes.data_as(PD),
                                   Mo.ctypes.data_as(PD))
    assert rc == 0
    Mo = Mo.reshape(m, m, order="F")
    assert np.allclose(Mo, M, rtol=1e-11, atol=1e-12 * np.abs(M).max())


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 40])
def test_oracle_lammin(n):
    """The smallest eigenvalue (Householder tridiagonalisation + bisection) is a lower bound
    within 1e-10 of numpy's."""
    rng = np.random.default_rng(n)
    for trial in range(4):
        G = rng.standard_normal((n, n))
        S = (G + G.T) / 2 + (trial - 1.5) * np.eye(n)
        lo = np.linalg.eigvalsh(S)[0]
        B = np.asfortranarray(S).reshape(-1, order="F").copy()
        got = oracle_lib().oracle_lammin(n, B.ctypes.data_as(PD))
        sc = np.abs(S).max()
        assert got <= lo + 1e-14 * sc and got >= lo - 1e-10 * (1 + sc)


@pytest.mark.parametrize("name", ["file_example2", "file_truss1", "multi_4_10_20_60_s0", "theta_30_0.3_s3"])
def test_oracle_dimacs_definitions(name):
    """DIMACS errors at an arbitrary point from their definitions (Mittelmann's benchmark):
    err1 = ||A(X) - b|| / (1 + ||b||_1), err2 = max(0, -lmin(X)) / (1 + ||b||_1),
    err3 = ||C - A'y - Z|| / (1 + ||C||_1) with Z = C - A'y, err4 = max(0, -lmin(Z)) / (1 + ||C||_1),
    err5 = (<C,X> - b'y) / (1 + |<C,X>| + |b'y|), err6 = <X,Z> / (1 + |<C,X>| + |b'y|);
    ||C||_1 sums |entries| of every block, lmin over all blocks."""
    rec = GOLD[name]
    inst = golden_instance(rec)
    a = inst.arrays()
    m, bs = a["m"], [abs(int(b)) for b in a["bs"]]
    lp = [int(b) < 0 for b in a["bs"]]
    Araw = [[np.zeros((n, n)) for n in bs] for _ in range(m + 1)]
    for t, k, i, j, v in zip(a["mat"], a["blk"], a["i"], a["j"], a["v"]):
        Araw[t][k - 1][i - 1, j - 1] += v
        if i != j:
            Araw[t][k - 1][j - 1, i - 1] += v
    C = [-B for B in Araw[0]]
    b = np.array(a["c"])
    rng = np.random.default_rng(3)
    X = [random_spd(n, rng, l) - 0.3 * np.eye(n) for n, l in zip(bs, lp)]
    y = rng.standard_normal(m)
    Z = [C[k] - sum(y[i] * Araw[i + 1][k] for i in range(m)) for k in range(len(bs))]
    AX = np.array([sum(float(np.sum(Araw[i + 1][k] * X[k])) for k in range(len(bs))) for i in range(m)])
    nb1 = np.abs(b).sum()
    nc1 = sum(float(np.abs(np.diag(B)).sum() if l else np.abs(B).sum()) for B, l in zip(C, lp))
    lx = min(float(np.linalg.eigvalsh(B)[0]) for B in X)
    lz = min(float(np.linalg.eigvalsh(B)[0]) for B in Z)
    po = sum(float(np.sum(C[k] * X[k])) for k in range(len(bs)))
    do = float(b @ y)
    xz = sum(float(np.sum(X[k] * Z[k])) for k in range(len(bs)))
    want = [np.linalg.norm(AX - b) / (1 + nb1), max(0.0, -lx) / (1 + nb1), 0.0, max(0.0, -lz) / (1 + nc1),
            (po - do) / (1 + abs(po) + abs(do)), xz / (1 + abs(po) + abs(do))]
    err, gpo, gdo = oracle_dimacs(inst, pack(X, lp), y)
    assert gpo == pytest.approx(po, rel=1e-12) and gdo == pytest.approx(do, rel=1e-12)
    assert np.allclose(err, want, rtol=1e-8, atol=1e-14), (err, want)


@pytest.mark.parametrize("name", sorted(k for k in GOLD if k != "file_arch0"))
def test_oracle_hsd_vs_reference(name):
    """The self-dual embedding variant (method 1) reaches the reference's optimum too: the
    normalised iterate (X, y) / tau, DIMACS errors <= 1e-8, objectives as the standard method's."""
    rec = GOLD[name]
    inst = golden_instance(rec)
    o = oracle_solve(inst, method=1)
    assert o["status"] == 0, o["status"]
    assert objective_ok(o["pobj"], rec, "pobj"), (o["pobj"], rec["pobj"], rec["dd"]["pobj"])
    assert objective_ok(o["dobj"], rec, "dobj"), (o["dobj"], rec["dobj"], rec["dd"]["dobj"])
    assert max(abs(e) for e in o["err"]) <= TOL_DIMACS, o["err"]


# ---- the Lanczos bound of the step-length test (reference solver.c:6029-6225) ------------------

def _krylov_ritz(T, k):
    """Rayleigh-Ritz on the k-dimensional Krylov space of T from the reference's start vector,
    by a QR of the Krylov matrix (independent of the Lanczos recurrence): the smallest Ritz value
    and the residual norm of its Ritz vector."""
    n = T.shape[0]
    q = 1.0 + 0.5 * np.sin(1.7 * np.arange(n) + 0.3)
    q /= np.linalg.norm(q)
    Kr = [q]
    for _ in range(k - 1):
        v = T @ Kr[-1]
        Kr.append(v / np.linalg.norm(v))
    Q, _ = np.linalg.qr(np.array(Kr).T)
    w, V = np.linalg.eigh(Q.T @ T @ Q)
    y = Q @ V[:, 0]
    return w[0], np.linalg.norm(T @ y - w[0] * y)


@pytest.mark.parametrize("with_l", [False, True])
@pytest.mark.parametrize("n", [40, 90])
def test_oracle_lanczos_rayleigh_ritz(n, with_l):
    """After k = 6 steps (no early stop) the restated Lanczos gives the smallest Ritz value of T on
    the Krylov space span{q, Tq, .., T^5 q} and |beta z_last| = the residual of its Ritz vector;
    with L, T = L^{-1} D L^{-T}."""
    rng = np.random.default_rng(n + with_l)
    G = rng.standard_normal((n, n))
    D_ = (G + G.T) / 2
    L = None
    T = D_
    if with_l:
        H = rng.standard_normal((n, n))
        L = np.linalg.cholesky(H @ H.T / n + np.eye(n))
        Li = np.linalg.inv(L)
        T = Li @ D_ @ Li.T
    b, k, theta, res = oracle_lanczos(D_, L, K=6, tol=0.0)
    th_rr, res_rr = _krylov_ritz(T, 6)
    sc = np.abs(T).max()
    assert k == 6
    assert abs(theta - th_rr) <= 1e-9 * sc
    assert abs(res - res_rr) <= 1e-7 * sc
    assert b == pytest.approx(theta - res, rel=1e-15, abs=1e-300)


@pytest.mark.parametrize("n", [3, 12, 30])
def test_oracle_lanczos_full_space_exact(n):
    """K = n steps without early stop span the whole space: the bound is lambda_min."""
    rng = np.random.default_rng(100 + n)
    G = rng.standard_normal((n, n))
    D_ = (G + G.T) / 2
    b, k, theta, res = oracle_lanczos(D_, None, K=n, tol=0.0)
    lam = np.linalg.eigvalsh(D_)[0]
    assert abs(b - lam) <= 1e-9 * (1 + abs(lam))


@pytest.mark.parametrize("n", [33, 64, 200])
def test_oracle_lanczos_conservative(n):
    """With the reference's 30 steps and accuracy 0.02, on shifted Wishart matrices (a clustered
    lower spectrum, the step test's usual case) the bound theta - res stays below lambda_min,
    and theta (a Ritz value) is never below it; for order <= 32 the reference's bound is the
    exact eigenvalue (small_lammin)."""
    rng = np.random.default_rng(200 + n)
    for trial in range(4):
        G = rng.standard_normal((n, n))
        D_ = G @ G.T / n + (trial - 1.5) * np.eye(n)
        lam = np.linalg.eigvalsh(D_)[0]
        b, k, theta, res = oracle_lanczos(D_, None)
        sc = np.abs(D_).max()
        assert b <= lam + 1e-12 * sc and theta >= lam - 1e-12 * sc
        assert oracle_step_lambda(D_) == b
    for m in (5, 32):
        G = rng.standard_normal((m, m))
        D_ = (G + G.T) / 2
        assert abs(oracle_step_lambda(D_) - np.linalg.eigvalsh(D_)[0]) <= 1e-9 * (1 + np.abs(D_).max())


# ---- the end phase's metric corrections -------------------------------------------------------

def _sqrtm(B):
    w, V = np.linalg.eigh(B)
    return (V * np.sqrt(np.maximum(w, 0))) @ V.T


def _correct_setup(case, shift, scale):
    """(instance, dense A_i, C, orders, LP flags, b, the point X = scale X* + shift I, y = y*)"""
    inst = Instance(case[0], *case[1])
    A, C, bs, lp, _ = dense_constraints(inst, scaled=False)
    b = np.array(inst.arrays()["c"], dtype=np.float64)
    o = oracle_solve(inst)
    assert o["status"] == 0
    X = [scale * B + shift * np.eye(B.shape[0]) for B in unpack(o["X"], bs, lp)]
    return inst, A, C, bs, lp, b, X, o["y"]


def _aop(A, X):
    return np.array([sum(float(np.sum(Ai[k] * X[k])) for k in range(len(X))) for Ai in A])


CORRECT_CASES = [("theta", (9, 0.4, 1)), ("maxcut", (10, 0.3, 2)), ("multi", (2, 6, 5, 12, 4)), ("dense", (6, 10, 2))]


@pytest.mark.parametrize("point", [(1e-3, 1.0), (0.1, 3.0), (1.0, 30.0)])
@pytest.mark.parametrize("case", CORRECT_CASES, ids=["%s%r" % c for c in CORRECT_CASES])
def test_oracle_sqrt_restore_definition(case, point):
    """One pass of the square-root-metric restoration against what defines it: with D = X^{1/2}
    (numpy's eigendecomposition) and a the step taken, dX = (X' - X) / a satisfies A(dX) = b - A(X)
    (so A(X') - b = (1 - a)(A(X) - b)) and D^-1 dX D^-1 lies in the span of the A_i; a is the
    largest of 1, 1/2, ... for which 0.99 I + a D^-1 dX D^-1 is positive definite in every block.
    Those conditions fix dX and a uniquely (A (D (x) D) A' is positive definite)."""
    inst, A, C, bs, lp, b, X, y = _correct_setup(case, *point)
    rc, xo, steps = oracle_correct(inst, pack(X, lp), y, 1, 1)
    assert rc == 0 and steps[0] > 0
    a = steps[0]
    Xo = unpack(xo, bs, lp)
    r0 = b - _aop(A, X)
    dX = [(Xo[k] - X[k]) / a for k in range(len(X))]
    assert np.linalg.norm(_aop(A, dX) - r0) <= 1e-9 * np.linalg.norm(r0)
    Dh = [_sqrtm(B) for B in X]
    E = [np.linalg.solve(Dh[k], np.linalg.solve(Dh[k], dX[k]).T).T for k in range(len(X))]
    basis = np.array([np.concatenate([Ai[k].ravel() for k in range(len(X))]) for Ai in A]).T
    e = np.concatenate([B.ravel() for B in E])
    coef = np.linalg.lstsq(basis, e, rcond=None)[0]
    assert np.linalg.norm(basis @ coef - e) <= 1e-8 * np.linalg.norm(e)
    lam = lambda t: min(np.linalg.eigvalsh(0.99 * np.eye(B.shape[0]) + t * B)[0] for B in E)
    assert lam(a) > 0
    assert a == 1.0 or lam(2 * a) <= 1e-12
    assert a in [0.5 ** h for h in range(8)]


@pytest.mark.parametrize("point", [(1e-3, 1.0), (0.1, 3.0)])
@pytest.mark.parametrize("case", [("theta", (9, 0.4, 1)), ("maxcut", (10, 0.3, 2))], ids=["theta", "maxcut"])
def test_oracle_xmetric_polish_definition(case, point):
    """One pass of the X-metric polish against what defines it: at y = y* - c (A*(c) = 0.05 I, so
    Z = C - A*(y) is positive definite), dX = (X' - X) / a satisfies A(dX) = b - A(X) and is
    sym(Zi W X) for some W in the span of the A_i; a is the largest of 1, 1/2, ... keeping X + a dX
    positive definite."""
    inst, A, C, bs, lp, b, X, y = _correct_setup(case, *point)
    basis = np.array([np.concatenate([Ai[k].ravel() for k in range(len(X))]) for Ai in A]).T
    eye = np.concatenate([np.eye(n).ravel() for n in bs])
    c = np.linalg.lstsq(basis, 0.05 * eye, rcond=None)[0]
    assert np.linalg.norm(basis @ c - 0.05 * eye) <= 1e-10
    y = y - c
    Z = [C[k] - sum(y[i] * A[i][k] for i in range(len(A))) for k in range(len(X))]
    assert min(np.linalg.eigvalsh(B)[0] for B in Z) > 0
    rc, xo, steps = oracle_correct(inst, pack(X, lp), y, 0, 1)
    assert rc == 0 and steps[0] > 0
    a = steps[0]
    Xo = unpack(xo, bs, lp)
    r0 = b - _aop(A, X)
    dX = [(Xo[k] - X[k]) / a for k in range(len(X))]
    assert np.linalg.norm(_aop(A, dX) - r0) <= 1e-9 * np.linalg.norm(r0)
    Zi = [np.linalg.inv(B) for B in Z]
    cols = []
    for Ai in A:
        G = [Zi[k] @ Ai[k] @ X[k] for k in range(len(X))]
        cols.append(np.concatenate([((B + B.T) / 2).ravel() for B in G]))
    cols = np.array(cols).T
    e = np.concatenate([B.ravel() for B in dX])
    coef = np.linalg.lstsq(cols, e, rcond=None)[0]
    assert np.linalg.norm(cols @ coef - e) <= 1e-8 * np.linalg.norm(e)
    lam = lambda t: min(np.linalg.eigvalsh(X[k] + t * dX[k])[0] for k in range(len(X)))
    assert lam(a) > 0 and (a == 1.0 or lam(2 * a) <= 1e-12)
