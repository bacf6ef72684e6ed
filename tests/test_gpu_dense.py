"""Kernel-level parity of the dense batched linear algebra (MFMA GEMM, blocked Cholesky,
triangular inverse, matrix-vector products) against numpy/LAPACK on the same inputs."""
import numpy as np
import pytest
from _gpu_util import lib, DBuf, GemmItem, MItem, GemvItem, SymvItem, TrsvItem, EigItem, items_array

pytestmark = pytest.mark.gpu


def _spd(n, rng, cond=1e3):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    e = np.logspace(0, np.log10(cond), n)
    return (Q * e) @ Q.T


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("shapes", [[(70, 45, 33), (1, 1, 1), (64, 64, 64), (130, 17, 200), (5, 129, 7)],
                                    [(300, 257, 129), (192, 200, 64), (640, 384, 31)],
                                    [(2011, 1999, 77), (200, 300, 50), (150, 400, 21)]],
                         ids=["small-tiles", "large-tiles", "128-tiles"])
def test_gemm_batched_ragged(dt, ta, tb, shapes):
    rng = np.random.default_rng(1 + ta + 2 * tb)
    A_parts, B_parts, C_parts, rows, refs = [], [], [], [], []
    oa = ob = oc = 0
    for (M, N, K) in shapes:
        A = rng.standard_normal((K, M) if ta else (M, K))
        B = rng.standard_normal((N, K) if tb else (K, N))
        C = rng.standard_normal((M, N))
        alpha, beta = 0.7, -0.3
        ref = alpha * (A.T if ta else A) @ (B.T if tb else B) + beta * C
        lda = A.shape[0]; ldb = B.shape[0]
        rows.append((M, N, K, lda, ldb, M, oa, ob, oc, -1, 0))
        A_parts.append(A.flatten(order="F")); B_parts.append(B.flatten(order="F")); C_parts.append(C.flatten(order="F"))
        refs.append((oc, M, N, ref))
        oa += A.size; ob += B.size; oc += C.size
    dA = DBuf(np.concatenate(A_parts).astype(dt)); dB = DBuf(np.concatenate(B_parts).astype(dt))
    dC = DBuf(np.concatenate(C_parts).astype(dt))
    it = items_array(GemmItem, rows)
    L = lib()
    f = L.bk_gemm_f64 if dt == np.float64 else L.bk_gemm_f32
    assert f(ta, tb, 0.7, -0.3, dA.p, dB.p, dC.p, it, len(rows), None) == 0
    out = dC.get()
    tol = 1e-12 if dt == np.float64 else 2e-4
    for (oc, M, N, ref) in refs:
        got = out[oc:oc + M * N].reshape((M, N), order="F")
        assert np.allclose(got, ref, rtol=tol, atol=tol * np.abs(ref).max()), (M, N)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("n", [300, 2100])
def test_gemm_triangular_flags(dt, n):
    """Zi = Linv' Linv with the K range trimmed (2|8) and a lower-only update (1); n = 2100 runs
    the trimmed product on 128 x 128 tiles."""
    rng = np.random.default_rng(7)
    Li = np.tril(rng.standard_normal((n, n)))
    dA = DBuf(Li.flatten(order="F").astype(dt))
    dC = DBuf(np.zeros(n * n, dt))
    it = items_array(GemmItem, [(n, n, n, n, n, n, 0, 0, 0, -1, 2 | 8)])
    f = lib().bk_gemm_f64 if dt == np.float64 else lib().bk_gemm_f32
    assert f(1, 0, 1.0, 0.0, dA.p, dA.p, dC.p, it, 1, None) == 0
    got = dC.get().reshape((n, n), order="F")
    ref = Li.T @ Li
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert np.allclose(got, ref, rtol=tol, atol=tol * np.abs(ref).max())
    # the same product's lower triangle only (1|2|8): the upper triangle left untouched
    C0 = rng.standard_normal((n, n))
    dC1 = DBuf(C0.flatten(order="F").astype(dt))
    it = items_array(GemmItem, [(n, n, n, n, n, n, 0, 0, 0, -1, 1 | 2 | 8)])
    assert f(1, 0, 1.0, 0.0, dA.p, dA.p, dC1.p, it, 1, None) == 0
    got = dC1.get().reshape((n, n), order="F")
    low = np.tril(np.ones((n, n), bool))
    assert np.allclose(got[low], ref[low], rtol=tol, atol=tol * np.abs(ref).max())
    assert np.array_equal(got[~low], C0.astype(dt)[~low])
    # lower-only: C = -W W' + C, upper triangle left untouched
    W = rng.standard_normal((n, 64))
    C0 = rng.standard_normal((n, n))
    dW = DBuf(W.flatten(order="F").astype(dt)); dC2 = DBuf(C0.flatten(order="F").astype(dt))
    it = items_array(GemmItem, [(n, n, 64, n, n, n, 0, 0, 0, -1, 1)])
    assert f(0, 1, -1.0, 1.0, dW.p, dW.p, dC2.p, it, 1, None) == 0
    got = dC2.get().reshape((n, n), order="F")
    ref = C0 - W @ W.T
    low = np.tril(np.ones((n, n), bool))
    assert np.allclose(got[low], ref[low], rtol=tol, atol=tol * 50)
    assert np.array_equal(got[~low], C0.astype(dt)[~low])


def _pack(mats, dt):
    offs, tot = [], 0
    for M in mats:
        offs.append(tot); tot += M.size
    buf = np.zeros(tot, dt)
    for M, o in zip(mats, offs):
        buf[o:o + M.size] = M.flatten(order="F")
    return buf, offs


SPLIT = 1024      # BK_POTRF_SPLIT (include/brisk_dev.h): the diagonal block and the panel in two launches


@pytest.mark.parametrize("pw", [128, 256], ids=["pw128", "pw256"])
@pytest.mark.parametrize("split", [0, SPLIT], ids=["fused", "split"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_potrf_trtri_batched(dt, split, pw):
    """The blocked Cholesky (both forms of its 64-column step: the diagonal block and its panel in
    one launch, with the step's update from the panel's earlier columns, K = 64 .. 192, inside it,
    or in two launches after a GEMM) at both panel widths, and the triangular inverse of a ragged
    batch against numpy; orders 700 and 1030 take several panels' trailing updates at both
    widths."""
    rng = np.random.default_rng(3)
    sizes = [1, 5, 63, 64, 65, 130, 200, 300, 700, 1030]
    mats = [_spd(n, rng, 1e2) for n in sizes]
    mats.append(-np.eye(7))                                # not positive definite: fails at column 1
    sizes.append(7)
    buf, offs = _pack(mats, dt)
    nb = [((n + 63) // 64) * 4096 for n in sizes]
    doff = np.concatenate([[0], np.cumsum(nb)[:-1]]).astype(np.int64)
    woff = np.array(offs, np.int64)
    rows = [(n, k, offs[k], int(doff[k]), int(woff[k]), offs[k], split | pw) for k, n in enumerate(sizes)]
    it = items_array(MItem, rows)
    dA = DBuf(buf); dD = DBuf(nbytes=int(sum(nb)) * np.dtype(dt).itemsize, dtype=dt)
    dW = DBuf(nbytes=buf.nbytes, dtype=dt)
    dinfo = DBuf(np.zeros(len(sizes), np.int32)); dpmin = DBuf(np.zeros(len(sizes), dt))
    L = lib()
    potrf = L.bk_potrf_f64 if dt == np.float64 else L.bk_potrf_f32
    trtri = L.bk_trtri_f64 if dt == np.float64 else L.bk_trtri_f32
    assert potrf(it, len(rows), dA.p, dD.p, dW.p, dinfo.p, dpmin.p) == 0
    info = dinfo.get(); pmin = dpmin.get()
    assert list(info[:-1]) == [0] * (len(sizes) - 1)
    assert info[-1] == 1
    out = dA.get()
    tol = 1e-11 if dt == np.float64 else 2e-4
    for k, n in enumerate(sizes[:-1]):
        Lg = out[offs[k]:offs[k] + n * n].reshape((n, n), order="F")
        Lr = np.linalg.cholesky(mats[k])
        assert np.allclose(Lg, Lr, rtol=tol, atol=tol * np.abs(Lr).max()), n
        assert abs(pmin[k] - np.min(np.diag(Lr)) ** 2) <= tol * 10 * max(1, pmin[k])
    dLi = DBuf(nbytes=buf.nbytes, dtype=dt)
    assert trtri(it, len(rows), dA.p, dD.p, dLi.p, dW.p, dinfo.p) == 0
    li = dLi.get()
    for k, n in enumerate(sizes[:-1]):
        Lg = out[offs[k]:offs[k] + n * n].reshape((n, n), order="F").astype(np.float64)
        Lig = li[offs[k]:offs[k] + n * n].reshape((n, n), order="F").astype(np.float64)
        assert np.allclose(np.triu(Lig, 1), 0)
        err = np.abs(Lig @ Lg - np.eye(n)).max()
        assert err < (1e-10 if dt == np.float64 else 5e-3), (n, err)


@pytest.mark.parametrize("split", [0, SPLIT], ids=["fused", "split"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_potrf_graph_replay(dt, split):
    """The same factorization (plan and operands) four times: the first call runs eagerly, the
    second captures its launch sequence into a graph and the later ones replay it. Unrelated
    eager work (a GEMM) and a synchronization run between the calls; every call's factor, info
    and pivot minima equal the first call's bit for bit."""
    rng = np.random.default_rng(11)
    sizes = [40, 64, 129, 300, 513]
    mats = [_spd(n, rng, 1e2) for n in sizes]
    buf, offs = _pack(mats, dt)
    nb = [((n + 63) // 64) * 4096 for n in sizes]
    doff = np.concatenate([[0], np.cumsum(nb)[:-1]]).astype(np.int64)
    rows = [(n, k, offs[k], int(doff[k]), offs[k], offs[k], split) for k, n in enumerate(sizes)]
    it = items_array(MItem, rows)
    dA = DBuf(buf); dD = DBuf(nbytes=int(sum(nb)) * np.dtype(dt).itemsize, dtype=dt)
    dW = DBuf(nbytes=buf.nbytes, dtype=dt)
    zi = np.zeros(len(sizes), np.int32); zp = np.zeros(len(sizes), dt)
    dinfo = DBuf(zi); dpmin = DBuf(zp)
    L = lib()
    potrf = L.bk_potrf_f64 if dt == np.float64 else L.bk_potrf_f32
    gemm = L.bk_gemm_f64 if dt == np.float64 else L.bk_gemm_f32
    n = 200
    X = rng.standard_normal((n, n)).astype(dt)
    dX = DBuf(X.flatten(order="F")); dY = DBuf(nbytes=X.nbytes, dtype=dt)
    git = items_array(GemmItem, [(n, n, n, n, n, n, 0, 0, 0, -1, 0)])
    ref = None
    for call in range(4):
        L.bk_h2d(dA.p, buf.ctypes.data, buf.nbytes)
        L.bk_h2d(dinfo.p, zi.ctypes.data, zi.nbytes); L.bk_h2d(dpmin.p, zp.ctypes.data, zp.nbytes)
        assert potrf(it, len(rows), dA.p, dD.p, dW.p, dinfo.p, dpmin.p) == 0, call
        assert gemm(0, 1, 1.0, 0.0, dX.p, dX.p, dY.p, git, 1, None) == 0
        got = (dA.get(), dD.get(), dinfo.get(), dpmin.get())
        y = dY.get().reshape((n, n), order="F")
        assert np.allclose(y, X.astype(np.float64) @ X.T.astype(np.float64), rtol=1e-4, atol=1e-3 * n)
        if ref is None:
            ref = got
            assert list(got[2]) == [0] * len(sizes)
            for k, m in enumerate(sizes):
                Lg = got[0][offs[k]:offs[k] + m * m].reshape((m, m), order="F").astype(np.float64)
                assert np.allclose(Lg, np.linalg.cholesky(mats[k]), rtol=2e-4, atol=2e-4 * np.abs(Lg).max()), m
        else:
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), call


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("sizes", [[1000, 1, 63, 64, 65, 255, 256, 257, 600, 5], [2100] * 16 + [300, 70]],
                         ids=["ragged", "wide"])
def test_trsv_batched(dt, sizes):
    """Factor, inverted 256 x 256 diagonal blocks, then the two blocked triangular solves for a
    ragged batch (orders around the 64 and 256 block edges). Checked against the same block
    algorithm in numpy on the device's own L and diagonal inverses (1e-12: it pins every index),
    and against a direct solve with L L' (the f32 factor's diagonal inverses carry f32 error).
    The wide batch has more than 512 chunks in its first forward and last backward launches, whose
    workgroups then take four chunks each."""
    rng = np.random.default_rng(13)
    if len(sizes) > 12:
        mats = []
        for n in sizes:
            B = rng.standard_normal((n, n))
            mats.append(B @ B.T / n + np.eye(n))
    else:
        mats = [_spd(n, rng, 1e3) for n in sizes]
    buf, offs = _pack(mats, dt)
    tot = buf.size
    big = np.concatenate([buf, np.zeros(2 * tot, dt)])          # L | diagonal inverses | transposes
    nb = [((n + 63) // 64) * 4096 for n in sizes]
    doff = np.concatenate([[0], np.cumsum(nb)[:-1]]).astype(np.int64)
    rows = [(n, k, offs[k], int(doff[k]), offs[k], tot + offs[k]) for k, n in enumerate(sizes)]
    it = items_array(MItem, rows)
    isz = np.dtype(dt).itemsize
    dA = DBuf(big); dD = DBuf(nbytes=int(sum(nb)) * isz, dtype=dt); dW = DBuf(nbytes=tot * isz, dtype=dt)
    dinfo = DBuf(np.zeros(len(sizes), np.int32)); dpmin = DBuf(np.zeros(len(sizes), dt))
    L = lib()
    f64 = dt == np.float64
    assert (L.bk_potrf_f64 if f64 else L.bk_potrf_f32)(it, len(rows), dA.p, dD.p, dW.p, dinfo.p, dpmin.p) == 0
    assert list(dinfo.get()) == [0] * len(sizes)
    assert (L.bk_trtri_diag_f64 if f64 else L.bk_trtri_diag_f32)(it, len(rows), dA.p, dD.p, dA.p, dW.p, dinfo.p) == 0
    voff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    totm = int(sum(sizes))
    r = rng.standard_normal(totm)
    dV = DBuf(np.concatenate([r, np.zeros(totm)]))
    tv = items_array(TrsvItem, [(n, 0, 0, offs[k], tot + offs[k], 2 * tot + offs[k], int(voff[k]), totm + int(voff[k]))
                                for k, n in enumerate(sizes)])
    assert (L.bk_trsv_prep_f64 if f64 else L.bk_trsv_prep_f32)(tv, len(sizes), dA.p) == 0
    assert (L.bk_trsv_f64 if f64 else L.bk_trsv_f32)(tv, len(sizes), dA.p, dV.p) == 0
    z = dV.get()[:totm]
    g = dA.get().astype(np.float64)
    for k, n in enumerate(sizes):
        Lk = g[offs[k]:offs[k] + n * n].reshape((n, n), order="F")
        Dk = g[tot + offs[k]:tot + offs[k] + n * n].reshape((n, n), order="F")
        rk = r[voff[k]:voff[k] + n]
        y = rk.copy(); out = np.zeros(n)
        for kb in range(0, n, 256):                                 # L out = y, block by block
            e = min(n, kb + 256)
            out[kb:e] = Dk[kb:e, kb:e] @ y[kb:e]
            y[e:] -= Lk[e:, kb:e] @ out[kb:e]
        zz = np.zeros(n)
        for kb in reversed(range(0, n, 256)):                       # L' zz = out
            e = min(n, kb + 256)
            zz[kb:e] = Dk[kb:e, kb:e].T @ out[kb:e]
            out[:kb] -= Lk[kb:e, :kb].T @ zz[kb:e]
        zk = z[voff[k]:voff[k] + n]
        assert np.abs(zk - zz).max() <= 1e-12 * np.abs(zz).max(), n
        ref = np.linalg.solve(Lk @ Lk.T, rk)
        assert np.abs(zk - ref).max() <= (1e-9 if f64 else 2e-3) * np.abs(ref).max(), n


@pytest.mark.parametrize("n", [333, 700, 1500])
def test_gemv_triangular_ranges(n):
    """trans = 1 with tri 1 (A lower: rows >= j) and tri 2 (A upper: rows <= j); orders 700 and
    1500 split each column over two and four waves (segments of the triangular row range)."""
    rng = np.random.default_rng(9)
    Lw = np.tril(rng.standard_normal((n, n)))
    x = rng.standard_normal(n)
    for tri, A in ((0, Lw), (1, Lw), (2, Lw.T.copy())):
        dA = DBuf(A.flatten(order="F")); dx = DBuf(x); dy = DBuf(np.zeros(n))
        it = items_array(GemvItem, [(n, n, n, 0, 0, 0, tri)])
        assert lib().bk_gemv_f64(1, 1.0, 0.0, dA.p, dx.p, dy.p, it, 1) == 0
        assert np.allclose(dy.get(), A.T @ x, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("trans", [0, 1])
def test_gemv_batched(trans):
    rng = np.random.default_rng(5)
    shapes = [(1, 1), (100, 37), (257, 513), (64, 64), (600, 9), (1100, 45), (5000, 3)]
    As, xs, ys, rows, refs = [], [], [], [], []
    oa = ox = oy = 0
    for (M, N) in shapes:
        A = rng.standard_normal((M, N)); x = rng.standard_normal(M if trans else N); y = rng.standard_normal(N if trans else M)
        ref = 1.5 * (A.T @ x if trans else A @ x) - 0.5 * y
        rows.append((M, N, M, oa, ox, oy)); refs.append((oy, ref))
        As.append(A.flatten(order="F")); xs.append(x); ys.append(y)
        oa += A.size; ox += x.size; oy += y.size
    dA = DBuf(np.concatenate(As)); dx = DBuf(np.concatenate(xs)); dy = DBuf(np.concatenate(ys))
    it = items_array(GemvItem, rows)
    assert lib().bk_gemv_f64(trans, 1.5, -0.5, dA.p, dx.p, dy.p, it, len(rows)) == 0
    out = dy.get()
    for (o, ref) in refs:
        assert np.allclose(out[o:o + ref.size], ref, rtol=1e-12, atol=1e-12)


def test_symv_lower_batched():
    """Symmetric products from the lower triangle only (the upper holds garbage), ragged batch,
    ld > n, against numpy; each item's partial slots are its own (ceil(n/64)^2 * 64 doubles)."""
    rng = np.random.default_rng(11)
    sizes = [1, 7, 63, 64, 65, 130, 300]
    As, xs, ys, rows, refs = [], [], [], [], []
    oa = ox = oy = op = 0
    for n in sizes:
        ld = n + 3
        S = rng.standard_normal((n, n)); S = S + S.T
        A = np.full((ld, n), 1e30)
        A[:n, :] = np.where(np.tril(np.ones((n, n), bool)), S, 1e30)
        x = rng.standard_normal(n); y = rng.standard_normal(n)
        refs.append((oy, 0.7 * S @ x - 0.3 * y))
        t = (n + 63) // 64
        rows.append((n, ld, oa, ox, oy, op))
        As.append(A.flatten(order="F")); xs.append(x); ys.append(y)
        oa += A.size; ox += n; oy += n; op += max(n * n, t * t * 64)
    dA = DBuf(np.concatenate(As)); dx = DBuf(np.concatenate(xs)); dy = DBuf(np.concatenate(ys)); dp = DBuf(np.zeros(op))
    it = items_array(SymvItem, rows)
    assert lib().bk_symv_f64(0.7, -0.3, dA.p, dx.p, dy.p, dp.p, it, len(rows)) == 0
    out = dy.get()
    for (o, ref) in refs:
        assert np.allclose(out[o:o + ref.size], ref, rtol=1e-12, atol=1e-11)


def _sym_cases(rng):
    """Symmetric test matrices: random of several orders (the in-LDS form up to 64, in place above),
    indefinite, a tight cluster around mu (the central path's X Z), a graded positive definite one,
    repeated eigenvalues, and orders 1 and 2."""
    out = []
    for n in (1, 2, 3, 17, 63, 64, 65, 130, 301):
        B = rng.standard_normal((n, n))
        out.append((B + B.T) / 2)
    n = 60
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    out.append((Q * (1.0 + 1e-6 * rng.standard_normal(n))) @ Q.T)          # cluster at 1
    out.append((Q * np.logspace(-10, 2, n)) @ Q.T)                          # graded, condition 1e12
    out.append((Q * np.repeat([-2.0, 0.5, 3.0], n // 3)) @ Q.T)             # three repeated eigenvalues
    n = 200
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    out.append((Q * np.linspace(-1, 1, n) ** 3) @ Q.T)
    return out


def test_syev_jacobi():
    """The batched Jacobi eigendecomposition against LAPACK (numpy.linalg.eigvalsh): eigenvalues to
    a few units of eps |A|, V orthogonal, A V = V diag(w); every matrix in one call."""
    rng = np.random.default_rng(11)
    mats = _sym_cases(rng)
    parts, rows, off = [], [], 0
    for A in mats:
        n = A.shape[0]
        parts.append(A.flatten(order="F"))
        rows.append((n, 0, off, 0, 0))
        off += n * n
    vo = off
    for k, A in enumerate(mats):
        n = A.shape[0]
        rows[k] = (n, 0, rows[k][2], vo, 0)
        vo += n * n
    wo = vo
    for k, A in enumerate(mats):
        n = A.shape[0]
        rows[k] = (n, 0, rows[k][2], rows[k][3], wo)
        wo += n
    H = DBuf(np.concatenate(parts + [np.zeros(wo - off)]))
    assert lib().bk_syev_f64(items_array(EigItem, rows), len(rows), H.p) == 0
    out = H.get()
    for k, A in enumerate(mats):
        n = A.shape[0]
        V = out[rows[k][3]:rows[k][3] + n * n].reshape((n, n), order="F")
        w = out[rows[k][4]:rows[k][4] + n]
        nrm = np.abs(np.linalg.eigvalsh(A)).max() if n else 1.0
        ref = np.linalg.eigvalsh(A)
        print("n", n, "eig err %.1e" % (np.abs(np.sort(w) - ref).max() / nrm), "orth %.1e" % np.abs(V.T @ V - np.eye(n)).max(),
              "resid %.1e" % (np.abs(A @ V - V * w).max() / nrm), flush=True)
        assert np.abs(np.sort(w) - ref).max() <= 1e-13 * n * nrm + 1e-300
        assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-13 * max(n, 8)
        assert np.abs(A @ V - V * w).max() <= 1e-13 * max(n, 8) * nrm
