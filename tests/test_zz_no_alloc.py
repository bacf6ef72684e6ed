"""The symmetric product without device allocation (bk_symv_f64_noalloc): its partial sums in the
library's static device array instead of a caller workspace. Through the C ABI, against the
allocating entry bk_symv_f64 on the same inputs (bit for bit: the same kernels' arithmetic) and
against numpy (1e-12), on two calls in a row, the second with a smaller input; calls beyond the
array's limits are refused on the host before any launch."""
import ctypes
import numpy as np
import pytest
from _gpu_util import lib, DBuf, SymvItem, items_array

pytestmark = pytest.mark.gpu


def _na():
    L = lib()
    P = ctypes.c_void_p
    L.bk_symv_f64_noalloc.argtypes = [ctypes.c_double, ctypes.c_double, P, P, P, P, ctypes.c_int]
    return L


def batch(sizes, seed):
    """A ragged batch of symmetric matrices (lower triangle significant, the upper holding
    garbage, ld > n), vectors and initial y, with the numpy reference of 0.7 A x - 0.3 y."""
    rng = np.random.default_rng(seed)
    As, xs, ys, rows, refs = [], [], [], [], []
    oa = ox = oy = op = 0
    for n in sizes:
        ld = n + 5
        S = rng.standard_normal((n, n))
        S = S + S.T
        A = np.full((ld, n), 1e30)
        A[:n, :] = np.where(np.tril(np.ones((n, n), bool)), S, 1e30)
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        refs.append((oy, 0.7 * S @ x - 0.3 * y))
        t = (n + 63) // 64
        rows.append((n, ld, oa, ox, oy, op))
        As.append(A.flatten(order="F"))
        xs.append(x)
        ys.append(y)
        oa += A.size
        ox += n
        oy += n
        op += t * t * 64
    return np.concatenate(As), np.concatenate(xs), np.concatenate(ys), rows, refs, op


def run_both(sizes, seed):
    A, x, y0, rows, refs, op = batch(sizes, seed)
    L = _na()
    dA, dx = DBuf(A), DBuf(x)
    dy_ref, dy_na, dp = DBuf(y0), DBuf(y0), DBuf(np.zeros(op))
    it = items_array(SymvItem, rows)
    print("sizes", sizes, "seed", seed, flush=True)
    assert L.bk_symv_f64(0.7, -0.3, dA.p, dx.p, dy_ref.p, dp.p, it, len(rows)) == 0
    assert L.bk_symv_f64_noalloc(0.7, -0.3, dA.p, dx.p, dy_na.p, it, len(rows)) == 0, L.bk_error()
    got, want = dy_na.get(), dy_ref.get()
    assert np.array_equal(got, want)
    for (o, ref) in refs:
        assert np.allclose(got[o:o + ref.size], ref, rtol=1e-12, atol=1e-11)


def test_noalloc_symv_equals_allocating_entry_twice():
    """Two calls in a row on the shared array: a batch of six (orders up to 2000), then a
    smaller batch of two; each equals the allocating entry's result bit for bit."""
    run_both([2000, 1, 63, 64, 130, 700], 21)
    run_both([300, 65], 22)


def test_noalloc_symv_refuses_oversize():
    """More than 16 items, or slots beyond the 32 MiB array (order 16385), are refused before
    any launch, with the reason; the library stays usable."""
    L = _na()
    it = items_array(SymvItem, [(8, 8, 0, 0, 0, 0)] * 17)
    assert L.bk_symv_f64_noalloc(1.0, 0.0, None, None, None, it, 17) == -1
    assert b"16 items" in L.bk_error()
    big = items_array(SymvItem, [(16385, 16385, 0, 0, 0, 0)])
    assert L.bk_symv_f64_noalloc(1.0, 0.0, None, None, None, big, 1) == -1
    assert b"32 MiB" in L.bk_error()
    run_both([100], 23)
