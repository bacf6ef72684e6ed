"""ctypes access to the device layer of libbrisk_amd.so for kernel-level GPU tests."""
import ctypes, os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_lib = None
I64 = ctypes.c_int64


class GemmItem(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("K", ctypes.c_int), ("lda", ctypes.c_int), ("ldb", ctypes.c_int),
                ("ldc", ctypes.c_int), ("a", I64), ("b", I64), ("c", I64), ("guard", ctypes.c_int), ("tri", ctypes.c_int)]


class MItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("slot", ctypes.c_int), ("off", I64), ("off2", I64), ("off3", I64), ("off4", I64),
                ("opt", ctypes.c_int), ("pad", ctypes.c_int)]


class GemvItem(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("lda", ctypes.c_int), ("a", I64), ("x", I64), ("y", I64), ("tri", ctypes.c_int)]


class SymvItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ld", ctypes.c_int), ("a", I64), ("x", I64), ("y", I64), ("p", I64)]


class EigItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("pad", ctypes.c_int), ("a", I64), ("v", I64), ("w", I64)]


class TrsvItem(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("c", ctypes.c_int), ("pre", I64), ("L", I64), ("Di", I64), ("DiT", I64), ("v0", I64), ("v1", I64)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ROOT, "brisk_amd", "libbrisk_amd.so")
        if not os.path.exists(path):
            raise RuntimeError("libbrisk_amd.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.bk_init.restype = ctypes.c_int
        L.bk_malloc.restype = P
        L.bk_malloc.argtypes = [ctypes.c_size_t]
        L.bk_free.argtypes = [P]
        L.bk_h2d.argtypes = [P, P, ctypes.c_size_t]
        L.bk_d2h.argtypes = [P, P, ctypes.c_size_t]
        L.bk_zero.argtypes = [P, ctypes.c_size_t]
        L.bk_error.restype = ctypes.c_char_p
        L.bk_device_name.restype = ctypes.c_char_p
        for nm in ("bk_gemm_f64", "bk_gemm_f32"):
            f = getattr(L, nm)
            sc = ctypes.c_double if nm.endswith("64") else ctypes.c_float
            f.argtypes = [ctypes.c_int, ctypes.c_int, sc, sc, P, P, P, P, ctypes.c_int, P]
        for nm in ("bk_potrf_f64", "bk_potrf_f32"):
            getattr(L, nm).argtypes = [P, ctypes.c_int, P, P, P, P, P]
        for nm in ("bk_trtri_f64", "bk_trtri_f32"):
            getattr(L, nm).argtypes = [P, ctypes.c_int, P, P, P, P, P]
        for nm in ("bk_trtri_diag_f64", "bk_trtri_diag_f32"):
            getattr(L, nm).argtypes = [P, ctypes.c_int, P, P, P, P, P]
        for nm in ("bk_trsv_prep_f64", "bk_trsv_prep_f32"):
            getattr(L, nm).argtypes = [P, ctypes.c_int, P]
        for nm in ("bk_trsv_f64", "bk_trsv_f32"):
            getattr(L, nm).argtypes = [P, ctypes.c_int, P, P]
        for nm in ("bk_gemv_f64", "bk_gemv_f32", "bk_gemv_mixed"):
            sc = ctypes.c_float if nm == "bk_gemv_f32" else ctypes.c_double
            getattr(L, nm).argtypes = [ctypes.c_int, sc, sc, P, P, P, P, ctypes.c_int]
        L.bk_symv_f64.argtypes = [ctypes.c_double, ctypes.c_double, P, P, P, P, P, ctypes.c_int]
        L.bk_syev_f64.argtypes = [P, ctypes.c_int, P]
        if L.bk_init() != 0:
            raise RuntimeError("bk_init failed: %s" % L.bk_error())
        _lib = L
    return _lib


class DBuf:
    """A device buffer holding a numpy array's bytes."""

    def __init__(self, arr=None, nbytes=None, dtype=np.float64):
        L = lib()
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            self.dtype, self.nbytes = arr.dtype, arr.nbytes
        else:
            self.dtype, self.nbytes = np.dtype(dtype), nbytes
        self.p = L.bk_malloc(max(self.nbytes, 8))
        if not self.p:
            raise MemoryError(L.bk_error())
        if arr is not None:
            L.bk_h2d(self.p, arr.ctypes.data, self.nbytes)
        else:
            L.bk_zero(self.p, max(self.nbytes, 8))
            L.bk_d2h(None, None, 0)

    def get(self, count=None):
        n = self.nbytes // self.dtype.itemsize if count is None else count
        out = np.empty(n, dtype=self.dtype)
        lib().bk_d2h(out.ctypes.data, self.p, out.nbytes)
        return out

    def __del__(self):
        try:
            lib().bk_free(self.p)
        except Exception:
            pass


def items_array(cls, rows):
    arr = (cls * max(len(rows), 1))()
    for k, r in enumerate(rows):
        arr[k] = cls(*r)
    return arr
