"""Host side on the CPU: the libraries load and export every declared entry point, data and
option errors return the reference's codes before any GPU work, the SDPA reader, the instance
generators, and the Python binding's contract (it refuses to run without the library)."""
import ctypes
import os
import re
import numpy as np
import pytest

from _ctypes_util import ROOT, Instance, read_sdpa, golden, PI, PD, I, D

INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "brisk_amd", "libbrisk_amd.so")
GEN = os.path.join(ROOT, "brisk_amd", "libbrisk_gen.so")


def declared(header):
    """Function names declared in a header (prototypes ending in ';' at file scope)."""
    txt = open(os.path.join(INC, header)).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", " ", txt, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", txt)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return", "sizeof")))


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(LIB)
    return L


@pytest.mark.parametrize("header,so", [("brisk_amd.h", LIB), ("brisk_dev.h", LIB), ("brisk_probe.h", LIB), ("brisk_gen.h", GEN)])
def test_exports(header, so):
    names = declared(header)
    assert len(names) >= 5
    L = ctypes.CDLL(so)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def opts(*a):
    return (ctypes.c_char_p * len(a))(*[x.encode() for x in a]), len(a)


def setup_sig(L):
    L.brisk_result_new.restype = ctypes.c_void_p
    L.brisk_result_delete.argtypes = [ctypes.c_void_p]
    L.brisk_solve_data.argtypes = [I, I, PI, PD, ctypes.c_int64, PI, PI, PI, PI, PD, I, ctypes.POINTER(ctypes.c_char_p),
                                   ctypes.c_void_p]
    L.brisk_solve_file.argtypes = [ctypes.c_char_p, I, ctypes.POINTER(ctypes.c_char_p), ctypes.c_void_p]
    L.brisk_result_status.argtypes = [ctypes.c_void_p]
    L.brisk_result_status_str.argtypes = [ctypes.c_void_p]
    L.brisk_result_status_str.restype = ctypes.c_char_p
    L.brisk_result_exit_code.argtypes = [ctypes.c_void_p]
    L.brisk_version.restype = ctypes.c_char_p
    L.brisk_set_output.argtypes = [I, ctypes.c_void_p]
    L.brisk_set_output(1, None)


def test_result_defaults(lib):
    setup_sig(lib)
    r = lib.brisk_result_new()
    assert lib.brisk_result_status(r) == -1
    assert lib.brisk_result_status_str(r) == b"NOT SOLVED"
    assert lib.brisk_result_status(None) == -1 and lib.brisk_result_exit_code(None) == -1
    assert lib.brisk_version().startswith(b"1.7")
    lib.brisk_result_delete(r)


EX1 = dict(m=3, nblk=1, bs=[2], c=[48.0, -8.0, 20.0], mat=[0, 0, 1, 1, 2, 3, 3], blk=[1] * 7,
           i=[1, 2, 1, 1, 2, 1, 2], j=[1, 2, 1, 2, 2, 2, 2], v=[-11.0, 23.0, 10.0, 4.0, -8.0, -8.0, -2.0])


def call_data(L, d, *o):
    arr = {k: np.ascontiguousarray(d[k], dtype=np.int32 if k in ("bs", "mat", "blk", "i", "j") else np.float64)
           for k in ("bs", "c", "mat", "blk", "i", "j", "v")}
    ov, no = opts(*o)
    r = L.brisk_result_new()
    rc = L.brisk_solve_data(d["m"], d["nblk"], arr["bs"].ctypes.data_as(PI), arr["c"].ctypes.data_as(PD), len(d["v"]),
                            arr["mat"].ctypes.data_as(PI), arr["blk"].ctypes.data_as(PI), arr["i"].ctypes.data_as(PI),
                            arr["j"].ctypes.data_as(PI), arr["v"].ctypes.data_as(PD), no, ov, r)
    L.brisk_result_delete(r)
    return rc


@pytest.mark.parametrize("o", [("-nosuch",), ("-prec",), ("-prec", "quad"), ("-tol", "0"), ("-tol", "-1"),
                               ("-maxit", "-3"), ("-threads",), ("-schurapply",), ("-schurapply", "lu")])
def test_bad_options(lib, o):
    """An unknown option or a bad value is exit code 1 (capi.c:175-192), before any GPU work."""
    setup_sig(lib)
    assert call_data(lib, EX1, *o) == 1


def bad(**kw):
    d = {k: (list(v) if isinstance(v, list) else v) for k, v in EX1.items()}
    d.update(kw)
    return d


@pytest.mark.parametrize("d", [
    bad(blk=[2] + [1] * 6),                      # block index beyond nblk
    bad(i=[3, 2, 1, 1, 2, 1, 2]),                # row beyond the block size
    bad(mat=[0, 0, 1, 1, 2, 3, 4]),              # constraint index beyond m
    bad(bs=[0]),                                 # empty block
    bad(bs=[-2], i=[1, 2, 1, 1, 2, 1, 2], j=[1, 2, 1, 2, 2, 2, 2]),   # off-diagonal entry in an LP block
    bad(m=-1),
], ids=["blk", "row", "con", "bs0", "lp_offdiag", "m"])
def test_bad_data(lib, d):
    """Malformed data is exit code 2, before any GPU work."""
    setup_sig(lib)
    assert call_data(lib, d) == 2


def test_missing_file(lib, tmp_path):
    setup_sig(lib)
    r = lib.brisk_result_new()
    ov, no = opts()
    assert lib.brisk_solve_file(str(tmp_path / "none.dat-s").encode(), no, ov, r) == 2
    lib.brisk_result_delete(r)


class BProb(ctypes.Structure):
    _fields_ = [("m", I), ("nblk", I), ("bs", PI), ("c", PD), ("nnz", ctypes.c_int64),
                ("mat", PI), ("blk", PI), ("i", PI), ("j", PI), ("v", PD)]


def product_read(L, path):
    L.brisk_read_sdpa.argtypes = [ctypes.c_char_p, ctypes.POINTER(BProb), ctypes.POINTER(ctypes.c_void_p)]
    L.brisk_sdpa_free.argtypes = [ctypes.c_void_p]
    p, own = BProb(), ctypes.c_void_p()
    rc = L.brisk_read_sdpa(str(path).encode(), ctypes.byref(p), ctypes.byref(own))
    if rc:
        return rc, None
    n = p.nnz
    out = dict(m=p.m, nblk=p.nblk, bs=np.ctypeslib.as_array(p.bs, (p.nblk,)).copy(),
               c=np.ctypeslib.as_array(p.c, (p.m,)).copy() if p.m else np.zeros(0))
    for k in ("mat", "blk", "i", "j", "v"):
        out[k] = np.ctypeslib.as_array(getattr(p, k), (n,)).copy() if n else np.zeros(0)
    L.brisk_sdpa_free(own)
    return 0, out


SDPA = sorted(f for f in os.listdir(os.path.join(ROOT, "tests", "golden", "sdpa")) if f.endswith(".dat-s"))


@pytest.mark.parametrize("f", SDPA)
def test_sdpa_reader(lib, f):
    """The library's reader and an independent one agree on every example file."""
    path = os.path.join(ROOT, "tests", "golden", "sdpa", f)
    rc, got = product_read(lib, path)
    assert rc == 0
    want = read_sdpa(path)
    for k in ("m", "nblk"):
        assert got[k] == want[k]
    for k in ("bs", "c", "mat", "blk", "i", "j", "v"):
        assert np.array_equal(np.asarray(got[k], dtype=np.float64), np.asarray(want[k], dtype=np.float64)), k


def test_sdpa_reader_forms(lib, tmp_path):
    """Comments, "=mDIM"-style words and brace separators in the header; a truncated entry line
    is an error."""
    f = tmp_path / "a.dat-s"
    f.write_text('"a comment\n* another\n3 =mDIM\n1 =nBLOCK\n{2}\n{48, -8, 20}\n'
                 "0 1 1 1 -11\n0 1 2 2 23\n1 1 1 1 10\n1 1 1 2 4\n2 1 2 2 -8\n3 1 1 2 -8\n3 1 2 2 -2\n")
    rc, got = product_read(lib, f)
    assert rc == 0 and got["m"] == 3 and list(got["bs"]) == [2] and list(got["c"]) == [48, -8, 20]
    assert list(got["v"]) == EX1["v"] and list(got["mat"]) == EX1["mat"]
    g = tmp_path / "b.dat-s"
    g.write_text("1\n1\n2\n1\n0 1 1 1 1\n1 1 1\n")
    rc, got = product_read(lib, g)          # an incomplete last entry is ignored (problem.c:531-535)
    assert rc == 0 and got["mat"].tolist() == [0]


EX1_TEXT = ("3 =mDIM\n1 =nBLOCK\n{2}\n{48, -8, 20}\n"
            "0 1 1 1 -11\n0 1 2 2 23\n1 1 1 1 10\n1 1 1 2 4\n2 1 2 2 -8\n3 1 1 2 -8\n3 1 2 2 -2\n")
# (file text, read as example 1: True, or rejected: False) for the reader rules of the reference's
# read_sdpa (problem.c:426-566)
READER_RULES = [
    ("fortran exponents", EX1_TEXT.replace("-11", "-1.1D+01").replace("23", "2.3d1").replace(" 10", " 10.D0"), True),
    ("separators in the data", EX1_TEXT.replace("0 1 1 1 -11", "0,1,1,1,-11").replace("1 1 1 2 4", "(1 1 1 2 4)"), True),
    ("comment lines in the data", EX1_TEXT.replace("1 1 1 1 10\n", '* a comment\n1 1 1 1 10\n" another\n'), True),
    ("text after the last entry", EX1_TEXT + "BEGIN.COMMENT\nsome words 1 2\nEND.COMMENT\n", True),
    ("incomplete last entry", EX1_TEXT + "1 1 2\n", True),
    ("zero entries", EX1_TEXT + "2 1 1 1 0\n", True),
    ("text before more entries", EX1_TEXT.replace("2 1 2 2 -8\n", "word\n2 1 2 2 -8\n"), False),
    ("non-integer index", EX1_TEXT.replace("1 1 1 2 4", "1 1 1.5 2 4"), False),
    ("number glued to text", EX1_TEXT.replace("-8\n3 1 1", "-8x\n3 1 1"), False),
    ("non-finite value", EX1_TEXT.replace(" 23", " inf"), False),
]


def _ref_solve(path):
    """The reference library (oracle/_ref, built from the reference's sources): its return code and
    primal objective on a file, or None when that library is not built here."""
    so = os.path.join(ROOT, "oracle", "_ref", "libbrisk.so")
    if not os.path.exists(so):
        return None
    R = ctypes.CDLL(so)
    R.brisk_result_new.restype = ctypes.c_void_p
    R.brisk_result_delete.argtypes = [ctypes.c_void_p]
    R.brisk_set_output.argtypes = [I, ctypes.c_void_p]
    R.brisk_solve_file.argtypes = [ctypes.c_char_p, I, ctypes.POINTER(ctypes.c_char_p), ctypes.c_void_p]
    R.brisk_result_pobj.argtypes = [ctypes.c_void_p]
    R.brisk_result_pobj.restype = D
    R.brisk_set_output(1, None)
    r = R.brisk_result_new()
    no = (ctypes.c_char_p * 1)()
    rc = R.brisk_solve_file(str(path).encode(), 0, no, r)
    pobj = R.brisk_result_pobj(r)
    R.brisk_result_delete(r)
    return rc, pobj


@pytest.mark.parametrize("case", READER_RULES, ids=[c[0] for c in READER_RULES])
def test_sdpa_reader_rules(lib, tmp_path, case):
    """The reader rules of the reference's read_sdpa: Fortran D exponents, separators and comment
    lines anywhere, text after the last entry and an incomplete last entry ignored, zero entries
    dropped; text before more entries, non-integer indices, numbers glued to text and non-finite
    values rejected. Each accepted file reads as the canonical example 1; the reference library,
    where it is built, accepts and rejects the same files and solves an accepted one to the
    canonical file's objective."""
    name, text, ok = case
    f = tmp_path / "case.dat-s"
    f.write_text(text)
    rc, got = product_read(lib, f)
    if not ok:
        assert rc != 0, name
    else:
        assert rc == 0, name
        assert got["m"] == 3 and got["bs"].tolist() == [2] and got["c"].tolist() == [48, -8, 20]
        for k in ("mat", "blk", "i", "j", "v"):
            assert np.asarray(got[k], dtype=np.float64).tolist() == [float(x) for x in EX1[k]], (name, k)
    canon = tmp_path / "canon.dat-s"
    canon.write_text(EX1_TEXT)
    ref, ref0 = _ref_solve(f), _ref_solve(canon)
    if ref is None:
        return
    if ok:
        assert ref[0] == ref0[0] == 0 and abs(ref[1] - ref0[1]) <= 1e-12 * (1 + abs(ref0[1])), (name, ref, ref0)
    else:
        assert ref[0] != 0, (name, ref)
    assert product_read(lib, tmp_path / "missing.dat-s")[0] != 0


# ---- generators (the synthetic workloads) ----

def test_generator_determinism():
    a = Instance("theta", 40, 0.2, 5).arrays()
    b = Instance("theta", 40, 0.2, 5).arrays()
    c = Instance("theta", 40, 0.2, 6).arrays()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k]))
    assert a["m"] != c["m"] or not np.array_equal(a["i"], c["i"])


def test_generator_shapes():
    t = Instance("theta", 30, 0.3, 1).arrays()
    edges = (t["m"] - 1)
    # F0 = J (C = -J): n(n+1)/2 upper-triangle entries of matrix 0; tr(X) = 1; one entry per edge
    assert np.sum(t["mat"] == 0) == 30 * 31 // 2 and np.all(t["v"][t["mat"] == 0] == 1)
    assert np.sum(t["mat"] == 1) == 30 and np.all(t["c"][1:] == 0) and t["c"][0] == 1
    assert np.sum(t["mat"] >= 2) == edges
    mc = Instance("maxcut", 50, 0.2, 0).arrays()
    assert mc["m"] == 50 and list(mc["bs"]) == [50] and np.all(mc["c"] == 1)
    d = Instance("dense", 10, 20, 0).arrays()
    assert d["m"] == 20 and list(d["bs"]) == [10]
    # every constraint of the dense family is a full symmetric matrix
    assert np.sum(d["mat"] >= 1) == 20 * 10 * 11 // 2
    mu = Instance("multi", 3, 8, 5, 20, 0).arrays()
    assert list(mu["bs"]) == [8, 8, 8, -5] and mu["m"] == 20


def test_generator_u01():
    g = ctypes.CDLL(GEN)
    g.brisk_gen_u01.argtypes = [ctypes.c_uint64] * 3
    g.brisk_gen_u01.restype = ctypes.c_double
    u = np.array([g.brisk_gen_u01(7, 3, k) for k in range(20000)])
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005


def test_binding_requires_library(tmp_path):
    """The Python binding raises when the native library is absent (no fallback)."""
    import brisk_amd
    with pytest.raises(RuntimeError):
        brisk_amd.load_library(str(tmp_path / "nolib.so"))


def test_acceptance_rules_nan(tmp_path):
    """The acceptance error, status rule and termination score of the host solver (solver.h) on
    a C driver: a NaN error never passes (reference brisk_nmax, brisk.h:182-189; status from the
    measured errors, main.c:2349-2359)."""
    import subprocess
    exe = str(tmp_path / "host_rules")
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-I" + os.path.join(ROOT, "brisk_amd", "csrc"), "-I" + INC,
                    os.path.join(ROOT, "tests", "host_rules.c"), "-o", exe, "-lm"], check=True)
    out = dict(line.split() for line in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines())
    inf = float("inf")
    assert float(out["nmax_nan_first"]) == inf and float(out["nmax_nan_second"]) == inf
    assert float(out["nmax_plain"]) == 3.0
    assert float(out["nmin_nan"]) == -inf and float(out["nmin_plain"]) == 2.0
    assert float(out["acc_plain"]) == 3e-9 and int(out["status_plain"]) == 0
    assert float(out["acc_nan_cone"]) == inf
    # an optimal claim with a NaN cone or complementarity error ends in numerical difficulties
    assert int(out["status_nan_cone"]) == 4 and int(out["status_nan_compl"]) == 4
    assert int(out["status_nan_err3"]) == 0
    assert int(out["status_reduced"]) == 5 and int(out["status_limit_reduced"]) == 5
    assert int(out["status_numeric"]) == 4 and int(out["status_limit_kept"]) == 3
    assert int(out["status_infeasible_kept"]) == 1
    assert float(out["score_nan"]) == inf and float(out["score_plain"]) == 3e-9
