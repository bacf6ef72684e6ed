"""Writes the golden fixtures of the interior-point path: the outputs of the reference library
(libbrisk.so compiled from the reference's own C sources by oracle/Makefile.ref into
oracle/_ref/) on seeded generated instances and on the SDPA example files the reference's tests
use (examples/*.dat-s, matlab/examples/example1.dat-s), which are copied here as data.

Runs only in a container that holds the reference tree (the fixtures travel, the reference does
not).  Usage:
    python tests/golden/make_golden.py small          # -> tests/golden/small.json (+ sdpa/*.dat-s)
    python tests/golden/make_golden.py large c2 c4    # -> tests/golden/large.json (entries merged)
    python tests/golden/make_golden.py stream         # -> tests/golden/stream.json (serving requests)
    python tests/golden/make_golden.py methods        # -> small.json: the infeasible-start method with NT / HKM
Each record: rc, status, pobj, dobj, iterations, DIMACS errors, y (for m <= 2000), X (blocks
concatenated, SDP column-major, when it has at most 10000 numbers), the options
and the seconds the reference took on this container's host; under "dd" the same problem solved
by the reference in double-double arithmetic ("-prec dd", DIMACS errors near 1e-25), which
gives the optimal value to far more digits than an FP64 solve at 1e-8, and its y and X: their
distance to the FP64 solve's point shows how far apart two 1e-8 solutions of the problem can lie
(large where the optimum is not unique).
"""
import ctypes, json, os, shutil, sys, time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from _ctypes_util import Instance  # noqa: E402

REF = "/root/reference"
FILES = [("examples", "example1"), ("examples", "example2"), ("examples", "truss1"), ("examples", "control1"),
         ("examples", "theta1"), ("examples", "arch0"), ("examples", "mcp250-1"), ("examples", "theta3")]
SRC = {"example1": "matlab/examples", "example2": "matlab/examples"}

# generated instances: (name, generator, arguments); the parity cases of the tests
SMALL = [
    ("theta_50_0.2_s0", "theta", (50, 0.2, 0)),
    ("theta_30_0.3_s3", "theta", (30, 0.3, 3)),
    ("theta_80_0.1_s5", "theta", (80, 0.1, 5)),
    ("maxcut_100_0.1_s0", "maxcut", (100, 0.1, 0)),
    ("maxcut_200_0.05_s1", "maxcut", (200, 0.05, 1)),
    ("maxcut_cycle_7", "maxcut_cycle", (7,)),
    ("maxcut_cycle_8", "maxcut_cycle", (8,)),
    ("multi_4_10_20_60_s0", "multi", (4, 10, 20, 60, 0)),
    ("multi_6_30_40_300_s2", "multi", (6, 30, 40, 300, 2)),
    ("dense_12_30_s0", "dense", (12, 30, 0)),
    ("dense_40_120_s1", "dense", (40, 120, 1)),
]
LARGE = {
    "c2": ("maxcut_2000_0.01_s0", "maxcut", (2000, 0.01, 0), ["-bm", "0"]),
    "c4": ("multi_64_60_500_5000_s0", "multi", (64, 60, 500, 5000, 0), []),
    "c3": ("dense_400_2000_s0", "dense", (400, 2000, 0), []),
}

# requests of the serving stream (theta / max-cut, n ~ U{20..200}, p ~ U[0.05, 0.3]; the stream's
# own seeded draws, request k built with instance seed 1000 + k): the first 26, which hold two
# theta requests with m > 1500 (m = 1529, 1707), and request 31 (theta, m = 4985)
STREAM_IDS = list(range(26)) + [31]

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libbrisk.so"))
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        pd, pi, ppc = ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(ctypes.c_char_p)
        L.brisk_result_new.restype = P
        L.brisk_result_delete.argtypes = [P]
        L.brisk_solve_data.argtypes = [I, I, pi, pd, ctypes.c_int64, pi, pi, pi, pi, pd, I, ppc, P]
        L.brisk_solve_file.argtypes = [ctypes.c_char_p, I, ppc, P]
        for n, t in [("pobj", D), ("dobj", D), ("time", D)]:
            getattr(L, "brisk_result_" + n).argtypes = [P]
            getattr(L, "brisk_result_" + n).restype = t
        for n in ("iterations", "status", "m", "exit_code", "nblk", "have_x"):
            getattr(L, "brisk_result_" + n).argtypes = [P]
        L.brisk_result_blocksize.argtypes = [P, I]
        L.brisk_result_X.argtypes = [P, I]
        L.brisk_result_X.restype = pd
        L.brisk_result_y.argtypes = [P]
        L.brisk_result_y.restype = pd
        L.brisk_result_dimacs.argtypes = [P, pd]
        L.brisk_set_output.argtypes = [I, P]
        L.brisk_set_output(1, None)
        _lib = L
    return _lib


def record(call, opts, ymax=2000, xmax=10000):
    L = lib()
    r = L.brisk_result_new()
    o = (ctypes.c_char_p * len(opts))(*[x.encode() for x in opts])
    t = time.time()
    rc = call(len(opts), o, r)
    t = time.time() - t
    e = (ctypes.c_double * 6)()
    L.brisk_result_dimacs(r, e)
    m = L.brisk_result_m(r)
    yp = L.brisk_result_y(r)
    out = dict(rc=rc, status=L.brisk_result_status(r), exit_code=L.brisk_result_exit_code(r),
               pobj=L.brisk_result_pobj(r), dobj=L.brisk_result_dobj(r), iterations=L.brisk_result_iterations(r),
               dimacs=list(e), m=m, options=opts, seconds=round(t, 3))
    if yp and 0 < m <= ymax:
        out["y"] = [yp[k] for k in range(m)]
    nb = L.brisk_result_nblk(r)
    bs = [L.brisk_result_blocksize(r, k) for k in range(nb)]
    xlen = sum(b * b if b > 0 else -b for b in bs)
    if L.brisk_result_have_x(r) and 0 < xlen <= xmax:
        X = []
        for k, b in enumerate(bs):
            xp = L.brisk_result_X(r, k)
            X += [xp[q] for q in range(b * b if b > 0 else -b)]
        out["X"] = X
        out["blocksizes"] = bs
    L.brisk_result_delete(r)
    return out


def gen_record(kind, args, extra, dd=True, ymax=2000):
    inst = Instance(kind, *args)
    L = lib()
    opts = ["-threads", "8"] + list(extra)
    rec = record(lambda n, o, r: L.brisk_solve_data(*inst.data_args(), n, o, r), opts, ymax)
    if dd:
        hp = record(lambda n, o, r: L.brisk_solve_data(*inst.data_args(), n, o, r), opts + ["-prec", "dd"])
        rec["dd"] = {k: hp[k] for k in ("status", "pobj", "dobj", "dimacs", "seconds", "y", "X") if k in hp}
    rec.update(generator=kind, args=list(args))
    return rec


def file_record(name):
    path = os.path.join(HERE, "sdpa", name + ".dat-s")
    L = lib()
    rec = record(lambda n, o, r: L.brisk_solve_file(path.encode(), n, o, r), ["-threads", "8"])
    hp = record(lambda n, o, r: L.brisk_solve_file(path.encode(), n, o, r), ["-threads", "8", "-prec", "dd"])
    rec["dd"] = {k: hp[k] for k in ("status", "pobj", "dobj", "dimacs", "seconds", "y", "X") if k in hp}
    rec.update(file="sdpa/%s.dat-s" % name)
    return rec


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "small"
    if which == "small":
        for _, name in FILES:
            src = os.path.join(REF, SRC.get(name, "examples"), name + ".dat-s")
            shutil.copyfile(src, os.path.join(HERE, "sdpa", name + ".dat-s"))
        out = {}
        for name, kind, args in SMALL:
            out[name] = gen_record(kind, args, [])
            print(name, out[name]["status"], out[name]["pobj"], out[name]["seconds"], flush=True)
        for _, name in FILES:
            out["file_" + name] = file_record(name)
            print(name, out["file_" + name]["status"], out["file_" + name]["pobj"], flush=True)
        dst = os.path.join(HERE, "small.json")
    elif which == "stream":
        from brisk_amd.generators import stream
        reqs, _ = stream()
        out = {}
        for k in STREAM_IDS:
            kind, n, p, seed = reqs[k]
            name = "req%04d_%s_%d" % (k, kind, n)
            m = Instance(kind, n, p, seed).m
            out[name] = gen_record(kind, (n, p, seed), [], dd=m <= 2000, ymax=2000)
            out[name]["request"] = k
            print(name, out[name]["m"], out[name]["status"], out[name]["pobj"], out[name]["seconds"], flush=True)
        dst = os.path.join(HERE, "stream.json")
    elif which == "methods":
        # the reference's infeasible-start method in either direction (-nohsd -dir nt | hkm) on the
        # small set, added to small.json under "std_nt" / "std_hkm": status, objectives, iterations
        dst = os.path.join(HERE, "small.json")
        out = json.load(open(dst))
        L = lib()
        for key, rec in sorted(out.items()):
            for d in ("nt", "hkm"):
                opts = ["-threads", "8", "-nohsd", "-dir", d]
                if "file" in rec:
                    path = os.path.join(HERE, rec["file"])
                    r = record(lambda n, o, r_: L.brisk_solve_file(path.encode(), n, o, r_), opts)
                else:
                    inst = Instance(rec["generator"], *rec["args"])
                    r = record(lambda n, o, r_: L.brisk_solve_data(*inst.data_args(), n, o, r_), opts)
                rec["std_" + d] = {k: r[k] for k in ("status", "pobj", "dobj", "iterations", "dimacs", "options")}
            print(key, rec["std_nt"]["iterations"], rec["std_hkm"]["iterations"], flush=True)
    else:
        dst = os.path.join(HERE, "large.json")
        out = json.load(open(dst)) if os.path.exists(dst) else {}
        for key in sys.argv[2:]:
            name, kind, args, extra = LARGE[key]
            out[name] = gen_record(kind, args, extra, dd=False, ymax=5000)
            out[name]["workload"] = key
            print(name, out[name]["status"], out[name]["pobj"], out[name]["seconds"], flush=True)
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
