#!/usr/bin/env python3
"""bench.py - time to a 1e-8 DIMACS solution of the max-cut SDP relaxation of a random
+-1-weighted graph (n = 2000, p = 0.01, seed 0; C = -L/4, diag X = 1, m = 2000) on one MI355X,
mixed precision (FP32-MFMA Schur factor + FP64 refinement) with the FP64 path beside it.

A step is one complete solve of that instance with its data already resident on the device:
the start point, the interior-point iterations and the DIMACS evaluation (the library's phase
clock, synchronised at both ends). The reference CPU solver (the reference library built from
its own sources into oracle/_ref/, where present) solves the same instance in a subprocess on
the host cores as the reported baseline. One JSON line on stdout.

Other workloads (--workload): c1 (Lovasz theta of G(50, 0.2)), c3 (dense n=400, m=2000),
c4 (64 blocks of 60 + LP 500, m=5000), c5 (the serving stream: solves/s with every request
queued at once, then p50/p99 latency under Poisson arrivals at --load times that rate; the
reference's CPU baseline on every 16th request, one thread per solve in a pool of processes).
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PEAK = {"f64": 78.6, "f32": 157.3}      # MI355X dense MFMA TFLOP/s
HBM_PEAK = 8000.0                        # GB/s
KCLASS = {0: "k_gemm<double, 0, 0>", 1: "k_gemm<double, 0, 1>", 2: "k_gemm<double, 1, 0>", 3: "k_gemm<double, 1, 1>",
          4: "k_gemm<float, 0, 0>", 5: "k_gemm<float, 0, 1>", 6: "k_gemm<float, 1, 0>", 7: "k_gemm<float, 1, 1>",
          8: "k_potrf_diag", 9: "k_gemv", 10: "Schur assembly", 11: "k_aop", 12: "k_lz_step", 13: "k_dot_chunks",
          14: "k_lincomb", 15: "k_symv", 16: "k_trsv"}

WORKLOADS = {
    "c1": dict(kind="theta", args=(50, 0.2, 0), desc="Lovasz theta SDP of G(50, 0.2), seed 0 (C=-J, tr X=1, X_ij=0 on edges)"),
    "c2": dict(kind="maxcut", args=(2000, 0.01, 0), desc="max-cut SDP of G(2000, 0.01), weights +-1, seed 0 (C=-L/4, diag X=1, m=2000)"),
    "c3": dict(kind="dense", args=(400, 2000, 0), desc="dense random SDP, one block n=400, m=2000, seed 0, feasible by X0/Z0/y0"),
    "c4": dict(kind="multi", args=(64, 60, 500, 5000, 0), desc="64 PSD blocks of 60 + LP 500, m=5000, 1-3 blocks per row, seed 0"),
}


# the serving stream: theta / max-cut requests of uneven sizes with Poisson arrivals
C5 = dict(count=4096, seed=0, nmin=20, nmax=200, pmin=0.05, pmax=0.3,
          desc="serving stream: 4096 theta/max-cut SDPs (50/50), n ~ U{20..200}, p ~ U[0.05,0.3], Poisson arrivals, seed 0")


def c5_stream(count=C5["count"], seed=C5["seed"]):
    """Request k: (kind, n, p, instance seed) and a unit-rate exponential inter-arrival gap
    (scaled by 1/lambda at run time)."""
    rng = np.random.default_rng(seed)
    theta = rng.random(count) < 0.5
    n = rng.integers(C5["nmin"], C5["nmax"] + 1, count)
    p = rng.uniform(C5["pmin"], C5["pmax"], count)
    gaps = rng.exponential(1.0, count)
    reqs = [("theta" if theta[k] else "maxcut", int(n[k]), float(p[k]), 1000 + k) for k in range(count)]
    return reqs, gaps


# measured HBM bytes per dispatch by kernel class (PMC FETCH_SIZE + WRITE_SIZE, copy-calibrated),
# kept with the rocprofv3 summaries; keyed by workload
TRAFFIC_FILES = {"c2": "maxcut2000_mixed_hbm_traffic_pmc.md"}


def measured_traffic(workload, kernel):
    fn = TRAFFIC_FILES.get(workload)
    if not fn:
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", fn)):
            if line.startswith("traffic_bytes_per_launch "):
                k, v = line[len("traffic_bytes_per_launch "):].rsplit("=", 1)
                if k.strip() == kernel:
                    return float(v)
    except OSError:
        return None
    return None


def instance(name):
    from _ctypes_util import Instance
    w = WORKLOADS[name]
    return Instance(w["kind"], *w["args"])


def solve(a, prec):
    import brisk_amd
    return brisk_amd.solve_sdpa(a["bs"], a["c"], a["mat"], a["blk"], a["i"], a["j"], a["v"], options=["-prec", prec])


def cpu_reference(name, threads, timeout):
    """Solve the instance with the reference library in a subprocess; returns a dict or None."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libbrisk.so")
    if not os.path.exists(ref):
        return None
    code = r"""
import ctypes, json, sys, time
sys.path.insert(0, %r); sys.path.insert(0, %r)
import bench
from _ctypes_util import Inst
inst = bench.instance(%r)
L = ctypes.CDLL(%r)
P = ctypes.c_void_p; I = ctypes.c_int; D = ctypes.c_double
pd = ctypes.POINTER(D); pi = ctypes.POINTER(I); ppc = ctypes.POINTER(ctypes.c_char_p)
L.brisk_result_new.restype = P
L.brisk_solve_data.argtypes = [I, I, pi, pd, ctypes.c_int64, pi, pi, pi, pi, pd, I, ppc, P]
for n, t in (("pobj", D), ("dobj", D), ("time", D)):
    getattr(L, "brisk_result_" + n).argtypes = [P]; getattr(L, "brisk_result_" + n).restype = t
L.brisk_result_status.argtypes = [P]; L.brisk_result_iterations.argtypes = [P]
L.brisk_result_dimacs.argtypes = [P, pd]
L.brisk_set_output.argtypes = [I, P]; L.brisk_set_output(1, None)
opts = ["-threads", str(%d), "-q"] + (["-bm", "0"] if %r == "c2" else [])
o = (ctypes.c_char_p * len(opts))(*[x.encode() for x in opts])
r = L.brisk_result_new()
s = inst.s
t0 = time.time()
rc = L.brisk_solve_data(s.m, s.nblk, s.bs, s.c, s.nnz, s.mat, s.blk, s.i, s.j, s.v, len(opts), o, r)
t = time.time() - t0
e = (D * 6)(); L.brisk_result_dimacs(r, e)
print(json.dumps(dict(rc=rc, status=L.brisk_result_status(r), pobj=L.brisk_result_pobj(r), dobj=L.brisk_result_dobj(r),
                      iters=L.brisk_result_iterations(r), wall=t, time=L.brisk_result_time(r), dimacs=list(e), opts=opts)))
""" % (ROOT, os.path.join(ROOT, "tests"), name, ref, threads, name)
    env = dict(os.environ, BRISK_NOBUSY="1", OMP_NUM_THREADS=str(threads))
    try:
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout, env=env)
    except subprocess.TimeoutExpired:
        return dict(error="timeout after %ds" % timeout)
    for line in out.stdout.splitlines()[::-1]:
        if line.startswith("{"):
            return json.loads(line)
    return dict(error=(out.stderr or out.stdout)[-400:])


def cpu_serving_sample(stride, workers):
    """The reference library on every stride-th request of the stream, one thread per solve
    (the library is not reentrant), a pool of worker processes; prints one JSON line."""
    import multiprocessing as mp
    reqs, _ = c5_stream()
    sample = reqs[::stride]
    t0 = time.time()
    with mp.get_context("fork").Pool(workers, initializer=_ref_worker_init) as pool:
        out = list(pool.imap_unordered(_ref_worker_solve, sample, chunksize=1))
    wall = time.time() - t0
    print(json.dumps(dict(count=len(sample), wall=wall, solves_per_s=len(sample) / wall, workers=workers,
                          optimal=sum(1 for o in out if o[0] == 0), cpu_s=sum(o[1] for o in out))))


_REF = None


def _ref_worker_init():
    global _REF
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libbrisk.so"))
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pd, pi, ppc = ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(ctypes.c_char_p)
    L.brisk_result_new.restype = P
    L.brisk_result_delete.argtypes = [P]
    L.brisk_solve_data.argtypes = [I, I, pi, pd, ctypes.c_int64, pi, pi, pi, pi, pd, I, ppc, P]
    L.brisk_result_status.argtypes = [P]
    L.brisk_set_output.argtypes = [I, P]
    L.brisk_set_output(1, None)
    _REF = L


def _ref_worker_solve(req):
    from _ctypes_util import Instance
    kind, n, p, seed = req
    inst = Instance(kind, n, p, seed)
    opts = ["-threads", "1", "-q"]
    o = (ctypes.c_char_p * len(opts))(*[x.encode() for x in opts])
    r = _REF.brisk_result_new()
    t = time.time()
    _REF.brisk_solve_data(*inst.data_args(), len(opts), o, r)
    t = time.time() - t
    st = _REF.brisk_result_status(r)
    _REF.brisk_result_delete(r)
    return st, t


def run_stream(insts, arrivals, opts):
    """Drive the serving loop: submit each request at its arrival time, step, collect; returns
    (makespan s, latencies s (finish - arrival), statuses, iterations)."""
    import brisk_amd
    sv = brisk_amd.Server(opts)
    N = len(insts)
    rid_of = {}
    lat = np.zeros(N)
    status = np.full(N, -9)
    iters = np.zeros(N, dtype=int)
    nxt, fin, left = 0, 0, 0
    t0 = time.perf_counter()
    while fin < N:
        now = time.perf_counter() - t0
        while nxt < N and arrivals[nxt] <= now:
            rid_of[sv.submit(*insts[nxt])] = nxt
            nxt += 1
            left += 1
        if left == 0 and nxt < N:
            time.sleep(max(0.0, min(arrivals[nxt] - now, 1e-3)))
            continue
        left = sv.step()
        for rid in sv.done():
            res, _ = sv.take(rid)
            k = rid_of[rid]
            lat[k] = time.perf_counter() - t0 - arrivals[k]
            status[k] = res.status
            iters[k] = res.iterations
            fin += 1
    span = time.perf_counter() - t0
    sv.close()
    return span, lat, status, iters


def serving(args):
    """The serving workload: solves/s with every request queued at t = 0 (saturation), then
    p50/p99 latency under Poisson arrivals at a fixed fraction of that rate."""
    from _ctypes_util import Instance
    count = args.requests or C5["count"]
    reqs, gaps = c5_stream()
    reqs, gaps = reqs[:count], gaps[:count]
    insts = []
    for kind, n, p, seed in reqs:
        a = Instance(kind, n, p, seed).arrays()
        insts.append((a["bs"], a["c"], a["mat"], a["blk"], a["i"], a["j"], a["v"]))
    ms = [int(x[1].shape[0]) for x in insts]
    print("bench: %d requests, m from %d to %d (mean %.0f)" % (count, min(ms), max(ms), np.mean(ms)), file=sys.stderr, flush=True)
    opts = ["-prec", "mixed", "-slots", str(args.slots), "-schurapply", args.schur_apply]
    for w in range(args.warmup):
        span, _, st, _ = run_stream(insts[:64], np.zeros(64), opts)
        print("bench: warmup %d: 64 requests in %.3fs" % (w, span), file=sys.stderr, flush=True)
    runs = []
    for k in range(args.steps):
        span, lat, st, it = run_stream(insts, np.zeros(count), opts)
        runs.append(span)
        print("bench: saturation run %d: %d requests in %.3fs = %.1f solves/s; optimal %d, p50 %.3fs p99 %.3fs, mean iterations %.1f"
              % (k, count, span, count / span, int(np.sum(st == 0)), np.percentile(lat, 50), np.percentile(lat, 99), it.mean()),
              file=sys.stderr, flush=True)
    span = float(np.median(runs))
    rate = count / span
    lam = args.load * rate
    arr = np.cumsum(gaps) / lam
    pspan, plat, pst, pit = run_stream(insts, arr, opts)
    print("bench: Poisson arrivals at %.1f/s (%.2f x the saturation rate): %d requests in %.3fs, optimal %d, p50 %.3fs, p99 %.3fs"
          % (lam, args.load, count, pspan, int(np.sum(pst == 0)), np.percentile(plat, 50), np.percentile(plat, 99)),
          file=sys.stderr, flush=True)
    cpu = None
    if not args.skip_cpu and os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libbrisk.so")):
        workers = args.cpu_threads or min(16, os.cpu_count() or 1)
        env = dict(os.environ, BRISK_NOBUSY="1", OMP_NUM_THREADS="1")
        try:
            out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-serving-sample", str(args.cpu_stride),
                                  "--cpu-threads", str(workers)], capture_output=True, text=True, timeout=900, env=env)
            line = [x for x in out.stdout.splitlines() if x.startswith("{")]
            ref = json.loads(line[-1]) if line else dict(error=out.stderr[-400:])
        except subprocess.TimeoutExpired:
            ref = dict(error="timeout")
        if "error" not in ref:
            # the rate of perfectly packed workers (solves x workers / CPU seconds): an upper bound of
            # the pool's throughput on the whole stream, above the sample's own wall-clock rate
            packed = ref["count"] * workers / max(ref["cpu_s"], 1e-9)
            cpu = dict(value=round(packed, 3), unit="solves/s", cores=workers, kind="reference",
                       sample="every %dth request of the stream (%d solves, %d optimal) by the reference library, one thread per "
                              "solve in %d worker processes: %.1f CPU s, %.1f s wall (%.2f solves/s measured; value = solves x "
                              "workers / CPU s), %s" % (args.cpu_stride, ref["count"], ref["optimal"], workers, ref["cpu_s"],
                                                       ref["wall"], ref["solves_per_s"], cpu_model()))
        else:
            cpu = dict(value=None, unit="solves/s", cores=workers, kind="reference", sample="not run: %s" % ref["error"])
    import brisk_amd
    out = {
        "metric": "time-to-1e-8 DIMACS (FP64 vs mixed, 1 GPU); batched solves/s + p50/p99 latency",
        "value": round(rate, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * span, 2), "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded generators, %s)" % C5["desc"],
        "config": {"workload": C5["desc"], "requests": count, "precision": "mixed", "tol": 1e-8, "slots": args.slots,
                   "m_range": [min(ms), max(ms)]},
        "optimal": int(np.sum(st == 0)), "latency_saturation_s": {"p50": round(float(np.percentile(lat, 50)), 4),
                                                                 "p99": round(float(np.percentile(lat, 99)), 4)},
        "poisson": {"rate_per_s": round(lam, 2), "load": args.load, "span_s": round(pspan, 3), "optimal": int(np.sum(pst == 0)),
                    "p50_s": round(float(np.percentile(plat, 50)), 4), "p99_s": round(float(np.percentile(plat, 99)), 4)},
        "cpu_baseline": cpu,
        "gpu_over_cpu": None if not cpu or not cpu.get("value") else round(rate / cpu["value"], 2),
        "device": brisk_amd.load_library().bk_device_name().decode(),
    }
    print(json.dumps(out))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS) + ["c5"])
    ap.add_argument("--requests", type=int, default=0, help="c5: requests of the stream to serve (default 4096)")
    ap.add_argument("--slots", type=int, default=512, help="c5: most solves in flight")
    ap.add_argument("--load", type=float, default=0.8, help="c5: Poisson arrival rate as a fraction of the saturation rate")
    ap.add_argument("--cpu-stride", type=int, default=16, help="c5: the CPU baseline solves every k-th request")
    ap.add_argument("--cpu-serving-sample", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--schur-apply", default="auto", choices=["auto", "inverse", "solves"],
                    help="c5: how the Schur factor is applied in the PCG (-schurapply)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.cpu_serving_sample:
        cpu_serving_sample(args.cpu_serving_sample, args.cpu_threads or min(16, os.cpu_count() or 1))
        return
    if args.gpus != 1:
        print("bench.py: the solve runs on one GPU; --gpus %d is not supported (the multi-GPU sharding is a design only)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "c5":
        serving(args)
        return
    import torch
    import brisk_amd
    from _ctypes_util import Instance  # noqa: F401
    name = args.workload
    inst = instance(name)
    a = inst.arrays()
    print("bench: %s: m=%d, blocks %s" % (WORKLOADS[name]["desc"], a["m"], list(a["bs"])[:4]), file=sys.stderr, flush=True)
    L = brisk_amd.load_library()

    def timed(prec, k):
        ts, res = [], None
        for _ in range(k):
            torch.cuda.synchronize()
            res = solve(a, prec)
            torch.cuda.synchronize()
            ts.append(res.phases[1] + res.phases[2])
        return ts, res

    for w in range(args.warmup):
        r = solve(a, "mixed")
        print("bench: warmup %d: %s in %.3fs (%d iterations)" % (w, r.status_str, r.phases[1] + r.phases[2], r.iterations),
              file=sys.stderr, flush=True)
    ts, rm = timed("mixed", args.steps)
    t_mixed = float(np.median(ts))
    print("bench: mixed: %s, %d iterations, %.4fs per solve (median of %d), DIMACS max %.1e, phases %s, counts %s"
          % (rm.status_str, rm.iterations, t_mixed, len(ts), np.max(np.abs(rm.dimacs)), np.round(rm.phases, 4), rm.counts),
          file=sys.stderr, flush=True)
    t64, r64 = None, None
    if not args.skip_fp64:
        ts64, r64 = timed("double", max(1, min(args.steps, 2)))
        t64 = float(np.median(ts64))
        print("bench: fp64: %s, %d iterations, %.4fs per solve, DIMACS max %.1e" % (r64.status_str, r64.iterations, t64,
              np.max(np.abs(r64.dimacs))), file=sys.stderr, flush=True)
    # live per-kernel timing (events on the library's stream) of one more mixed solve
    brisk_amd.profile(True)
    brisk_amd.profile_reset()
    rp = solve(a, "mixed")
    prof = {c: brisk_amd.profile_read(c) for c in KCLASS}
    brisk_amd.profile(False)
    tot = sum(p["ms"] for p in prof.values())
    dom = max(prof, key=lambda c: prof[c]["ms"])
    for c in sorted(prof, key=lambda c: -prof[c]["ms"])[:8]:
        p = prof[c]
        if p["launches"]:
            print("bench: kernel %-22s %8.2f ms over %5d launches (%.1f us avg), %.2f TFLOP/s, %.0f GB/s" % (
                KCLASS[c], p["ms"], p["launches"], 1e3 * p["ms"] / p["launches"],
                p["flops"] / max(p["ms"], 1e-9) / 1e9, p["bytes"] / max(p["ms"], 1e-9) / 1e6), file=sys.stderr, flush=True)
    d = prof[dom]
    avg_s = d["ms"] / max(d["launches"], 1) / 1e3
    if dom <= 8 and d["flops"] > 0:           # the MFMA classes: GEMMs and the diagonal-block factor
        prec = "f64" if dom < 4 or dom in (8,) else "f32"
        ach = d["flops"] / max(d["launches"], 1) / avg_s / 1e12
        roof = dict(bound="mfma", achieved=round(ach, 3), peak=PEAK[prec], unit="TFLOP/s", frac=round(ach / PEAK[prec], 4),
                    traffic=measured_traffic(name, KCLASS[dom]), kernel=KCLASS[dom], launches=int(d["launches"]), avg_us=round(avg_s * 1e6, 2),
                    flops_per_launch=d["flops"] / max(d["launches"], 1), share_of_kernel_time=round(d["ms"] / max(tot, 1e-9), 3))
    else:
        ach = d["bytes"] / max(d["launches"], 1) / avg_s / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK, unit="GB/s", frac=round(ach / HBM_PEAK, 4),
                    traffic=measured_traffic(name, KCLASS[dom]),
                    kernel=KCLASS[dom], launches=int(d["launches"]), avg_us=round(avg_s * 1e6, 2),
                    bytes_per_launch=d["bytes"] / max(d["launches"], 1), share_of_kernel_time=round(d["ms"] / max(tot, 1e-9), 3))
    cpu = None
    vs_ref = None
    if not args.skip_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ref = cpu_reference(name, threads, timeout=600)
        if ref and "error" not in ref:
            cpu = dict(value=round(ref["wall"], 4), unit="s", cores=threads, kind="reference",
                       sample="one full solve of the same instance by the reference library (%s), %s" % (" ".join(ref["opts"]), cpu_model()),
                       status=ref["status"], iterations=ref["iters"], pobj=ref["pobj"], dimacs_max=max(abs(x) for x in ref["dimacs"]))
            vs_ref = dict(pobj_rel_mixed=abs(rm.pobj - ref["pobj"]) / (1 + abs(ref["pobj"])),
                          dobj_rel_mixed=abs(rm.dobj - ref["dobj"]) / (1 + abs(ref["dobj"])),
                          same_status=rm.status == ref["status"], cpu_over_gpu=round(ref["wall"] / t_mixed, 2))
        else:
            cpu = dict(value=None, unit="s", cores=threads, kind="reference", sample="not run: %s" % (ref or {}).get("error", "oracle/_ref/libbrisk.so absent"))
    out = {
        "metric": "time-to-1e-8 DIMACS (FP64 vs mixed, 1 GPU); batched solves/s + p50/p99 latency",
        "value": round(t_mixed, 5), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded generator, %s)" % WORKLOADS[name]["desc"],
        "config": {"workload": WORKLOADS[name]["desc"], "precision": "mixed (FP32-MFMA Schur factor, FP64 refinement and iterates)",
                   "tol": 1e-8, "m": int(a["m"]), "n": [int(x) for x in a["bs"]][:8]},
        "status": rm.status_str, "iterations": rm.iterations, "dimacs": [float("%.3e" % x) for x in rm.dimacs],
        "pobj": rm.pobj, "dobj": rm.dobj, "solves_per_s": round(1.0 / t_mixed, 3),
        "fp64_time_s": None if t64 is None else round(t64, 5), "fp64_status": None if r64 is None else r64.status_str,
        "fp64_iterations": None if r64 is None else r64.iterations, "mixed_over_fp64": None if t64 is None else round(t64 / t_mixed, 3),
        "phases_s": {"setup_upload": round(rm.phases[0], 4), "iterations": round(rm.phases[1], 4), "dimacs": round(rm.phases[2], 4),
                     "download": round(rm.phases[3], 4)},
        "factor_counts": {"f32": rm.counts[0], "f64": rm.counts[1], "pcg_steps": rm.counts[2]},
        "roofline": roof, "cpu_baseline": cpu, "vs_reference": vs_ref,
        "device": L.bk_device_name().decode(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
