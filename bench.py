#!/usr/bin/env python3
"""bench.py - time to a 1e-8 DIMACS solution on one MI355X, FP64 and mixed precision, and the
serving loop's batched solves/s with p50/p99 latency.

Headline (`value`): the max-cut SDP relaxation of a +-1-weighted random graph (n = 2000,
p = 0.01, seed 0; C = -L/4, diag X = 1, m = 2000) in mixed precision (FP32-MFMA Schur factor +
FP64 refinement). A step is one complete solve with the problem already resident on the device:
the start point, the interior-point iterations and the DIMACS evaluation (the library's phase
clock, synchronised at both ends). Beside it: the whole call as a user sees it (host setup,
upload, download: `end_to_end_s`), on which the CPU ratio is taken, since the reference CPU
solver is timed the same way (a wall clock around brisk_solve_data); the reference's CPU solve
of the same instance, both by its IPM (-bm 0) and by its default for this instance
(Burer-Monteiro); the dominant kernel's roofline.

The same line carries the other workloads as sub-objects ("workloads"): c1 the Lovasz theta SDP
of G(50, 0.2) (FP64), c3 the dense random SDP n = 400, m = 2000 (mixed and FP64; the 10x
target's config), c4 64 blocks of 60 + LP 500 with m = 5000, and c5 the serving stream (4096
theta / max-cut requests: solves/s with every request queued at once, p50/p99 latency under
Poisson arrivals, the reference's CPU rate, and sum(flops) / (t peak) over the stream).

The reference CPU solver is the reference library compiled from its own sources into
oracle/_ref/ (where present), run in a subprocess on the host cores. One JSON line on stdout.
`--workload c1|c2|c3|c4|c5` runs one workload alone with it as the headline (for profiling).
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

from brisk_amd.generators import Instance, STREAM, stream  # noqa: E402

METRIC = "time-to-1e-8 DIMACS (FP64 vs mixed, 1 GPU); batched solves/s + p50/p99 latency"
PEAK = {"f64": 78.6, "f32": 157.3}      # MI355X dense MFMA TFLOP/s
HBM_PEAK = 8000.0                        # GB/s
KCLASS = {0: "k_gemm<double, 0, 0>", 1: "k_gemm<double, 0, 1>", 2: "k_gemm<double, 1, 0>", 3: "k_gemm<double, 1, 1>",
          4: "k_gemm<float, 0, 0>", 5: "k_gemm<float, 0, 1>", 6: "k_gemm<float, 1, 0>", 7: "k_gemm<float, 1, 1>",
          8: "k_potrf_diag", 9: "k_gemv", 10: "Schur assembly", 11: "k_aop", 12: "k_lz_step", 13: "k_dot_chunks",
          14: "k_lincomb", 15: "k_symv", 16: "k_trsv"}
F32_CLASSES = (4, 5, 6, 7)

WORKLOADS = {
    "c1": dict(kind="theta", args=(50, 0.2, 0), prec="double", fixture=("small", "theta_50_0.2_s0"),
               desc="Lovasz theta SDP of G(50, 0.2), seed 0 (C=-J, tr X=1, X_ij=0 on edges)"),
    "c2": dict(kind="maxcut", args=(2000, 0.01, 0), prec="mixed", fixture=("large", "maxcut_2000_0.01_s0"),
               desc="max-cut SDP of G(2000, 0.01), weights +-1, seed 0 (C=-L/4, diag X=1, m=2000)"),
    "c3": dict(kind="dense", args=(400, 2000, 0), prec="mixed", fixture=("large", "dense_400_2000_s0"),
               desc="dense random SDP, one block n=400, m=2000, seed 0, feasible by X0/Z0/y0"),
    "c4": dict(kind="multi", args=(64, 60, 500, 5000, 0), prec="mixed", fixture=("large", "multi_64_60_500_5000_s0"),
               desc="64 PSD blocks of 60 + LP 500, m=5000, 1-3 blocks per row, seed 0"),
}
# measured HBM bytes per dispatch by kernel class (PMC FETCH_SIZE + WRITE_SIZE, copy-calibrated),
# kept with the rocprofv3 summaries; keyed by workload
TRAFFIC_FILES = {"c2": "maxcut2000_mixed_hbm_traffic_pmc.md", "c4": "blocks64x60_lp500_m5000_mixed_hbm_traffic_pmc.md"}


def log(*a):
    print("bench:", *a, file=sys.stderr, flush=True)


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
    w = WORKLOADS[name]
    return Instance(w["kind"], *w["args"])


def fixture(name):
    kind, key = WORKLOADS[name]["fixture"]
    path = os.path.join(ROOT, "tests", "golden", kind + ".json")
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except OSError:
        return None


def solve(a, prec):
    import brisk_amd
    return brisk_amd.solve_sdpa(*a, options=["-prec", prec])


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libbrisk.so")


def cpu_reference(name, threads, timeout, extra=()):
    """Solve the instance with the reference library in a subprocess (wall clock around
    brisk_solve_data, the same terms as end_to_end_s); returns a dict or None."""
    if not os.path.exists(REF_LIB):
        return None
    code = r"""
import ctypes, json, sys, time
sys.path.insert(0, %r)
import bench
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
opts = ["-threads", str(%d), "-q"] + %r
o = (ctypes.c_char_p * len(opts))(*[x.encode() for x in opts])
r = L.brisk_result_new()
t0 = time.time()
rc = L.brisk_solve_data(*inst.data_args(), len(opts), o, r)
t = time.time() - t0
e = (D * 6)(); L.brisk_result_dimacs(r, e)
print(json.dumps(dict(rc=rc, status=L.brisk_result_status(r), pobj=L.brisk_result_pobj(r), dobj=L.brisk_result_dobj(r),
                      iters=L.brisk_result_iterations(r), wall=t, time=L.brisk_result_time(r), dimacs=list(e), opts=opts)))
""" % (ROOT, name, REF_LIB, threads, list(extra))
    env = dict(os.environ, BRISK_NOBUSY="1", OMP_NUM_THREADS=str(threads))
    try:
        out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout, env=env)
    except subprocess.TimeoutExpired:
        return dict(error="timeout after %ds" % timeout)
    for line in out.stdout.splitlines()[::-1]:
        if line.startswith("{"):
            return json.loads(line)
    return dict(error=(out.stderr or out.stdout)[-400:])


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def vs_fixture(r, rec):
    """Our result against the reference's recorded solve of the instance (tests/golden)."""
    if not rec:
        return None
    out = dict(same_status=r.status == rec["status"], pobj_rel=rel(r.pobj, rec["pobj"]), dobj_rel=rel(r.dobj, rec["dobj"]),
               ref_iterations=rec["iterations"], ref_dimacs_max=max(abs(x) for x in rec["dimacs"]))
    if "y" in rec and r.y is not None and len(rec["y"]) == len(r.y):
        yr = np.asarray(rec["y"])
        out["y_distance"] = float(np.linalg.norm(r.y - yr) / (1 + np.linalg.norm(yr)))
    return out


def profile_solve(a, prec, name):
    """One more solve with the per-class kernel timers (HIP events on the library stream): the
    dominant class's achieved rate against its roofline, and the class table on stderr."""
    import brisk_amd
    brisk_amd.profile(True)
    brisk_amd.profile_reset()
    solve(a, prec)
    prof = {c: brisk_amd.profile_read(c) for c in KCLASS}
    brisk_amd.profile(False)
    tot = sum(p["ms"] for p in prof.values())
    # the dominant class among those with an algorithmic count (the Lanczos steps, whose count
    # depends on where each block's bound converged, carry none)
    counted = [c for c in prof if prof[c]["flops"] > 0 or prof[c]["bytes"] > 0] or list(prof)
    dom = max(counted, key=lambda c: prof[c]["ms"])
    for c in sorted(prof, key=lambda c: -prof[c]["ms"])[:8]:
        p = prof[c]
        if p["launches"]:
            log("%s kernel %-22s %8.2f ms over %5d launches (%.1f us avg), %.2f TFLOP/s, %.0f GB/s" % (
                name, KCLASS[c], p["ms"], p["launches"], 1e3 * p["ms"] / p["launches"],
                p["flops"] / max(p["ms"], 1e-9) / 1e9, p["bytes"] / max(p["ms"], 1e-9) / 1e6))
    d = prof[dom]
    avg_s = d["ms"] / max(d["launches"], 1) / 1e3
    if dom <= 8 and d["flops"] > 0:           # the MFMA classes: GEMMs and the diagonal-block factor
        pk = PEAK["f32" if dom in F32_CLASSES else "f64"]
        ach = d["flops"] / max(d["launches"], 1) / avg_s / 1e12
        roof = dict(bound="mfma", achieved=round(ach, 3), peak=pk, unit="TFLOP/s", frac=round(ach / pk, 4),
                    traffic=measured_traffic(name, KCLASS[dom]), kernel=KCLASS[dom], launches=int(d["launches"]),
                    avg_us=round(avg_s * 1e6, 2), flops_per_launch=d["flops"] / max(d["launches"], 1),
                    share_of_kernel_time=round(d["ms"] / max(tot, 1e-9), 3))
    else:
        ach = d["bytes"] / max(d["launches"], 1) / avg_s / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK, unit="GB/s", frac=round(ach / HBM_PEAK, 4),
                    traffic=measured_traffic(name, KCLASS[dom]), kernel=KCLASS[dom], launches=int(d["launches"]),
                    avg_us=round(avg_s * 1e6, 2), bytes_per_launch=d["bytes"] / max(d["launches"], 1),
                    share_of_kernel_time=round(d["ms"] / max(tot, 1e-9), 3))
    if dom == 8:
        roof["note"] = ("latency-bound: each launch is one 64-column step of the blocked Cholesky (the diagonal block "
                        "factored and inverted in LDS by every workgroup, its panel tile multiplied on MFMA); the "
                        "diagonal block's 64 pivots form a dependency chain (an in-kernel timer puts it at 13 of "
                        "the step's 22 us), so the step's time does not shrink with its flops")
    # the next classes by time, each against its own roofline
    nxt = []
    for c in sorted(counted, key=lambda c: -prof[c]["ms"]):
        if c == dom or len(nxt) == 2:
            continue
        p = prof[c]
        if c <= 8 and p["flops"] > 0:
            pk = PEAK["f32" if c in F32_CLASSES else "f64"]
            a_ = p["flops"] / max(p["ms"], 1e-9) / 1e9
            nxt.append(dict(kernel=KCLASS[c], bound="mfma", achieved=round(a_, 3), peak=pk, unit="TFLOP/s", frac=round(a_ / pk, 4),
                            share_of_kernel_time=round(p["ms"] / max(tot, 1e-9), 3)))
        else:
            a_ = p["bytes"] / max(p["ms"], 1e-9) / 1e6
            nxt.append(dict(kernel=KCLASS[c], bound="hbm", achieved=round(a_, 1), peak=HBM_PEAK, unit="GB/s", frac=round(a_ / HBM_PEAK, 4),
                            share_of_kernel_time=round(p["ms"] / max(tot, 1e-9), 3)))
    roof["next_classes"] = nxt
    f64 = sum(prof[c]["flops"] for c in prof if c <= 3 or c == 8)
    f32 = sum(prof[c]["flops"] for c in F32_CLASSES)
    return roof, dict(flops_f64=f64, flops_f32=f32, kernel_ms=tot)


def timed(a, prec, k):
    """k solves: their device-resident times (start point + iterations + DIMACS), whole-call
    times (brisk_result_time: host setup, upload, iterations, DIMACS, download) and wall clocks."""
    import torch
    res, ts, te, tw = None, [], [], []
    for _ in range(k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = solve(a, prec)
        torch.cuda.synchronize()
        tw.append(time.perf_counter() - t0)
        ts.append(res.phases[1] + res.phases[2])
        te.append(res.time)
    return float(np.median(ts)), float(np.median(te)), float(np.median(tw)), res


def workload(name, steps, warmup, fp64_steps, cpu, threads, cpu_timeout=600):
    """One of c1-c4: the precision the workload is quoted in and the other beside it."""
    w = WORKLOADS[name]
    inst = instance(name)
    a = inst.sdpa()
    log("%s: %s: m=%d, blocks %s" % (name, w["desc"], inst.m, inst.bs[:4]))
    main_prec = w["prec"]
    for k in range(warmup):
        r = solve(a, main_prec)
        log("%s warmup %d: %s in %.3fs (%d iterations)" % (name, k, r.status_str, r.phases[1] + r.phases[2], r.iterations))
    t, te, tw, r = timed(a, main_prec, steps)
    log("%s %s: %s, %d iterations, %.4fs per solve resident (median of %d), %.4fs whole call, phases %s, counts %s, DIMACS max %.1e"
        % (name, main_prec, r.status_str, r.iterations, t, steps, te, np.round(r.phases, 4), r.counts, np.max(np.abs(r.dimacs))))
    out = dict(workload=w["desc"], m=int(inst.m), precision=main_prec, time_s=round(t, 5), end_to_end_s=round(te, 5),
               wall_s=round(tw, 5), status=r.status_str, iterations=r.iterations, dimacs=[float("%.3e" % x) for x in r.dimacs],
               pobj=r.pobj, dobj=r.dobj,
               phases_s=dict(setup_upload=round(r.phases[0], 4), iterations=round(r.phases[1], 4),
                             dimacs=round(r.phases[2], 4), download=round(r.phases[3], 4)),
               factor_counts=dict(f32=r.counts[0], f64=r.counts[1], pcg_steps=r.counts[2]),
               vs_reference_fixture=vs_fixture(r, fixture(name)))
    other = "double" if main_prec == "mixed" else "mixed"
    if fp64_steps:
        t2, te2, _, r2 = timed(a, other, fp64_steps)
        log("%s %s: %s, %d iterations, %.4fs per solve resident, %.4fs whole call" % (name, other, r2.status_str, r2.iterations, t2, te2))
        out["other_precision"] = dict(precision=other, time_s=round(t2, 5), end_to_end_s=round(te2, 5), status=r2.status_str,
                                      iterations=r2.iterations, dimacs_max=float(np.max(np.abs(r2.dimacs))),
                                      vs_reference_fixture=vs_fixture(r2, fixture(name)))
        t_f64, t_mixed = (t2, t) if main_prec == "mixed" else (t, t2)
        out["mixed_over_fp64"] = round(t_f64 / t_mixed, 3)
    out["roofline"], tot = profile_solve(a, main_prec, name)
    out["kernel_flops"] = tot
    if cpu:
        runs = [("ipm", ["-bm", "0"] if name == "c2" else [])]
        if name == "c2":
            runs.append(("default", []))          # the reference's default on max-cut n >= 500: Burer-Monteiro
        cb = {}
        for label, extra in runs:
            ref = cpu_reference(name, threads, cpu_timeout, extra)
            if ref and "error" not in ref:
                cb[label] = dict(value=round(ref["wall"], 4), unit="s", cores=threads, kind="reference",
                                 sample="one full solve of the same instance by the reference library (%s), %s"
                                        % (" ".join(ref["opts"]), cpu_model()),
                                 status=ref["status"], iterations=ref["iters"], pobj=ref["pobj"],
                                 dimacs_max=max(abs(x) for x in ref["dimacs"]),
                                 cpu_over_gpu_end_to_end=round(ref["wall"] / te, 2), cpu_over_gpu_resident=round(ref["wall"] / t, 2))
                log("%s reference CPU (%s): %.3fs, status %d" % (name, " ".join(ref["opts"]), ref["wall"], ref["status"]))
            else:
                cb[label] = dict(value=None, unit="s", cores=threads, kind="reference",
                                 sample="not run: %s" % (ref or {}).get("error", "oracle/_ref/libbrisk.so absent"))
        out["cpu_baseline"] = cb
    else:
        rec = fixture(name)
        if rec:
            out["cpu_baseline"] = {"ipm": dict(value=rec["seconds"], unit="s", cores=8, kind="reference", timed_on_this_host=False,
                                               sample="the reference library's full solve of the same instance recorded with the "
                                                      "fixture (tests/golden/make_golden.py, %s, 8 threads of an Intel Xeon @ 2.70 GHz); "
                                                      "not re-timed here (a solve takes ~10 min)" % " ".join(rec["options"]),
                                               cpu_over_gpu_end_to_end=round(rec["seconds"] / te, 2))}
    return out, r


# ---- the serving stream ----------------------------------------------------------------------

def cpu_serving_sample(stride, workers):
    """The reference library on every stride-th request of the stream, one thread per solve
    (the library is not reentrant), a pool of worker processes; prints one JSON line."""
    import multiprocessing as mp
    reqs, _ = stream()
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
    L = ctypes.CDLL(REF_LIB)
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


def serving(count, slots, steps, warmup, load, cpu, workers, stride, schur_apply="auto"):
    """The serving workload: solves/s with every request queued at t = 0 (saturation), p50/p99
    latency under Poisson arrivals at a fixed fraction of that rate, the reference's CPU rate, and
    sum(flops) / (t peak) of the stream's kernels (counted in a profiled run)."""
    import brisk_amd
    reqs, gaps = stream()
    reqs, gaps = reqs[:count], gaps[:count]
    insts = [Instance(kind, n, p, seed).sdpa() for kind, n, p, seed in reqs]
    ms = [int(x[1].shape[0]) for x in insts]
    log("c5: %d requests, m from %d to %d (mean %.0f)" % (count, min(ms), max(ms), np.mean(ms)))
    opts = ["-prec", "mixed", "-slots", str(slots), "-schurapply", schur_apply]
    for w in range(warmup):
        span, _, st, _ = run_stream(insts[:64], np.zeros(64), opts)
        log("c5 warmup %d: 64 requests in %.3fs" % (w, span))
    runs = []
    for k in range(steps):
        span, lat, st, it = run_stream(insts, np.zeros(count), opts)
        runs.append(span)
        log("c5 saturation run %d: %d requests in %.3fs = %.1f solves/s; optimal %d, p50 %.3fs p99 %.3fs, mean iterations %.1f"
            % (k, count, span, count / span, int(np.sum(st == 0)), np.percentile(lat, 50), np.percentile(lat, 99), it.mean()))
    span = float(np.median(runs))
    rate = count / span
    lam = load * rate
    arr = np.cumsum(gaps) / lam
    pspan, plat, pst, pit = run_stream(insts, arr, opts)
    log("c5 Poisson arrivals at %.1f/s (%.2f x the saturation rate): %d requests in %.3fs, optimal %d, p50 %.3fs, p99 %.3fs"
        % (lam, load, count, pspan, int(np.sum(pst == 0)), np.percentile(plat, 50), np.percentile(plat, 99)))
    # the stream's flops: the same saturation run with the per-class kernel counters on
    brisk_amd.profile(True)
    brisk_amd.profile_reset()
    pr_span, _, _, _ = run_stream(insts, np.zeros(count), opts)
    prof = {c: brisk_amd.profile_read(c) for c in KCLASS}
    brisk_amd.profile(False)
    f64 = sum(prof[c]["flops"] for c in prof if c <= 3 or c == 8)
    f32 = sum(prof[c]["flops"] for c in F32_CLASSES)
    # the stream's arithmetic at the two precisions' MFMA peaks: a machine that ran only these
    # flops at peak would take t_peak; frac = t_peak / t
    t_peak = f64 / (PEAK["f64"] * 1e12) + f32 / (PEAK["f32"] * 1e12)
    roof = dict(bound="mfma", achieved=round((f64 + f32) / span / 1e12, 3), peak=round((f64 + f32) / max(t_peak, 1e-30) / 1e12, 2),
                unit="TFLOP/s", frac=round(t_peak / span, 4), traffic=None, kernel="all MFMA classes over the stream",
                flops_f64=f64, flops_f32=f32, profiled_span_s=round(pr_span, 3),
                note="achieved = the stream's MFMA flops (counted with the kernel timers on) / the unprofiled saturation "
                     "makespan; peak = those flops' rate at the FP64 and FP32 MFMA peaks")
    log("c5 flops: %.3e f64 + %.3e f32 over %.3fs: %.2f TFLOP/s, %.4f of the mixed peak" % (f64, f32, span, (f64 + f32) / span / 1e12, t_peak / span))
    cpu_b = None
    if cpu and os.path.exists(REF_LIB):
        env = dict(os.environ, BRISK_NOBUSY="1", OMP_NUM_THREADS="1")
        try:
            out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-serving-sample", str(stride),
                                  "--cpu-threads", str(workers)], capture_output=True, text=True, timeout=900, env=env)
            line = [x for x in out.stdout.splitlines() if x.startswith("{")]
            ref = json.loads(line[-1]) if line else dict(error=out.stderr[-400:])
        except subprocess.TimeoutExpired:
            ref = dict(error="timeout")
        if "error" not in ref:
            # the rate of perfectly packed workers (solves x workers / CPU seconds): an upper bound of
            # the pool's throughput on the whole stream, above the sample's own wall-clock rate
            packed = ref["count"] * workers / max(ref["cpu_s"], 1e-9)
            cpu_b = dict(value=round(packed, 3), unit="solves/s", cores=workers, kind="reference",
                         sample="every %dth request of the stream (%d solves, %d optimal) by the reference library, one thread per "
                                "solve in %d worker processes: %.1f CPU s, %.1f s wall (%.2f solves/s measured; value = solves x "
                                "workers / CPU s), %s" % (stride, ref["count"], ref["optimal"], workers, ref["cpu_s"],
                                                         ref["wall"], ref["solves_per_s"], cpu_model()))
            log("c5 reference CPU: %.2f solves/s packed (%d workers)" % (packed, workers))
        else:
            cpu_b = dict(value=None, unit="solves/s", cores=workers, kind="reference", sample="not run: %s" % ref["error"])
    return dict(workload=STREAM["desc"], requests=count, precision="mixed", slots=slots, solves_per_s=round(rate, 3),
                makespan_s=round(span, 3), optimal=int(np.sum(st == 0)),
                latency_saturation_s=dict(p50=round(float(np.percentile(lat, 50)), 4), p99=round(float(np.percentile(lat, 99)), 4)),
                poisson=dict(rate_per_s=round(lam, 2), load=load, span_s=round(pspan, 3), optimal=int(np.sum(pst == 0)),
                             p50_s=round(float(np.percentile(plat, 50)), 4), p99_s=round(float(np.percentile(plat, 99)), 4)),
                m_range=[min(ms), max(ms)], roofline=roof, cpu_baseline=cpu_b,
                gpu_over_cpu=None if not cpu_b or not cpu_b.get("value") else round(rate / cpu_b["value"], 2))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="all", choices=["all"] + sorted(WORKLOADS) + ["c5"],
                    help="all: the c2 headline with c1, c3, c4, c5 beside it; else one workload as the headline")
    ap.add_argument("--requests", type=int, default=0, help="c5: requests of the stream to serve (default 4096)")
    ap.add_argument("--slots", type=int, default=2048, help="c5: most solves in flight (admission also waits on device memory)")
    ap.add_argument("--load", type=float, default=0.8, help="c5: Poisson arrival rate as a fraction of the saturation rate")
    ap.add_argument("--cpu-stride", type=int, default=16, help="c5: the CPU baseline solves every k-th request")
    ap.add_argument("--cpu-serving-sample", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--schur-apply", default="auto", choices=["auto", "inverse", "solves"],
                    help="c5: how the Schur factor is applied in the PCG (-schurapply)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    if args.cpu_serving_sample:
        cpu_serving_sample(args.cpu_serving_sample, threads)
        return
    if args.gpus != 1:
        print("bench.py: the solve runs on one GPU; --gpus %d is not supported (the multi-GPU sharding is a design only)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    # torch first: its HIP runtime is then the one the library binds to (loaded the other way
    # round, two runtimes share the process and torch finds no device)
    import torch  # noqa: F401
    import brisk_amd
    L = brisk_amd.load_library()
    count = args.requests or STREAM["count"]
    cpu = not args.skip_cpu
    if args.workload == "c5":
        s = serving(count, args.slots, max(1, args.steps), args.warmup, args.load, cpu, threads, args.cpu_stride, args.schur_apply)
        out = {"metric": METRIC, "value": s["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": max(1, args.steps),
               "warmup": args.warmup, "ms_per_step": round(1e3 * s["makespan_s"], 2), "higher_is_better": True,
               "scaling": "strong", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded generators, %s)" % STREAM["desc"],
               "config": {"workload": STREAM["desc"], "requests": count, "precision": "mixed", "tol": 1e-8, "slots": args.slots},
               "roofline": s["roofline"], "cpu_baseline": s["cpu_baseline"], "serving": s,
               "device": L.bk_device_name().decode()}
        print(json.dumps(out))
        return
    head = "c2" if args.workload == "all" else args.workload
    h, r = workload(head, args.steps, args.warmup, max(1, min(args.steps, 2)), cpu, threads)
    extras = {}
    if args.workload == "all":
        extras["c1"], _ = workload("c1", 5, 1, 2, cpu, threads, cpu_timeout=120)
        extras["c3"], _ = workload("c3", 2, 1, 2, False, threads)
        extras["c4"], _ = workload("c4", 3, 1, 2, cpu, threads, cpu_timeout=300)
        extras["c5"] = serving(count, args.slots, 1, 1, args.load, cpu, threads, args.cpu_stride)
    cb = h.get("cpu_baseline", {})
    main_cpu = cb.get("ipm")
    out = {
        "metric": METRIC, "value": h["time_s"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * h["time_s"], 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64" if h["precision"] == "mixed" else "f64",
        "data": "synthetic (seeded generator, %s)" % WORKLOADS[head]["desc"],
        "config": {"workload": WORKLOADS[head]["desc"], "precision": h["precision"], "tol": 1e-8, "m": h["m"]},
        "status": h["status"], "iterations": h["iterations"], "dimacs": h["dimacs"], "pobj": h["pobj"], "dobj": h["dobj"],
        "end_to_end_s": h["end_to_end_s"], "phases_s": h["phases_s"], "factor_counts": h["factor_counts"],
        "other_precision": h.get("other_precision"), "mixed_over_fp64": h.get("mixed_over_fp64"),
        "roofline": h["roofline"], "cpu_baseline": main_cpu, "cpu_baseline_default_method": cb.get("default"),
        "vs_reference": dict(fixture=h["vs_reference_fixture"],
                             cpu_over_gpu_end_to_end=None if not main_cpu else main_cpu.get("cpu_over_gpu_end_to_end"),
                             cpu_over_gpu_resident=None if not main_cpu else main_cpu.get("cpu_over_gpu_resident")),
        "workloads": extras, "device": L.bk_device_name().decode(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
