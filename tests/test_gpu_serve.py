"""Batched solves and the serving loop on the GPU: every result is the same alone and in the
mix (same status and iteration count, objectives within 1e-10 relative); requests that arrive
while others are in flight are admitted as slots free up; limits end a solve with the
reference's statuses."""
import numpy as np
import pytest
import brisk_amd
from _ctypes_util import Instance, golden, golden_instance
from _parity import assert_identical, check_vs_reference

pytestmark = pytest.mark.gpu

MIX = [("theta", 30, 0.3, 3), ("maxcut", 100, 0.1, 0), ("multi", 4, 10, 20, 60, 0), ("theta", 50, 0.2, 0),
       ("dense", 12, 30, 0), ("maxcut_cycle", 7), ("multi", 6, 30, 40, 300, 2), ("theta", 80, 0.1, 5)]


def tup(inst):
    a = inst.arrays()
    return (a["bs"], a["c"], a["mat"], a["blk"], a["i"], a["j"], a["v"])


@pytest.fixture(scope="module")
def alone():
    out = {}
    for case in MIX:
        out[case] = brisk_amd.solve_sdpa(*tup(Instance(*case)), options=["-prec", "double"])
        assert out[case].status == 0
    return out


def same(a, b):
    assert a.status == b.status and a.iterations == b.iterations, (a.status, b.status, a.iterations, b.iterations)
    assert abs(a.pobj - b.pobj) <= 1e-10 * (1 + abs(b.pobj))
    assert abs(a.dobj - b.dobj) <= 1e-10 * (1 + abs(b.dobj))
    assert np.allclose(a.dimacs, b.dimacs, rtol=1e-4, atol=1e-13)


def test_batch_equals_alone(alone):
    print("batch of", len(MIX), flush=True)
    rs = brisk_amd.solve_batch([tup(Instance(*c)) for c in MIX], options=["-prec", "double"])
    for case, r in zip(MIX, rs):
        print(case, r.status_str, r.iterations, r.pobj, flush=True)
        same(r, alone[case])


@pytest.mark.parametrize("prec", ["double", "mixed"])
def test_schur_apply_modes_agree(prec):
    """The factor applied by its explicit inverse (alone) and by the blocked triangular solves (in
    a batch): the same statuses, iteration counts and objectives to the batch tolerance."""
    cases = MIX[:5]
    rs = brisk_amd.solve_batch([tup(Instance(*c)) for c in cases], options=["-prec", prec, "-schurapply", "solves"])
    for case, r in zip(cases, rs):
        a = brisk_amd.solve_sdpa(*tup(Instance(*case)), options=["-prec", prec, "-schurapply", "inverse"])
        print(case, prec, r.status_str, r.iterations, r.pobj, a.iterations, a.pobj, flush=True)
        if prec == "double":    # the same bar as alone against batched, DIMACS errors (near 1e-9) to 1e-2
            assert r.status == a.status and r.iterations == a.iterations
            assert abs(r.pobj - a.pobj) <= 1e-10 * (1 + abs(a.pobj)) and abs(r.dobj - a.dobj) <= 1e-10 * (1 + abs(a.dobj))
            assert np.allclose(r.dimacs, a.dimacs, rtol=1e-2, atol=1e-13)
        else:                   # the f32 preconditioners differ: the PCG may take other rounds
            assert r.status == a.status == 0
            assert abs(r.pobj - a.pobj) <= 1e-7 * (1 + abs(a.pobj)) and abs(r.dobj - a.dobj) <= 1e-7 * (1 + abs(a.dobj))


def test_serving_equals_alone(alone):
    """Continuous batching: three slots, requests arriving in waves while others run."""
    sv = brisk_amd.Server(["-prec", "double", "-slots", "3", "-heapgb", "2"])
    ids = {}
    pending = list(MIX)
    steps, left, max_inflight = 0, 1, 0
    while pending or left > 0:
        if pending and steps % 3 == 0:          # two new requests every third step
            for c in pending[:2]:
                ids[c] = sv.submit(*tup(Instance(*c)))
            pending = pending[2:]
        left = sv.step()
        max_inflight = max(max_inflight, sv.inflight())
        steps += 1
        assert steps < 2000
    assert 1 < max_inflight <= 3
    finished = sv.done()
    assert sorted(finished) == sorted(ids.values())
    for case, rid in ids.items():
        got = sv.take(rid)
        assert got is not None
        r, lat = got
        print(case, r.status_str, r.iterations, r.pobj, "latency %.3f s" % lat, flush=True)
        assert lat > 0
        same(r, alone[case])
    sv.close()


def test_iteration_limit():
    r = brisk_amd.solve_sdpa(*tup(Instance("theta", 50, 0.2, 0)), options=["-maxit", "3"])
    assert r.status_str == "ITERATION LIMIT" and r.exit_code == 13 and r.iterations == 3


def test_time_limit():
    r = brisk_amd.solve_sdpa(*tup(Instance("maxcut", 200, 0.05, 1)), options=["-timelimit", "1e-6"])
    assert r.status_str == "TIME LIMIT" and r.exit_code == 15


def test_mixed_precision_counts():
    """A mixed-precision solve factors in FP32 while mu is large and in FP64 at the end."""
    rec = golden("small")["maxcut_200_0.05_s1"]
    r = brisk_amd.solve_sdpa(*tup(golden_instance(rec)), options=["-prec", "mixed"])
    assert r.status == 0
    f32, f64, pcg = r.counts
    assert f32 > 0 and f64 > 0 and f32 + f64 >= r.iterations


STREAM = golden("stream")


def test_serving_mixed_stream_equals_alone():
    """The serving stream's requests in mixed precision (its stated precision): 27 requests of the
    stream (requests 0-25 and 31: theta and max-cut with m from 32 to 4985, three theta requests
    with m > 1500, so that f32 factors, their refinement and the 256-column panels of the largest
    are used), arriving three every second step into 8 slots and a 1.5 GB heap, on which the
    largest must wait for memory. Every served result equals the same request solved alone with
    the same options (the serving loop's default -throughput 1), bit for bit (status, iterations,
    objectives, DIMACS errors, y, X, Z), and matches the reference library's solve of that request
    (tests/golden/stream.json)."""
    names = sorted(STREAM, key=lambda k: STREAM[k]["request"])
    tups = {k: tup(golden_instance(STREAM[k])) for k in names}
    solo = {}
    for k in names:
        print("alone", k, "m =", STREAM[k]["m"], flush=True)
        # the serving loop's default plan choices are -throughput 1
        solo[k] = brisk_amd.solve_sdpa(*tups[k], options=["-prec", "mixed", "-throughput", "1"])
    sv = brisk_amd.Server(["-prec", "mixed", "-slots", "8", "-heapgb", "1.5"])
    ids, pending = {}, list(names)
    steps, left, max_inflight, waited = 0, 1, 0, False
    while pending or left > 0:
        if pending and steps % 2 == 0:
            for k in pending[:3]:
                ids[k] = sv.submit(*tups[k])
            pending = pending[3:]
        left = sv.step()
        fly = sv.inflight()
        max_inflight = max(max_inflight, fly)
        waited |= left - fly > 0 and fly < 8          # a request waits with a slot free: memory
        steps += 1
        assert steps < 5000
    print("steps", steps, "max in flight", max_inflight, "waited on memory", waited, flush=True)
    assert max_inflight > 3 and waited
    assert sorted(sv.done()) == sorted(ids.values())
    for k in names:
        r, lat = sv.take(ids[k])
        print(k, r.status_str, r.iterations, r.counts, repr(r.pobj), "latency %.3f s" % lat, flush=True)
        assert_identical(r, solo[k])
        check_vs_reference(r, STREAM[k], k)
    sv.close()


def test_serving_setup_failure_ends_unsolved(monkeypatch):
    """A request whose device setup fails (forced from the host with BRISK_FAULT=setup, the
    reference's fault-injection hook) with nothing in flight ends unsolved with its cause, and the
    loop drains instead of reporting it as waiting forever; without the fault the next request
    is solved."""
    sv = brisk_amd.Server(["-prec", "double", "-slots", "2", "-heapgb", "0.5"])
    monkeypatch.setenv("BRISK_FAULT", "setup")
    bad = sv.submit(*tup(Instance("theta", 30, 0.3, 3)))
    assert sv.step() == 0
    r, _ = sv.take(bad)
    assert r.status == -1 and r.exit_code == -1 and "setup" in r.cause
    monkeypatch.delenv("BRISK_FAULT")
    good = sv.submit(*tup(Instance("theta", 30, 0.3, 3)))
    for _ in range(200):
        if sv.step() == 0:
            break
    r, _ = sv.take(good)
    assert r.status == 0
    sv.close()


def test_serving_bad_data_and_build_failure(monkeypatch):
    """Submission checks the data itself (an entry outside its block is refused at once, as the
    solve refuses it); a request whose host build fails on the worker threads (forced with
    BRISK_FAULT=build) ends unsolved with its cause while the requests around it are solved."""
    sv = brisk_amd.Server(["-prec", "double", "-slots", "4"])
    bs, c, mat, blk, i, j, v = tup(Instance("theta", 30, 0.3, 3))
    i_bad = np.array(i, copy=True)
    i_bad[3] = 31
    with pytest.raises(ValueError):
        sv.submit(bs, c, mat, blk, i_bad, j, v)
    first = sv.submit(*tup(Instance("maxcut", 100, 0.1, 0)))
    monkeypatch.setenv("BRISK_FAULT", "build")
    bad = sv.submit(*tup(Instance("theta", 30, 0.3, 3)))
    monkeypatch.delenv("BRISK_FAULT")
    last = sv.submit(*tup(Instance("theta", 50, 0.2, 0)))
    for _ in range(400):
        if sv.step() == 0:
            break
    r, _ = sv.take(bad)
    assert r.status == -1 and "host structures" in r.cause
    for rid in (first, last):
        r, _ = sv.take(rid)
        assert r.status == 0
    sv.close()
