"""Comparisons of a GPU result with a reference record (golden fixture) or with another GPU result."""
import numpy as np

TOL_OBJ, TOL_DIMACS = 1e-7, 1e-8


def rel(a, b):
    return abs(a - b) / (1.0 + abs(b))


def objective_ok(ours, rec, key):
    """Within 1e-7 of the reference's FP64 solve, or (where that solve is itself further than
    that from the optimum its double-double solve gives) at least as close to the optimum."""
    if rel(ours, rec[key]) <= TOL_OBJ:
        return True
    if "dd" not in rec:
        return False
    dd = rec["dd"][key]
    return rel(ours, dd) <= max(rel(rec[key], dd), 1e-9)


def flat_x(r):
    """X of a result as the fixtures store it: the blocks concatenated, SDP blocks column-major."""
    if r.X is None:
        return None
    return np.concatenate([np.asarray(x).reshape(-1, order="F") for x in r.X])


def dist(a, b):
    """||a - b|| / (1 + ||b||) (Frobenius for the stacked X blocks)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / (1 + np.linalg.norm(b)))


# Iterate distance: a 1e-8 solution of an SDP is not a point but a neighbourhood of the optimal
# face. The bar is set per problem by the reference's own spread: our point lies within
# max(1e-4, 10 x d_ref) of the reference's double-double optimum, d_ref being the distance of the
# reference's own FP64 solve to that optimum (large where the optimum is not unique: theta1's
# dual 5e-2). Where the record has no dd point (the full-size workloads, generic random data with
# a unique optimum), y lies within 1e-4 of the reference's FP64 solve.
DIST_FLOOR, DIST_FACTOR = 1e-4, 10.0


def iterate_distance(r, rec):
    """Distances of our y and X to the reference's returned point ("y", "X"), to its double-double
    optimum ("y_dd", "X_dd"), the reference's own distance to that optimum ("y_ref_dd",
    "X_ref_dd"), and the bounds asserted on them ("y_bound", "X_bound"); None where absent."""
    out = dict.fromkeys(("y", "X", "y_dd", "X_dd", "y_ref_dd", "X_ref_dd", "y_bound", "X_bound"))
    ours = {"y": None if r.y is None else np.asarray(r.y), "X": flat_x(r)}
    dd = rec.get("dd", {})
    for key in ("y", "X"):
        v = ours[key]
        if v is None:
            continue
        if key in rec and len(rec[key]) == len(v):
            out[key] = dist(v, rec[key])
        if key in dd and len(dd[key]) == len(v):
            out[key + "_dd"] = dist(v, dd[key])
            if key in rec:
                out[key + "_ref_dd"] = dist(rec[key], dd[key])
                out[key + "_bound"] = max(DIST_FLOOR, DIST_FACTOR * out[key + "_ref_dd"])
        elif out[key] is not None:
            out[key + "_bound"] = DIST_FLOOR
    return out


def check_distance(d):
    for key in ("y", "X"):
        b = d[key + "_bound"]
        if b is None:
            continue
        got = d[key + "_dd"] if d[key + "_dd"] is not None else d[key]
        assert got <= b, "%s distance %.2e above %.2e" % (key, got, b)


def check_vs_reference(r, rec, label=""):
    """Same status (OPTIMAL), objectives within 1e-7 relative (or as close to the dd optimum as the
    reference's own solve), DIMACS errors <= 1e-8, y and X within the iterate-distance bar above.
    Returns the distances."""
    d = iterate_distance(r, rec)
    print("gpu", label, r.status_str, r.iterations, repr(r.pobj), repr(r.dobj), r.dimacs, r.counts, flush=True)
    print("ref", label, rec["status"], rec["iterations"], repr(rec["pobj"]), repr(rec["dobj"]), rec["dimacs"], flush=True)
    print("iterate distance", label, {k: ("%.2e" % v if v is not None else None) for k, v in d.items()}, flush=True)
    assert r.status == rec["status"] == 0, r.status_str
    assert objective_ok(r.pobj, rec, "pobj")
    assert objective_ok(r.dobj, rec, "dobj")
    assert np.max(np.abs(r.dimacs)) <= TOL_DIMACS
    check_distance(d)
    return d


def assert_identical(a, b):
    """Two results of the same problem bit for bit: status, iterations, objectives, DIMACS errors,
    y and every block of X and Z."""
    assert a.status == b.status and a.iterations == b.iterations, (a.status, b.status, a.iterations, b.iterations)
    assert a.pobj == b.pobj and a.dobj == b.dobj, (a.pobj, b.pobj, a.dobj, b.dobj)
    assert np.array_equal(a.dimacs, b.dimacs), (a.dimacs, b.dimacs)
    assert np.array_equal(a.y, b.y)
    assert len(a.X) == len(b.X)
    for xa, xb, za, zb in zip(a.X, b.X, a.Z, b.Z):
        assert np.array_equal(xa, xb) and np.array_equal(za, zb)
