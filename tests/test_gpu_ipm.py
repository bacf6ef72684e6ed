"""End-to-end parity of the GPU interior-point solve with the oracle (the CPU restatement of the
reference's iteration) on seeded instances: same status, objectives within 1e-7 relative, DIMACS
errors <= 1e-8 measured independently on the returned point."""
import numpy as np
import pytest
import brisk_amd
from _ctypes_util import Instance, oracle_solve, oracle_dimacs

pytestmark = pytest.mark.gpu

CASES = [
    ("theta", 50, 0.2, 0),          # the C1 instance: Lovasz theta of G(50, 0.2), seed 0
    ("theta", 30, 0.3, 3),
    ("maxcut", 100, 0.1, 0),
    ("maxcut_cycle", 7),
    ("maxcut_cycle", 8),
    ("multi", 4, 10, 20, 60, 0),
    ("dense", 12, 30, 0),           # dense route: P-SYRK assembly, operator-form residual
    ("dense", 40, 120, 1),
    ("multi", 6, 30, 40, 300, 2),   # several blocks + an LP block
]


def _solve(inst, prec):
    a = inst.arrays()
    return brisk_amd.solve_sdpa(a["bs"], a["c"], a["mat"], a["blk"], a["i"], a["j"], a["v"], options=["-prec", prec])


def _packX(res):
    return np.concatenate([x.flatten(order="F") for x in res.X])


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
@pytest.mark.parametrize("prec", ["double", "mixed"])
def test_gpu_matches_oracle(case, prec):
    inst = Instance(*case)
    print("case", case, prec, flush=True)
    r = _solve(inst, prec)
    o = oracle_solve(inst)
    print("gpu   ", r.status_str, r.iterations, r.pobj, r.dobj, r.dimacs, r.counts, flush=True)
    print("oracle", o["status"], o["iters"], o["pobj"], o["dobj"], o["err"], flush=True)
    assert o["status"] == 0
    assert r.status == 0, r.status_str
    assert abs(r.pobj - o["pobj"]) <= 1e-7 * (1 + abs(o["pobj"]))
    assert abs(r.dobj - o["dobj"]) <= 1e-7 * (1 + abs(o["dobj"]))
    assert np.max(np.abs(r.dimacs)) <= 1e-8
    # the returned point measured by the oracle's own DIMACS code
    err, po, do = oracle_dimacs(inst, _packX(r), r.y)
    assert np.max(np.abs(err)) <= 1e-8, err
    assert abs(po - r.pobj) <= 1e-9 * (1 + abs(po))
