"""Dense matrices from a problem's SDPA triplets, for checks against definitions: the
constraint matrices A_i and C = -F0 as dense blocks, random SPD blocks, packing."""
import math
import numpy as np


def dense_constraints(inst, scaled=True):
    """A_i (i = 1..m) and C = -F0 as lists of dense blocks (LP blocks as diagonal matrices);
    with scaled, each A_i divided by ||A_i||_F (the solver's row equilibration). Returns
    (A, C, block orders, LP flags, norms ||A_i||_F)."""
    a = inst.arrays()
    m, bs = a["m"], [abs(int(b)) for b in a["bs"]]
    A = [[np.zeros((n, n)) for n in bs] for _ in range(m + 1)]
    for t, k, i, j, v in zip(a["mat"], a["blk"], a["i"], a["j"], a["v"]):
        B = A[t][k - 1]
        B[i - 1, j - 1] += v
        if i != j:
            B[j - 1, i - 1] += v
    C = [-B for B in A[0]]
    out, norms = [], []
    for t in range(1, m + 1):
        nrm = math.sqrt(sum(float(np.sum(B * B)) for B in A[t]))
        norms.append(nrm)
        out.append([B / nrm if (scaled and nrm > 0) else B for B in A[t]])
    return out, C, bs, [int(b) < 0 for b in a["bs"]], np.array(norms)


def random_spd(n, rng, lp):
    if lp:
        return np.diag(rng.uniform(0.5, 2.0, n))
    G = rng.standard_normal((n, n))
    return G @ G.T / n + 0.5 * np.eye(n)


def random_sym(n, rng, lp):
    if lp:
        return np.diag(rng.standard_normal(n))
    G = rng.standard_normal((n, n))
    return (G + G.T) / 2


def pack(blocks, lp):
    return np.concatenate([np.diag(B) if l else B.reshape(-1, order="F") for B, l in zip(blocks, lp)])


def unpack(vec, bs, lp):
    out, o = [], 0
    for n, l in zip(bs, lp):
        if l:
            out.append(np.diag(vec[o:o + n]))
            o += n
        else:
            out.append(vec[o:o + n * n].reshape(n, n, order="F"))
            o += n * n
    return out
