"""Test helpers: tiny text readers for the fixtures and seeded random inputs.

The readers here are deliberately plain Python so that they stay independent of
the product's C++ readers, which the format tests compare against them.
"""
import os

import numpy as np

INF = np.iinfo(np.int64).max
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_graph(path):
    """Graph file: 'a b w' per line (README.md:22-27)."""
    rows = [line.split() for line in open(path) if line.strip()]
    a = np.array([int(r[0]) for r in rows], np.int32)
    b = np.array([int(r[1]) for r in rows], np.int32)
    w = np.array([int(r[2]) for r in rows], np.int32)
    return a, b, w


def read_changes(path):
    """Change file: 'a b w inst' per line (README.md:33-37).  Returns int32 (k,4)."""
    rows = [line.split() for line in open(path) if line.strip()]
    return np.array([[int(t) for t in r[:4]] for r in rows], np.int32).reshape(-1, 4)


def read_ref_sssp(path, n):
    """op_ssspSeq output 'v parent dist ' (seqSSSPwithDist_undir.cpp:170-173).

    dist is printed as a double with 6 significant digits; parent -1 or
    dist >= 1000*n (the reference's cap) means unreachable."""
    par = np.full(n, -1, np.int32)
    dist = np.full(n, INF, np.int64)
    for line in open(path):
        t = line.split()
        if len(t) < 3:
            continue
        v, p, d = int(t[0]), int(t[1]), float(t[2])
        if p == -1 or d >= 1000 * n:
            continue
        par[v] = p
        dist[v] = int(round(d))
    return par, dist


def golden_case(name):
    a, b, w = read_graph(os.path.join(GOLDEN, f"{name}_graph.txt"))
    chg = read_changes(os.path.join(GOLDEN, f"{name}_cE.txt"))
    n = int(max(a.max(), b.max(), chg[:, :2].max() if len(chg) else 0)) + 1
    return a, b, w, chg, n


GOLDEN_N = {"toy": 6, "r300": 300, "r2000": 2000, "sparse1500": 1500}


def random_connected(rng, n, m, wlo, whi):
    """Random recursive spanning tree + uniform distinct extra pairs (m is capped
    at the number of vertex pairs)."""
    m = min(m, n * (n - 1) // 2)
    a = [int(rng.integers(0, i)) for i in range(1, n)]
    b = list(range(1, n))
    seen = {(min(x, y), max(x, y)) for x, y in zip(a, b)}
    while len(a) < m:
        x, y = (int(v) for v in rng.integers(0, n, 2))
        if x == y or (min(x, y), max(x, y)) in seen:
            continue
        seen.add((min(x, y), max(x, y)))
        a.append(x)
        b.append(y)
    w = rng.integers(wlo, whi + 1, len(a)).astype(np.int32)
    return np.array(a, np.int32), np.array(b, np.int32), w


def random_multigraph(rng, n, m, wlo, whi):
    """Random edges with parallel copies, self-loops and possibly several components."""
    a = rng.integers(0, n, m).astype(np.int32)
    b = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(wlo, whi + 1, m).astype(np.int32)
    return a, b, w


def random_batch(rng, n, a, b, w, k, ins_frac, wlo, whi):
    """Mixed batch: deletions copy existing edges (either orientation, repeats allowed),
    insertions are random pairs.  Returns int32 (k,4)."""
    out = np.zeros((k, 4), np.int32)
    for i in range(k):
        if len(a) == 0 or rng.random() < ins_frac:
            x, y = (int(v) for v in rng.integers(0, n, 2))
            out[i] = (x, y, int(rng.integers(wlo, whi + 1)), 1)
        else:
            e = int(rng.integers(0, len(a)))
            if rng.random() < 0.5:
                out[i] = (a[e], b[e], w[e], 0)
            else:
                out[i] = (b[e], a[e], w[e], 0)
    return out
