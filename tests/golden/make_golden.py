"""Write the golden vectors under tests/golden/ from the reference's own programs.

Run in a container that has /root/reference (it builds oracle/_ref first):

    python tests/golden/make_golden.py

For each case it writes
  <case>_graph.txt     graph file (a b w), each undirected edge once
  <case>_SSSP.txt      op_ssspSeq <graph> <n>           (reference Dijkstra, source 0)
  <case>_cE.txt        change file (a b w inst)
  <case>_updated.txt   op_ssspSeq on the graph after the batch
The updated graph is formed here by plain multiset arithmetic, independent of
the oracle: insertions are appended, then each deletion (x,y,w) removes every
edge between x and y with weight w (all_structure_undir.cuh:106-117,
gpuFunctions_undir.cuh:200-213).  One case takes its change file from the
reference's own generator op_cE (createChangedEdges.cpp), the others from a
seeded numpy generator with the same record format.

The toy case uses the reference's graph.txt / SSSP.txt / cE.txt as they are.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def run_sssp(graph_path, n, out_path):
    out = subprocess.check_output([os.path.join(REF_BIN, "op_ssspSeq"), graph_path, str(n)])
    with open(out_path, "wb") as f:
        f.write(out)


def write_edges(path, a, b, w):
    with open(path, "w") as f:
        f.write("\n".join(f"{x} {y} {z}" for x, y, z in zip(a, b, w)))


def random_graph(rng, n, m, wmax, wmin=0):
    a, b = [], []
    seen = set()
    for i in range(1, n):  # random recursive spanning tree
        j = int(rng.integers(0, i))
        a.append(j); b.append(i); seen.add((j, i))
    while len(a) < m:
        x, y = (int(v) for v in rng.integers(0, n, 2))
        if x == y:
            continue
        key = (min(x, y), max(x, y))
        if key in seen:
            continue
        seen.add(key)
        a.append(x); b.append(y)
    w = rng.integers(wmin, wmax + 1, len(a))
    return np.array(a), np.array(b), w


def random_batch(rng, n, a, b, w, k, ins_frac, wmax):
    recs = []
    for _ in range(k):
        if rng.random() < ins_frac:
            x, y = (int(v) for v in rng.integers(0, n, 2))
            if x == y:
                y = (x + 1) % n
            recs.append((x, y, int(rng.integers(0, wmax + 1)), 1))
        else:
            e = int(rng.integers(0, len(a)))
            if rng.random() < 0.5:
                recs.append((int(a[e]), int(b[e]), int(w[e]), 0))
            else:
                recs.append((int(b[e]), int(a[e]), int(w[e]), 0))
    return recs


def updated_edges(a, b, w, recs):
    edges = [(int(x), int(y), int(z)) for x, y, z in zip(a, b, w)]
    edges += [(x, y, z) for x, y, z, op in recs if op == 1]
    dead = {(min(x, y), max(x, y), z) for x, y, z, op in recs if op == 0}
    return [e for e in edges if (min(e[0], e[1]), max(e[0], e[1]), e[2]) not in dead]


def read_cE(path):
    recs = []
    for line in open(path):
        t = line.split()
        if len(t) >= 4:
            recs.append(tuple(int(v) for v in t[:4]))
    return recs


def make_case(name, n, a, b, w, recs=None, ce_from_ref=None):
    gpath = os.path.join(HERE, f"{name}_graph.txt")
    write_edges(gpath, a, b, w)
    run_sssp(gpath, n, os.path.join(HERE, f"{name}_SSSP.txt"))
    cpath = os.path.join(HERE, f"{name}_cE.txt")
    if ce_from_ref is not None:
        k, pct = ce_from_ref
        out = subprocess.check_output([os.path.join(REF_BIN, "op_cE"), gpath, str(n), str(k), str(pct)])
        with open(cpath, "wb") as f:
            f.write(out)
        recs = read_cE(cpath)
    else:
        with open(cpath, "w") as f:
            f.write("\n".join(f"{x} {y} {z} {op}" for x, y, z, op in recs) + "\n")
    ue = updated_edges(a, b, w, recs)
    upath = os.path.join(HERE, f"{name}_upd_graph.tmp")
    write_edges(upath, [e[0] for e in ue], [e[1] for e in ue], [e[2] for e in ue])
    run_sssp(upath, n, os.path.join(HERE, f"{name}_updated.txt"))
    os.remove(upath)


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    # toy: the reference's own fixture trio
    ta, tb, tw = np.loadtxt(os.path.join(HERE, "toy_graph.txt"), dtype=np.int64, ndmin=2).T
    recs = read_cE(os.path.join(HERE, "toy_cE.txt"))
    ue = updated_edges(ta, tb, tw, recs)
    upath = os.path.join(HERE, "toy_upd_graph.tmp")
    write_edges(upath, [e[0] for e in ue], [e[1] for e in ue], [e[2] for e in ue])
    run_sssp(upath, 6, os.path.join(HERE, "toy_updated.txt"))
    os.remove(upath)

    rng = np.random.default_rng(20240611)
    # small, with zero weights and op_cE's own change file (weights rand()%100 incl. 0)
    a, b, w = random_graph(rng, 300, 1200, 99)
    make_case("r300", 300, a, b, w, ce_from_ref=(200, 50))
    # medium, positive weights, mixed batch
    a, b, w = random_graph(rng, 2000, 9000, 100, wmin=1)
    make_case("r2000", 2000, a, b, w, recs=random_batch(rng, 2000, a, b, w, 600, 0.5, 100))
    # sparse (long paths, many invalidations), deletion-heavy batch
    a, b, w = random_graph(rng, 1500, 1800, 50, wmin=1)
    make_case("sparse1500", 1500, a, b, w, recs=random_batch(rng, 1500, a, b, w, 300, 0.3, 50))
    print("golden vectors written to", HERE, file=sys.stderr)


if __name__ == "__main__":
    main()
