"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Golden vectors come from the reference's own programs (tests/golden/make_golden.py:
op_ssspSeq and op_cE compiled from seqSSSPwithDist_undir.cpp and
createChangedEdges.cpp).  The rest are closed forms, brute force and planted
mistakes.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests._util import (GOLDEN, GOLDEN_N, INF, golden_case, random_batch, random_connected,
                         random_multigraph, read_graph, read_ref_sssp)


# --------------------------------------------------------------------------- toy

def test_toy_djk_matches_reference_fixture():
    """SSSP.txt is op_ssspSeq's output on graph.txt (README.md:39-48): parents too."""
    a, b, w = read_graph(os.path.join(GOLDEN, "toy_graph.txt"))
    par, dist, _ = oracle.djk(oracle.CSR.from_edges(6, a, b, w), 0)
    rp, rd = read_ref_sssp(os.path.join(GOLDEN, "toy_SSSP.txt"), 6)
    assert par.tolist() == rp.tolist() == [0, 0, 4, 1, 1, 2]
    assert dist.tolist() == rd.tolist() == [0, 2, 5, 4, 3, 8]


def test_toy_update_known_answer():
    """Appendix-B answer of the toy batch: dist [0,4,6,6,4,9]."""
    a, b, w, chg, n = golden_case("toy")
    ua, ub, uw = oracle.apply_changes(n, a, b, w, chg)
    # 0-1 w2 deleted; 0-4 w4 and 1-4 w0 appended after the file's lines
    assert list(zip(ua, ub, uw)) == [(0, 2, 6), (1, 3, 2), (1, 4, 1), (2, 5, 3), (4, 2, 2), (0, 4, 4), (1, 4, 0)]
    csr = oracle.CSR.from_edges(n, ua, ub, uw)
    par, dist, _ = oracle.djk(csr, 0)
    assert dist.tolist() == [0, 4, 6, 6, 4, 9]
    rp, rd = read_ref_sssp(os.path.join(GOLDEN, "toy_updated.txt"), n)
    assert par.tolist() == rp.tolist() and dist.tolist() == rd.tolist()
    # the reference's own update kernels, run sequentially, reach the same distances
    p0, d0 = read_ref_sssp(os.path.join(GOLDEN, "toy_SSSP.txt"), n)
    rpar, rdist = oracle.ref_update(n, a, b, w, p0, d0, chg)
    assert rdist.tolist() == [0, 4, 6, 6, 4, 9]
    assert oracle.certificate(csr, rdist, rpar)[0] == 0
    # canonical tree (smallest-id tight in-neighbour): 1<-4 via the w0 edge, 2<-0 (tie 0/4)
    assert oracle.canonical_parents(csr, dist, rule=0).tolist() == [0, 4, 0, 1, 0, 2]
    assert oracle.canonical_parents(csr, dist, rule=1).tolist() == [0, 4, 0, 1, 0, 2]


# ------------------------------------------------------------------ golden cases

@pytest.mark.parametrize("name", ["r300", "r2000", "sparse1500"])
def test_golden_djk_and_update(name):
    n = GOLDEN_N[name]
    a, b, w, chg, _ = golden_case(name)
    csr = oracle.CSR.from_edges(n, a, b, w)
    par, dist, _ = oracle.djk(csr, 0)
    gp, gd = read_ref_sssp(os.path.join(GOLDEN, f"{name}_SSSP.txt"), n)
    assert np.array_equal(dist, gd)
    assert np.array_equal(par, gp)  # same queue discipline => same parents

    ua, ub, uw = oracle.apply_changes(n, a, b, w, chg)
    ucsr = oracle.CSR.from_edges(n, ua, ub, uw)
    upar, udist, _ = oracle.djk(ucsr, 0)
    ep, ed = read_ref_sssp(os.path.join(GOLDEN, f"{name}_updated.txt"), n)
    assert np.array_equal(udist, ed)
    assert np.array_equal(upar, ep)

    # the reference's update rules reach Dijkstra's distances on the updated graph
    rpar, rdist = oracle.ref_update(n, a, b, w, gp, gd, chg)
    assert np.array_equal(rdist, ed)
    assert oracle.certificate(ucsr, rdist, rpar)[0] == 0


def test_apply_changes_multiset_arithmetic():
    rng = np.random.default_rng(7)
    for _ in range(20):
        n = int(rng.integers(2, 30))
        a, b, w = random_multigraph(rng, n, int(rng.integers(0, 60)), 0, 4)
        chg = random_batch(rng, n, a, b, w, int(rng.integers(0, 40)), 0.5, 0, 4)
        chg[rng.random(len(chg)) < 0.1, 3] = 7  # other inst values are ignored
        ua, ub, uw = oracle.apply_changes(n, a, b, w, chg)
        from collections import Counter
        edges = Counter()
        for x, y, z in zip(a, b, w):
            edges[(min(x, y), max(x, y), z)] += 1
        for x, y, z, op in chg:
            if op == 1:
                edges[(min(x, y), max(x, y), z)] += 1
        for x, y, z, op in chg:
            if op == 0:
                edges.pop((min(x, y), max(x, y), z), None)
        got = Counter((min(x, y), max(x, y), z) for x, y, z in zip(ua, ub, uw))
        assert got == +edges


# ------------------------------------------------------- cross-checks, closed forms

def test_path_star_grid_closed_forms():
    n = 50
    w = np.arange(1, n, dtype=np.int32)
    a = np.arange(0, n - 1, dtype=np.int32)
    csr = oracle.CSR.from_edges(n, a, a + 1, w)
    _, d, _ = oracle.djk(csr, 0, oracle.INF)
    assert d.tolist() == np.concatenate([[0], np.cumsum(w)]).tolist()
    # star with centre 3
    leaves = np.array([i for i in range(n) if i != 3], np.int32)
    csr = oracle.CSR.from_edges(n, np.full(n - 1, 3, np.int32), leaves, leaves + 1)
    _, d = oracle.dijkstra(csr, [0])
    exp = np.where(np.arange(n) == 3, 1, np.arange(n) + 1 + 1)
    exp[0] = 0
    assert d.tolist() == exp.tolist()
    # unit grid 7x9: distance = Manhattan distance to the corner
    R, C = 7, 9
    ea, eb = [], []
    for r in range(R):
        for c in range(C):
            v = r * C + c
            if c + 1 < C:
                ea.append(v); eb.append(v + 1)
            if r + 1 < R:
                ea.append(v); eb.append(v + C)
    csr = oracle.CSR.from_edges(R * C, ea, eb, np.ones(len(ea), np.int32))
    _, d = oracle.dijkstra(csr, [0])
    assert d.tolist() == [r + c for r in range(R) for c in range(C)]


def test_dijkstras_agree_with_bellman_ford():
    rng = np.random.default_rng(11)
    for t in range(40):
        n = int(rng.integers(1, 40))
        a, b, w = random_multigraph(rng, n, int(rng.integers(0, 100)), 0 if t % 2 else 1, 20)
        csr = oracle.CSR.from_edges(n, a, b, w)
        bf = oracle.bellman_ford(n, a, b, w, [0])
        _, d1, _ = oracle.djk(csr, 0, oracle.INF)
        p2, d2 = oracle.dijkstra(csr, [0])
        assert np.array_equal(d1, bf) and np.array_equal(d2, bf)
        assert oracle.certificate(csr, d2, p2)[0] == 0


def test_djk_cap_quirk():
    """Vertices at distance >= 1000*n are unreachable for the reference's Dijkstra."""
    csr = oracle.CSR.from_edges(4, [0, 1], [1, 2], [600000, 1])
    p, d, _ = oracle.djk(csr, 0)
    assert d.tolist() == [0, oracle.INF, oracle.INF, oracle.INF] and p.tolist() == [0, -1, -1, -1]
    p, d, _ = oracle.djk(csr, 0, oracle.INF)
    assert d.tolist() == [0, 600000, 600001, oracle.INF]


@pytest.mark.skipif(oracle.ref_sssp_binary() is None, reason="oracle/_ref not built")
def test_djk_against_reference_binary(tmp_path):
    rng = np.random.default_rng(3)
    for t in range(6):
        n = int(rng.integers(20, 400))
        a, b, w = random_connected(rng, n, int(n * rng.uniform(1.0, 4.0)), 0, 99)
        g = tmp_path / f"g{t}.txt"
        g.write_text("\n".join(f"{x} {y} {z}" for x, y, z in zip(a, b, w)))
        out = subprocess.check_output([oracle.ref_sssp_binary(), str(g), str(n)])
        s = tmp_path / f"s{t}.txt"
        s.write_bytes(out)
        rp, rd = read_ref_sssp(str(s), n)
        par, dist, _ = oracle.djk(oracle.CSR.from_edges(n, a, b, w), 0)
        assert np.array_equal(dist, rd) and np.array_equal(par, rp)


def test_ref_update_random_batches():
    rng = np.random.default_rng(5)
    for t in range(25):
        n = int(rng.integers(5, 120))
        a, b, w = random_connected(rng, n, int(n * rng.uniform(1.0, 3.0)), 1, 30)
        csr = oracle.CSR.from_edges(n, a, b, w)
        p0, d0, _ = oracle.djk(csr, 0)
        chg = random_batch(rng, n, a, b, w, int(rng.integers(1, 3 * n)), rng.uniform(0, 1), 1, 30)
        rpar, rdist = oracle.ref_update(n, a, b, w, p0, d0, chg)
        ua, ub, uw = oracle.apply_changes(n, a, b, w, chg)
        ucsr = oracle.CSR.from_edges(n, ua, ub, uw)
        _, ud = oracle.dijkstra(ucsr, [0])
        assert np.array_equal(rdist, ud), t


# ------------------------------------------------------- canonical tree, certificate

def test_canonical_parents_unique_paths_equal_dijkstra():
    """With weights from a huge range shortest paths are unique, so every rule must
    return Dijkstra's own tree."""
    rng = np.random.default_rng(13)
    for _ in range(20):
        n = int(rng.integers(2, 200))
        a, b, w = random_connected(rng, n, int(n * 2.5), 1, 1 << 24)
        csr = oracle.CSR.from_edges(n, a, b, w)
        p, d = oracle.dijkstra(csr, [0])
        assert np.array_equal(oracle.canonical_parents(csr, d, rule=0), p)
        assert np.array_equal(oracle.canonical_parents(csr, d, rule=1), p)


def test_canonical_parents_with_ties_and_zero_weights():
    rng = np.random.default_rng(17)
    for t in range(40):
        n = int(rng.integers(2, 60))
        a, b, w = random_multigraph(rng, n, int(rng.integers(n, 4 * n)), 0, 2)
        csr = oracle.CSR.from_edges(n, a, b, w)
        _, d = oracle.dijkstra(csr, [0])
        p1 = oracle.canonical_parents(csr, d, rule=1)
        assert oracle.certificate(csr, d, p1)[0] == 0  # acyclic even with zero weights
        p0 = oracle.canonical_parents(csr, d, rule=0)
        for v in range(n):  # brute force: smallest id among all tight in-neighbours
            if v == 0 or d[v] == INF:
                continue
            tight = [int(csr.col[j]) for j in range(csr.rp[v], csr.rp[v + 1])
                     if d[csr.col[j]] != INF and d[csr.col[j]] + csr.wt[j] == d[v]]
            assert p0[v] == min(tight)
    # a zero-weight pair where rule 0 is cyclic and rule 1 is not
    csr = oracle.CSR.from_edges(4, [0, 0, 1, 2], [3, 2, 2, 3], [5, 5, 0, 0])
    _, d = oracle.dijkstra(csr, [0])
    csr2 = oracle.CSR.from_edges(4, [0, 0, 1, 2, 3], [3, 2, 2, 3, 1], [5, 9, 0, 0, 0])
    _, d2 = oracle.dijkstra(csr2, [0])
    p0 = oracle.canonical_parents(csr2, d2, rule=0)
    p1 = oracle.canonical_parents(csr2, d2, rule=1)
    assert d2.tolist() == [0, 5, 5, 5]
    assert p0.tolist() == [0, 2, 1, 0]  # 1<->2 cycle
    assert oracle.certificate(csr2, d2, p0)[0] == 4
    assert oracle.certificate(csr2, d2, p1)[0] == 0 and p1.tolist() == [0, 3, 3, 0]


def test_certificate_rejects_planted_mistakes():
    rng = np.random.default_rng(19)
    a, b, w = random_connected(rng, 60, 150, 1, 9)
    a = np.append(a, [60, 61]).astype(np.int32)   # vertices 60,61 form their own component
    b = np.append(b, [61, 60]).astype(np.int32)
    w = np.append(w, [3, 4]).astype(np.int32)
    n = 62
    csr = oracle.CSR.from_edges(n, a, b, w)
    p, d = oracle.dijkstra(csr, [0])
    assert oracle.certificate(csr, d, p) == (0, -1)
    v = int(np.argmax(np.where(d == INF, -1, d)))  # a far vertex (a leaf of the tree)
    for delta, code in ((-1, 2), (+1, 3)):
        d2 = d.copy(); d2[v] += delta
        assert oracle.certificate(csr, d2, p)[0] in (2, 3), delta
    d2 = d.copy(); d2[0] = 1
    assert oracle.certificate(csr, d2, p)[0] == 1
    p2 = p.copy(); p2[v] = (p[v] + 1) % 60
    if p2[v] == v:
        p2[v] = (p2[v] + 1) % 60
    assert oracle.certificate(csr, d, p2)[0] == 3
    d3 = d.copy(); d3[60] = 5
    assert oracle.certificate(csr, d3, p)[0] in (2, 3)   # unreachable vertex claimed reachable
    p3 = p.copy(); p3[60] = 61
    assert oracle.certificate(csr, d, p3)[0] == 5
    # a zero-weight 2-cycle with consistent distances is caught by the walk
    csr0 = oracle.CSR.from_edges(3, [0, 1], [1, 2], [0, 0])
    assert oracle.certificate(csr0, [0, 0, 0], [0, 2, 1])[0] == 4


def test_unordered_csr_builder_matches_ordered():
    """The all-threads CSR builder (rows in an unspecified order) holds exactly
    the file-order builder's rows as multisets, and every order-independent
    check gives the same answers on it: Dijkstra distances, canonical parents,
    the certificate (also on planted mistakes)."""
    rng = np.random.default_rng(23)
    for trial in range(4):
        n = int(rng.integers(50, 3000))
        a, b, w = random_multigraph(rng, n, int(rng.integers(n, 8 * n)), 0 if trial % 2 else 1, 50)
        c1 = oracle.CSR.from_edges(n, a, b, w)
        c2 = oracle.CSR.from_edges(n, a, b, w, ordered=False)
        assert np.array_equal(c1.rp, c2.rp)
        for v in range(n):
            lo, hi = c1.rp[v], c1.rp[v + 1]
            r1 = sorted(zip(c1.col[lo:hi].tolist(), c1.wt[lo:hi].tolist()))
            r2 = sorted(zip(c2.col[lo:hi].tolist(), c2.wt[lo:hi].tolist()))
            assert r1 == r2, (trial, v)
        _, d1 = oracle.dijkstra(c1, [0])
        _, d2 = oracle.dijkstra(c2, [0])
        assert np.array_equal(d1, d2)
        for rule in (0, 1):
            p1 = oracle.canonical_parents(c1, d1, [0], rule=rule)
            assert np.array_equal(p1, oracle.canonical_parents(c2, d1, [0], rule=rule))
        p1 = oracle.canonical_parents(c1, d1, [0], rule=1)
        assert oracle.certificate(c1, d1, p1) == oracle.certificate(c2, d1, p1) == (0, -1)
        reach = np.flatnonzero((d1 != INF) & (np.arange(n) != 0))
        if len(reach):
            v = int(reach[-1])
            d3 = d1.copy()
            d3[v] += 1
            assert oracle.certificate(c1, d3, p1) == oracle.certificate(c2, d3, p1) != (0, -1)
    with pytest.raises(ValueError):
        oracle.CSR.from_edges(4, [0, 5], [1, 2], [1, 1], ordered=False)


def test_apply_changes_order_and_parallel_compaction():
    """The edge filter runs on all threads and compacts in the original order:
    survivors keep the graph's order, then the batch's insertions in batch
    order (checked against a direct Python restatement on a large input)."""
    rng = np.random.default_rng(29)
    n = 400
    a, b, w = random_multigraph(rng, n, 200_000, 1, 5)
    chg = random_batch(rng, n, a, b, w, 3000, 0.4, 1, 5)
    ua, ub, uw = oracle.apply_changes(n, a, b, w, chg)
    dels = {(min(x, y), max(x, y), z) for x, y, z, op in chg.tolist() if op == 0}
    keep = [(x, y, z) for x, y, z in zip(a.tolist(), b.tolist(), w.tolist())
            if (min(x, y), max(x, y), z) not in dels]
    keep += [(x, y, z) for x, y, z, op in chg.tolist() if op == 1 and (min(x, y), max(x, y), z) not in dels]
    assert list(zip(ua.tolist(), ub.tolist(), uw.tolist())) == keep
