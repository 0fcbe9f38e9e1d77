"""GPU parity: the HIP update path (through the C ABI) against the CPU oracle.

Bar: distances bit-exact to Dijkstra on the updated graph; parents bit-exact to
the canonical tree (smallest-id tight in-neighbour; for graphs that ever held a
zero weight, the smallest-id tight in-neighbour one BFS level up); the updated
CSR equal, as a multiset, to the reference's graph-after-batch semantics.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle
from tests._util import (GOLDEN, GOLDEN_N, INF, golden_case, random_batch, random_connected,
                         random_multigraph, read_ref_sssp)

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def edge_multiset(a, b, w):
    a, b, w = np.asarray(a, np.int64), np.asarray(b, np.int64), np.asarray(w, np.int64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keys = (lo << 40) | (hi << 8) | 0
    order = np.lexsort((w, hi, lo))
    return np.stack([lo[order], hi[order], w[order]], 1)


class Mirror:
    """CPU copy of the graph driven through the oracle's change semantics."""

    def __init__(self, n, a, b, w):
        self.n = n
        self.a, self.b, self.w = (np.asarray(x, np.int32) for x in (a, b, w))
        self.zero = bool(np.any(self.w == 0))

    def apply(self, chg):
        chg = np.asarray(chg, np.int32).reshape(-1, 4)
        if np.any((chg[:, 3] == 1) & (chg[:, 2] == 0)):
            self.zero = True
        self.a, self.b, self.w = oracle.apply_changes(self.n, self.a, self.b, self.w, chg)

    def csr(self):
        return oracle.CSR.from_edges(self.n, self.a, self.b, self.w)


def check_tree(mirror, par, dist, srcs, what=""):
    csr = mirror.csr()
    _, dref = oracle.dijkstra(csr, srcs)
    bad = np.flatnonzero(dist != dref)
    assert len(bad) == 0, f"{what}: {len(bad)} distances differ, first {bad[:5]}: {dist[bad[:5]]} vs {dref[bad[:5]]}"
    pref = oracle.canonical_parents(csr, dref, srcs, rule=1 if mirror.zero else 0)
    badp = np.flatnonzero(par != pref)
    assert len(badp) == 0, f"{what}: {len(badp)} parents differ, first {badp[:5]}: {par[badp[:5]]} vs {pref[badp[:5]]}"
    assert oracle.certificate(csr, dist, par, srcs)[0] == 0


def check_graph(mirror, g):
    ga, gb, gw = g.export()
    assert np.array_equal(edge_multiset(ga, gb, gw), edge_multiset(mirror.a, mirror.b, mirror.w))


# ----------------------------------------------------------------------------- known answers

def test_toy_known_answer():
    say("toy known answer")
    a, b, w, chg, n = golden_case("toy")
    g = ds.Graph(n, a, b, w)
    p0, d0 = ds.read_sssp(os.path.join(GOLDEN, "toy_SSSP.txt"), n)
    t = ds.Tree(g, parent=p0, dist=d0)
    st = ds.update(g, t, chg)
    par, dist = t.read()
    assert dist.tolist() == [0, 4, 6, 6, 4, 9]
    # the batch inserts a zero-weight edge: BFS-level canonical parents, 1<-4 via w0
    assert par.tolist() == [0, 4, 0, 1, 0, 2]
    assert st["n_insert"] == 2 and st["n_delete"] == 1 and st["n_invalidated"] == 5
    m = Mirror(n, a, b, w)
    m.apply(chg)
    check_graph(m, g)
    check_tree(m, par, dist, [0], "toy")


@pytest.mark.parametrize("name", ["r300", "r2000", "sparse1500"])
def test_golden_reference_vectors(name):
    """Input tree = the reference Dijkstra's SSSP file; expected distances = the
    reference Dijkstra on the updated graph (tests/golden)."""
    say(f"golden {name}")
    n = GOLDEN_N[name]
    a, b, w, chg, _ = golden_case(name)
    g = ds.Graph(n, a, b, w)
    p0, d0 = ds.read_sssp(os.path.join(GOLDEN, f"{name}_SSSP.txt"), n)
    t = ds.Tree(g, parent=p0, dist=d0)
    ds.update(g, t, chg)
    par, dist = t.read()
    _, ed = read_ref_sssp(os.path.join(GOLDEN, f"{name}_updated.txt"), n)
    assert np.array_equal(dist, ed)
    m = Mirror(n, a, b, w)
    m.apply(chg)
    check_graph(m, g)
    csr = m.csr()
    assert oracle.certificate(csr, dist, par)[0] == 0
    if m.zero:  # every parent is re-derived: canonical even from a non-canonical input tree
        assert np.array_equal(par, oracle.canonical_parents(csr, dist, rule=1))
    # the reference's own update rules, run sequentially, reach the same distances
    _, rdist = oracle.ref_update(n, a, b, w, p0, d0, chg)
    assert np.array_equal(rdist, dist)


# ----------------------------------------------------------------------------- from scratch

@pytest.mark.parametrize("seed", range(6))
def test_tree_from_scratch(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 3000))
    if seed % 3 == 0:
        a, b, w = random_multigraph(rng, n, int(rng.integers(0, 6 * n)), 0, 5)  # zero weights, parallel, loops
    else:
        a, b, w = random_connected(rng, n, int(n * rng.uniform(1, 8)), 1, 1000)
    say(f"from scratch seed={seed} n={n} m={len(a)}")
    g = ds.Graph(n, a, b, w)
    src = int(rng.integers(0, n))
    t = ds.Tree(g, src)
    par, dist = t.read()
    check_tree(Mirror(n, a, b, w), par, dist, [src], f"scratch {seed}")


def test_canonicalize_given_tree():
    rng = np.random.default_rng(7)
    n = 2000
    a, b, w = random_connected(rng, n, 8000, 1, 5)  # many ties
    csr = oracle.CSR.from_edges(n, a, b, w)
    p, d, _ = oracle.djk(csr, 0, oracle.INF)
    say("canonicalize")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, parent=p, dist=d, canonical=True)
    par, dist = t.read()
    assert np.array_equal(dist, d)
    assert np.array_equal(par, oracle.canonical_parents(csr, d, rule=0))
    t2 = ds.Tree(g, parent=p, dist=d)
    assert np.array_equal(t2.read()[0], p)  # kept as given without the flag


# ----------------------------------------------------------------------------- updates

def run_sequence(n, a, b, w, batches, src=0, what=""):
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, src)
    m = Mirror(n, a, b, w)
    check_tree(m, *t.read(), [src], f"{what} initial")
    for i, chg in enumerate(batches):
        say(f"{what} batch {i} k={len(chg)}")
        st = ds.update(g, t, chg)
        m.apply(chg)
        par, dist = t.read()
        check_tree(m, par, dist, [src], f"{what} batch {i}")
        assert st["n_insert"] == int(np.sum(np.asarray(chg).reshape(-1, 4)[:, 3] == 1))
    check_graph(m, g)
    return g, t


@pytest.mark.parametrize("seed", range(8))
def test_random_update_sequences(seed):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(5, 4000))
    zero = seed % 4 == 3
    a, b, w = random_connected(rng, n, int(n * rng.uniform(1, 6)), 0 if zero else 1, 50)
    batches = []
    ca, cb, cw = a, b, w
    for _ in range(4):
        k = int(rng.integers(1, max(2, n // 2)))
        chg = random_batch(rng, n, ca, cb, cw, k, rng.uniform(0, 1), 0 if zero else 1, 50)
        batches.append(chg)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
    run_sequence(n, a, b, w, batches, src=int(rng.integers(0, n)), what=f"seq{seed}")


def test_edge_cases():
    n = 12
    # path 0-1-...-7, a triangle 8-9-10, vertex 11 isolated; parallel edge 2-3
    a = [0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 2]
    b = [1, 2, 3, 4, 5, 6, 7, 9, 10, 8, 3]
    w = [1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 5]
    batches = [
        np.zeros((0, 4), np.int32),                          # empty batch
        [[5, 9, 3, 7]],                                      # unknown op: ignored
        [[2, 3, 1, 0], [3, 2, 1, 0]],                        # duplicate deletion of a tree edge (parallel 2-3 w5 stays)
        [[2, 3, 4, 0]],                                      # deletion whose weight matches nothing
        [[2, 3, 5, 0]],                                      # last 2-3 copy: 3..7 disconnected
        [[7, 8, 1, 1], [4, 4, 0, 1]],                        # reconnect through the triangle side; self-loop
        [[4, 4, 0, 0], [0, 9, 3, 1], [0, 9, 3, 0]],          # delete self-loop; insert+delete same edge in one batch
        [[11, 0, 1, 1], [11, 0, 1, 1], [11, 0, 1, 0]],       # parallel inserts then deletion of both
        [[0, 1, 1, 0]],                                      # cut at the source: almost everything unreachable
        [[0, 11, 2, 1], [11, 1, 2, 1]],                      # reconnect via the isolated vertex
    ]
    g, t = run_sequence(n, a, b, w, [np.asarray(x, np.int32).reshape(-1, 4) for x in batches], 0, "edge")
    par, dist = t.read()
    assert dist[0] == 0 and par[0] == 0


def test_errors_leave_state_untouched():
    n = 5
    g = ds.Graph(n, [0, 1, 2], [1, 2, 3], [1, 1, 1])
    t = ds.Tree(g, 0)
    before = t.read()
    for bad, code in (([[0, 7, 1, 1]], ds.ERR_RANGE), ([[0, 1, -3, 0]], ds.ERR_WEIGHT), ([[-1, 1, 1, 0]], ds.ERR_RANGE)):
        with pytest.raises(ds.DssspError) as e:
            ds.update(g, t, bad)
        assert e.value.code == code
    after = t.read()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    with pytest.raises(ds.DssspError) as e:
        ds.Tree(g, parent=[0, 0, 1, 2, 9], dist=[0, 1, 2, 3, 4])
    assert e.value.code == ds.ERR_RANGE
    with pytest.raises(ds.DssspError) as e:
        ds.Tree(g, parent=[0, 1, 1, 2, -1], dist=[0, 5, 2, 3, 0])
    assert e.value.code == ds.ERR_TREE  # a root (parent itself) with nonzero distance


def test_large_host_batches_from_host_memory():
    """Host records of a large batch (checked on the host, staged in pinned
    memory, copied to the device): a valid batch gives the tree of the same
    batch uploaded first, and of several invalid records the first one is
    reported, by its index, with nothing changed."""
    a, b, w = ds.gen_rmat(13, 16, seed=51)
    n = 1 << 13
    chg = ds.gen_changes(n, a, b, w, 90_000, 0.5, 1, 1000, seed=52)
    assert len(chg) >= 1 << 16
    g1, g2 = ds.Graph(n, a, b, w), ds.Graph(n, a, b, w)
    t1, t2 = ds.Tree(g1, 0), ds.Tree(g2, 0)
    try:
        bt = ds.Batch(g1, chg)
        s1 = ds.update(g1, t1, bt)
        bt.close()
        s2 = ds.update(g2, t2, chg)
        assert (s1["n_insert"], s1["n_delete"]) == (s2["n_insert"], s2["n_delete"])
        assert s1["n_insert"] + s1["n_delete"] == len(chg)
        p1, d1 = t1.read()
        p2, d2 = t2.read()
        assert np.array_equal(p1, p2) and np.array_equal(d1, d2)
        m = Mirror(n, a, b, w)
        m.apply(chg)
        check_tree(m, p2, d2, [0], "large host batch")
        before = t2.read()
        for first, code, rec in ((70_001, ds.ERR_RANGE, [3, n + 5, 7, 1]), (66_000, ds.ERR_WEIGHT, [3, 4, -7, 0])):
            bad = np.array(chg, np.int32, copy=True)
            bad[first] = rec
            bad[first + 15_000] = [n, 1, 1, 0]  # a later invalid record, in another range
            with pytest.raises(ds.DssspError) as e:
                ds.update(g2, t2, bad)
            assert e.value.code == code
            assert f"change {first} " in str(e.value), str(e.value)
        after = t2.read()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    finally:
        for o in (t1, t2, g1, g2):
            o.close()


def test_delta_growth_past_the_folded_group_count():
    """A graph of more than 2048 x 256 bitmap words (33.5 M vertices): the
    pool offsets of grown delta blocks come from the scan launch, not from
    each workgroup's own sum of the group sizes.  Batches that grow thousands
    of rows at the top of the id range leave the graph and the tree equal to
    the oracle's."""
    n = (1 << 25) + (1 << 21)
    rng = np.random.default_rng(61)
    core = np.concatenate([np.arange(100_000), np.arange(n - 100_000, n)]).astype(np.int64)
    par = core[(rng.random(len(core) - 1) * np.arange(1, len(core))).astype(np.int64)]  # a random tree
    extra = rng.integers(0, len(core), (100_000, 2))
    a = np.concatenate([par, core[extra[:, 0]]]).astype(np.int32)
    b = np.concatenate([core[1:], core[extra[:, 1]]]).astype(np.int32)
    keep = a != b
    a, b = a[keep], b[keep]
    w = rng.integers(1, 101, len(a)).astype(np.int32)
    say(f"big-n graph n={n} m={len(a)}")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, 0)
    m = Mirror(n, a, b, w)
    try:
        for step in range(2):
            hot = core[len(core) - 5_000:]  # high ids: rows in bitmap groups past 2048
            ins_a = np.repeat(hot, 12)
            ins_b = core[rng.integers(0, len(core), len(ins_a))]
            ok = ins_a != ins_b
            ins = np.stack([ins_a[ok], ins_b[ok], rng.integers(1, 101, int(ok.sum())), np.ones(int(ok.sum()))], 1)
            cur = np.stack([m.a, m.b, m.w], 1)
            dele = cur[rng.choice(len(cur), 20_000, replace=False)]
            dele = np.concatenate([dele, np.zeros((len(dele), 1))], 1)
            chg = np.concatenate([ins, dele]).astype(np.int32)
            chg = chg[rng.permutation(len(chg))]
            say(f"big-n batch {step}: k={len(chg)}")
            ds.update(g, t, chg)
            m.apply(chg)
            assert np.array_equal(edge_multiset(*g.export()), edge_multiset(m.a, m.b, m.w))
            check_tree(m, *t.read(), [0], f"big-n step {step}")
    finally:
        t.close()
        g.close()


def test_overflow_is_reported():
    n = 4
    g = ds.Graph(n, [0, 1, 2], [1, 2, 3], [2_000_000_000] * 3)
    with pytest.raises(ds.DssspError) as e:
        ds.Tree(g, 0)
    assert e.value.code == ds.ERR_OVERFLOW


def test_large_distances_below_the_limit():
    """Distances up to 2^32-2 are exact; candidates past the limit that improve
    nothing (here 2 -> 1 at 6e9) are no overflow."""
    n = 4
    g = ds.Graph(n, [0, 1, 0], [1, 2, 3], [2_000_000_000, 2_000_000_000, 2**31 - 1])
    t = ds.Tree(g, 0)
    p, d = t.read()
    assert list(p) == [0, 0, 1, 0] and list(d) == [0, 2_000_000_000, 4_000_000_000, 2**31 - 1]
    ds.update(g, t, np.array([[3, 2, 2**31 - 1, 1]], dtype=np.int32))  # 2^32 - 2 via 3: not shorter
    p, d = t.read()
    assert list(d) == [0, 2_000_000_000, 4_000_000_000, 2**31 - 1] and list(p) == [0, 0, 1, 0]
    ds.update(g, t, np.array([[1, 2, 2_000_000_000, 0]], dtype=np.int32))  # now 2 hangs off 3
    p, d = t.read()
    assert list(p) == [0, 0, 3, 0] and d[2] == 2**32 - 2


def test_update_after_overflow_error():
    """An update that overflows reports it and leaves the device scratch clean:
    deleting the offending edge afterwards gives the exact tree again."""
    n = 4
    g = ds.Graph(n, [0], [1], [2_000_000_000])
    t = ds.Tree(g, 0)
    ds.update(g, t, np.array([[1, 2, 2_000_000_000, 1]], dtype=np.int32))
    with pytest.raises(ds.DssspError) as e:
        ds.update(g, t, np.array([[2, 3, 2_000_000_000, 1]], dtype=np.int32))
    assert e.value.code == ds.ERR_OVERFLOW
    ds.update(g, t, np.array([[2, 3, 2_000_000_000, 0]], dtype=np.int32))
    p, d = t.read()
    assert list(p) == [0, 0, 1, -1]
    assert list(d[:3]) == [0, 2_000_000_000, 4_000_000_000] and d[3] == ds.INF


def test_pool_compaction_and_map_rebuild():
    """Forty insertion-heavy batches on a 500-vertex ring: the delta pool is
    compacted and grown and the edge map rebuilt several times; every batch is
    checked."""
    rng = np.random.default_rng(5)
    n = 500
    a = np.arange(n, dtype=np.int32)
    b = ((a + 1) % n).astype(np.int32)
    w = rng.integers(1, 20, n).astype(np.int32)
    batches, ca, cb, cw = [], a, b, w
    for _ in range(40):
        chg = random_batch(rng, n, ca, cb, cw, 400, 0.8, 1, 20)
        batches.append(chg)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
    g, _ = run_sequence(n, a, b, w, batches, src=0, what="compaction")
    assert g.info()["device_bytes"] > 0


def test_long_rows_tombstones():
    """Deletions in rows longer than 64 entries (base and delta), including a key
    with 150 parallel copies that spans several 64-entry scans."""
    rng = np.random.default_rng(11)
    n = 1000
    a = [0] * (n - 1) + [0] * 150 + list(rng.integers(1, n, 3000))
    b = list(range(1, n)) + [5] * 150 + list(rng.integers(1, n, 3000))
    w = list(rng.integers(1, 50, n - 1)) + [7] * 150 + list(rng.integers(1, 50, 3000))
    ins = [[0, int(x), int(rng.integers(1, 50)), 1] for x in rng.integers(1, n, 300)]
    ins += [[0, 9, 3, 1]] * 70 + [[9, 0, 3, 1]] * 10  # 80 parallel copies in the delta rows of 0 and 9
    dels = [[0, 5, 7, 0], [5, 0, 7, 0], [0, 9, 3, 0]]
    dels += [[0, int(x), int(y), 0] for x, y in zip(range(1, n, 7), np.asarray(w[: n - 1])[::7])]
    dels += [[int(p), int(q), int(r), 0] for p, q, r, _ in ins[:120:3]]
    run_sequence(n, a, b, w, [np.array(ins, np.int32), np.array(dels, np.int32)], src=0, what="long-rows")


@pytest.mark.parametrize("ept", [1, 2, 4])
def test_heavy_rows_and_chunks(ept):
    """Tiny heavy threshold so that most rows go through the chunk path; every
    edges-in-flight variant of the frontier kernel."""
    old = ds.get_tuning()
    ds.set_tuning(heavy_degree=64, chunk=64, grid=64, edges_per_thread=ept)
    try:
        a, b, w = ds.gen_rmat(13, 16, wlo=1, whi=1000, seed=77)
        n = 1 << 13
        rng = np.random.default_rng(3)
        batches = []
        ca, cb, cw = a, b, w
        for ins in (0.0, 0.5, 1.0):
            chg = ds.gen_changes(n, ca, cb, cw, 3000, ins, 1, 1000, seed=int(rng.integers(1 << 30)))
            batches.append(chg)
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        deg = np.bincount(np.concatenate([a, b]), minlength=n)
        run_sequence(n, a, b, w, batches, src=int(np.argmax(deg)), what=f"rmat13-heavy-ept{ept}")
    finally:
        ds.set_tuning(**old)


@pytest.mark.parametrize("ept", [1, 4])
def test_budget_deferred_rows(ept):
    """Smallest tile budget and chunk (64 entries) under a heavy threshold no
    row reaches: most rows leave their tiles for the budget and are processed
    as items by the next launch.  Every frontier mode goes through that path:
    the edge-map build, the tree from scratch, invalidation, seeding,
    relaxation, and (zero weights) the parent repair."""
    old = ds.get_tuning()
    ds.set_tuning(heavy_degree=1 << 20, tile_entries=64, chunk=64, grid=64, edges_per_thread=ept)
    try:
        a, b, w = ds.gen_rmat(13, 16, wlo=1, whi=1000, seed=91)
        n = 1 << 13
        rng = np.random.default_rng(92)
        batches, ca, cb, cw = [], a, b, w
        for ins in (0.5, 0.0, 1.0):
            chg = ds.gen_changes(n, ca, cb, cw, 3000, ins, 1, 1000, seed=int(rng.integers(1 << 30)))
            batches.append(chg)
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        deg = np.bincount(np.concatenate([a, b]), minlength=n)
        run_sequence(n, a, b, w, batches, src=int(np.argmax(deg)), what=f"rmat13-budget-ept{ept}")
        za, zb, zw = random_multigraph(rng, 3000, 24000, 0, 4)
        zbatches, ca, cb, cw = [], za, zb, zw
        for _ in range(2):
            chg = random_batch(rng, 3000, ca, cb, cw, 800, 0.5, 0, 4)
            zbatches.append(chg)
            ca, cb, cw = oracle.apply_changes(3000, ca, cb, cw, chg)
        run_sequence(3000, za, zb, zw, zbatches, src=1, what=f"zero-weights-budget-ept{ept}")
    finally:
        ds.set_tuning(**old)


def test_config1_full_size():
    """n=100,000, m=1,000,000, w U[1,100], src 0, batch 10,000 (50/50), deletions
    from the non-spanning-tree edges; the oracle is Dijkstra on the updated graph."""
    n, m = 100_000, 1_000_000
    a, b, w = ds.gen_tree_plus_random(n, m, 1, 100, seed=1001)
    chg = ds.gen_changes(n, a, b, w, 10_000, 0.5, 1, 100, seed=1002, first_deletable=n - 1)
    say("config1 full size")
    g, t = run_sequence(n, a, b, w, [chg], 0, "config1")
    # and the reference's own update rules agree on this input
    csr = oracle.CSR.from_edges(n, a, b, w)
    p0, d0, _ = oracle.djk(csr, 0)
    _, rdist = oracle.ref_update(n, a, b, w, p0, d0, chg)
    assert np.array_equal(rdist, t.read()[1])


def test_grid_deep_updates():
    side = 300
    a, b, w = ds.gen_grid(side, 0.01, 1, 10000, seed=31)
    n = side * side
    rng = np.random.default_rng(4)
    chg = random_batch(rng, n, a, b, w, 2000, 0.3, 1, 10000)
    run_sequence(n, a, b, w, [chg], 0, "grid300")


def test_deterministic_bytes():
    a, b, w = ds.gen_rmat(12, 16, seed=8)
    n = 1 << 12
    chg = ds.gen_changes(n, a, b, w, 2000, 0.5, 1, 1000, seed=9)
    outs = []
    for _ in range(3):
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, 0)
        ds.update(g, t, chg)
        outs.append(t.read())
    for p, d in outs[1:]:
        assert np.array_equal(p, outs[0][0]) and np.array_equal(d, outs[0][1])


def test_resident_batch_and_copies():
    a, b, w = ds.gen_tree_plus_random(5000, 30000, 1, 100, seed=3)
    n = 5000
    chg = ds.gen_changes(n, a, b, w, 3000, 0.5, 1, 100, seed=4, first_deletable=n - 1)
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, 0)
    g0, t0 = g.clone(), ds.Tree(g, 0)
    t0.copy_from(t)
    batch = ds.Batch(g, chg)
    ds.update(g, t, batch)
    r1 = t.read()
    # restore and repeat: identical result
    g.copy_from(g0)
    t.copy_from(t0)
    ds.update(g, t, batch)
    r2 = t.read()
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    m = Mirror(n, a, b, w)
    m.apply(chg)
    check_tree(m, *r2, [0], "resident")


def test_cli_drop_in(tmp_path):
    out = tmp_path / "out.txt"
    cmd = [ds.CLI_PATH, os.path.join(GOLDEN, "toy_graph.txt"), "6", "6", os.path.join(GOLDEN, "toy_SSSP.txt"),
           os.path.join(GOLDEN, "toy_cE.txt"), "--out", str(out)]
    say("cli toy")
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "Total Time taken for SSSP update" in res.stdout
    p, d = ds.read_sssp(str(out), 6)
    assert d.tolist() == [0, 4, 6, 6, 4, 9]


def test_hipgraph_rounds_match_eager():
    """Round groups replayed from captured hipGraphs give the eager result on
    every update of a sequence (each update replays several captured groups,
    with host synchronisations and other kernels between the replays)."""
    a, b, w = ds.gen_rmat(12, 16, seed=31)
    n = 1 << 12
    rng = np.random.default_rng(31)
    batches, ca, cb, cw = [], a, b, w
    for i in range(6):
        c = ds.gen_changes(n, ca, cb, cw, 1500, [0.0, 0.5, 1.0][i % 3], 1, 1000, seed=int(rng.integers(1 << 30)))
        batches.append(c)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
    old = ds.get_tuning()
    results = {}
    try:
        for use_graph in (0, 1):
            ds.set_tuning(use_graph=use_graph, round_batch=2)
            say(f"hipgraph use_graph={use_graph}")
            g = ds.Graph(n, a, b, w)
            t = ds.Tree(g, 0)
            seq = [t.read()]
            for c in batches:
                ds.update(g, t, c)
                seq.append(t.read())
            results[use_graph] = seq
    finally:
        ds.set_tuning(**old)
    for (p0, d0), (p1, d1) in zip(results[0], results[1]):
        assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
    m = Mirror(n, a, b, w)
    for c in batches:
        m.apply(c)
    check_tree(m, *results[1][-1], [0], "hipgraph")


def test_phase_timing_changes_only_the_timers():
    """The events between an update's phases (phase_timing) fill the phase
    times and nothing else: the same batches with and without them give the
    same trees, the phase times are positive only with them and sum to at most
    the update's time, and the update's own time is measured either way."""
    a, b, w = ds.gen_rmat(12, 16, seed=41)
    n = 1 << 12
    rng = np.random.default_rng(41)
    batches, ca, cb, cw = [], a, b, w
    for i in range(4):
        c = ds.gen_changes(n, ca, cb, cw, 1500, [0.0, 0.5, 1.0, 0.5][i], 1, 1000, seed=int(rng.integers(1 << 30)))
        batches.append(c)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
    old = ds.get_tuning()
    phases = ("t_upkeep_ms", "t_insert_ms", "t_classify_ms", "t_invalidate_ms", "t_seed_ms", "t_relax_ms")
    results = {}
    try:
        for pt in (0, 1):
            ds.set_tuning(phase_timing=pt)
            say(f"phase_timing={pt}")
            g = ds.Graph(n, a, b, w)
            t = ds.Tree(g, 0)
            seq = []
            for i, c in enumerate(batches):
                st = ds.update(g, t, c)
                assert st["t_update_ms"] > 0
                if pt:
                    assert st["t_relax_ms"] > 0 and st["t_insert_ms"] >= 0
                    if i != 2:  # the batch of insertions only has no deletions to classify
                        assert st["t_classify_ms"] > 0 and st["t_invalidate_ms"] > 0
                    assert sum(st[f] for f in phases) <= st["t_update_ms"] * 1.001 + 1e-3
                else:
                    assert all(st[f] == 0 for f in phases)
                seq.append(t.read())
            results[pt] = seq
            t.close()
            g.close()
    finally:
        ds.set_tuning(**old)
    for (p0, d0), (p1, d1) in zip(results[0], results[1]):
        assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
    m = Mirror(n, a, b, w)
    for c in batches:
        m.apply(c)
    check_tree(m, *results[1][-1], [0], "phase timing")


def test_op_tools_sssp_matches_reference_dijkstra(tmp_path):
    """`op_tools sssp` (the op_ssspSeq role on the GPU) against the reference
    Dijkstra's own output on the golden graphs: same distances, exact integers."""
    tools = os.path.join(os.path.dirname(ds.LIB_PATH), "op_tools")
    for name in ("toy", "r2000"):
        n = GOLDEN_N[name]
        say(f"op_tools sssp {name}")
        out = tmp_path / f"{name}.txt"
        out.write_bytes(subprocess.check_output([tools, "sssp", os.path.join(GOLDEN, f"{name}_graph.txt"), str(n)],
                                                timeout=120))
        p, d = ds.read_sssp(str(out), n)
        _, ed = read_ref_sssp(os.path.join(GOLDEN, f"{name}_SSSP.txt"), n)
        assert np.array_equal(d, ed)


def test_autotune_keeps_graph_tree_and_results():
    """The autotuner leaves graph and tree as they were, returns a valid tuning,
    and the batches then give the oracle's tree."""
    old = ds.get_tuning()
    try:
        a, b, w = ds.gen_rmat(12, 16, wlo=1, whi=1000, seed=21)
        n = 1 << 12
        deg = np.bincount(np.concatenate([a, b]), minlength=n)
        src = int(np.argmax(deg))
        batches, ca, cb, cw = [], a, b, w
        for i in range(2):
            c = ds.gen_changes(n, ca, cb, cw, 1500, 0.5, 1, 1000, seed=30 + i)
            batches.append(c)
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
        say("autotune rmat12")
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, src)
        p0, d0 = t.read()
        e0 = edge_multiset(*g.export())
        res = [ds.Batch(g, c) for c in batches]
        tun, ms = ds.autotune(g, t, res)
        assert ms > 0 and tun["grid"] >= 1 and tun["edges_per_thread"] in (1, 2, 4)
        assert ds.get_tuning() == tun
        p1, d1 = t.read()
        assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
        assert np.array_equal(edge_multiset(*g.export()), e0)
        m = Mirror(n, a, b, w)
        for r, c in zip(res, batches):
            ds.update(g, t, r)
            m.apply(c)
        check_tree(m, *t.read(), [src], "after autotune")
        for r in res:
            r.close()
    finally:
        ds.set_tuning(**old)


@pytest.mark.parametrize("block", [128, 512])
def test_block_sizes_byte_equal(block):
    """The frontier kernels at 128 and 512 threads per workgroup give the bytes
    of the default 256: trees from scratch, update sequences, a forest, and the
    full edge pass."""
    a, b, w = ds.gen_rmat(12, 16, seed=61)
    n = 1 << 12
    src = int(np.argmax(np.bincount(np.concatenate([a, b]), minlength=n)))
    batches, ca, cb, cw = [], a, b, w
    for i in range(3):
        c = ds.gen_changes(n, ca, cb, cw, 1500, 0.5, 1, 1000, seed=62 + i)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
        batches.append(c)
    old = ds.get_tuning()

    def run():
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, src)
        outs = [np.stack(t.read())]
        for c in batches:
            ds.update(g, t, c)
            outs.append(np.stack(t.read()))
        st = t.relax_full()
        outs.append(np.stack(t.read()))
        t.close()
        g.close()
        g = ds.Graph(n, a, b, w)  # the forest on a graph of its own
        f = ds.Tree(g, sources=[src, 1, 2, 3, 4])
        ds.update(g, f, batches[0])
        outs += [np.stack(f.read(j)) for j in range(5)]
        f.close()
        g.close()
        return outs, st["relax_improved"]

    try:
        ref, imp0 = run()
        ds.set_tuning(block=block)
        say(f"block {block}")
        got, imp1 = run()
    finally:
        ds.set_tuning(**{k: old[k] for k in ("block",)})
    assert imp0 == 0 and imp1 == 0
    for x, y in zip(ref, got):
        assert np.array_equal(x, y)
    csr = oracle.CSR.from_edges(n, ca, cb, cw)
    _, dref = oracle.dijkstra(csr, [src])
    assert np.array_equal(got[3][1], dref)


def test_resident_csr_upkeep_bounded():
    """Forty-eight deletion-heavy batches: tombstoned entries are reclaimed on
    the device (repacks inside the updates), so the entries that row scans read
    stay within a fixed factor of the live ones instead of growing with every
    batch; the pool is compacted on the device as insertions outgrow it; every
    tree is exact and the resident CSR equals the oracle's graph."""
    rng = np.random.default_rng(91)
    n = 6000
    a, b, w = random_connected(rng, n, 60_000, 1, 100)
    src = 0
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, src)
    mirror = Mirror(n, a, b, w)
    peak = 0
    inserted = 0
    try:
        for step in range(48):
            # 60 % deletions of current edges, 40 % insertions of random pairs
            chg = random_batch(rng, n, mirror.a, mirror.b, mirror.w, 2500, 0.4, 1, 100)
            inserted += 2 * int(np.sum(chg[:, 3] == 1))
            st = ds.update(g, t, chg)
            mirror.apply(chg)
            u = g.upkeep()
            live = 2 * len(mirror.a)
            scanned = u["base_entries"] + u["delta_entries"]
            peak = max(peak, scanned / max(1, live))
            # repacked before a batch once tombstones pass 1/8 of the scanned
            # entries, so after any batch: scanned <= live + 1/8 scanned + the
            # batch's own deletions (2 copies each, parallel copies included)
            # the tombstone count is exact unless duplicate records of one
            # batch kill a copy together (then it counts it twice)
            assert u["dead_entries"] <= scanned
            assert scanned - u["dead_entries"] <= live <= scanned, (step, u, live)
            if step % 8 == 7:
                check_tree(mirror, *t.read(), [src], f"upkeep step {step}")
                check_graph(mirror, g)
        u = g.upkeep()
        scanned = u["base_entries"] + u["delta_entries"]
        say(f"upkeep: {u['repacks']} repacks, peak scanned/live {peak:.2f}, scanned {scanned} "
            f"vs {2 * len(a) + inserted} without reclamation")
        assert u["repacks"] >= 3
        assert peak < 1.5
        assert scanned < 0.6 * (2 * len(a) + inserted)
        g.repack()  # an explicit repack drops every tombstone
        u = g.upkeep()
        assert u["dead_entries"] == 0 and u["base_entries"] + u["delta_entries"] == 2 * len(mirror.a)
        chg = random_batch(rng, n, mirror.a, mirror.b, mirror.w, 2000, 0.5, 1, 100)
        ds.update(g, t, chg)
        mirror.apply(chg)
        check_tree(mirror, *t.read(), [src], "after explicit repack")
        check_graph(mirror, g)
    finally:
        t.close()
        g.close()


def test_round_loop_gives_up_cleanly():
    """A loop of rounds still running past give_up_ms is abandoned between
    round groups (every kernel has returned by then): the call fails with
    DSSSP_ERR_STATE, and the library, the graph and later trees work as before.
    Forced from the host with a 1 ms limit on a deep grid built from scratch."""
    side = 400
    a, b, w = ds.gen_grid(side, 0.01, 1, 10000, seed=33)
    n = side * side
    g = ds.Graph(n, a, b, w)
    old = ds.get_tuning()
    try:
        ds.set_tuning(give_up_ms=1, delta=-1, local_tiles=-1, round_batch=1)
        with pytest.raises(ds.DssspError) as e:
            ds.Tree(g, 0)
        assert e.value.code == ds.ERR_STATE
    finally:
        ds.set_tuning(**{k: old[k] for k in ("give_up_ms", "delta", "local_tiles", "round_batch")})
    t = ds.Tree(g, 0)
    chg = random_batch(np.random.default_rng(34), n, a, b, w, 500, 0.5, 1, 10000)
    ds.update(g, t, chg)
    mirror = Mirror(n, a, b, w)
    mirror.apply(chg)
    check_tree(mirror, *t.read(), [0], "after a give-up")
    t.close()
    g.close()
