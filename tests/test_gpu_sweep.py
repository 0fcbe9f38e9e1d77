"""GPU parity of full-sweep relaxation rounds (tuning `sweep`, the
push-vs-full-sweep threshold) and of dsssp_tree_relax_full.

A sweep round relaxes every reachable slot instead of the marked ones.  Extra
relaxations only offer candidates that are lengths of real paths, and the packed
atomicMin keeps the smallest (distance, parent), so every result must be
byte-identical to the push-only path and to the oracle: random update
sequences (zero weights included), hub rows split into pieces, forests of
trees (the 64-tree layout never sweeps), and trees built from scratch.
"""
import contextlib
import sys

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle
from tests._util import random_batch, random_connected, random_multigraph
from tests.test_gpu_update import Mirror, check_tree, run_sequence

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@contextlib.contextmanager
def tuned(**kw):
    old = ds.get_tuning()
    ds.set_tuning(**kw)
    try:
        yield
    finally:
        ds.set_tuning(**old)


@pytest.mark.parametrize("seed", range(4))
def test_sweep_update_sequences(seed):
    """sweep=1 (per mille): nearly every round after a marking round is a full
    sweep; results must match the oracle exactly."""
    rng = np.random.default_rng(1300 + seed)
    n = int(rng.integers(50, 4000))
    zero = seed == 3
    a, b, w = random_connected(rng, n, int(n * rng.uniform(1, 6)), 0 if zero else 1, 1000)
    batches, ca, cb, cw = [], a, b, w
    for _ in range(3):
        k = int(rng.integers(1, max(2, n // 2)))
        chg = random_batch(rng, n, ca, cb, cw, k, rng.uniform(0, 1), 0 if zero else 1, 1000)
        batches.append(chg)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
    # odd seeds: no in-launch continuation, so every improvement is a mark
    with tuned(sweep=1, delta=-1, local_tiles=-1 if seed % 2 else 2):
        run_sequence(n, a, b, w, batches, src=int(rng.integers(0, n)), what=f"sweep seq{seed}")


def test_sweep_equals_push_on_rmat_with_split_rows():
    """R-MAT hub rows split into pieces (heavy 64, chunk 64, small grid): the
    sweep path (without in-launch continuation, so that every improvement is a
    mark) is byte-equal to the push path, from scratch and after updates, and
    it did sweep."""
    a, b, w = ds.gen_rmat(13, 16, wlo=1, whi=1000, seed=21)
    n = 1 << 13
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    src = int(np.argmax(deg))
    chgs = [ds.gen_changes(n, a, b, w, 4000, ins, 1, 1000, seed=22 + i) for i, ins in enumerate((0.5, 0.0, 1.0))]

    def run(**kw):
        with tuned(delta=-1, heavy_degree=64, chunk=64, grid=64, **kw):
            g = ds.Graph(n, a, b, w)
            t = ds.Tree(g, src)
            out = [t.read()]
            sweeps = 0
            for c in chgs:
                # each batch is generated on the original graph: deletions of
                # edges already gone are no-ops, as in the reference
                st = ds.update(g, t, c)
                sweeps += st["relax_sweeps"]
                out.append(t.read())
            return out, sweeps

    say("sweep vs push rmat13")
    push, s0 = run(sweep=-1)
    full, s1 = run(sweep=1, local_tiles=-1)
    assert s0 == 0 and s1 > 0, (s0, s1)
    for i, ((p0, d0), (p1, d1)) in enumerate(zip(push, full)):
        assert np.array_equal(d0, d1) and np.array_equal(p0, p1), i


def test_sweep_rounds_from_captured_graphs():
    """Round groups with the sweep threshold set, replayed from captured
    hipGraphs (the sweep choice is made on the device from the previous
    round's mark counts), equal the eager push-only result on every update of
    a sequence."""
    a, b, w = ds.gen_rmat(12, 16, seed=41)
    n = 1 << 12
    rng = np.random.default_rng(41)
    batches, ca, cb, cw = [], a, b, w
    for i in range(5):
        c = ds.gen_changes(n, ca, cb, cw, 1500, [0.0, 0.5, 1.0][i % 3], 1, 1000, seed=int(rng.integers(1 << 30)))
        batches.append(c)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
    results, sweeps = {}, 0
    for label, kw in (("eager push", dict(sweep=-1, use_graph=0)),
                      ("graph sweep", dict(sweep=1, use_graph=1, round_batch=2, local_tiles=-1))):
        say(f"sweep rounds {label}")
        with tuned(delta=-1, **kw):
            g = ds.Graph(n, a, b, w)
            t = ds.Tree(g, 0)
            seq = [t.read()]
            for c in batches:
                st = ds.update(g, t, c)
                if kw["sweep"] > 0:
                    sweeps += st["relax_sweeps"]
                seq.append(t.read())
            results[label] = seq
    assert sweeps > 0
    for (p0, d0), (p1, d1) in zip(results["eager push"], results["graph sweep"]):
        assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
    m = Mirror(n, a, b, w)
    for c in batches:
        m.apply(c)
    check_tree(m, *results["graph sweep"][-1], [0], "graph sweep")


@pytest.mark.parametrize("K", [3, 8, 64])
def test_sweep_forest(K):
    """Forests: K' = 4, 8 sweep over slots vertex*K'+tree; K = 64 (tree-vector
    path) never sweeps.  Every tree against Dijkstra and canonical parents."""
    rng = np.random.default_rng(1400 + K)
    a, b, w = ds.gen_rmat(11, 16, wlo=1, whi=1000, seed=K + 3)
    n = 1 << 11
    srcs = rng.integers(0, n, K).astype(np.int32)
    c = ds.gen_changes(n, a, b, w, 1500, 0.5, 1, 1000, seed=int(rng.integers(1 << 30)))
    ua, ub, uw = oracle.apply_changes(n, a, b, w, c)
    say(f"sweep forest K={K}")
    with tuned(sweep=1, delta=-1):
        g = ds.Graph(n, a, b, w)
        f = ds.Tree(g, sources=srcs)
        st = ds.update(g, f, c)
    if K == 64:
        assert st["relax_sweeps"] == 0
    csr = oracle.CSR.from_edges(n, ua, ub, uw)
    for j in range(K):
        par, dist = f.read(j)
        _, dref = oracle.dijkstra(csr, [int(srcs[j])])
        assert np.array_equal(dist, dref), j
        assert np.array_equal(par, oracle.canonical_parents(csr, dref, [int(srcs[j])], rule=0)), j


def test_relax_full_on_converged_tree_changes_nothing():
    rng = np.random.default_rng(1500)
    n = 3000
    a, b, w = random_connected(rng, n, 12000, 1, 100)
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, 5)
    chg = random_batch(rng, n, a, b, w, 800, 0.5, 1, 100)
    ds.update(g, t, chg)
    before = t.read()
    say("relax_full converged")
    st = t.relax_full()
    after = t.read()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert st["relax_improved"] == 0, st
    assert st["frontier_vertices"] == n and st["relax_sweeps"] >= 1, st
    assert st["t_sweep_ms"] > 0 and st["alg_bytes_sweep"] > 0, st
    # the rows of the reachable vertices were scanned
    m = Mirror(n, a, b, w)
    m.apply(chg)
    assert st["relax_edges"] >= len(m.a), st
    # and updates still work afterwards
    chg2 = random_batch(rng, n, m.a, m.b, m.w, 500, 0.5, 1, 100)
    ds.update(g, t, chg2)
    m.apply(chg2)
    check_tree(m, *t.read(), [5], "after relax_full")


def test_relax_full_multigraph_and_forests():
    """The edge pass on a multigraph (zero weights, parallel edges, self-loops,
    several components) and on forests of 3 trees (slots vertex*4+tree) and 64
    trees (the push path): nothing changes on converged trees, which match the
    oracle.  With zero weights the packed minimum also moves equal-distance
    parents to smaller ids, which the parent repair undoes, so there the tree,
    not the improvement count, is what stays the same."""
    rng = np.random.default_rng(1700)
    n = 2000
    a, b, w = random_multigraph(rng, n, 12000, 0, 5)
    say("relax_full multigraph")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, 4)
    chg = random_batch(rng, n, a, b, w, 600, 0.5, 0, 5)
    ds.update(g, t, chg)
    before = t.read()
    t.relax_full()
    after = t.read()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    m = Mirror(n, a, b, w)
    m.apply(chg)
    check_tree(m, *after, [4], "relax_full multigraph")
    ra, rb, rw = ds.gen_rmat(11, 16, wlo=1, whi=1000, seed=1701)
    rn = 1 << 11
    for K in (3, 64):
        say(f"relax_full forest K={K}")
        srcs = rng.integers(0, rn, K).astype(np.int32)
        gf = ds.Graph(rn, ra, rb, rw)
        f = ds.Tree(gf, sources=srcs)
        st = f.relax_full()
        assert st["relax_improved"] == 0, (K, st)
        csr = oracle.CSR.from_edges(rn, ra, rb, rw)
        for j in range(K):
            par, dist = f.read(j)
            _, dref = oracle.dijkstra(csr, [int(srcs[j])])
            assert np.array_equal(dist, dref), (K, j)


@pytest.mark.parametrize("zero", [False, True])
def test_relax_full_converges_upper_bounds_zero_weights(zero):
    """Upper bounds from a spanning subgraph on a multigraph-like graph (zero
    weights allowed): the pass plus rounds (and the parent repair) give the
    oracle's distances and canonical parents."""
    rng = np.random.default_rng(1800 + zero)
    n = 1500
    a, b, w = random_connected(rng, n, 7000, 0 if zero else 1, 50)
    src = 3
    sub = np.arange(len(a)) < n - 1
    csr_sub = oracle.CSR.from_edges(n, a[sub], b[sub], w[sub])
    _, dsub = oracle.dijkstra(csr_sub, [src])
    psub = oracle.canonical_parents(csr_sub, dsub, [src], rule=1 if zero else 0)
    say(f"relax_full upper bounds zero={zero}")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, src, parent=psub, dist=dsub)
    st = t.relax_full()
    assert st["relax_improved"] > 0
    par, dist = t.read()
    m = Mirror(n, a, b, w)
    if zero:
        check_tree(m, par, dist, [src], "relax_full upper bounds zero")  # repaired: canonical parents
    else:
        csr = m.csr()
        _, dref = oracle.dijkstra(csr, [src])
        assert np.array_equal(dist, dref)
        assert oracle.certificate(csr, dist, par, [src])[0] == 0


def test_relax_full_ordered_on_a_deep_grid():
    """From a tree holding only its root, relax_full on a deep grid with a
    distance step scans fewer entries than without one (its rounds start from
    a threshold at distance 0), with the same exact result."""
    side = 200
    a, b, w = ds.gen_grid(side, 0.01, 1, 10000, seed=19)
    n = side * side
    par = np.full(n, -1, np.int32)
    dist = np.full(n, ds.INF, np.int64)
    par[0], dist[0] = 0, 0
    out = {}
    for delta in (-1, 5000):
        say(f"relax_full grid delta={delta}")
        with tuned(delta=delta):
            g = ds.Graph(n, a, b, w)
            t = ds.Tree(g, 0, parent=par, dist=dist)
            st = t.relax_full()
            out[delta] = (t.read(), st["relax_edges"])
    (p0, d0), e0 = out[-1]
    (p1, d1), e1 = out[5000]
    assert np.array_equal(d0, d1) and np.array_equal(p0, p1)
    assert e1 < e0, (e1, e0)
    check_tree(Mirror(n, a, b, w), p1, d1, [0], "relax_full grid")


@pytest.mark.parametrize("sweep", [-1, 1])
def test_relax_full_converges_upper_bounds(sweep):
    """A given tree whose distances are lengths of real paths but not shortest
    (the shortest paths of a spanning subgraph) converges to Dijkstra on the
    whole graph, with a valid tree."""
    rng = np.random.default_rng(1600 + sweep)
    n = 2500
    a, b, w = random_connected(rng, n, 10000, 1, 1000)
    src = 7
    sub = np.arange(len(a)) < n - 1  # the random spanning tree comes first
    csr_sub = oracle.CSR.from_edges(n, a[sub], b[sub], w[sub])
    _, dsub = oracle.dijkstra(csr_sub, [src])
    psub = oracle.canonical_parents(csr_sub, dsub, [src], rule=0)
    assert np.all(dsub < ds.INF)
    say(f"relax_full upper bounds sweep={sweep}")
    with tuned(sweep=sweep, delta=-1):
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, src, parent=psub, dist=dsub)
        st = t.relax_full()
    par, dist = t.read()
    csr = oracle.CSR.from_edges(n, a, b, w)
    _, dref = oracle.dijkstra(csr, [src])
    assert np.array_equal(dist, dref)
    assert st["relax_improved"] > 0
    assert oracle.certificate(csr, dist, par, [src])[0] == 0


def test_edge_pass_pieces_of_long_rows():
    """Rows with more entries below their own id than the pass's piece size
    (1024) go to the pass's item launch in pieces: high-id vertices joined to
    2500-5000 lower-id vertices.  A tree built from scratch (which starts with
    edge passes) and a full pass over an upper-bound tree both give the
    oracle's distances and canonical parents."""
    rng = np.random.default_rng(2100)
    n = 12000
    a, b, w = random_connected(rng, n, 30000, 1, 100)
    hubs = [n - 1, n - 2, n - 7]
    ea, eb = [a], [b]
    for h, cnt in zip(hubs, (5000, 3000, 2500)):
        nb = rng.choice(h, cnt, replace=False).astype(np.int32)  # all below the hub's id
        ea.append(np.full(cnt, h, np.int32))
        eb.append(nb)
    a = np.concatenate(ea)
    b = np.concatenate(eb)
    w = np.concatenate([w, rng.integers(1, 101, len(a) - len(w)).astype(np.int32)])
    m = Mirror(n, a, b, w)
    say("edge pass pieces: tree from scratch")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, 0)
    check_tree(m, *t.read(), [0], "from scratch with long rows")
    say("edge pass pieces: relax_full from upper bounds")
    sub = np.arange(len(a)) < n - 1  # the spanning tree of random_connected
    csr_sub = oracle.CSR.from_edges(n, a[sub], b[sub], w[sub])
    _, dsub = oracle.dijkstra(csr_sub, [0])
    psub = oracle.canonical_parents(csr_sub, dsub, [0], rule=0)
    t2 = ds.Tree(g, 0, parent=psub, dist=dsub)
    st = t2.relax_full()
    assert st["relax_improved"] > 0
    check_tree(m, *t2.read(), [0], "relax_full with long rows")
