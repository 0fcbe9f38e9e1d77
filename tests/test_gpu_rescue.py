"""GPU parity of the invalidation rescue (through the C ABI).

A vertex about to be reset by the invalidation (a child endpoint of a deleted
tree edge, or a child found under a reset vertex) whose row is long enough is
first offered to a rescue stage: if another live neighbour that is tight for
its distance is still valid, the vertex keeps its distance under the
smallest-id such neighbour and its subtree is left alone.  The final tree must
be exactly the one without rescue: distances bit-exact to Dijkstra, parents
equal to the canonical tree, and the bytes equal to the same updates with the
rescue switched off.  A rescue degree of 1 offers every candidate (the most
stages and commits); larger ones offer only long rows.
"""
import sys

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle
from tests._util import random_batch, random_connected

pytestmark = pytest.mark.gpu

OFF = -1


def say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(autouse=True)
def restore_tuning():
    saved = ds.get_tuning()
    yield
    ds.set_tuning(**saved)


def exact(n, a, b, w, par, dist, srcs, what):
    csr = oracle.CSR.from_edges(n, a, b, w)
    _, dref = oracle.dijkstra(csr, srcs)
    bad = np.flatnonzero(dist != dref)
    assert len(bad) == 0, f"{what}: {len(bad)} distances differ from Dijkstra, first {bad[:5]}"
    canon = oracle.canonical_parents(csr, dref, srcs, rule=0)
    badp = np.flatnonzero(par != canon)
    assert len(badp) == 0, f"{what}: {len(badp)} parents not canonical, first {badp[:5]}"
    assert oracle.certificate(csr, dist, par, srcs)[0] == 0, what


def deletion_heavy(rng, n, a, b, w, k, whi):
    return random_batch(rng, n, a, b, w, k, 0.2, 1, whi)


def run_sequence(n, a, b, w, srcs, batches, tuning, forest=False):
    """Applies the batches in turn; returns the trees read after each batch and
    the summed rescue count."""
    ds.set_tuning(**tuning)
    g = ds.Graph(n, a, b, w)
    if forest:
        t = ds.Tree(g, sources=srcs)
    else:
        t = ds.Tree(g, srcs[0]) if len(srcs) == 1 else ds.Tree(g, roots=srcs)
    out, rescued = [], 0
    for chg in batches:
        st = ds.update(g, t, chg)
        rescued += st["n_rescued"]
        if forest:
            out.append([t.read(j) for j in range(len(srcs))])
        else:
            out.append(t.read())
    t.close()
    g.close()
    return out, rescued


def make_batches(rng, n, a, b, w, count, k, whi):
    ca, cb, cw = a, b, w
    out = []
    for _ in range(count):
        chg = deletion_heavy(rng, n, ca, cb, cw, k, whi)
        out.append(chg)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
    return out


@pytest.mark.parametrize("deg,levels", [(1, 1), (1, 2), (1, 64), (4, 3), (32, 64)])
def test_rescue_rmat_equals_plain(deg, levels):
    say(f"rescue rmat deg={deg} levels={levels}")
    a, b, w = ds.gen_rmat(12, 16, (0.57, 0.19, 0.19), 1, 20, seed=11)
    n = 1 << 12
    src = int(np.argmax(np.bincount(np.concatenate([a, b]), minlength=n)))
    rng = np.random.default_rng(5)
    batches = make_batches(rng, n, a, b, w, 4, 3000, 20)
    plain, r0 = run_sequence(n, a, b, w, [src], batches, {"rescue_degree": OFF})
    res, r1 = run_sequence(n, a, b, w, [src], batches, {"rescue_degree": deg, "rescue_levels": levels})
    assert r0 == 0
    assert r1 > 0, "narrow weights on a hub-heavy graph must leave some candidate a second tight parent"
    ca, cb, cw = a, b, w
    for i, chg in enumerate(batches):
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        assert np.array_equal(plain[i][0], res[i][0]) and np.array_equal(plain[i][1], res[i][1]), f"batch {i}"
        exact(n, ca, cb, cw, res[i][0], res[i][1], [src], f"batch {i}")


def test_rescue_connected_random_sequences():
    rng = np.random.default_rng(21)
    for case in range(6):
        n = int(rng.integers(50, 600))
        m = int(rng.integers(n, 6 * n))
        whi = int(rng.choice([3, 10, 1000]))
        a, b, w = random_connected(rng, n, m, 1, whi)
        batches = make_batches(rng, n, a, b, w, 5, int(rng.integers(5, m // 3 + 6)), whi)
        say(f"rescue random case {case} n={n} m={m} whi={whi}")
        plain, _ = run_sequence(n, a, b, w, [0], batches, {"rescue_degree": OFF})
        res, _ = run_sequence(n, a, b, w, [0], batches, {"rescue_degree": 1, "rescue_levels": 1024})
        ca, cb, cw = a, b, w
        for i, chg in enumerate(batches):
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
            assert np.array_equal(plain[i][0], res[i][0]) and np.array_equal(plain[i][1], res[i][1]), (case, i)
            exact(n, ca, cb, cw, res[i][0], res[i][1], [0], f"case {case} batch {i}")


def test_rescue_forest_and_roots():
    say("rescue forest K=8 and a many-root tree")
    a, b, w = ds.gen_rmat(11, 8, (0.57, 0.19, 0.19), 1, 8, seed=4)
    n = 1 << 11
    rng = np.random.default_rng(8)
    srcs = [int(x) for x in rng.choice(n, 8, replace=False)]
    batches = make_batches(rng, n, a, b, w, 3, 1500, 8)
    plain, _ = run_sequence(n, a, b, w, srcs, batches, {"rescue_degree": OFF}, forest=True)
    res, r1 = run_sequence(n, a, b, w, srcs, batches, {"rescue_degree": 1, "rescue_levels": 16}, forest=True)
    assert r1 > 0
    ca, cb, cw = a, b, w
    for i, chg in enumerate(batches):
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        for j, s in enumerate(srcs):
            assert np.array_equal(plain[i][j][0], res[i][j][0]) and np.array_equal(plain[i][j][1], res[i][j][1])
            exact(n, ca, cb, cw, res[i][j][0], res[i][j][1], [s], f"batch {i} tree {j}")
    roots = srcs[:3]
    plain, _ = run_sequence(n, a, b, w, roots, batches, {"rescue_degree": OFF})
    res, _ = run_sequence(n, a, b, w, roots, batches, {"rescue_degree": 1, "rescue_levels": 16})
    ca, cb, cw = a, b, w
    for i, chg in enumerate(batches):
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        assert np.array_equal(plain[i][0], res[i][0]) and np.array_equal(plain[i][1], res[i][1])
        exact(n, ca, cb, cw, res[i][0], res[i][1], roots, f"roots batch {i}")


def test_rescue_grid_deep():
    say("rescue on a deep grid")
    side = 120
    a, b, w = ds.gen_grid(side, 0.01, 1, 4, seed=6)
    n = side * side
    rng = np.random.default_rng(13)
    batches = make_batches(rng, n, a, b, w, 3, 800, 4)
    plain, _ = run_sequence(n, a, b, w, [0], batches, {"rescue_degree": OFF})
    res, r1 = run_sequence(n, a, b, w, [0], batches, {"rescue_degree": 1, "rescue_levels": 1024})
    assert r1 > 0
    ca, cb, cw = a, b, w
    for i, chg in enumerate(batches):
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        assert np.array_equal(plain[i][0], res[i][0]) and np.array_equal(plain[i][1], res[i][1])
        exact(n, ca, cb, cw, res[i][0], res[i][1], [0], f"grid batch {i}")


def test_rescue_off_with_zero_weights():
    say("rescue stays off once a zero weight is present")
    rng = np.random.default_rng(3)
    n = 300
    a, b, w = random_connected(rng, n, 1200, 0, 5)
    batches = make_batches(rng, n, a, b, w, 3, 200, 5)
    res, r1 = run_sequence(n, a, b, w, [0], batches, {"rescue_degree": 1, "rescue_levels": 64})
    assert r1 == 0
    ca, cb, cw = a, b, w
    for i, chg in enumerate(batches):
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
        csr = oracle.CSR.from_edges(n, ca, cb, cw)
        _, dref = oracle.dijkstra(csr, [0])
        assert np.array_equal(res[i][1], dref)
        assert np.array_equal(res[i][0], oracle.canonical_parents(csr, dref, [0], rule=1))


def test_rescue_auto_policy():
    """The default (automatic) rescue runs for batches with at least n/8
    deletions and only offers rows of at least 1024 entries; results equal the
    rescue switched off either way."""
    say("rescue automatic policy")
    scale = 14
    n = 1 << scale
    a, b, w = ds.gen_rmat(scale, 16, (0.57, 0.19, 0.19), 1, 10, seed=17)
    src = int(np.argmax(np.bincount(np.concatenate([a, b]), minlength=n)))
    rng = np.random.default_rng(29)
    # The tree-parent edge of every vertex with a row of 1024 or more entries is
    # deleted, so those rows are certainly offered to the rescue when it runs.
    csr = oracle.CSR.from_edges(n, a, b, w)
    _, d0 = oracle.dijkstra(csr, [src])
    p0 = oracle.canonical_parents(csr, d0, [src], rule=0)
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    hubs = [h for h in np.flatnonzero(deg >= 1024) if h != src and p0[h] >= 0]
    assert hubs
    cut = np.array([[p0[h], h, d0[h] - d0[p0[h]], 0] for h in hubs], np.int32)
    big = [np.concatenate([cut, random_batch(rng, n, a, b, w, n // 4, 0.0, 1, 10)])]  # > n/8 deletions: on
    small = [np.concatenate([cut, random_batch(rng, n, a, b, w, n // 16, 0.0, 1, 10)])]  # < n/8: off
    for batches, expect in ((big, True), (small, False)):
        plain, _ = run_sequence(n, a, b, w, [src], batches, {"rescue_degree": OFF})
        auto, r = run_sequence(n, a, b, w, [src], batches, {"rescue_degree": ds.RESCUE_AUTO, "rescue_levels": 16})
        assert (r > 0) == expect, f"rescued {r}"
        assert np.array_equal(plain[0][0], auto[0][0]) and np.array_equal(plain[0][1], auto[0][1])
        ua, ub, uw = oracle.apply_changes(n, a, b, w, batches[0])
        exact(n, ua, ub, uw, auto[0][0], auto[0][1], [src], f"auto expect={expect}")


def test_rescue_in_allocation_free_update():
    """Once a tree holds the rescue buffers (an earlier update made them), the
    allocation-free update rescues too, with the same result as the update
    that uploads its records."""
    from tests.test_zz_no_alloc import device_records
    say("rescue in the allocation-free update")
    scale = 12
    n = 1 << scale
    a, b, w = ds.gen_rmat(scale, 16, (0.57, 0.19, 0.19), 1, 20, seed=31)
    src = int(np.argmax(np.bincount(np.concatenate([a, b]), minlength=n)))
    rng = np.random.default_rng(37)
    batches = make_batches(rng, n, a, b, w, 2, 3000, 20)
    ds.set_tuning(rescue_degree=1, rescue_levels=16)
    g1, g2 = ds.Graph(n, a, b, w), ds.Graph(n, a, b, w)
    t1, t2 = ds.Tree(g1, src), ds.Tree(g2, src)
    keep = []
    try:
        ds.update(g1, t1, batches[0])
        ds.update(g2, t2, batches[0])  # allocates the rescue buffers of t2
        s1 = ds.update(g1, t1, batches[1])
        d = device_records(batches[1])
        keep.append(d)
        s2 = ds.update_noalloc(g2, t2, d.data_ptr(), len(batches[1]))
        assert s2["n_rescued"] > 0 and s2["n_rescued"] == s1["n_rescued"]
        p1, d1 = t1.read()
        p2, d2 = t2.read()
        assert np.array_equal(p1, p2) and np.array_equal(d1, d2)
        ca, cb, cw = oracle.apply_changes(n, a, b, w, batches[0])
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, batches[1])
        exact(n, ca, cb, cw, p2, d2, [src], "allocation-free with rescue")
    finally:
        for o in keep + [t1, t2, g1, g2]:
            o.close()


def test_rescue_random_stress():
    """Random connected graphs and multigraphs (parallel edges, self-loops,
    several components), single trees, forests and many-root trees, narrow
    and wide weight ranges, sequences of mixed batches: every candidate offered
    (degree 1, 2 or 4; 1 to 64 stages) against the rescue off and the oracle."""
    from tests._util import random_multigraph
    rng = np.random.default_rng(7)
    rescued = 0
    for case in range(40):
        kind = str(rng.choice(["tree", "tree", "forest", "roots"]))
        n = int(rng.integers(30, 2000))
        whi = int(rng.choice([2, 5, 20, 1000]))
        if rng.random() < 0.7:
            a, b, w = random_connected(rng, n, int(rng.integers(n, 8 * n)), 1, whi)
        else:
            a, b, w = random_multigraph(rng, n, int(rng.integers(n, 8 * n)), 1, whi)
        K = int(rng.integers(2, 9)) if kind != "tree" else 1
        srcs = [int(x) for x in rng.choice(n, K, replace=False)]
        batches, ca, cb, cw = [], a, b, w
        for _ in range(int(rng.integers(1, 4))):
            c = random_batch(rng, n, ca, cb, cw, int(rng.integers(1, max(2, len(ca) // 2))),
                             float(rng.uniform(0, 0.6)), 1, whi)
            batches.append(c)
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
        say(f"rescue stress case {case}: {kind} n={n} m={len(a)} K={K} w<={whi} batches={len(batches)}")
        tun = {"rescue_degree": int(rng.choice([1, 2, 4])), "rescue_levels": int(rng.choice([1, 2, 3, 8, 64]))}
        forest = kind == "forest"
        off, _ = run_sequence(n, a, b, w, srcs, batches, {"rescue_degree": OFF}, forest=forest)
        on, r = run_sequence(n, a, b, w, srcs, batches, tun, forest=forest)
        rescued += r
        ca, cb, cw = a, b, w
        for i, c in enumerate(batches):
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
            if forest:
                for j, s in enumerate(srcs):
                    assert np.array_equal(off[i][j][0], on[i][j][0]) and np.array_equal(off[i][j][1], on[i][j][1])
                    exact(n, ca, cb, cw, on[i][j][0], on[i][j][1], [s], f"case {case} batch {i} tree {j}")
            else:
                assert np.array_equal(off[i][0], on[i][0]) and np.array_equal(off[i][1], on[i][1]), (case, i)
                exact(n, ca, cb, cw, on[i][0], on[i][1], srcs, f"case {case} batch {i}")
    assert rescued > 0
