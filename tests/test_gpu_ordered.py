"""GPU parity of distance-ordered relaxation (tuning `delta` > 0): a round visits
only the marked slots below a threshold, which advances by delta past the
smallest marked distance whenever nothing below it is left.  Any visiting order
converges to the same fixed point, so every result must be byte-identical to the
unordered path and to the oracle: tiny steps (many threshold moves), steps near
the edge weights, steps above every distance (one phase, the plain schedule),
zero weights, heavy rows, deep grids, forests and trees built from scratch."""
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


@pytest.mark.parametrize("delta", [1, 37, 5000, 1 << 30])
@pytest.mark.parametrize("seed", range(3))
def test_ordered_update_sequences(delta, seed):
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(50, 3000))
    zero = seed == 2
    a, b, w = random_connected(rng, n, int(n * rng.uniform(1, 6)), 0 if zero else 1, 1000)
    batches, ca, cb, cw = [], a, b, w
    for _ in range(3):
        k = int(rng.integers(1, max(2, n // 2)))
        chg = random_batch(rng, n, ca, cb, cw, k, rng.uniform(0, 1), 0 if zero else 1, 1000)
        batches.append(chg)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
    with tuned(delta=delta):
        run_sequence(n, a, b, w, batches, src=int(rng.integers(0, n)), what=f"ordered d={delta} seq{seed}")


@pytest.mark.parametrize("delta", [1, 700])
def test_ordered_from_scratch_multigraph(delta):
    rng = np.random.default_rng(41)
    n = 1500
    a, b, w = random_multigraph(rng, n, 6000, 0, 5)  # zero weights, parallel edges, loops
    say(f"ordered from scratch d={delta}")
    with tuned(delta=delta):
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, 3)
        check_tree(Mirror(n, a, b, w), *t.read(), [3], f"ordered scratch d={delta}")


def test_ordered_heavy_rows_grid_and_plain_equal():
    """Chunked hub rows and a deep grid, each byte-equal to the unordered run."""
    a, b, w = ds.gen_rmat(12, 16, wlo=1, whi=1000, seed=5)
    n = 1 << 12
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    src = int(np.argmax(deg))
    chg = ds.gen_changes(n, a, b, w, 2500, 0.5, 1, 1000, seed=6)
    side = 200
    ga, gb, gw = ds.gen_grid(side, 0.01, 1, 10000, seed=7)
    gchg = random_batch(np.random.default_rng(8), side * side, ga, gb, gw, 1500, 0.3, 1, 10000)

    def run(nn, aa, bb, ww, s, c):
        g = ds.Graph(nn, aa, bb, ww)
        t = ds.Tree(g, s)
        ds.update(g, t, c)
        return t.read()

    for what, args in (("rmat12", (n, a, b, w, src, chg)), ("grid200", (side * side, ga, gb, gw, 0, gchg))):
        say(f"ordered vs plain {what}")
        with tuned(delta=-1):
            p0, d0 = run(*args)
        with tuned(delta=300, heavy_degree=64, chunk=64, grid=64):
            p1, d1 = run(*args)
        assert np.array_equal(d0, d1) and np.array_equal(p0, p1), what
    # and the grid against the oracle
    with tuned(delta=2500):
        run_sequence(side * side, ga, gb, gw, [gchg], 0, "ordered grid200")


@pytest.mark.parametrize("K", [8, 64])
def test_ordered_forest(K):
    rng = np.random.default_rng(950 + K)
    a, b, w = ds.gen_rmat(11, 16, wlo=1, whi=1000, seed=K + 1)
    n = 1 << 11
    srcs = rng.integers(0, n, K).astype(np.int32)
    batches, ca, cb, cw = [], a, b, w
    for ins in (0.5, 0.0):
        c = ds.gen_changes(n, ca, cb, cw, 1500, ins, 1, 1000, seed=int(rng.integers(1 << 30)))
        batches.append(c)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
    say(f"ordered forest K={K}")
    with tuned(delta=150):
        g = ds.Graph(n, a, b, w)
        f = ds.Tree(g, sources=srcs)
        for c in batches:
            ds.update(g, f, c)
        csr = oracle.CSR.from_edges(n, ca, cb, cw)
        for j in range(K):
            par, dist = f.read(j)
            _, dref = oracle.dijkstra(csr, [int(srcs[j])])
            assert np.array_equal(dist, dref), j
            assert np.array_equal(par, oracle.canonical_parents(csr, dref, [int(srcs[j])], rule=0)), j


def test_automatic_step_only_for_deep_trees():
    """Automatic mode (the default): a deep grid tree relaxes in distance order
    (far fewer entries scanned than unordered), an R-MAT tree does not (the
    same rounds as unordered); results byte-equal either way."""
    side = 400
    ga, gb, gw = ds.gen_grid(side, 0.01, 1, 10000, seed=11)
    gchg = random_batch(np.random.default_rng(12), side * side, ga, gb, gw, 3000, 0.3, 1, 10000)
    a, b, w = ds.gen_rmat(12, 16, wlo=1, whi=1000, seed=13)
    chg = ds.gen_changes(1 << 12, a, b, w, 2000, 0.5, 1, 1000, seed=14)

    def run(nn, aa, bb, ww, c):
        g = ds.Graph(nn, aa, bb, ww)
        with tuned(delta=-1):
            p0, d0 = ds.Tree(g, 0).read()
        t = ds.Tree(g, parent=p0, dist=d0)  # deep or not from its distances
        st = ds.update(g, t, c)
        return t.read(), st

    for what, args, deep in (("grid400", (side * side, ga, gb, gw, gchg), True),
                             ("rmat12", (1 << 12, a, b, w, chg), False)):
        say(f"automatic step {what}")
        with tuned(delta=-1):
            (p0, d0), s0 = run(*args)
        with tuned(delta=ds.DELTA_AUTO):
            (p1, d1), s1 = run(*args)
        assert np.array_equal(d0, d1) and np.array_equal(p0, p1), what
        if deep:
            assert s1["relax_edges"] < s0["relax_edges"], (what, s1["relax_edges"], s0["relax_edges"])
        else:
            assert s1["relax_rounds"] == s0["relax_rounds"], what
