"""GPU parity of the batched form: K trees on one graph, and many graphs in one
device graph (one root per component), each bit-exact to the plain path and to
the oracle."""
import sys

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle
from tests._util import random_batch, random_connected

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def plain_result(n, a, b, w, src, batches):
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, src)
    for c in batches:
        ds.update(g, t, c)
    return t.read()


@pytest.mark.parametrize("K", [1, 3, 8, 64])
def test_forest_matches_plain_and_oracle(K):
    rng = np.random.default_rng(500 + K)
    a, b, w = ds.gen_rmat(11, 16, wlo=1, whi=1000, seed=K)
    n = 1 << 11
    srcs = rng.integers(0, n, K).astype(np.int32)
    batches, ca, cb, cw = [], a, b, w
    for ins in (0.5, 0.0, 1.0):
        c = ds.gen_changes(n, ca, cb, cw, 1500, ins, 1, 1000, seed=int(rng.integers(1 << 30)))
        batches.append(c)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
    say(f"forest K={K}")
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
        if j < 3:
            pp, pd = plain_result(n, a, b, w, int(srcs[j]), batches)
            assert np.array_equal(pp, par) and np.array_equal(pd, dist)


@pytest.mark.parametrize("K,wlo", [(40, 0), (64, 1)])
def test_forest64_layout(K, wlo):
    """K' = 64 (one vertex per frontier word): the tree-vector path, with zero
    weights (level repair and parent pick run as forests too) and with hub rows
    split into pieces."""
    rng = np.random.default_rng(700 + K)
    a, b, w = ds.gen_rmat(10, 16, wlo=wlo, whi=6 if wlo == 0 else 1000, seed=K)
    n = 1 << 10
    srcs = rng.integers(0, n, K).astype(np.int32)
    batches, ca, cb, cw = [], a, b, w
    for ins in (0.5, 0.0, 1.0, 0.3):
        c = ds.gen_changes(n, ca, cb, cw, 900, ins, wlo, 6 if wlo == 0 else 1000, seed=int(rng.integers(1 << 30)))
        batches.append(c)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
    say(f"forest64 K={K} wlo={wlo}")
    g = ds.Graph(n, a, b, w)
    f = ds.Tree(g, sources=srcs)
    for c in batches:
        ds.update(g, f, c)
    csr = oracle.CSR.from_edges(n, ca, cb, cw)
    zero = wlo == 0
    for j in range(K):
        par, dist = f.read(j)
        _, dref = oracle.dijkstra(csr, [int(srcs[j])])
        assert np.array_equal(dist, dref), j
        assert np.array_equal(par, oracle.canonical_parents(csr, dref, [int(srcs[j])], rule=1 if zero else 0)), j
    pp, pd = plain_result(n, a, b, w, int(srcs[0]), batches)
    assert np.array_equal(pp, f.read(0)[0]) and np.array_equal(pd, f.read(0)[1])


def test_forest_with_invalidation_heavy_batches():
    rng = np.random.default_rng(9)
    n = 3000
    a, b, w = random_connected(rng, n, 4500, 1, 20)  # sparse: deep trees, big invalidations
    K = 16
    srcs = rng.integers(0, n, K).astype(np.int32)
    c = random_batch(rng, n, a, b, w, 800, 0.2, 1, 20)
    g = ds.Graph(n, a, b, w)
    f = ds.Tree(g, sources=srcs)
    say("forest sparse K=16")
    st = ds.update(g, f, c)
    assert st["n_invalidated"] > 0
    ua, ub, uw = oracle.apply_changes(n, a, b, w, c)
    csr = oracle.CSR.from_edges(n, ua, ub, uw)
    for j in range(K):
        par, dist = f.read(j)
        _, dref = oracle.dijkstra(csr, [int(srcs[j])])
        assert np.array_equal(dist, dref)
        assert oracle.certificate(csr, dist, par, [int(srcs[j])])[0] == 0


def test_many_graphs_one_device_graph():
    """Graphs of uneven size concatenated with vertex offsets, one root each."""
    rng = np.random.default_rng(11)
    parts = []
    off = 0
    for i in range(12):
        n_i = int(np.exp(rng.uniform(np.log(50), np.log(20000))))
        deg = float(np.exp(rng.uniform(np.log(2), np.log(16))))
        a, b, w = random_connected(rng, n_i, max(n_i - 1, int(n_i * deg / 2)), 1, 1000)
        c = random_batch(rng, n_i, a, b, w, max(1, len(a) // 100), 0.5, 1, 1000)
        src = int(rng.integers(0, n_i))
        parts.append((n_i, a, b, w, c, src, off))
        off += n_i
    N = off
    A = np.concatenate([p[1] + p[6] for p in parts])
    B = np.concatenate([p[2] + p[6] for p in parts])
    W = np.concatenate([p[3] for p in parts])
    C = np.concatenate([p[4] + np.array([p[6], p[6], 0, 0], np.int32) for p in parts])
    roots = np.array([p[5] + p[6] for p in parts], np.int32)
    say(f"union of {len(parts)} graphs N={N}")
    g = ds.Graph(N, A, B, W)
    t = ds.Tree(g, roots=roots)
    ds.update(g, t, C)
    par, dist = t.read()
    for n_i, a, b, w, c, src, o in parts:
        pp, pd = plain_result(n_i, a, b, w, src, [c])
        sp = par[o:o + n_i].astype(np.int64)
        sp[sp >= 0] -= o
        assert np.array_equal(sp, pp) and np.array_equal(dist[o:o + n_i], pd)
