"""GPU parity of the served form: every instance's tree after a mixed, merged
request stream equals the plain path applied to that instance alone, in the
same per-instance order, with all instances in one lane, in a few lanes whose
updates overlap, and in one lane each."""
import sys
import threading

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle
from tests._util import random_batch, random_connected

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("lanes", [1, 2, 0])
def test_served_mix_equals_alone(lanes):
    rng = np.random.default_rng(21)
    insts = []
    for i in range(6):
        if i % 2 == 0:
            a, b, w = ds.gen_rmat(int(rng.integers(9, 13)), 16, wlo=1, whi=1000, seed=100 + i)
            n = int(max(a.max(), b.max())) + 1
        else:
            n = int(rng.integers(200, 5000))
            a, b, w = random_connected(rng, n, int(n * 3), 1, 1000)
        insts.append((n, a, b, w, int(rng.integers(0, n))))
    srv = ds.Server(mem_budget_bytes=1 << 34, lanes=lanes)
    ids = [srv.add_instance(n, a, b, w, s) for n, a, b, w, s in insts]
    srv.start()
    # request stream: 60 batches, log-uniform sizes, random instance, random mix
    reqs = []
    cur = [(a, b, w) for _, a, b, w, _ in insts]
    for r in range(60):
        i = int(rng.integers(0, len(insts)))
        n = insts[i][0]
        k = int(np.exp(rng.uniform(0, np.log(2000))))
        ca, cb, cw = cur[i]
        c = random_batch(rng, n, ca, cb, cw, k, rng.uniform(0, 1), 1, 1000)
        cur[i] = oracle.apply_changes(n, ca, cb, cw, c)
        reqs.append((i, c))
    say(f"served stream of 60 requests, lanes={lanes}")
    tickets = [None] * len(reqs)

    def producer(lo, hi):
        for j in range(lo, hi):
            tickets[j] = srv.submit(ids[reqs[j][0]], reqs[j][1])

    # submit in order (per-instance order matters), in two bursts
    producer(0, 30)
    stats = [srv.wait(t) for t in tickets[:30]]
    producer(30, 60)
    stats += [srv.wait(t) for t in tickets[30:]]
    info = srv.info()
    assert info["requests"] == 60
    if 0 < lanes < len(insts):
        assert max(s["merged"] for s in stats) > 1  # the loops did merge requests
        assert info["launches"] < 60
    for i, (n, a, b, w, s) in enumerate(insts):
        par, dist = srv.read(ids[i])
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, s)
        for j, c in reqs:
            if j == i:
                ds.update(g, t, c)
        pp, pd = t.read()
        assert np.array_equal(pd, dist) and np.array_equal(pp, par), i
        ca, cb, cw = cur[i]
        _, dref = oracle.dijkstra(oracle.CSR.from_edges(n, ca, cb, cw), [s])
        assert np.array_equal(dist, dref)
    srv.close()


def test_served_concurrent_submitters():
    """Several host threads submit to disjoint instances at once."""
    rng = np.random.default_rng(5)
    insts = []
    for i in range(4):
        n = 3000
        a, b, w = random_connected(rng, n, 9000, 1, 100)
        insts.append((n, a, b, w, 0))
    srv = ds.Server(mem_budget_bytes=1 << 32, max_merge=4, lanes=2)
    ids = [srv.add_instance(n, a, b, w, s) for n, a, b, w, s in insts]
    srv.start()
    batches = {}
    for i, (n, a, b, w, _) in enumerate(insts):
        seq, ca, cb, cw = [], a, b, w
        for _ in range(8):
            c = random_batch(rng, n, ca, cb, cw, 300, 0.5, 1, 100)
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
            seq.append(c)
        batches[i] = seq
    errs = []

    def worker(i):
        try:
            for c in batches[i]:
                srv.wait(srv.submit(ids[i], c))
        except Exception as e:  # pragma: no cover - reported below
            errs.append(e)

    say("served concurrent submitters")
    th = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs
    for i, (n, a, b, w, s) in enumerate(insts):
        par, dist = srv.read(ids[i])
        g = ds.Graph(n, a, b, w)
        t = ds.Tree(g, s)
        for c in batches[i]:
            ds.update(g, t, c)
        pp, pd = t.read()
        assert np.array_equal(pd, dist) and np.array_equal(pp, par)
    srv.close()
