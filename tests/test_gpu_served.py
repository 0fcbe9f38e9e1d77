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


def plain_tree(n, a, b, w, s, seq):
    """The plain path: the requests applied one update each to a graph of its own."""
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, s)
    for c in seq:
        ds.update(g, t, c)
    out = t.read()
    t.close()
    g.close()
    return out


def test_served_merges_one_instance_exactly():
    """Requests of one instance merge into shared updates except across a key
    that an earlier request deletes and a later one inserts again; the result
    equals the requests applied one by one."""
    rng = np.random.default_rng(31)
    a, b, w = ds.gen_rmat(12, 16, wlo=1, whi=50, seed=310)
    n = 1 << 12
    src = int(np.argmax(np.bincount(np.concatenate([a, b]), minlength=n)))
    srv = ds.Server(mem_budget_bytes=1 << 32)
    iid = srv.add_instance(n, a, b, w, src)
    srv.start()
    seq, ca, cb, cw = [], a, b, w
    for j in range(40):
        if j % 7 == 3 and seq:
            # re-insert some keys the previous request deleted: must not merge with it
            prev = seq[-1]
            dels = prev[prev[:, 3] == 0][:5].copy()
            dels[:, 3] = 1
            c = np.concatenate([dels, random_batch(rng, n, ca, cb, cw, 50, 0.5, 1, 50)]).astype(np.int32)
        else:
            c = random_batch(rng, n, ca, cb, cw, int(rng.integers(1, 400)), rng.uniform(0, 1), 1, 50)
        ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, c)
        seq.append(c)
    say("served: 40 requests to one instance, all at once")
    tickets = [srv.submit(iid, c) for c in seq]
    stats = [srv.wait(t) for t in tickets]
    assert max(s["merged"] for s in stats) > 1  # consecutive requests of the instance merged
    par, dist = srv.read(iid)
    pp, pd = plain_tree(n, a, b, w, src, seq)
    assert np.array_equal(pd, dist) and np.array_equal(pp, par)
    _, dref = oracle.dijkstra(oracle.CSR.from_edges(n, ca, cb, cw), [src])
    assert np.array_equal(dist, dref)
    srv.close()


def test_served_admit_and_remove_while_serving():
    """Instances join while requests of others are in flight and leave after
    theirs; the budget admits a new instance once a late one is removed; every
    served instance equals the plain path alone."""
    rng = np.random.default_rng(41)

    def inst(scale, seed):
        a, b, w = ds.gen_rmat(scale, 16, wlo=1, whi=1000, seed=seed)
        n = 1 << scale
        return n, a, b, w, int(np.argmax(np.bincount(np.concatenate([a, b]), minlength=n)))

    A, B, C, D = inst(12, 411), inst(11, 412), inst(13, 413), inst(13, 414)
    # the budget: what A and B hold, plus room for one of C, D (measured on an
    # unstarted server, which charges the same bytes at admission)
    probe = ds.Server(mem_budget_bytes=1 << 40)
    probe.add_instance(*A)
    probe.add_instance(*B)
    u = probe.info()["used_bytes"]
    probe.add_instance(*C)
    c_bytes = probe.info()["used_bytes"] - u
    probe.close()
    srv = ds.Server(mem_budget_bytes=u + c_bytes + c_bytes // 2)
    ia, ib = srv.add_instance(*A), srv.add_instance(*B)
    srv.start()
    cur = {ia: A[1:4], ib: B[1:4]}
    seqs = {ia: [], ib: []}

    def next_batch(i, n, k):
        ca, cb, cw = cur[i]
        c = random_batch(rng, n, ca, cb, cw, k, rng.uniform(0, 1), 1, 1000)
        cur[i] = oracle.apply_changes(n, ca, cb, cw, c)
        seqs[i].append(c)
        return c

    tickets = []
    for j in range(30):  # a burst to A and B, in flight while C is admitted
        i = ia if j % 2 else ib
        tickets.append(srv.submit(i, next_batch(i, (A if i == ia else B)[0], int(rng.integers(50, 2000)))))
    say("served: admitting an instance while requests are in flight")
    ic = srv.add_instance(*C)
    cur[ic], seqs[ic] = C[1:4], []
    for j in range(20):
        i = (ia, ib, ic)[j % 3]
        n = (A, B, C)[(ia, ib, ic).index(i)][0]
        tickets.append(srv.submit(i, next_batch(i, n, int(rng.integers(50, 2000)))))
    with pytest.raises(ds.DssspError) as e:  # no room for D while C is held
        srv.add_instance(*D)
    assert e.value.code == ds.ERR_BUDGET
    for t in tickets:
        srv.wait(t)
    pc = srv.read(ic)
    srv.remove_instance(ic)  # C had a lane of its own: its budget comes back
    with pytest.raises(ds.DssspError):
        srv.submit(ic, seqs[ic][0])
    u_before_d = srv.info()["used_bytes"]
    assert u_before_d == u
    id_d = srv.add_instance(*D)
    d_bytes = srv.info()["used_bytes"] - u_before_d
    cur[id_d], seqs[id_d] = D[1:4], []
    for j in range(10):
        srv.wait(srv.submit(id_d, next_batch(id_d, D[0], int(rng.integers(50, 2000)))))
    for i, I in ((ia, A), (ib, B), (id_d, D)):
        par, dist = srv.read(i)
        pp, pd = plain_tree(*I, seqs[i])
        assert np.array_equal(pd, dist) and np.array_equal(pp, par), i
    pp, pd = plain_tree(*C, seqs[ic])
    assert np.array_equal(pd, pc[1]) and np.array_equal(pp, pc[0])
    srv.remove_instance(ia)
    assert srv.info()["used_bytes"] == u + d_bytes  # A's lane still holds B
    srv.remove_instance(ib)  # the start-time lane of A and B is released with its last instance
    assert srv.info()["used_bytes"] == d_bytes
    srv.close()
