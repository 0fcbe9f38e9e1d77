"""GPU parity at full size: every workload configuration at its own size.

Each configuration runs through the C ABI at the size it is specified at and is
checked against the oracle with size-independent checks:
  * the O(n + m) certificate (every live edge satisfies d[v] <= d[u] + w, every
    reachable vertex has a tight parent edge, parents reach a root, unreachable
    vertices have parent -1), which holds exactly when the distances are the
    shortest-path distances of the updated graph;
  * parents equal the canonical tree (smallest-id tight neighbour) of those
    distances, rule 0 (no graph here holds a zero weight);
  * Dijkstra on the updated graph where it finishes in seconds, and once per
    configuration otherwise;
  * for the batched and served forms, byte equality with the plain path.
The updated graph on the host is the oracle's restatement of the reference's
change semantics (oracle.apply_changes), applied batch after batch, so
deletions are always sampled from the current edges.

Configurations (sizes as specified):
  C2  R-MAT scale 24, edge factor 16, w U[1,1000], source = highest degree;
      the nine batch shapes k in {1e5, 1e6, 1e7} x insertions {0, 50, 100 %}
      applied in sequence to the resident graph and tree.
  C3  4096 x 4096 grid, 1 % of the edges removed, w U[1,1e4], source (0,0);
      batches of 1e5 records, 70 % deletions (half from tree edges), 30 %
      insertions at Manhattan distance <= 3.
  C4a R-MAT scale 22, K = 64 trees from uniform sources, one shared 1e6-record
      50/50 batch.
  C4b 32 random graphs, n log-uniform in [2^14, 2^21], average degree
      log-uniform in [4, 32], each with its own 50/50 batch of 1 % of its edges,
      all updated in one device graph.
  C5  16 resident R-MAT instances of scale 20 to 24 under a 160 GiB budget,
      a Poisson stream of batches (k log-uniform in [1, 1e5], insertion ratio
      U[0,1], uniform target instance).
"""
import math
import sys
import time

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle

pytestmark = pytest.mark.gpu

RMAT = (0.57, 0.19, 0.19)


def say(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def max_degree_vertex(n, a, b):
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    return int(np.argmax(deg))  # ties -> smallest id


def check_exact(csr, par, dist, srcs, what, dijkstra=False):
    """Certificate + canonical parents (+ Dijkstra) of one tree on the host CSR."""
    srcs = [int(s) for s in srcs]
    code, where = oracle.certificate(csr, dist, par, srcs)
    assert code == 0, f"{what}: certificate condition {code} fails at vertex {where}"
    canon = oracle.canonical_parents(csr, dist, srcs, rule=0)
    bad = np.flatnonzero(par != canon)
    assert len(bad) == 0, f"{what}: {len(bad)} parents not canonical, first {bad[:5]}: {par[bad[:5]]} vs {canon[bad[:5]]}"
    if dijkstra:
        _, dref = oracle.dijkstra(csr, srcs)
        badd = np.flatnonzero(dist != dref)
        assert len(badd) == 0, f"{what}: {len(badd)} distances differ from Dijkstra, first {badd[:5]}"


def edge_digest(a, b, w):
    """Order-independent digest of an undirected edge multiset."""
    a, b, w = (np.asarray(x).astype(np.uint64) for x in (a, b, w))
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    with np.errstate(over="ignore"):
        h = ((lo << np.uint64(32)) | hi) * np.uint64(0x9E3779B97F4A7C15)
        h ^= (w + np.uint64(1)) * np.uint64(0xC2B2AE3D27D4EB4F)
        h ^= h >> np.uint64(29)
        h *= np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(32)
        return int(np.sum(h, dtype=np.uint64)), len(a)


# ----------------------------------------------------------------------------- C2: R-MAT scale 24

C2_SHAPES = [(k, ins) for k in (100_000, 1_000_000, 10_000_000) for ins in (0.0, 0.5, 1.0)]


@pytest.fixture(scope="module")
def c2():
    t0 = time.time()
    a, b, w = ds.gen_rmat(24, 16, RMAT, 1, 1000, seed=2024)
    n = 1 << 24
    src = max_degree_vertex(n, a, b)
    say(f"C2: rmat s24 n={n} m={len(a)} src={src} generated in {time.time() - t0:.1f}s")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, src)
    say(f"C2: graph + tree from scratch on the GPU after {time.time() - t0:.1f}s")
    state = {"n": n, "src": src, "g": g, "t": t, "edges": (a, b, w), "done": 0, "csr": None}
    yield state
    t.close()
    g.close()


def test_c2_initial_tree(c2):
    n, (a, b, w) = c2["n"], c2["edges"]
    csr = oracle.CSR.from_edges(n, a, b, w, ordered=False)
    par, dist = c2["t"].read()
    check_exact(csr, par, dist, [c2["src"]], "C2 initial tree")


@pytest.mark.parametrize("shape", range(len(C2_SHAPES)))
def test_c2_rmat_s24_batch_shapes(c2, shape):
    """The nine shapes in sequence on the same resident graph and tree."""
    assert c2["done"] == shape, "shapes apply in sequence"
    n, src = c2["n"], c2["src"]
    k, ins = C2_SHAPES[shape]
    a, b, w = c2["edges"]
    t0 = time.time()
    chg = ds.gen_changes(n, a, b, w, k, ins, 1, 1000, seed=7000 + shape)
    say(f"C2 shape {shape}: k={k} ins={ins} (batch generated in {time.time() - t0:.1f}s)")
    st = ds.update(c2["g"], c2["t"], chg)
    say(f"C2 shape {shape}: update {st['t_update_ms']:.2f} ms, invalidated {st['n_invalidated']}, "
        f"rounds {st['inv_rounds']}/{st['relax_rounds']}")
    a, b, w = oracle.apply_changes(n, a, b, w, chg)
    c2["edges"] = (a, b, w)
    csr = oracle.CSR.from_edges(n, a, b, w, ordered=False)
    par, dist = c2["t"].read()
    check_exact(csr, par, dist, [src], f"C2 shape {shape} (k={k}, ins={ins})")
    c2["csr"] = csr
    c2["done"] += 1
    say(f"C2 shape {shape}: checked ({time.time() - t0:.1f}s)")


def test_c2_dijkstra_and_graph_after_all_shapes(c2):
    assert c2["done"] == len(C2_SHAPES)
    par, dist = c2["t"].read()
    t0 = time.time()
    _, dref = oracle.dijkstra(c2["csr"], [c2["src"]])
    say(f"C2: Dijkstra on the final graph {time.time() - t0:.1f}s")
    assert np.array_equal(dist, dref)
    # the resident CSR holds exactly the oracle's edge multiset
    ga, gb, gw = c2["g"].export()
    assert edge_digest(ga, gb, gw) == edge_digest(*c2["edges"])


# ----------------------------------------------------------------------------- C3: 4096^2 grid

def test_c3_grid_4096_batches():
    side = 4096
    n = side * side
    t0 = time.time()
    a, b, w = ds.gen_grid(side, 0.01, 1, 10000, seed=3)
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, 0)
    say(f"C3: grid {side}^2 m={len(a)}, graph + tree after {time.time() - t0:.1f}s")
    try:
        for step in range(2):
            par0, _ = t.read()
            chg = ds.gen_changes_grid(side, a, b, w, par0, 100_000, 0.7, 0.5, 3, 1, 10000, seed=31 + step)
            st = ds.update(g, t, chg)
            say(f"C3 batch {step}: update {st['t_update_ms']:.1f} ms, invalidated {st['n_invalidated']}, "
                f"rounds {st['inv_rounds']}/{st['relax_rounds']}")
            a, b, w = oracle.apply_changes(n, a, b, w, chg)
            csr = oracle.CSR.from_edges(n, a, b, w, ordered=False)
            par, dist = t.read()
            check_exact(csr, par, dist, [0], f"C3 batch {step}", dijkstra=True)
    finally:
        t.close()
        g.close()


# ----------------------------------------------------------------------------- C4a: 64 trees, scale 22

def test_c4a_forest_64_trees_s22():
    scale, K, k = 22, 64, 1_000_000
    n = 1 << scale
    t0 = time.time()
    a, b, w = ds.gen_rmat(scale, 16, RMAT, 1, 1000, seed=2024)
    srcs = np.random.default_rng(2024).integers(0, n, K).astype(np.int32)
    chg = ds.gen_changes(n, a, b, w, k, 0.5, 1, 1000, seed=2025)
    g = ds.Graph(n, a, b, w)
    f = ds.Tree(g, sources=srcs)
    say(f"C4a: s22 m={len(a)}, forest of {K} built after {time.time() - t0:.1f}s")
    try:
        st = ds.update(g, f, chg)
        say(f"C4a: shared update {st['t_update_ms']:.1f} ms, invalidated {st['n_invalidated']}")
        ua, ub, uw = oracle.apply_changes(n, a, b, w, chg)
        csr = oracle.CSR.from_edges(n, ua, ub, uw, ordered=False)
        trees = [f.read(j) for j in range(K)]
        for j in range(K):
            check_exact(csr, trees[j][0], trees[j][1], [srcs[j]], f"C4a tree {j}", dijkstra=j < 4)
        say(f"C4a: {K} trees certified, 4 against Dijkstra ({time.time() - t0:.1f}s)")
    finally:
        f.close()
        g.close()
    # the plain path, one tree per graph, gives the same bytes
    for j in range(2):
        gp = ds.Graph(n, a, b, w)
        tp = ds.Tree(gp, int(srcs[j]))
        ds.update(gp, tp, chg)
        pp, pd = tp.read()
        tp.close()
        gp.close()
        assert np.array_equal(pp, trees[j][0]) and np.array_equal(pd, trees[j][1]), f"C4a tree {j} != plain"


# ----------------------------------------------------------------------------- C4b: 32 graphs in one

def c4b_parts(seed=2024, graphs=32):
    """The 32 random graphs of C4b with their batches, roots and id offsets."""
    rng = np.random.default_rng(seed)
    parts, off = [], 0
    for i in range(graphs):
        n_i = int(math.exp(rng.uniform(math.log(2 ** 14), math.log(2 ** 21))))
        deg = math.exp(rng.uniform(math.log(4), math.log(32)))
        m_i = max(n_i - 1, int(n_i * deg / 2))
        a, b, w = ds.gen_tree_plus_random(n_i, m_i, 1, 1000, seed=seed + 10 * i)
        c = ds.gen_changes(n_i, a, b, w, max(1, m_i // 100), 0.5, 1, 1000, seed=seed + 10 * i + 1)
        parts.append((n_i, a, b, w, c, int(rng.integers(0, n_i)), off))
        off += n_i
    return parts, off


def test_c4b_32_graphs_one_device_graph():
    t0 = time.time()
    parts, N = c4b_parts()
    A = np.concatenate([p[1] + p[6] for p in parts])
    B = np.concatenate([p[2] + p[6] for p in parts])
    W = np.concatenate([p[3] for p in parts])
    C = np.concatenate([p[4] + np.array([p[6], p[6], 0, 0], np.int32) for p in parts])
    roots = np.array([p[5] + p[6] for p in parts], np.int32)
    g = ds.Graph(N, A, B, W)
    t = ds.Tree(g, roots=roots)
    say(f"C4b: {len(parts)} graphs, N={N} M={len(A)} K={len(C)} ({time.time() - t0:.1f}s)")
    try:
        st = ds.update(g, t, C)
        say(f"C4b: merged update {st['t_update_ms']:.2f} ms")
        par, dist = t.read()
    finally:
        t.close()
        g.close()
    for i, (n_i, a, b, w, c, s, o) in enumerate(parts):
        sp = par[o:o + n_i].astype(np.int64)
        sp[sp >= 0] -= o
        sp = sp.astype(np.int32)
        sd = dist[o:o + n_i]
        ua, ub, uw = oracle.apply_changes(n_i, a, b, w, c)
        csr = oracle.CSR.from_edges(n_i, ua, ub, uw, ordered=False)
        check_exact(csr, sp, sd, [s], f"C4b graph {i}", dijkstra=n_i <= 1 << 18)
        gp = ds.Graph(n_i, a, b, w)
        tp = ds.Tree(gp, s)
        ds.update(gp, tp, c)
        pp, pd = tp.read()
        tp.close()
        gp.close()
        assert np.array_equal(pp, sp) and np.array_equal(pd, sd), f"C4b graph {i} != plain"
    say(f"C4b: every graph certified and equal to plain ({time.time() - t0:.1f}s)")


# ----------------------------------------------------------------------------- C5: served stream

C5_INSTANCES = 16
C5_REQUESTS = 160
C5_RATE = 2000.0  # arrivals per second, below the measured capacity


def c5_requests(rng, insts, count):
    """count requests: k log-uniform in [1, 1e5], insertion ratio U[0, 1],
    uniform target instance.  Per instance, all of its deletions are distinct
    edges of the initial graph and all of its insertions distinct absent pairs
    (one generator call), so the graph after its requests is the same in any
    order and the oracle applies them as one batch."""
    plan = []
    for _ in range(count):
        i = int(rng.integers(0, len(insts)))
        k = int(math.exp(rng.uniform(0, math.log(100_000))))
        ki = int(round(k * float(rng.uniform(0, 1))))
        plan.append((i, k - ki, ki))
    reqs = [None] * count
    for i, inst in enumerate(insts):
        mine = [j for j, p in enumerate(plan) if p[0] == i]
        kd, ki = sum(plan[j][1] for j in mine), sum(plan[j][2] for j in mine)
        if kd + ki == 0:
            continue
        a, b, w = inst["edges"]
        allc = ds.gen_changes(inst["n"], a, b, w, kd + ki, ki / (kd + ki), 1, 1000, seed=9000 + i)
        dels, ins = allc[allc[:, 3] == 0], allc[allc[:, 3] == 1]
        assert len(dels) == kd and len(ins) == ki
        od = oi = 0
        for j in mine:
            c = np.concatenate([dels[od:od + plan[j][1]], ins[oi:oi + plan[j][2]]])
            od += plan[j][1]
            oi += plan[j][2]
            rng.shuffle(c)
            reqs[j] = (i, np.ascontiguousarray(c))
        inst["all"] = allc
    return [r for r in reqs if r is not None and len(r[1])]


@pytest.fixture(scope="module")
def c5():
    t0 = time.time()
    rng = np.random.default_rng(55)
    scales = [20 + (i * 5) // C5_INSTANCES for i in range(C5_INSTANCES)]
    srv = ds.Server(mem_budget_bytes=160 << 30)
    insts = []
    for i, sc in enumerate(scales):
        a, b, w = ds.gen_rmat(sc, 16, RMAT, 1, 1000, seed=500 + i)
        n = 1 << sc
        src = max_degree_vertex(n, a, b)
        assert srv.add_instance(n, a, b, w, src) == i
        insts.append({"n": n, "edges": (a, b, w), "src": src, "reqs": [], "all": None})
    say(f"C5: {len(insts)} instances s{min(scales)}-s{max(scales)} admitted ({time.time() - t0:.1f}s)")
    srv.start()
    say(f"C5: server started ({time.time() - t0:.1f}s)")
    reqs = c5_requests(rng, insts, C5_REQUESTS)
    for i, c in reqs:
        insts[i]["reqs"].append(c)
    say(f"C5: {len(reqs)} requests generated ({time.time() - t0:.1f}s)")
    yield {"srv": srv, "insts": insts, "rng": rng, "reqs": reqs, "streamed": False}
    srv.close()


def test_c5_poisson_stream(c5):
    srv, reqs, rng = c5["srv"], c5["reqs"], c5["rng"]
    gaps = rng.exponential(1.0 / C5_RATE, len(reqs))
    say(f"C5: {len(reqs)} requests, {sum(len(c) for _, c in reqs)} records, Poisson rate {C5_RATE}/s")
    tickets, tnext = [], time.perf_counter()
    for (i, c), gap in zip(reqs, gaps):
        tnext += gap
        d = tnext - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(srv.submit(i, c))
    rs = [srv.wait(tk) for tk in tickets]
    lat = sorted(r["latency_ms"] for r in rs)
    c5["streamed"] = True
    say(f"C5: applied; p50 {lat[len(lat) // 2]:.2f} ms, p99 {lat[int(len(lat) * 0.99)]:.2f} ms, "
        f"mean merged {np.mean([r['merged'] for r in rs]):.2f}")


@pytest.mark.parametrize("inst", range(C5_INSTANCES))
def test_c5_instance_exact_and_equal_to_plain(c5, inst):
    assert c5["streamed"]
    srv, I = c5["srv"], c5["insts"][inst]
    n, src = I["n"], I["src"]
    a, b, w = I["edges"]
    sp, sd = srv.read(inst)
    ua, ub, uw = (a, b, w) if I["all"] is None else oracle.apply_changes(n, a, b, w, I["all"])
    csr = oracle.CSR.from_edges(n, ua, ub, uw, ordered=False)
    check_exact(csr, sp, sd, [src], f"C5 instance {inst}", dijkstra=n <= 1 << 20)
    del csr, ua, ub, uw
    # plain: the same requests, in the same order, one update each on a graph of its own
    gp = ds.Graph(n, a, b, w)
    tp = ds.Tree(gp, src)
    for c in I["reqs"]:
        ds.update(gp, tp, c)
    pp, pd = tp.read()
    tp.close()
    gp.close()
    assert np.array_equal(pp, sp) and np.array_equal(pd, sd), f"C5 instance {inst}: served != plain"
    say(f"C5 instance {inst}: s{n.bit_length() - 1}, {len(I['reqs'])} requests, exact and equal to plain")
