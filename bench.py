"""Benchmark of the dynamic SSSP batch update on one MI355X.

Default workload (configuration 2, the headline): R-MAT scale 24, edge factor
16, (a,b,c,d) = (0.57, 0.19, 0.19, 0.05), deduplicated, symmetrised, self-loops
dropped, int32 weights U[1,1000]; source = the highest-degree vertex; change
batches of k = 1,000,000 records, 50% insertions of absent pairs, 50% deletions
of current edges.  A step is one batch applied to the resident graph and tree
(in-HBM insert, classify + tombstone, invalidate, seed, relax) with the batch
already in HBM; steps apply distinct batches one after another to the evolving
graph.  The timed region is the K steps between two device synchronisations.

Other workloads (--config): 1 small random graph; 3 high-diameter grid;
4a K=64 trees on one R-MAT graph; 4b 32 graphs in one launch; 5 served stream.
Each prints one JSON line of the same shape.

The JSON line carries the changed-edges/s throughput, p50/p99 per-batch latency,
the relaxation kernel's roofline figures (algorithmic bytes / summed launch
durations, measured with HIP events on the graph's stream), a bit-exactness
check against a CPU Dijkstra of the final graph, and the CPU baseline: one
from-scratch sequential Dijkstra (the reference's CPU path, restated in
oracle/oracle.c) on the final graph, on one core.

    python bench.py [--steps K] [--warmup W] [--config 2] [--scale 24] [--k 1000000] [--ins 0.5]
"""
import argparse
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "changed edges/s + HBM GB/s per update; p50/p99 batch latency; dist bit-exact"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--config", default="2", choices=["1", "2", "3", "4a", "4b", "5"])
    p.add_argument("--scale", type=int, default=24)
    p.add_argument("--edge-factor", type=int, default=16)
    p.add_argument("--k", type=int, default=1_000_000)
    p.add_argument("--ins", type=float, default=0.5)
    p.add_argument("--side", type=int, default=4096)
    p.add_argument("--trees", type=int, default=64)
    p.add_argument("--graphs", type=int, default=32)
    p.add_argument("--instances", type=int, default=16)
    p.add_argument("--requests", type=int, default=1600, help="served: requests per run (Poisson, then saturated)")
    p.add_argument("--rate", type=float, default=5000.0,
                   help="served: Poisson arrivals per second of the first half of the requests (below capacity)")
    p.add_argument("--serve-scale", type=int, default=0, help="served: every instance at this R-MAT scale")
    p.add_argument("--lanes", type=int, default=0, help="served: concurrent lanes (0: the library's default)")
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU Dijkstra baseline / exactness check")
    p.add_argument("--no-timing", action="store_true", help="no per-launch events (roofline unmeasured)")
    p.add_argument("--autotune", action="store_true",
                   help="plain configs: tune the launch knobs on the warm-up batches first (dsssp_autotune)")
    p.add_argument("--probe", action="store_true",
                   help="only run the library's probe launches of known footprint (a 1 GiB copy, 8-byte gathers "
                        "over 8 GiB and over 64 MiB) for calibrating HBM byte counters, then exit")
    for knob in ("block", "grid", "heavy_degree", "chunk", "round_batch", "use_graph", "edges_per_thread",
                 "tile_entries", "local_tiles", "delta", "sweep", "rescue_degree", "rescue_levels"):
        p.add_argument("--" + knob.replace("_", "-"), type=int, default=None, help="tuning knob (dsssp_tuning)")
    return p.parse_args()


def pct(x, q):
    import numpy as np
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def evolving_batches(ds, np, n, a, b, w, k, ins, count, seed, wlo=1, whi=1000, first_deletable=0):
    """count batches of exactly k records (round(k*ins) insertions) for one graph
    that they change in turn.  Insertions are distinct pairs absent from the
    graph (never deleted later, so absent when inserted); the deletions of batch
    i are sampled uniformly, without replacement, from the edges present before
    it: the deletable initial edges [first_deletable, m) and the insertions of
    batches 0..i-1, less everything deleted by batches 0..i-1."""
    ki = int(round(k * ins))
    kd = k - ki
    m = len(a)
    ins_all = (ds.gen_changes(n, a, b, w, ki * count, 1.0, wlo, whi, seed=seed) if ki
               else np.zeros((0, 4), np.int32))
    A = np.concatenate([a, ins_all[:, 0]])
    B = np.concatenate([b, ins_all[:, 1]])
    W = np.concatenate([w, ins_all[:, 2]])
    dead = np.zeros(len(A), bool)
    dead[:first_deletable] = True
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        present = m + ki * i  # index space of the edges present before batch i
        pick = np.zeros(0, np.int64)
        while len(pick) < kd:
            cand = np.unique(rng.integers(0, present, 2 * (kd - len(pick)) + 64))
            cand = cand[~dead[cand]]
            cand = rng.permutation(cand)[:kd - len(pick)]
            dead[cand] = True
            pick = np.concatenate([pick, cand])
        flip = rng.random(kd) < 0.5
        dels = np.stack([np.where(flip, B[pick], A[pick]), np.where(flip, A[pick], B[pick]), W[pick],
                         np.zeros(kd, np.int32)], 1).astype(np.int32)
        c = np.concatenate([ins_all[i * ki:(i + 1) * ki], dels])
        rng.shuffle(c)
        out.append(np.ascontiguousarray(c))
    return out


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_check(ds, np, g, n, srcs, par, dist, k, label, run_cpu=True):
    """Exactness on the final graph + the CPU baseline (one Dijkstra per source),
    with the textbook lazy-deletion Dijkstra timed beside it for context."""
    import oracle
    ea, eb, ew = g.export()
    csr = oracle.CSR.from_edges(n, ea, eb, ew)
    del ea, eb, ew
    tc = time.perf_counter()
    if len(srcs) == 1:
        _, cd, _ = oracle.djk(csr, int(srcs[0]), oracle.INF)
    else:
        _, cd = oracle.dijkstra(csr, [int(s) for s in srcs])
    t_cpu = time.perf_counter() - tc
    tc = time.perf_counter()
    _, ld = oracle.dijkstra(csr, [int(s) for s in srcs])
    t_lazy = time.perf_counter() - tc
    mism = int(np.sum(cd != dist)) + int(np.sum(ld != dist))
    code, _ = oracle.certificate(csr, dist, par, [int(s) for s in srcs])
    canon = bool(np.array_equal(par, oracle.canonical_parents(csr, cd, [int(s) for s in srcs], rule=0)))
    check = {"dijkstra_mismatches": mism, "certificate": int(code), "parents_canonical": canon}
    cpu = {"value": k / t_cpu, "unit": "changed edges/s", "cores": 1, "kind": "port",
           "sample": f"one from-scratch sequential Dijkstra (oracle/oracle.c, restating "
                     f"seqSSSPwithDist_undir.cpp:66-128) on the final {label} graph ({len(csr.col) // 2} edges) "
                     f"= the reference's CPU cost of one k={k} batch; {t_cpu:.2f} s",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(), "seconds": t_cpu,
           "lazy_dijkstra_seconds": t_lazy}
    log(f"cpu dijkstra {t_cpu:.2f}s (lazy {t_lazy:.2f}s), mismatches {mism}, certificate {code}, canonical {canon}")
    return check, cpu


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_frontier_relax_traffic_rmat_s24_k1e6.md")


def measured_traffic():
    """HBM bytes per relaxation dispatch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs of this workload, scaled by the factors that
    profiles/probe_counter_calibration.md measured for random reads)."""
    try:
        for ln in open(TRAFFIC_FILE):
            if ln.startswith("machine-readable:"):
                return json.loads(ln.split(":", 1)[1])["hbm_bytes_per_dispatch"]
    except (OSError, ValueError, KeyError):
        pass
    return None


def gather_ceiling(ds):
    """The device's rate of independent 8-byte reads at hashed positions of an
    8 GiB buffer (beyond every cache): dsssp_probe kind 1, 2^26 reads."""
    ms = ds.probe(1, 8 << 30, 1 << 26, reps=5)
    return (1 << 26) / (ms / 1e3)


def roofline(stats, traffic_applies=False, gather_rate=None):
    relax_bytes = sum(s["alg_bytes_relax"] for s in stats)
    relax_ms = sum(s["t_relax_kernel_ms"] for s in stats)
    launches = sum(s["relax_launches"] for s in stats)
    achieved = relax_bytes / (relax_ms / 1e3) / 1e9 if relax_ms > 0 else None
    traffic = measured_traffic() if traffic_applies else None
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": (achieved / HBM_PEAK_GBPS) if achieved else None, "traffic": traffic,
           "kernel": "k_frontier<M_RELAX>", "launches": launches,
           "avg_launch_us": relax_ms / max(1, launches) * 1e3,
           "alg_bytes_per_launch": relax_bytes / max(1, launches)}
    if traffic:
        out["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
        out["traffic_GBps"] = traffic * launches / (relax_ms / 1e3) / 1e9 if relax_ms > 0 else None
    if gather_rate and relax_ms > 0:
        # The kernel's accesses at unpredictable addresses: per frontier slot its
        # vertex record and own state, per scanned entry the neighbour's state,
        # per improvement the state atomicMin and the bitmap mark.
        acc = sum(2 * s["frontier_vertices"] + s["relax_edges"] + 2 * s["relax_improved"] for s in stats)
        rate = acc / (relax_ms / 1e3)
        out["random_access"] = {"bound": "hbm random 8-byte accesses", "achieved": rate, "peak": gather_rate,
                                "unit": "accesses/s", "frac": rate / gather_rate,
                                "accesses_per_launch": acc / max(1, launches),
                                "peak_source": "dsssp_probe kind 1: 2^26 reads at hashed positions of 8 GiB"}
    return out


SWEEP_PASSES = 5  # passes timed per measurement; the median pass is reported


def timed_passes(ds, t, timing):
    """SWEEP_PASSES full edge passes of a converged tree (each must be its
    fixed point); returns the stats of the median pass by its time, every
    pass's time and the states lowered by all of them (0 when exact)."""
    ds.set_tuning(kernel_timing=1 if timing else 0)
    runs = [t.relax_full() for _ in range(SWEEP_PASSES)]
    ds.set_tuning(kernel_timing=0)
    lowered = sum(r["relax_improved"] for r in runs)
    ms = [r["t_sweep_ms"] for r in runs]
    runs.sort(key=lambda r: r["t_sweep_ms"])
    return runs[len(runs) // 2], ms, lowered


def full_sweep(fs, n, m_live):
    """Roofline figures of the edge pass of dsssp_tree_relax_full: every live
    edge relaxed in both directions once, from the row of its higher-id
    endpoint (k_frontier<M_SWEEP>: a tile launch and an item launch, timed with
    HIP events on the graph's stream).

    Two byte counts.  "frac" prices the pass as the plain relaxation sweep it
    replaces: 16 B per vertex (row offset + own state) and 16 B per directed
    entry (entry + neighbour state), n*16 + 2m*16 for m live undirected edges,
    the per-sweep figure the path's measurement contract states.
    "frac_bytes_read" counts what this pass itself reads and writes, because it
    visits each edge once: 32 B per slot, 8 B per scanned entry, 8 B more per
    entry that reads its neighbour's state, 16 B per improvement."""
    ms = fs["t_sweep_ms"]
    ach_read = fs["alg_bytes_sweep"] / (ms / 1e3) / 1e9 if ms > 0 else None
    sweep_bytes = 16.0 * n + 32.0 * m_live
    ach = sweep_bytes / (ms / 1e3) / 1e9 if ms > 0 else None
    return {"kernel": "k_frontier<M_SWEEP> (each edge once, both directions)", "achieved": ach,
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBPS if ach else None,
            "sweep_alg_bytes": sweep_bytes, "achieved_bytes_read": ach_read,
            "frac_bytes_read": ach_read / HBM_PEAK_GBPS if ach_read else None, "pass_ms": ms,
            "alg_bytes": fs["alg_bytes_sweep"], "slots": fs["frontier_vertices"], "entries_scanned": fs["relax_edges"],
            "rounds_after": fs["relax_rounds"], "rounds_kernel_ms": fs["t_relax_kernel_ms"] - ms,
            "improved": fs["relax_improved"]}


def line(args, value, wall, steps, cfg, stats, check, cpu, unit="changed edges/s", extra=None, traffic=False,
         gather_rate=None, roof_stats=None):
    import numpy as np
    lat = [s["t_update_ms"] for s in stats]
    alg = sum(s["alg_bytes"] for s in stats)
    phases = {ph: sum(s[f"t_{ph}_ms"] for s in stats) / max(1, len(stats))
              for ph in ("upkeep", "insert", "classify", "invalidate", "seed", "relax")}
    out = {
        "metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": wall / max(1, steps) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic", "config": cfg,
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
        "update_alg_GBps": alg / (sum(lat) / 1e3) / 1e9 if sum(lat) > 0 else None,
        "dist_bitexact": bool(check.get("gpu_recompute_equal", True) and check.get("full_sweep_fixed_point", True)
                              and check.get("dijkstra_mismatches", 0) == 0 and check.get("certificate", 0) == 0),
        "check": check, "roofline": roofline(roof_stats or stats, traffic, gather_rate), "cpu_baseline": cpu,
        # (null when the updates ran without the events between phases)
        "phases_ms": phases if any(phases.values()) else None,
        "rounds": {"invalidate": sum(s["inv_rounds"] for s in stats) / max(1, len(stats)),
                   "relax": sum(s["relax_rounds"] for s in stats) / max(1, len(stats))},
        # per update: vertices invalidated (reset, then seeded), candidates kept by
        # the rescue, entries scanned by relaxation and by invalidation + seeding
        "work": {k: sum(s[k] for s in stats) / max(1, len(stats))
                 for k in ("n_invalidated", "n_rescued", "relax_edges", "other_edges", "frontier_vertices")},
    }
    if extra:
        out.update(extra)
    print(json.dumps(out), flush=True)


def timed_steps(ds, g, t, res, warmup, timing=True, warm=True):
    """Applies res[warmup:] (after res[:warmup] untimed when warm is set)
    between two device synchronisations; returns the stats and the wall time."""
    for i in range(warmup if warm else 0):
        ds.update(g, t, res[i])
    ds.set_tuning(kernel_timing=1 if timing else 0)
    stats = []
    t0 = time.perf_counter()
    for r in res[warmup:]:
        stats.append(ds.update(g, t, r))
    wall = time.perf_counter() - t0
    ds.set_tuning(kernel_timing=0)
    return stats, wall


def run_plain(args, ds, np, n, a, b, w, src, batches, label, cfg, traffic=False, reruns=False):
    gather_rate = gather_ceiling(ds)
    log(f"random 8-byte gather ceiling {gather_rate:.3e} accesses/s")
    g = ds.Graph(n, a, b, w)
    t = ds.Tree(g, src)
    extra = {}
    # The full edge pass on the graph as built (rows sorted, no delta entries),
    # beside the one on the final graph below; the tree is its fixed point.
    fs0, ms0, low0 = timed_passes(ds, t, not args.no_timing)
    extra["full_sweep_initial"] = full_sweep(fs0, n, g.info()["m_live"])
    extra["full_sweep_initial"].update(graph="as built, before any batch", passes_ms=ms0, improved=low0)
    log(f"{label}: full sweep of the graph as built {fs0['t_sweep_ms']:.3f} ms (median of {ms0}), "
        f"frac {extra['full_sweep_initial']['frac']}")
    res = [ds.Batch(g, c) for c in batches]
    if args.autotune and args.warmup > 0:
        t0 = time.perf_counter()
        tun, tms = ds.autotune(g, t, res[:args.warmup])
        log(f"{label}: autotuned in {time.perf_counter() - t0:.1f}s: {tun}, {tms:.3f} ms per warm-up batch")
        extra["autotuned"] = tun
    for r in res[:args.warmup]:
        ds.update(g, t, r)
    # the state before the timed steps, to run the same batches again
    snap = (g.clone(),)
    snap = (snap[0], ds.Tree(snap[0], src))
    snap[1].copy_from(t)
    stats, wall = timed_steps(ds, g, t, res, args.warmup, not args.no_timing, warm=False)
    k = len(batches[-1])
    value = k * args.steps / wall
    log(f"{label}: {args.steps} steps in {wall:.3f}s -> {value:.3e} changed edges/s, "
        f"p50 {pct([s['t_update_ms'] for s in stats], 50):.3f} ms")
    final = (g.clone(),)
    final = (final[0], ds.Tree(final[0], src))
    final[1].copy_from(t)
    same = []

    def rerun(rs, timing=False, phases=False):
        g.copy_from(snap[0])
        t.copy_from(snap[1])
        ds.set_tuning(phase_timing=1 if phases else 0)
        st_, wall_ = timed_steps(ds, g, t, rs, args.warmup, timing, warm=False)
        ds.set_tuning(phase_timing=0)
        same.append(bool(np.array_equal(np.stack(t.read()), np.stack(final[1].read()))))
        return st_, wall_

    if reruns:
        # The same batches from the same state without per-launch timing
        # events (whose dispatch-side recording adds a few microseconds to
        # every launch), then from host memory through dsssp_update, which adds
        # the PCIe copy of each batch (t_h2d_ms, outside t_update_ms).
        st2, wall2 = rerun(res)
        extra["value_without_launch_events"] = k * args.steps / wall2
        extra["p50_ms_without_launch_events"] = pct([s["t_update_ms"] for s in st2], 50)
        log(f"{label}: without launch events {extra['value_without_launch_events']:.3e} changed edges/s")
        st3, wall3 = rerun(batches)
        extra["value_pcie_inclusive"] = k * args.steps / wall3
        extra["t_h2d_ms"] = sum(s["t_h2d_ms"] for s in st3) / len(st3)
        log(f"{label}: from host memory {extra['value_pcie_inclusive']:.3e} changed edges/s, "
            f"H2D {extra['t_h2d_ms']:.3f} ms per batch")
    # The phases of the same updates: once more with an event between each
    # two phases (phase_timing; the timed steps above record only the start
    # and end of each update, as the library does by default).
    st4, wall4 = rerun(res, timing=not args.no_timing, phases=True)
    extra["phases_ms"] = {ph: sum(s[f"t_{ph}_ms"] for s in st4) / len(st4)
                          for ph in ("upkeep", "insert", "classify", "invalidate", "seed", "relax")}
    extra["value_with_phase_events"] = k * args.steps / wall4
    log(f"{label}: with phase events {extra['value_with_phase_events']:.3e} changed edges/s, "
        f"phases {extra['phases_ms']}")
    extra["reruns_equal"] = all(same)
    g.copy_from(final[0])
    t.copy_from(final[1])
    for o in (final[1], final[0], snap[1], snap[0]):
        o.close()
    for r in res:
        r.close()
    # One full relaxation sweep over the converged tree: every live entry of
    # every reachable vertex, rows in vertex order.  Nothing may change (a
    # device-side fixed-point check); its launch is the kernel's streaming case.
    fs, ms, low = timed_passes(ds, t, not args.no_timing)
    extra["full_sweep"] = full_sweep(fs, n, g.info()["m_live"])
    extra["full_sweep"].update(graph="after the warm-up and timed batches", passes_ms=ms, improved=low)
    log(f"{label}: full sweep {fs['t_sweep_ms']:.3f} ms (median of {ms}), "
        f"{extra['full_sweep']['achieved']} GB/s algorithmic, improved {fs['relax_improved']}")
    par, dist = t.read()
    t0 = time.perf_counter()
    ft = ds.Tree(g, src)  # from scratch on the GPU: the recompute the update replaces
    recompute_ms = (time.perf_counter() - t0) * 1e3
    fpar, fdist = ft.read()
    del ft
    check = {"gpu_recompute_equal": bool(np.array_equal(dist, fdist) and np.array_equal(par, fpar)),
             "full_sweep_fixed_point": low == 0 and low0 == 0}
    extra["gpu_recompute_ms"] = recompute_ms
    log(f"{label}: GPU tree from scratch on the final graph {recompute_ms:.2f} ms")
    cpu = None
    if not args.no_cpu:
        c2, cpu = cpu_check(ds, np, g, n, [src], par, dist, k, label)
        check.update(c2)
    line(args, value, wall, args.steps, cfg, stats, check, cpu, extra=extra, traffic=traffic, gather_rate=gather_rate)


def config1(args, ds, np):
    n, m, k = 100_000, 1_000_000, 10_000
    a, b, w = ds.gen_tree_plus_random(n, m, 1, 100, seed=args.seed)
    batches = evolving_batches(ds, np, n, a, b, w, k, 0.5, args.warmup + args.steps, args.seed + 1, 1, 100,
                               first_deletable=n - 1)
    run_plain(args, ds, np, n, a, b, w, 0, batches, "config1",
              {"workload": "random tree+pairs n=1e5 m=1e6 k=1e4 ins=0.5", "n": n, "m": m, "k": k, "src": 0})


def config2(args, ds, np):
    t0 = time.time()
    a, b, w = ds.gen_rmat(args.scale, args.edge_factor, (0.57, 0.19, 0.19), 1, 1000, seed=args.seed)
    n = 1 << args.scale
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    src = int(np.argmax(deg))  # ties -> smallest id
    log(f"rmat scale={args.scale} n={n} m={len(a)} max_deg={deg.max()} src={src} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    nb = args.warmup + args.steps
    batches = evolving_batches(ds, np, n, a, b, w, args.k, args.ins, nb, args.seed + 1)
    log(f"{nb} batches of k={args.k} generated ({time.time() - t0:.1f}s)")
    cfg = {"workload": f"rmat-s{args.scale}-ef{args.edge_factor} batch k={args.k} ins={args.ins:g}",
           "scale": args.scale, "n": n, "m": len(a), "k": args.k, "ins_frac": args.ins, "src": src,
           "weights": "U[1,1000]"}
    default = args.scale == 24 and args.edge_factor == 16 and args.k == 1_000_000 and args.ins == 0.5
    run_plain(args, ds, np, n, a, b, w, src, batches, f"rmat-s{args.scale}", cfg, traffic=default,
              reruns=not args.no_timing)


def config3(args, ds, np):
    side = args.side
    n = side * side
    a, b, w = ds.gen_grid(side, 0.01, 1, 10000, seed=args.seed)
    g0 = ds.Graph(n, a, b, w)
    t0 = ds.Tree(g0, 0)
    par0, _ = t0.read()
    k = 100_000
    batches = [ds.gen_changes_grid(side, a, b, w, par0, k, 0.7, 0.5, 3, 1, 10000, seed=args.seed + i)
               for i in range(args.warmup + args.steps)]
    # every step starts from the same graph and tree (restored outside the timing)
    g = g0.clone()
    t = ds.Tree(g, parent=par0, dist=t0.read()[1])  # a tree on the clone; its state is restored per step
    # Each timed batch runs twice from the restored state: without launch
    # events (the update's time, ~2000 rounds each) and with them (the
    # relaxation kernel's duration for the roofline); both must end equal.
    stats, tstats = [], []
    same = True
    ds.set_tuning(phase_timing=1)  # a few us of events on updates of ~0.1 s
    for i, c in enumerate(batches):
        timed = i >= args.warmup
        for with_events in ((False, True) if timed and not args.no_timing else (False,)):
            g.copy_from(g0)
            t.copy_from(t0)
            bt = ds.Batch(g, c)
            ds.set_tuning(kernel_timing=1 if with_events else 0)
            st = ds.update(g, t, bt)
            bt.close()
            if not timed:
                continue
            if with_events:
                tstats.append(st)
                same &= bool(np.array_equal(np.stack(t.read()), first))
            else:
                stats.append(st)
                first = np.stack(t.read())
    ds.set_tuning(kernel_timing=0, phase_timing=0)
    dev = sum(s["t_update_ms"] for s in stats) / 1e3
    value = k * args.steps / dev
    log(f"grid {side}: {value:.3e} changed edges/s (device time), p50 {pct([s['t_update_ms'] for s in stats], 50):.2f} ms"
        f", rounds inv {stats[-1]['inv_rounds']} relax {stats[-1]['relax_rounds']}")
    par, dist = t.read()
    check = {}
    cpu = None
    if not args.no_cpu:
        check, cpu = cpu_check(ds, np, g, n, [0], par, dist, k, f"grid-{side}")
    check["launch_event_rerun_equal"] = same
    line(args, value, dev, args.steps, {"workload": f"grid {side}x{side} drop 1% k=1e5 del 70% (half tree)",
                                          "n": n, "m": len(a), "k": k, "src": 0,
                                          "note": "value = k / device update time; state restored between steps; "
                                                  "roofline from a rerun of each batch with launch events"},
         stats, check, cpu, roof_stats=tstats or None)


def config4a(args, ds, np):
    scale, K, k = min(args.scale, 22), args.trees, 1_000_000 if args.k == 1_000_000 else args.k
    a, b, w = ds.gen_rmat(scale, 16, (0.57, 0.19, 0.19), 1, 1000, seed=args.seed)
    n = 1 << scale
    rng = np.random.default_rng(args.seed)
    srcs = rng.integers(0, n, K).astype(np.int32)
    batches = evolving_batches(ds, np, n, a, b, w, k, 0.5, args.warmup + args.steps, args.seed + 1)
    g = ds.Graph(n, a, b, w)
    f = ds.Tree(g, sources=srcs)
    res = [ds.Batch(g, c) for c in batches]
    tuned = None
    base_tuning = ds.get_tuning()  # the plain comparison runs with these, not the forest's tuned ones
    if args.autotune and args.warmup > 0:
        t0 = time.perf_counter()
        tuned, tms = ds.autotune(g, f, res[:args.warmup])
        log(f"forest: autotuned in {time.perf_counter() - t0:.1f}s: {tuned}, {tms:.2f} ms per warm-up batch")
    # upkeep is reported apart for the forest and the plain trees: both run
    # with the events between phases
    ds.set_tuning(phase_timing=1)
    stats, wall = timed_steps(ds, g, f, res, args.warmup, not args.no_timing)
    for r in res:
        r.close()
    # plain comparison: every tree alone (one tree per graph, a device copy of
    # the initial graph each), the same batches, device time per update
    fr = [f.read(j) for j in range(K)]
    g0 = ds.Graph(n, a, b, w)
    plain_ms = plain_upkeep_ms = 0.0
    same = True
    forest_tuning = ds.get_tuning()
    # a fresh graph per tree: captures would only add instantiation time
    ds.set_tuning(**{**base_tuning, "use_graph": 0, "kernel_timing": 0, "phase_timing": 1})
    for j in range(K):
        gp = g0.clone()
        tp = ds.Tree(gp, int(srcs[j]))
        for i, c in enumerate(batches):
            st = ds.update(gp, tp, c)
            if i >= args.warmup:
                plain_ms += st["t_update_ms"]
                plain_upkeep_ms += st["t_upkeep_ms"]
        pp, pd = tp.read()
        same &= bool(np.array_equal(pp, fr[j][0]) and np.array_equal(pd, fr[j][1]))
        tp.close()
        gp.close()
    g0.close()
    ds.set_tuning(**forest_tuning)
    plain_per_tree = plain_ms / (K * args.steps)
    tree_updates = k * K * args.steps
    value = tree_updates / wall
    log(f"forest K={K} s{scale}: {value:.3e} tree-edge updates/s; plain {plain_per_tree:.2f} ms/tree/batch")
    line(args, value, wall, args.steps,
         {"workload": f"rmat-s{scale} K={K} trees, shared batch k={k} ins=0.5", "n": n, "m": len(a), "k": k,
          "trees": K}, stats, {"every_tree_equals_plain": same}, None, unit="changed edges x trees /s",
         extra={"speedup_vs_plain": (plain_ms / 1e3) / wall, "plain_ms_per_step": plain_ms / args.steps,
                # the same without the resident-CSR upkeep (a repack of each
                # plain tree's own graph copy against one of the shared graph)
                "speedup_vs_plain_excl_upkeep": ((plain_ms - plain_upkeep_ms) / 1e3) /
                                                 (wall - sum(s["t_upkeep_ms"] for s in stats) / 1e3),
                "plain_upkeep_ms_per_step": plain_upkeep_ms / args.steps,
                "forest_upkeep_ms_per_step": sum(s["t_upkeep_ms"] for s in stats) / args.steps,
                "autotuned": tuned,
                "use_graph": ds.get_tuning()["use_graph"]})


def config4b(args, ds, np):
    rng = np.random.default_rng(args.seed)
    parts, off = [], 0
    for i in range(args.graphs):
        n_i = int(math.exp(rng.uniform(math.log(2 ** 14), math.log(2 ** 21))))
        deg = math.exp(rng.uniform(math.log(4), math.log(32)))
        m_i = max(n_i - 1, int(n_i * deg / 2))
        a, b, w = ds.gen_tree_plus_random(n_i, m_i, 1, 1000, seed=args.seed + 10 * i)
        c = ds.gen_changes(n_i, a, b, w, max(1, m_i // 100), 0.5, 1, 1000, seed=args.seed + 10 * i + 1)
        parts.append((n_i, a, b, w, c, int(rng.integers(0, n_i)), off))
        off += n_i
    N = off
    A = np.concatenate([p[1] + p[6] for p in parts])
    B = np.concatenate([p[2] + p[6] for p in parts])
    W = np.concatenate([p[3] for p in parts])
    C = np.concatenate([p[4] + np.array([p[6], p[6], 0, 0], np.int32) for p in parts])
    roots = np.array([p[5] + p[6] for p in parts], np.int32)
    g0 = ds.Graph(N, A, B, W)
    t0 = ds.Tree(g0, roots=roots)
    g = g0.clone()
    t = ds.Tree(g, roots=roots)  # a tree on the clone; its state is restored per step
    tuned = None
    base_tuning = ds.get_tuning()  # the plain comparison runs with these, not the merged run's tuned ones
    if args.autotune:
        bt = ds.Batch(g, C)
        tq = time.perf_counter()
        tuned, tms = ds.autotune(g, t, [bt])
        bt.close()
        log(f"32 graphs: autotuned in {time.perf_counter() - tq:.1f}s: {tuned}, {tms:.2f} ms per batch")
    stats = []
    for i in range(args.warmup + args.steps):
        g.copy_from(g0)
        t.copy_from(t0)
        bt = ds.Batch(g, C)
        ds.set_tuning(kernel_timing=0 if args.no_timing or i < args.warmup else 1)
        st = ds.update(g, t, bt)
        bt.close()
        if i >= args.warmup:
            stats.append(st)
    ds.set_tuning(kernel_timing=0)
    dev = sum(s["t_update_ms"] for s in stats) / 1e3
    # plain: each graph separately, one launch sequence per graph (no captured
    # round groups: on a fresh graph they would only add instantiation time)
    plain = 0.0
    par, dist = t.read()
    same = True
    merged_tuning = ds.get_tuning()
    ds.set_tuning(**{**base_tuning, "use_graph": 0, "kernel_timing": 0})
    for n_i, a, b, w, c, s, o in parts:
        gp = ds.Graph(n_i, a, b, w)
        tp = ds.Tree(gp, s)
        st = ds.update(gp, tp, c)
        plain += st["t_update_ms"]
        pp, pd = tp.read()
        sp = par[o:o + n_i].astype(np.int64)
        sp[sp >= 0] -= o
        same &= bool(np.array_equal(sp, pp) and np.array_equal(dist[o:o + n_i], pd))
        gp.close()
    ds.set_tuning(**merged_tuning)
    value = len(C) * args.steps / dev
    log(f"{args.graphs} graphs in one launch: {value:.3e} changed edges/s; plain total {plain:.2f} ms vs "
        f"{dev / args.steps * 1e3:.2f} ms")
    line(args, value, dev, args.steps, {"workload": f"{args.graphs} random graphs n~logU[2^14,2^21] 1% batches",
                                          "N": N, "M": len(A), "k_total": len(C)}, stats,
         {"each_equals_plain": same}, None,
         extra={"speedup_vs_plain": plain / (dev / args.steps * 1e3), "autotuned": tuned,
                "use_graph": ds.get_tuning()["use_graph"]})


def served_requests(ds, np, rng, insts, count):
    """count requests: k log-uniform in [1, 1e5], insertion ratio U[0, 1],
    uniform target instance.  Per instance, one generator call gives distinct
    deletions of its initial edges and distinct absent pairs to insert, split
    over its requests in order."""
    plan = []
    for _ in range(count):
        i = int(rng.integers(0, len(insts)))
        k = int(math.exp(rng.uniform(0, math.log(100_000))))
        ki = int(round(k * float(rng.uniform(0, 1))))
        plan.append((i, k - ki, ki))
    reqs = [None] * count
    for i, (n, a, b, w, _) in enumerate(insts):
        mine = [j for j, p in enumerate(plan) if p[0] == i]
        kd, ki = sum(plan[j][1] for j in mine), sum(plan[j][2] for j in mine)
        if kd + ki == 0:
            continue
        allc = ds.gen_changes(n, a, b, w, kd + ki, ki / (kd + ki), 1, 1000, seed=9000 + i)
        dels, ins = allc[allc[:, 3] == 0], allc[allc[:, 3] == 1]
        od = oi = 0
        for j in mine:
            c = np.concatenate([dels[od:od + plan[j][1]], ins[oi:oi + plan[j][2]]])
            od += plan[j][1]
            oi += plan[j][2]
            rng.shuffle(c)
            reqs[j] = (i, np.ascontiguousarray(c))
    return [r for r in reqs if r is not None and len(r[1])]


def serve_run(srv, np, reqs, rate, rng):
    """Submits reqs (Poisson arrivals at `rate` per second, or all at once when
    rate is 0) and waits for all; returns (wall seconds, request stats)."""
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate > 0 else np.zeros(len(reqs))
    tickets = []
    t0 = time.perf_counter()
    tnext = t0
    for (i, c), gap in zip(reqs, gaps):
        tnext += gap
        d = tnext - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(srv.submit(i, c))
    rs = [srv.wait(tk) for tk in tickets]
    return time.perf_counter() - t0, rs


def config5(args, ds, np):
    rng = np.random.default_rng(args.seed)
    if args.serve_scale:
        scales = [args.serve_scale] * args.instances
    else:
        scales = [20 + (i * 5) // args.instances for i in range(args.instances)]  # 20..24 spread evenly
        if args.scale < 24:
            scales = [min(s, args.scale) for s in scales]
    budget = 160 << 30
    srv = ds.Server(mem_budget_bytes=budget, lanes=args.lanes)
    insts = []
    for i, sc in enumerate(scales):
        a, b, w = ds.gen_rmat(sc, 16, (0.57, 0.19, 0.19), 1, 1000, seed=args.seed + i)
        n = 1 << sc
        deg = np.bincount(np.concatenate([a, b]), minlength=n)
        src = int(np.argmax(deg))
        try:
            srv.add_instance(n, a, b, w, src)
        except ds.DssspError as e:  # admission by memory budget
            log(f"instance {i}: s{sc} not admitted ({e})")
            continue
        insts.append((n, a, b, w, src))
        log(f"instance {i}: s{sc} m={len(a)}")
    t0 = time.time()
    srv.start()
    log(f"server started ({time.time() - t0:.1f}s)")
    reqs = served_requests(ds, np, rng, insts, 2 * args.requests)
    poisson, burst = reqs[:len(reqs) // 2], reqs[len(reqs) // 2:]
    log(f"{len(reqs)} requests generated")
    # 1. a Poisson stream below capacity; 2. the saturation run (all at once)
    wall_p, rs_p = serve_run(srv, np, poisson, args.rate, rng)
    wall_s, rs_s = serve_run(srv, np, burst, 0.0, rng)
    rec_p, rec_s = sum(len(c) for _, c in poisson), sum(len(c) for _, c in burst)
    lat_p, lat_s = [r["latency_ms"] for r in rs_p], [r["latency_ms"] for r in rs_s]
    served = rec_s / wall_s
    log(f"poisson {args.rate}/s: {rec_p / wall_p:.3e} changed edges/s, p50 {pct(lat_p, 50):.2f} p99 {pct(lat_p, 99):.2f} ms"
        f"; saturated: {served:.3e}, p50 {pct(lat_s, 50):.2f} p99 {pct(lat_s, 99):.2f} ms")
    # plain: the same requests, in order, one update each on per-instance graphs
    plain_ms = {"poisson": 0.0, "burst": 0.0}
    same = True
    for i in range(len(insts)):
        n, a, b, w, src = insts[i]
        gp = ds.Graph(n, a, b, w)
        tp = ds.Tree(gp, src)
        for part, rr in (("poisson", poisson), ("burst", burst)):
            for j, c in rr:
                if j == i:
                    plain_ms[part] += ds.update(gp, tp, c)["t_update_ms"]
        pp, pd = tp.read()
        sp, sd = srv.read(i)
        same &= bool(np.array_equal(pp, sp) and np.array_equal(pd, sd))
        tp.close()
        gp.close()
    plain = rec_s / (plain_ms["burst"] / 1e3)
    info = srv.info()
    srv.close()
    log(f"served {served:.3e} vs plain {plain:.3e} changed edges/s (x{served / plain:.2f}); equal to plain: {same}")
    out = {"metric": METRIC, "value": served, "unit": "changed edges/s", "n_gpus": 1, "steps": len(burst),
           "warmup": 0, "ms_per_step": wall_s / len(burst) * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32", "data": "synthetic",
           "config": {"workload": f"served {len(insts)} rmat instances s{min(scales)}-s{max(scales)}, "
                                  f"requests k~logU[1,1e5] ins~U[0,1]", "lanes": args.lanes, "scales": scales,
                      "budget_bytes": budget},
           "p50_ms": pct(lat_s, 50), "p99_ms": pct(lat_s, 99),
           "poisson": {"rate_per_s": args.rate, "requests": len(poisson), "records": rec_p,
                       "changed_edges_per_s": rec_p / wall_p, "p50_ms": pct(lat_p, 50), "p99_ms": pct(lat_p, 99),
                       "mean_merged": float(np.mean([r["merged"] for r in rs_p])),
                       "plain_changed_edges_per_s": rec_p / (plain_ms["poisson"] / 1e3)},
           "saturated": {"requests": len(burst), "records": rec_s, "changed_edges_per_s": served,
                         "p50_ms": pct(lat_s, 50), "p99_ms": pct(lat_s, 99),
                         "mean_merged": float(np.mean([r["merged"] for r in rs_s]))},
           "plain_changed_edges_per_s": plain, "speedup_vs_plain": served / plain,
           "dist_bitexact": same, "check": {"each_equals_plain": same}, "server": info,
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the update path is single-GPU; "
              "multi-GPU sharding is a design note, not an implemented path)", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import dynamicssspv2_amd as ds
    ds.set_device(0)
    if args.probe:
        # kind 0: 1 GiB read + 1 GiB written; kind 1: 2^26 reads of 8 bytes at
        # hashed positions of 8 GiB (beyond every cache) and of 64 MiB (held in
        # the last-level cache)
        for kind, nbytes, count in ((0, 1 << 30, 0), (1, 8 << 30, 1 << 26), (1, 64 << 20, 1 << 26)):
            ms = ds.probe(kind, nbytes, count, reps=5)
            log(f"probe kind {kind} bytes {nbytes} count {count}: {ms:.3f} ms per launch")
        return
    knobs = {k: getattr(args, k) for k in ("block", "grid", "heavy_degree", "chunk", "round_batch", "use_graph",
                                           "edges_per_thread", "tile_entries", "local_tiles", "delta", "sweep",
                                           "rescue_degree", "rescue_levels")
             if getattr(args, k) is not None}
    if knobs:
        ds.set_tuning(**knobs)
    log(f"tuning {ds.get_tuning()}")
    {"1": config1, "2": config2, "3": config3, "4a": config4a, "4b": config4b, "5": config5}[args.config](args, ds, np)


if __name__ == "__main__":
    main()
