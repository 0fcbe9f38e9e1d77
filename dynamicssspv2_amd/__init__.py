"""MI355X (gfx950) dynamic SSSP update -- Python binding of libdsssp.so.

The reference (ArindamK92/DynamicSSSPv2) exposes a CLI and three text formats
(``op_main <graph> <n> <m> <sssp> <changes>``, README.md:13-37).  This package
binds the C ABI in ``include/dsssp.h`` with ctypes and numpy:

* formats: :func:`read_graph`, :func:`read_changes`, :func:`read_sssp`,
  :func:`write_sssp` (all_structure_undir.cuh:89-198);
* workloads: :func:`gen_tree_plus_random`, :func:`gen_rmat`, :func:`gen_grid`,
  :func:`gen_changes`;
* device objects: :class:`Graph` (CSR resident in HBM), :class:`Tree` (SSSP
  tree, optionally computed on the GPU like op_ssspSeq,
  seqSSSPwithDist_undir.cpp:66-128), :class:`Batch` (change batch in HBM) and
  :func:`update` (STEP1 + STEP2 of SSSPmainForUndirected.cu:196-292).

The compute path is the HIP library only: if ``libdsssp.so`` is missing the
import of any device entry point raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsssp.so")
CLI_PATH = os.path.join(HERE, "op_main")
INF = np.iinfo(np.int64).max

OK, ERR_ARG, ERR_RANGE, ERR_WEIGHT, ERR_OVERFLOW, ERR_NOMEM, ERR_HIP, ERR_NODEV = range(8)
ERR_IO, ERR_FORMAT, ERR_TREE, ERR_BUDGET, ERR_STATE = range(8, 13)
TREE_CANONICAL = 1

CHANGE_DTYPE = np.dtype([("a", "<i4"), ("b", "<i4"), ("w", "<i4"), ("op", "<i4")])


class DssspError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in (
        "n_insert", "n_delete", "n_invalidated", "inv_rounds", "relax_rounds", "relax_launches",
        "frontier_vertices", "relax_edges", "relax_improved", "relax_enqueued", "other_edges", "probes")] + \
        [(f, ctypes.c_double) for f in (
            "t_h2d_ms", "t_update_ms", "t_insert_ms", "t_classify_ms", "t_invalidate_ms", "t_seed_ms",
            "t_relax_ms", "t_relax_kernel_ms", "alg_bytes", "alg_bytes_relax")] + \
        [("relax_sweeps", ctypes.c_int64), ("t_sweep_ms", ctypes.c_double), ("alg_bytes_sweep", ctypes.c_double),
         ("t_upkeep_ms", ctypes.c_double), ("n_rescued", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class GraphUpkeep(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("base_entries", ctypes.c_int64), ("delta_entries", ctypes.c_int64),
                ("dead_entries", ctypes.c_int64), ("pool_top", ctypes.c_uint64), ("pool_cap", ctypes.c_uint64),
                ("emap_slots", ctypes.c_uint64), ("repacks", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("mem_budget_bytes", ctypes.c_uint64), ("max_merge", ctypes.c_int32), ("lanes", ctypes.c_int32),
                ("max_records", ctypes.c_int64)]


class RequestStats(ctypes.Structure):
    _fields_ = [("instance", ctypes.c_int32), ("merged", ctypes.c_int32), ("records", ctypes.c_int64),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("update_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerTotals(ctypes.Structure):
    _fields_ = [("instances", ctypes.c_int32), ("pad0", ctypes.c_int32), ("budget_bytes", ctypes.c_uint64),
                ("used_bytes", ctypes.c_uint64), ("launches", ctypes.c_int64), ("records", ctypes.c_int64),
                ("requests", ctypes.c_int64), ("pending", ctypes.c_int64), ("busy_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad0"}


DELTA_AUTO = -2  # Tuning.delta: ordered relaxation chosen per tree (include/dsssp.h)
RESCUE_AUTO = -2  # Tuning.rescue_degree: invalidation rescue chosen per batch (include/dsssp.h)


class Tuning(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int32) for f in ("block", "grid", "heavy_degree", "chunk", "round_batch",
                                              "kernel_timing", "use_graph", "edges_per_thread",
                                              "tile_entries", "local_tiles", "delta", "sweep", "give_up_ms",
                                              "rescue_degree", "rescue_levels", "phase_timing")]


def build(force: bool = True) -> None:
    """Compile libdsssp.so and op_main for gfx950 (hipcc; no GPU needed)."""
    args = ["make", "-s", "-C", os.path.join(HERE, "csrc"), f"-j{min(8, os.cpu_count() or 1)}"]
    if force:
        args.insert(2, "-B")
    subprocess.check_call(args)


_lib = None


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with dynamicssspv2_amd.build() "
                           "(there is no fallback compute path)")
    L = ctypes.CDLL(LIB_PATH)
    i64, i32, u64, u32, p, pp = (ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint32,
                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p))
    P64 = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "dsssp_last_error": ([], ctypes.c_char_p),
        "dsssp_version": ([], ctypes.c_char_p),
        "dsssp_set_device": ([ctypes.c_int], ctypes.c_int),
        "dsssp_get_tuning": ([ctypes.POINTER(Tuning)], ctypes.c_int),
        "dsssp_set_tuning": ([ctypes.POINTER(Tuning)], ctypes.c_int),
        "dsssp_free": ([p], None),
        "dsssp_probe": ([i32, u64, u64, i32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "dsssp_read_graph": ([ctypes.c_char_p, i64, P64, pp, pp, pp], ctypes.c_int),
        "dsssp_read_changes": ([ctypes.c_char_p, i64, P64, pp], ctypes.c_int),
        "dsssp_read_sssp": ([ctypes.c_char_p, i64, p, p], ctypes.c_int),
        "dsssp_write_sssp": ([ctypes.c_char_p, i64, p, p], ctypes.c_int),
        "dsssp_gen_tree_plus_random": ([i64, i64, i32, i32, u64, p, p, p], ctypes.c_int),
        "dsssp_gen_rmat": ([i32, i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, i32, i32, u64, P64,
                            pp, pp, pp], ctypes.c_int),
        "dsssp_gen_grid": ([i32, ctypes.c_double, i32, i32, u64, P64, pp, pp, pp], ctypes.c_int),
        "dsssp_gen_changes": ([i64, i64, p, p, p, i64, ctypes.c_double, i32, i32, i64, u64, p], ctypes.c_int),
        "dsssp_gen_changes_grid": ([i32, i64, p, p, p, p, i64, ctypes.c_double, ctypes.c_double, i32, i32, i32,
                                    u64, p], ctypes.c_int),
        "dsssp_graph_create": ([i64, i64, p, p, p, u32, pp], ctypes.c_int),
        "dsssp_graph_destroy": ([p], ctypes.c_int),
        "dsssp_graph_info": ([p, P64, P64, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "dsssp_graph_export": ([p, P64, p, p, p], ctypes.c_int),
        "dsssp_graph_stats": ([p, ctypes.POINTER(GraphUpkeep)], ctypes.c_int),
        "dsssp_graph_repack": ([p], ctypes.c_int),
        "dsssp_graph_clone": ([p, pp], ctypes.c_int),
        "dsssp_graph_copy": ([p, p], ctypes.c_int),
        "dsssp_tree_create": ([p, i32, p, p, u32, pp], ctypes.c_int),
        "dsssp_tree_destroy": ([p], ctypes.c_int),
        "dsssp_tree_read": ([p, p, p], ctypes.c_int),
        "dsssp_tree_copy": ([p, p], ctypes.c_int),
        "dsssp_tree_relax_full": ([p, ctypes.POINTER(Stats)], ctypes.c_int),
        "dsssp_tree_create_roots": ([p, i32, p, u32, pp], ctypes.c_int),
        "dsssp_forest_create": ([p, i32, p, u32, pp], ctypes.c_int),
        "dsssp_tree_count": ([p, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "dsssp_forest_read": ([p, i32, p, p], ctypes.c_int),
        "dsssp_update": ([p, p, p, i64, ctypes.POINTER(Stats)], ctypes.c_int),
        "dsssp_batch_upload": ([p, p, i64, pp], ctypes.c_int),
        "dsssp_batch_destroy": ([p], ctypes.c_int),
        "dsssp_update_batch": ([p, p, p, ctypes.POINTER(Stats)], ctypes.c_int),
        "dsssp_update_noalloc": ([p, p, p, i64, ctypes.POINTER(Stats)], ctypes.c_int),
        "dsssp_autotune": ([p, p, pp, i32, i32, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_double)],
                           ctypes.c_int),
        "dsssp_server_create": ([ctypes.POINTER(ServerConfig), pp], ctypes.c_int),
        "dsssp_server_add_instance": ([p, i64, i64, p, p, p, i32, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
        "dsssp_server_remove_instance": ([p, i32], ctypes.c_int),
        "dsssp_server_start": ([p], ctypes.c_int),
        "dsssp_server_submit": ([p, i32, p, i64, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "dsssp_server_wait": ([p, ctypes.c_uint64, ctypes.POINTER(RequestStats)], ctypes.c_int),
        "dsssp_server_read": ([p, i32, p, p], ctypes.c_int),
        "dsssp_server_info": ([p, ctypes.POINTER(ServerTotals)], ctypes.c_int),
        "dsssp_server_destroy": ([p], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != OK:
        raise DssspError(rc, lib().dsssp_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _take(ptr: ctypes.c_void_p, count: int, dtype) -> np.ndarray:
    """Copies a library-allocated array into numpy and frees it."""
    try:
        if count == 0:
            return np.zeros(0, dtype)
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dtype).copy()
    finally:
        lib().dsssp_free(ptr)


def as_changes(chg) -> np.ndarray:
    """(k,4) int32 array or structured array -> contiguous CHANGE_DTYPE records."""
    if isinstance(chg, np.ndarray) and chg.dtype == CHANGE_DTYPE:
        return np.ascontiguousarray(chg)
    c = _i32(chg).reshape(-1, 4)
    return np.ascontiguousarray(c).view(CHANGE_DTYPE).reshape(-1)


def changes_array(chg) -> np.ndarray:
    """CHANGE_DTYPE records -> (k,4) int32."""
    return as_changes(chg).view(np.int32).reshape(-1, 4)


# ----------------------------------------------------------------------------- formats

def read_graph(path: str, n: int = 0):
    m = ctypes.c_int64()
    a, b, w = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _check(lib().dsssp_read_graph(path.encode(), n, ctypes.byref(m), ctypes.byref(a), ctypes.byref(b),
                                  ctypes.byref(w)))
    return _take(a, m.value, np.int32), _take(b, m.value, np.int32), _take(w, m.value, np.int32)


def read_changes(path: str, n: int = 0) -> np.ndarray:
    k = ctypes.c_int64()
    c = ctypes.c_void_p()
    _check(lib().dsssp_read_changes(path.encode(), n, ctypes.byref(k), ctypes.byref(c)))
    return _take(c, k.value * 4, np.int32).reshape(-1, 4)


def read_sssp(path: str, n: int):
    par = np.empty(n, np.int32)
    dist = np.empty(n, np.int64)
    _check(lib().dsssp_read_sssp(path.encode(), n, _ptr(par), _ptr(dist)))
    return par, dist


def write_sssp(path: str, parent, dist) -> None:
    par = _i32(parent)
    d = np.ascontiguousarray(np.asarray(dist, np.int64))
    _check(lib().dsssp_write_sssp(path.encode(), len(par), _ptr(par), _ptr(d)))


# ----------------------------------------------------------------------------- workloads

def gen_tree_plus_random(n: int, m: int, wlo: int, whi: int, seed: int):
    a = np.empty(m, np.int32)
    b = np.empty(m, np.int32)
    w = np.empty(m, np.int32)
    _check(lib().dsssp_gen_tree_plus_random(n, m, wlo, whi, seed, _ptr(a), _ptr(b), _ptr(w)))
    return a, b, w


def gen_rmat(scale: int, edge_factor: int = 16, abc=(0.57, 0.19, 0.19), wlo: int = 1, whi: int = 1000,
             seed: int = 2):
    m = ctypes.c_int64()
    a, b, w = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _check(lib().dsssp_gen_rmat(scale, edge_factor, abc[0], abc[1], abc[2], wlo, whi, seed, ctypes.byref(m),
                                ctypes.byref(a), ctypes.byref(b), ctypes.byref(w)))
    return _take(a, m.value, np.int32), _take(b, m.value, np.int32), _take(w, m.value, np.int32)


def gen_grid(side: int, drop: float = 0.01, wlo: int = 1, whi: int = 10000, seed: int = 3):
    m = ctypes.c_int64()
    a, b, w = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _check(lib().dsssp_gen_grid(side, drop, wlo, whi, seed, ctypes.byref(m), ctypes.byref(a), ctypes.byref(b),
                                ctypes.byref(w)))
    return _take(a, m.value, np.int32), _take(b, m.value, np.int32), _take(w, m.value, np.int32)


def gen_changes(n: int, a, b, w, k: int, ins_frac: float, wlo: int, whi: int, seed: int,
                first_deletable: int = 0) -> np.ndarray:
    a, b, w = _i32(a), _i32(b), _i32(w)
    out = np.empty((k, 4), np.int32)
    _check(lib().dsssp_gen_changes(n, len(a), _ptr(a), _ptr(b), _ptr(w), k, ins_frac, wlo, whi,
                                   first_deletable, seed, _ptr(out)))
    return out


def gen_changes_grid(side: int, a, b, w, parent, k: int, del_frac: float = 0.7, tree_frac: float = 0.5,
                     max_manhattan: int = 3, wlo: int = 1, whi: int = 10000, seed: int = 5) -> np.ndarray:
    a, b, w, p = _i32(a), _i32(b), _i32(w), _i32(parent)
    out = np.empty((k, 4), np.int32)
    _check(lib().dsssp_gen_changes_grid(side, len(a), _ptr(a), _ptr(b), _ptr(w), _ptr(p), k, del_frac, tree_frac,
                                        max_manhattan, wlo, whi, seed, _ptr(out)))
    return out


# ----------------------------------------------------------------------------- tuning

def get_tuning() -> dict:
    t = Tuning()
    _check(lib().dsssp_get_tuning(ctypes.byref(t)))
    return {f: getattr(t, f) for f, _ in Tuning._fields_}


def set_tuning(**kw) -> None:
    cur = get_tuning()
    cur.update(kw)
    t = Tuning(**cur)
    _check(lib().dsssp_set_tuning(ctypes.byref(t)))


def probe(kind: int, nbytes: int, count: int = 0, reps: int = 5) -> float:
    """Launches of known footprint for counter calibration (kind 0: copy of
    nbytes; kind 1: count 8-byte reads at hashed positions of nbytes); returns
    the device ms per launch."""
    ms = ctypes.c_double()
    _check(lib().dsssp_probe(kind, nbytes, count, reps, ctypes.byref(ms)))
    return ms.value


def set_device(device: int) -> None:
    _check(lib().dsssp_set_device(device))


# ----------------------------------------------------------------------------- device objects

class Graph:
    """Undirected graph resident in HBM (sorted CSR + delta rows for insertions)."""

    def __init__(self, n: int, a, b, w, _handle=None):
        self.n = int(n)
        if _handle is not None:
            self._h = _handle
            return
        a, b, w = _i32(a), _i32(b), _i32(w)
        if not (len(a) == len(b) == len(w)):
            raise ValueError("a, b, w must have the same length")
        h = ctypes.c_void_p()
        _check(lib().dsssp_graph_create(self.n, len(a), _ptr(a), _ptr(b), _ptr(w), 0, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().dsssp_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    def info(self) -> dict:
        n, m = ctypes.c_int64(), ctypes.c_int64()
        by = ctypes.c_uint64()
        _check(lib().dsssp_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(by)))
        return {"n": n.value, "m_live": m.value, "device_bytes": by.value}

    def upkeep(self) -> dict:
        """Row entries scanned (base + delta), tombstones since the last repack,
        pool and edge-map sizes, repacks so far."""
        u = GraphUpkeep()
        _check(lib().dsssp_graph_stats(self._h, ctypes.byref(u)))
        return u.as_dict()

    def repack(self) -> None:
        """Drops every tombstoned entry and compacts the delta pool (on the device)."""
        _check(lib().dsssp_graph_repack(self._h))

    def export(self):
        """Live undirected edges (a <= b, each once) as int32 arrays."""
        m = ctypes.c_int64()
        _check(lib().dsssp_graph_export(self._h, ctypes.byref(m), None, None, None))
        a = np.empty(m.value, np.int32)
        b = np.empty(m.value, np.int32)
        w = np.empty(m.value, np.int32)
        _check(lib().dsssp_graph_export(self._h, ctypes.byref(m), _ptr(a), _ptr(b), _ptr(w)))
        return a, b, w

    def clone(self) -> "Graph":
        h = ctypes.c_void_p()
        _check(lib().dsssp_graph_clone(self._h, ctypes.byref(h)))
        return Graph(self.n, None, None, None, _handle=h)

    def copy_from(self, other: "Graph") -> None:
        _check(lib().dsssp_graph_copy(self._h, other._h))


class Tree:
    """SSSP tree(s) of a Graph, resident in HBM.

    ``Tree(g, src)`` computes it on the GPU from ``src``; ``Tree(g, parent=p,
    dist=d)`` loads a given tree (``canonical=True`` re-derives parents);
    ``Tree(g, roots=[...])`` grows one forest from several roots;
    ``Tree(g, sources=[...])`` holds one tree per source (batched form)."""

    def __init__(self, graph: Graph, src: int = 0, parent=None, dist=None, canonical: bool = False,
                 roots=None, sources=None):
        self.graph = graph
        h = ctypes.c_void_p()
        flags = TREE_CANONICAL if canonical else 0
        self.ntrees = 1
        if sources is not None:
            s = _i32(sources)
            _check(lib().dsssp_forest_create(graph._h, len(s), _ptr(s), flags, ctypes.byref(h)))
            self.ntrees = len(s)
        elif roots is not None:
            r = _i32(roots)
            _check(lib().dsssp_tree_create_roots(graph._h, len(r), _ptr(r), flags, ctypes.byref(h)))
        elif parent is None:
            _check(lib().dsssp_tree_create(graph._h, int(src), None, None, flags, ctypes.byref(h)))
        else:
            p = _i32(parent)
            d = np.ascontiguousarray(np.asarray(dist, np.int64))
            if len(p) != graph.n or len(d) != graph.n:
                raise ValueError("parent/dist must have n entries")
            _check(lib().dsssp_tree_create(graph._h, int(src), _ptr(p), _ptr(d), flags, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().dsssp_tree_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close

    def read(self, tree: int = 0):
        par = np.empty(self.graph.n, np.int32)
        dist = np.empty(self.graph.n, np.int64)
        _check(lib().dsssp_forest_read(self._h, tree, _ptr(par), _ptr(dist)))
        return par, dist

    def copy_from(self, other: "Tree") -> None:
        _check(lib().dsssp_tree_copy(self._h, other._h))

    def relax_full(self) -> dict:
        """One full relaxation sweep (every reachable slot), then rounds until
        nothing changes; on a converged tree nothing changes
        (``relax_improved == 0``).  Returns the relaxation statistics."""
        st = Stats()
        _check(lib().dsssp_tree_relax_full(self._h, ctypes.byref(st)))
        return st.as_dict()


class Batch:
    """A change batch uploaded to HBM once (validated on the host)."""

    def __init__(self, graph: Graph, changes):
        self.graph = graph
        rec = as_changes(changes)
        self.k = len(rec)
        h = ctypes.c_void_p()
        _check(lib().dsssp_batch_upload(graph._h, _ptr(rec), self.k, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().dsssp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close


class Server:
    """Served form: resident instances + continuous-batching request loops, one
    per lane (a device graph of some of the instances, with its own stream)."""

    def __init__(self, mem_budget_bytes: int, max_merge: int = 0, max_records: int = 0, lanes: int = 0):
        cfg = ServerConfig(mem_budget_bytes, max_merge, lanes, max_records)
        h = ctypes.c_void_p()
        _check(lib().dsssp_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.sizes = {}

    def add_instance(self, n: int, a, b, w, src: int) -> int:
        """Admits an instance (before start, or while the server runs)."""
        a, b, w = _i32(a), _i32(b), _i32(w)
        i = ctypes.c_int32()
        _check(lib().dsssp_server_add_instance(self._h, n, len(a), _ptr(a), _ptr(b), _ptr(w), src, ctypes.byref(i)))
        self.sizes[i.value] = n
        return i.value

    def remove_instance(self, instance: int) -> None:
        """Retires an instance once its submitted requests are applied."""
        _check(lib().dsssp_server_remove_instance(self._h, instance))

    def start(self) -> None:
        _check(lib().dsssp_server_start(self._h))

    def submit(self, instance: int, changes) -> int:
        rec = as_changes(changes)
        t = ctypes.c_uint64()
        _check(lib().dsssp_server_submit(self._h, instance, _ptr(rec), len(rec), ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int) -> dict:
        st = RequestStats()
        _check(lib().dsssp_server_wait(self._h, ticket, ctypes.byref(st)))
        return st.as_dict()

    def read(self, instance: int):
        n = self.sizes[instance]
        par = np.empty(n, np.int32)
        dist = np.empty(n, np.int64)
        _check(lib().dsssp_server_read(self._h, instance, _ptr(par), _ptr(dist)))
        return par, dist

    def info(self) -> dict:
        t = ServerTotals()
        _check(lib().dsssp_server_info(self._h, ctypes.byref(t)))
        return t.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().dsssp_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    __del__ = close


def autotune(graph: Graph, tree: Tree, batches, passes: int = 1) -> tuple:
    """Coordinate descent over the launch knobs on resident sample batches
    (``Batch`` objects of ``graph``); graph and tree end as they started.  Sets
    the fastest tuning and returns it with its device ms per batch."""
    hs = (ctypes.c_void_p * len(batches))(*[b._h.value for b in batches])
    t = Tuning()
    ms = ctypes.c_double()
    _check(lib().dsssp_autotune(graph._h, tree._h, hs, len(batches), passes, ctypes.byref(t), ctypes.byref(ms)))
    return {f: getattr(t, f) for f, _ in Tuning._fields_}, ms.value


def update(graph: Graph, tree: Tree, changes) -> dict:
    """Apply one change batch (host array or Batch) and return the statistics."""
    st = Stats()
    if isinstance(changes, Batch):
        _check(lib().dsssp_update_batch(graph._h, tree._h, changes._h, ctypes.byref(st)))
    else:
        rec = as_changes(changes)
        _check(lib().dsssp_update(graph._h, tree._h, _ptr(rec), len(rec), ctypes.byref(st)))
    return st.as_dict()


def update_noalloc(graph: Graph, tree: Tree, device_records: int, k: int) -> dict:
    """Allocation-free update (dsssp_update_noalloc): device_records is the
    device address of k records in the dsssp_change layout (int32 a, b, w, op)."""
    st = Stats()
    _check(lib().dsssp_update_noalloc(graph._h, tree._h, ctypes.c_void_p(device_records), k, ctypes.byref(st)))
    return st.as_dict()


def header_symbols() -> list:
    """Every function name declared in include/dsssp.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "dsssp.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(dsssp_[a-z_0-9]+)\s*\(", txt)))
