"""CPU oracle for the dynamic SSSP update -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import it; the product package
``dynamicssspv2_amd`` never does.

It wraps ``oracle/liboracle.so`` (built from ``oracle.c`` by ``oracle/Makefile``)
with numpy arrays.  Each function's docstring names the reference file:line it
restates; how each one is pinned is listed at the top of ``oracle.c``.

Distances are int64 with ``INF`` (int64 max) for unreachable vertices, whose
parent is -1; a source's parent is itself.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
INF = np.iinfo(np.int64).max
REF_INVALID = 999999  # gpuFunctions_undir.cuh:190 -- the reference's "invalidated" distance

_lib = None


def build() -> None:
    """Compile liboracle.so with gcc (and oracle/_ref when the reference sources exist)."""
    subprocess.check_call(["make", "-s", "-B", "-C", HERE])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-B", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i64, i32, p = ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p
        L.orc_build_csr.argtypes = [i64, i64, p, p, p, p, p, p]
        L.orc_build_csr.restype = ctypes.c_int
        L.orc_build_csr_any.argtypes = [i64, i64, p, p, p, p, p, p]
        L.orc_build_csr_any.restype = ctypes.c_int
        L.orc_djk.argtypes = [i64, p, p, p, i32, i64, p, p]
        L.orc_djk.restype = i64
        L.orc_dijkstra_lazy.argtypes = [i64, p, p, p, p, i32, p, p]
        L.orc_dijkstra_lazy.restype = i64
        L.orc_bellman_ford.argtypes = [i64, i64, p, p, p, p, i32, p]
        L.orc_bellman_ford.restype = i64
        L.orc_apply_changes.argtypes = [i64, i64, p, p, p, i64, p, p, p, p]
        L.orc_apply_changes.restype = i64
        L.orc_ref_update.argtypes = [i64, p, p, p, p, p, p, i64, p, i64, p, i64]
        L.orc_ref_update.restype = i64
        L.orc_canonical_parents.argtypes = [i64, p, p, p, p, p, i32, ctypes.c_int, p]
        L.orc_canonical_parents.restype = i64
        L.orc_certificate.argtypes = [i64, p, p, p, p, p, p, i32, ctypes.POINTER(i64)]
        L.orc_certificate.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _i64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int64))


class CSR:
    """Adjacency in file order (all_structure_undir.cuh:178-198).  wt < 0 = tombstone."""

    def __init__(self, n, rp, col, wt):
        self.n = int(n)
        self.rp, self.col, self.wt = rp, col, wt

    @classmethod
    def from_edges(cls, n, a, b, w, ordered: bool = True) -> "CSR":
        """ordered=False: rows in an unspecified order, built by all threads; for
        the order-independent checks (distances, canonical parents, certificate)."""
        a, b, w = _i32(a), _i32(b), _i32(w)
        m = len(a)
        rp = np.empty(n + 1, np.int64)
        col = np.empty(max(2 * m, 1), np.int32)
        wt = np.empty(max(2 * m, 1), np.int32)
        build = lib().orc_build_csr if ordered else lib().orc_build_csr_any
        rc = build(n, m, _ptr(a), _ptr(b), _ptr(w), _ptr(rp), _ptr(col), _ptr(wt))
        if rc == -1:
            raise ValueError("edge endpoint outside [0, n)")
        if rc != 0:
            raise MemoryError
        return cls(n, rp, col[: 2 * m], wt[: 2 * m])


def djk(csr: CSR, src: int = 0, cap: int = 0):
    """The reference's Dijkstra (seqSSSPwithDist_undir.cpp:66-128).

    cap=0 uses the reference's 1000*n cap (:71-74); pass INF for exact distances.
    Returns (parent int32[n], dist int64[n], pops)."""
    n = csr.n
    par = np.empty(n, np.int32)
    dist = np.empty(n, np.int64)
    pops = lib().orc_djk(n, _ptr(csr.rp), _ptr(csr.col), _ptr(csr.wt), src, cap, _ptr(par), _ptr(dist))
    if pops < 0:
        raise MemoryError
    return par, dist, pops


def dijkstra(csr: CSR, srcs=(0,)):
    """Textbook lazy-deletion Dijkstra, several sources allowed.  Returns (parent, dist)."""
    n = csr.n
    s = _i32(srcs)
    par = np.empty(n, np.int32)
    dist = np.empty(n, np.int64)
    if lib().orc_dijkstra_lazy(n, _ptr(csr.rp), _ptr(csr.col), _ptr(csr.wt), _ptr(s), len(s), _ptr(par), _ptr(dist)) < 0:
        raise MemoryError
    return par, dist


def bellman_ford(n, a, b, w, srcs=(0,)) -> np.ndarray:
    a, b, w, s = _i32(a), _i32(b), _i32(w), _i32(srcs)
    dist = np.empty(n, np.int64)
    if lib().orc_bellman_ford(n, len(a), _ptr(a), _ptr(b), _ptr(w), _ptr(s), len(s), _ptr(dist)) < 0:
        raise RuntimeError("Bellman-Ford did not settle")
    return dist


def apply_changes(n, a, b, w, chg):
    """Edges after a batch (all_structure_undir.cuh:106-117, gpuFunctions_undir.cuh:200-213).

    chg is an int32 array of shape (k, 4): a, b, w, op (0 delete, 1 insert).
    Returns (a, b, w) of the surviving edges in adjacency order."""
    a, b, w = _i32(a), _i32(b), _i32(w)
    c = _i32(chg).reshape(-1, 4)
    cap = len(a) + len(c)
    oa = np.empty(max(cap, 1), np.int32)
    ob = np.empty(max(cap, 1), np.int32)
    ow = np.empty(max(cap, 1), np.int32)
    cnt = lib().orc_apply_changes(n, len(a), _ptr(a), _ptr(b), _ptr(w), len(c), _ptr(c), _ptr(oa), _ptr(ob), _ptr(ow))
    if cnt < 0:
        raise MemoryError
    return oa[:cnt].copy(), ob[:cnt].copy(), ow[:cnt].copy()


def ref_update(n, a, b, w, parent, dist, chg, max_rounds=None):
    """Sequential schedule of the reference's update kernels
    (gpuFunctions_undir.cuh:125-433, SSSPmainForUndirected.cu:196-281).

    (a, b, w) is the graph before the batch and (parent, dist) the tree as the
    reference's reader holds it (unreachable: parent -1).  Returns
    (parent int32, dist int64) with every distance >= 999999 (the reference's
    invalidated / unreachable sentinels) mapped to INF and its parent to -1."""
    c = _i32(chg).reshape(-1, 4)
    ins = c[c[:, 3] == 1]
    dele = c[c[:, 3] == 0]
    # Insertions are appended to the adjacency before the update
    # (all_structure_undir.cuh:106-117), after the graph's own lines.
    aa = np.concatenate([_i32(a), ins[:, 0]])
    bb = np.concatenate([_i32(b), ins[:, 1]])
    ww = np.concatenate([_i32(w), ins[:, 2]])
    csr = CSR.from_edges(n, aa, bb, ww)
    wt = csr.wt.copy()
    par = _i32(parent).copy()
    d = _i64(dist).copy()
    d32 = np.where((par < 0) | (d >= 9999999), 9999999, d).astype(np.int32)  # read_SSSP, :153-156
    upd = np.zeros(n, np.int32)
    insf = _i32(ins[:, :3]).copy()
    delf = _i32(dele[:, :3]).copy()
    mr = max_rounds if max_rounds is not None else 4 * n + 16
    r = lib().orc_ref_update(n, _ptr(csr.rp), _ptr(csr.col), _ptr(wt), _ptr(par), _ptr(d32), _ptr(upd),
                             len(insf), _ptr(insf), len(delf), _ptr(delf), mr)
    if r == -1:
        raise RuntimeError("reference update did not settle within max_rounds")
    if r < -1:
        raise MemoryError
    out_d = d32.astype(np.int64)
    bad = out_d >= REF_INVALID
    out_d[bad] = INF
    par[bad] = -1
    return par, out_d


def canonical_parents(csr: CSR, dist, srcs=(0,), rule: int = 0) -> np.ndarray:
    """Canonical shortest-path tree for exact distances (rule 0: smallest-id tight
    neighbour; rule 1: smallest-id tight neighbour one BFS level up in the tight graph)."""
    d = _i64(dist)
    s = _i32(srcs)
    par = np.empty(csr.n, np.int32)
    bad = lib().orc_canonical_parents(csr.n, _ptr(csr.rp), _ptr(csr.col), _ptr(csr.wt), _ptr(d),
                                      _ptr(s), len(s), rule, _ptr(par))
    if bad < 0:
        raise MemoryError
    if bad > 0:
        raise ValueError(f"{bad} reachable vertices have no tight in-edge: distances are not exact")
    return par


def certificate(csr: CSR, dist, parent, srcs=(0,)):
    """Exactness certificate; returns (code, vertex), code 0 = exact tree."""
    d = _i64(dist)
    p = _i32(parent)
    s = _i32(srcs)
    where = ctypes.c_int64(-1)
    code = lib().orc_certificate(csr.n, _ptr(csr.rp), _ptr(csr.col), _ptr(csr.wt), _ptr(d), _ptr(p),
                                 _ptr(s), len(s), ctypes.byref(where))
    if code < 0:
        raise MemoryError
    return code, where.value


def ref_sssp_binary() -> str | None:
    """Path of op_ssspSeq compiled from the reference's own source, if it was built."""
    p = os.path.join(REF_DIR, "op_ssspSeq")
    return p if os.path.exists(p) else None
