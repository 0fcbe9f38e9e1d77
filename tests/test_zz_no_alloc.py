"""The allocation-free update (dsssp_update_noalloc) against dsssp_update_batch.

Two identical graphs and trees take the same batches, one through
dsssp_update_batch (records uploaded by the library), the other through
dsssp_update_noalloc (records already in device memory, scratch in the
library's static device array).  After each of two batches in a row, the
second smaller than the first, the trees must be byte-identical, the graphs
must hold the same edge multiset, and the tree must be the exact one of the
oracle's graph after the batch (Dijkstra distances, canonical parents).
"""
import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle

pytestmark = pytest.mark.gpu


class DeviceRecords:
    """Records copied into device memory with the HIP runtime the library
    itself links (libamdhip64.so.7), through ctypes: the caller's buffer."""

    def __init__(self, chg):
        import ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        rec = np.ascontiguousarray(np.asarray(chg, np.int32).reshape(-1, 4))
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(rec.nbytes, 16))) == 0
        assert self.hip.hipMemcpy(self.ptr, rec.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(rec.nbytes),
                                  1) == 0  # hipMemcpyHostToDevice
        assert self.hip.hipDeviceSynchronize() == 0

    def data_ptr(self):
        return self.ptr.value

    def close(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr.value = None


def device_records(chg):
    return DeviceRecords(chg)


def edge_key_sorted(a, b, w):
    a, b, w = (np.asarray(x, np.int64) for x in (a, b, w))
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    return np.sort((lo << 40) | (hi << 16) | w)


def test_noalloc_update_matches_update_batch():
    ds.set_device(0)
    scale = 14
    n = 1 << scale
    a, b, w = ds.gen_rmat(scale, 16, (0.57, 0.19, 0.19), 1, 1000, seed=77)
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    src = int(np.argmax(deg))
    g1, g2 = ds.Graph(n, a, b, w), ds.Graph(n, a, b, w)
    t1, t2 = ds.Tree(g1, src), ds.Tree(g2, src)
    ca, cb, cw = a, b, w
    keep = []
    try:
        for step, (k, ins) in enumerate(((12_000, 0.5), (3_000, 0.3))):  # the second call is smaller
            chg = ds.gen_changes(n, ca, cb, cw, k, ins, 1, 1000, seed=78 + step)
            bt = ds.Batch(g1, chg)
            s1 = ds.update(g1, t1, bt)
            bt.close()
            d = device_records(chg)
            keep.append(d)
            s2 = ds.update_noalloc(g2, t2, d.data_ptr(), len(chg))
            assert (s1["n_insert"], s1["n_delete"]) == (s2["n_insert"], s2["n_delete"])
            p1, d1 = t1.read()
            p2, d2 = t2.read()
            assert np.array_equal(p1, p2) and np.array_equal(d1, d2), f"step {step}: trees differ"
            ca, cb, cw = oracle.apply_changes(n, ca, cb, cw, chg)
            assert np.array_equal(edge_key_sorted(*g2.export()), edge_key_sorted(ca, cb, cw))
            assert np.array_equal(edge_key_sorted(*g1.export()), edge_key_sorted(ca, cb, cw))
            csr = oracle.CSR.from_edges(n, ca, cb, cw)
            _, dref = oracle.dijkstra(csr, [src])
            assert np.array_equal(d2, dref)
            assert np.array_equal(p2, oracle.canonical_parents(csr, dref, [src], rule=0))
        # rejected inputs leave everything as it was: a vertex outside [0, n)
        # and a batch beyond the maximum size
        before = t2.read()
        bad = device_records(np.array([[0, 1, 5, 1], [3, n + 7, 2, 0]], np.int32))
        keep.append(bad)
        with pytest.raises(ds.DssspError) as e:
            ds.update_noalloc(g2, t2, bad.data_ptr(), 2)
        assert e.value.code == ds.ERR_RANGE
        with pytest.raises(ds.DssspError) as e:
            ds.update_noalloc(g2, t2, bad.data_ptr(), (1 << 21) + 1)
        assert e.value.code == ds.ERR_ARG
        after = t2.read()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    finally:
        for o in keep + [t1, t2, g1, g2]:
            o.close()
