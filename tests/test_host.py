"""CPU tests of the product library's host side: ABI exports, text formats,
workload generators and input validation (no GPU calls)."""
import os
import subprocess

import numpy as np
import pytest

import dynamicssspv2_amd as ds
import oracle
from tests._util import GOLDEN, GOLDEN_N, INF, golden_case, read_changes, read_graph, read_ref_sssp


def test_library_exports_every_header_symbol():
    syms = ds.header_symbols()
    assert len(syms) >= 25
    L = ds.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.check_output(["nm", "-D", "--defined-only", ds.LIB_PATH]).decode()
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(syms) <= exported
    assert ds.lib().dsssp_version().startswith(b"dsssp")


def test_tuning_knobs_validated():
    """Host-side checks of dsssp_set_tuning (no device call): the
    push-vs-full-sweep threshold is -1 (never) or 1..1000 per mille, phase
    timing 0 or 1, and a rejected setting leaves the tuning as it was."""
    old = ds.get_tuning()
    assert old["sweep"] == -1  # default: push rounds only
    try:
        ds.set_tuning(sweep=250)
        assert ds.get_tuning()["sweep"] == 250
        for bad in (-2, 1001):
            with pytest.raises(ds.DssspError) as e:
                ds.set_tuning(sweep=bad)
            assert e.value.code == ds.ERR_ARG
            assert ds.get_tuning()["sweep"] == 250
        ds.set_tuning(sweep=-1)
        assert ds.get_tuning() == old
        # phase timing: 0 (default, only the update's start and end) or 1
        assert old["phase_timing"] == 0
        ds.set_tuning(phase_timing=1)
        assert ds.get_tuning()["phase_timing"] == 1
        for bad in (-1, 2):
            with pytest.raises(ds.DssspError) as e:
                ds.set_tuning(phase_timing=bad)
            assert e.value.code == ds.ERR_ARG
        ds.set_tuning(phase_timing=0)
        assert ds.get_tuning() == old
    finally:
        ds.set_tuning(**old)


@pytest.mark.parametrize("name", ["toy", "r300", "r2000", "sparse1500"])
def test_readers_match_plain_python(name):
    n = GOLDEN_N[name]
    a, b, w = ds.read_graph(os.path.join(GOLDEN, f"{name}_graph.txt"), n)
    ea, eb, ew = read_graph(os.path.join(GOLDEN, f"{name}_graph.txt"))
    assert np.array_equal(a, ea) and np.array_equal(b, eb) and np.array_equal(w, ew)
    c = ds.read_changes(os.path.join(GOLDEN, f"{name}_cE.txt"), n)
    assert np.array_equal(c, read_changes(os.path.join(GOLDEN, f"{name}_cE.txt")))
    p, d = ds.read_sssp(os.path.join(GOLDEN, f"{name}_SSSP.txt"), n)
    ep, ed = read_ref_sssp(os.path.join(GOLDEN, f"{name}_SSSP.txt"), n)
    assert np.array_equal(p, ep) and np.array_equal(d, ed)


def test_sssp_reader_quirks(tmp_path):
    f = tmp_path / "s.txt"
    # gaps and a missing tail are unreachable; parent -1 ignores the (lossy) distance
    f.write_text("0 0 0 \n2 0 5\n3 -1 1e+07 \n")
    p, d = ds.read_sssp(str(f), 6)
    assert p.tolist() == [0, -1, 0, -1, -1, -1]
    assert d.tolist() == [0, INF, 5, INF, INF, INF]
    f.write_text("0 0 0\n1 0 1.23457e+06\n")
    with pytest.raises(ds.DssspError) as e:
        ds.read_sssp(str(f), 3)
    assert e.value.code == ds.ERR_FORMAT and "line 2" in str(e.value)


def test_write_read_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    n = 1000
    par = rng.integers(-1, n, n).astype(np.int32)
    dist = rng.integers(0, 1 << 40, n).astype(np.int64)
    dist[par == -1] = INF
    f = str(tmp_path / "o.txt")
    ds.write_sssp(f, par, dist)
    p, d = ds.read_sssp(f, n)
    assert np.array_equal(p, par) and np.array_equal(d, dist)


def test_reader_errors(tmp_path):
    g = tmp_path / "g.txt"
    g.write_text("0 1 2\n1 9 3\n")
    with pytest.raises(ds.DssspError) as e:
        ds.read_graph(str(g), 5)
    assert e.value.code == ds.ERR_RANGE and "line 2" in str(e.value)
    g.write_text("0 1 -2\n")
    with pytest.raises(ds.DssspError) as e:
        ds.read_graph(str(g), 5)
    assert e.value.code == ds.ERR_WEIGHT
    g.write_text("0 1\n")
    with pytest.raises(ds.DssspError) as e:
        ds.read_graph(str(g), 5)
    assert e.value.code == ds.ERR_FORMAT
    with pytest.raises(ds.DssspError) as e:
        ds.read_graph(str(tmp_path / "missing.txt"), 5)
    assert e.value.code == ds.ERR_IO
    c = tmp_path / "c.txt"
    c.write_text("0 1 2 0\n0 1 2 7\n3   4 5 1 \n")  # inst 7 ignored later, multiple blanks fine
    assert ds.read_changes(str(c), 5).tolist() == [[0, 1, 2, 0], [0, 1, 2, 7], [3, 4, 5, 1]]


def test_big_file_parallel_parse(tmp_path):
    rng = np.random.default_rng(2)
    m = 400_000
    a = rng.integers(0, 1 << 20, m)
    b = rng.integers(0, 1 << 20, m)
    w = rng.integers(0, 1000, m)
    f = tmp_path / "big.txt"
    f.write_text("\n".join(f"{x}  {y}\t{z}" for x, y, z in zip(a, b, w)))
    ga, gb, gw = ds.read_graph(str(f), 1 << 20)
    assert np.array_equal(ga, a) and np.array_equal(gb, b) and np.array_equal(gw, w)


def test_gen_tree_plus_random():
    n, m = 5000, 40000
    a, b, w = ds.gen_tree_plus_random(n, m, 1, 100, seed=11)
    assert np.array_equal(b[: n - 1], np.arange(1, n)) and np.all(a[: n - 1] < b[: n - 1])
    keys = np.minimum(a, b).astype(np.int64) << 32 | np.maximum(a, b)
    assert len(np.unique(keys)) == m and np.all(a != b)
    assert w.min() >= 1 and w.max() <= 100
    a2, b2, w2 = ds.gen_tree_plus_random(n, m, 1, 100, seed=11)
    assert np.array_equal(a, a2) and np.array_equal(w, w2)
    # connected: Dijkstra reaches everything
    _, d = oracle.dijkstra(oracle.CSR.from_edges(n, a, b, w), [0])
    assert np.all(d != INF)


def test_gen_rmat_small():
    a, b, w = ds.gen_rmat(12, 16, wlo=1, whi=1000, seed=5)
    n = 1 << 12
    assert np.all(a < b) and a.min() >= 0 and b.max() < n
    keys = a.astype(np.int64) << 32 | b
    assert np.all(np.diff(keys) > 0)  # sorted, distinct
    assert w.min() >= 1 and w.max() <= 1000
    assert 0.3 * 16 * n < len(a) <= 16 * n
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    assert deg[0] == deg.max()  # a=0.57: vertex 0 is the hub
    a2, b2, w2 = ds.gen_rmat(12, 16, wlo=1, whi=1000, seed=5)
    assert np.array_equal(a, a2) and np.array_equal(w, w2)


def test_gen_grid_small():
    side = 60
    a, b, w = ds.gen_grid(side, 0.05, 1, 10000, seed=3)
    E = 2 * side * (side - 1)
    assert len(a) <= E - round(0.05 * E)
    diff = np.abs(b - a)
    assert np.all((diff == 1) | (diff == side))
    _, d = oracle.dijkstra(oracle.CSR.from_edges(side * side, a, b, w), [0])
    reach = d != INF
    touched = np.zeros(side * side, bool)
    touched[a] = touched[b] = True
    assert np.array_equal(reach & touched, touched)  # every kept edge is in 0's component


def test_gen_changes_properties():
    n, m = 3000, 20000
    a, b, w = ds.gen_tree_plus_random(n, m, 1, 100, seed=1)
    k = 4000
    c = ds.gen_changes(n, a, b, w, k, 0.5, 1, 100, seed=9, first_deletable=n - 1)
    ins, dele = c[c[:, 3] == 1], c[c[:, 3] == 0]
    assert len(ins) == 2000 and len(dele) == 2000
    ek = {(min(x, y), max(x, y)): z for x, y, z in zip(a, b, w)}
    tree = {(min(x, y), max(x, y)) for x, y in zip(a[: n - 1], b[: n - 1])}
    dk = [(min(x, y), max(x, y)) for x, y, _, _ in dele]
    assert len(set(dk)) == len(dk)
    for (x, y, z, _), key in zip(dele, dk):
        assert ek[key] == z and key not in tree
    ik = [(min(x, y), max(x, y)) for x, y, _, _ in ins]
    assert len(set(ik)) == len(ik) and not (set(ik) & set(ek)) and all(x != y for x, y in ik)
    assert not np.all(c[:2000, 3] == c[0, 3])  # interleaved


def test_gpu_entry_points_fail_cleanly_without_gpu():
    """Here there is no GPU: device calls must return NODEV, never crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(ds.DssspError) as e:
        ds.Graph(3, [0], [1], [1])
    assert e.value.code in (ds.ERR_NODEV, ds.ERR_HIP)


def test_server_admission_by_budget():
    """Admission happens on the host before anything touches the GPU."""
    a, b, w = ds.gen_tree_plus_random(10_000, 50_000, 1, 10, seed=1)
    srv = ds.Server(mem_budget_bytes=6_000_000)
    assert srv.add_instance(10_000, a, b, w, 0) == 0
    with pytest.raises(ds.DssspError) as e:
        srv.add_instance(10_000, a, b, w, 0)
    assert e.value.code == ds.ERR_BUDGET
    info = srv.info()
    assert info["instances"] == 1 and 0 < info["used_bytes"] <= 6_000_000
    with pytest.raises(ds.DssspError) as e:
        srv.submit(0, [[0, 1, 1, 1]])
    assert e.value.code == ds.ERR_STATE  # not started
    with pytest.raises(ds.DssspError) as e:
        srv.add_instance(10, [0], [1], [1], 99)
    assert e.value.code == ds.ERR_RANGE
    srv.close()
    # an instance whose rows one device graph's 32-bit pool offsets cannot address
    big = ds.Server(mem_budget_bytes=1 << 50)
    with pytest.raises(ds.DssspError) as e:
        big.add_instance(200_000_000, [0], [1], [1], 0)
    assert e.value.code == ds.ERR_BUDGET and "too large" in str(e.value)
    big.close()


def test_gen_changes_grid():
    side = 80
    a, b, w = ds.gen_grid(side, 0.01, 1, 10000, seed=3)
    n = side * side
    par, dist = oracle.dijkstra(oracle.CSR.from_edges(n, a, b, w), [0])
    k = 1000
    c = ds.gen_changes_grid(side, a, b, w, par, k, 0.7, 0.5, 3, 1, 10000, seed=4)
    dele, ins = c[c[:, 3] == 0], c[c[:, 3] == 1]
    assert len(dele) == 700 and len(ins) == 300
    tree = {(min(v, int(p)), max(v, int(p))) for v, p in enumerate(par) if p >= 0 and p != v}
    dk = [(min(x, y), max(x, y)) for x, y, _, _ in dele]
    assert len(set(dk)) == 700 and sum(key in tree for key in dk) == 350
    ek = {(min(x, y), max(x, y)) for x, y in zip(a, b)}
    for x, y, z, _ in ins:
        md = abs(x // side - y // side) + abs(x % side - y % side)
        assert 1 <= md <= 3 and (min(x, y), max(x, y)) not in ek and 1 <= z <= 10000


TOOLS = os.path.join(os.path.dirname(ds.LIB_PATH), "op_tools")


def test_op_tools_apply_and_ce(tmp_path):
    """`op_tools apply` gives the oracle's graph-after-batch; `op_tools ce` writes
    op_cE-style change lines (createChangedEdges.cpp:104,116) that the reader takes."""
    a, b, w, chg, n = golden_case("r300")
    out = subprocess.check_output([TOOLS, "apply", os.path.join(GOLDEN, "r300_graph.txt"), "300",
                                   os.path.join(GOLDEN, "r300_cE.txt")]).decode()
    got = np.array([[int(x) for x in line.split()] for line in out.splitlines()], np.int32)
    ea, eb, ew = oracle.apply_changes(300, a, b, w, chg)
    assert np.array_equal(got, np.stack([ea, eb, ew], 1))
    ce = tmp_path / "ce.txt"
    ce.write_bytes(subprocess.check_output([TOOLS, "ce", os.path.join(GOLDEN, "r300_graph.txt"), "300", "100", "40",
                                            "3"]))
    c = ds.read_changes(str(ce), 300)
    assert len(c) == 100 and int(np.sum(c[:, 3] == 1)) == 40 and c[:, 2].max() <= 99


def test_op_main_argument_and_input_errors(tmp_path):
    """The drop-in CLI (SSSPmainForUndirected.cu:17-32 argv): a wrong argument
    count prints the usage with status 2; an unreadable input is named on
    stderr with a non-zero status (the reference prints and exits 0).  Both are
    decided on the host before any device call."""
    cli = os.path.join(os.path.dirname(ds.LIB_PATH), "op_main")
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage: op_main" in (r.stdout + r.stderr)
    missing = str(tmp_path / "no_such_cE.txt")
    r = subprocess.run([cli, os.path.join(GOLDEN, "toy_graph.txt"), "6", "6", os.path.join(GOLDEN, "toy_SSSP.txt"),
                        missing], capture_output=True, text=True)
    assert r.returncode != 0 and missing in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus above 1 exits with a message."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["python", os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
