"""GPU parity: the HIP path (through the C-ABI) against the oracle, bit-exact.

Every traversal mode x split threshold x golden vector; GPU chunk tables
against the oracle's NodeSplit restatement; the split known answers of the
reference's published runs on synthetic graphs with the published degree
sequences.
"""
import sys

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
import oracle
from conftest import golden_graph

pytestmark = pytest.mark.gpu

MODES = ["topo", "worklist", "persistent"]
INF = gga.INF
# frontier representation: default switch, every step a flag step, every step a queue step
DENSE = {"auto": 0, "flags": 1, "queue": 0xFFFFFFFF}


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_golden_vectors_all_modes(gpu, golden):
    index, arrays, _ = golden
    graphs = sorted({e["graph"] for e in index})
    for name in graphs:
        rp, dst, w = golden_graph(arrays, name)
        announce(f"golden {name} n={rp.size - 1} m={dst.size}")
        g = gga.csr_from_arrays(rp, dst, w)
        with gga.DeviceGraph(g) as dg:
            maxdeg = int(np.diff(rp.astype(np.int64)).max()) if dst.size else 0
            for mda in sorted({0, 1, 2, 3, 64, max(maxdeg, 1)}):
                if mda:
                    dg.split(mda=mda)
                else:
                    dg.split()
                for e in (e for e in index if e["graph"] == name):
                    for mode, rep in [("topo", "auto")] + [(m, r) for m in MODES[1:] for r in DENSE]:
                        announce(f"  {e['key']} mda={mda} mode={mode} frontier={rep}")
                        run = dg.bfs if e["algo"] == "bfs" else dg.sssp
                        d, st = run(e["src"], mode=mode, inf=e["inf"], count=True,
                                    dense_min=DENSE[rep])
                        np.testing.assert_array_equal(
                            d, arrays[e["key"]], err_msg=f"{e['key']} mda={mda} {mode} {rep}")
                        assert st["reached"] == int((arrays[e["key"]] != e["inf"]).sum())


def test_rmat14_bfs_config_all_modes(gpu, golden):
    _, arrays, meta = golden
    g = gga.gen_rmat(14, 8, seed=1, wlo=1, whi=1)
    want = arrays[meta["rmat14"]["key"]]
    with gga.DeviceGraph(g) as dg:
        for level in (0, 1, 4, 16, -1):
            st = dg.split(level=level) if level else dg.split()
            for mode in MODES:
                announce(f"rmat14 bfs level={level} mode={mode}")
                d, _ = dg.bfs(0, mode=mode)
                np.testing.assert_array_equal(d, want)
            assert st["ns_nnode"] == oracle.split_count(g.degrees(), st["mda"])


@pytest.mark.parametrize("T", [1, 3, 32, 64, 128, 256, 1024])
def test_split_chunks_equal_oracle_node_split(gpu, T):
    rng = np.random.default_rng(T)
    g = gga.gen_rmat(13, 16, seed=100 + T, wlo=1, whi=100)
    announce(f"split chunks T={T} n={g.n} m={g.m}")
    with gga.DeviceGraph(g) as dg:
        st = dg.split(mda=T)
        cb, own = dg.chunks()
        ocb, oown, _, _ = oracle.node_split(g.row_ptr, T)
        np.testing.assert_array_equal(cb, ocb)
        np.testing.assert_array_equal(own, oown)
        assert st["ns_nnode"] == own.size
        assert st["extranode"] == own.size - g.n
        assert st["maxdeg"] == g.maxdeg()[0]
        assert st["max_chunk"] == int(np.diff(ocb.astype(np.int64)).max())
    del rng


def synth_from_dd(dd, seed):
    """A graph with exactly the published degree sequence (destinations random)."""
    dd = np.asarray(dd, dtype=np.int64)
    deg = np.repeat(dd[:, 0], dd[:, 1]).astype(np.uint32)
    rng = np.random.default_rng(seed)
    rng.shuffle(deg)
    rp = np.concatenate([[0], np.cumsum(deg, dtype=np.int64)]).astype(np.uint32)
    dst = rng.integers(0, deg.size, int(rp[-1]), dtype=np.int64).astype(np.uint32)
    return gga.csr_from_arrays(rp, dst, None)


@pytest.mark.parametrize("graph", ["rmat20", "webs"])
def test_split_kat_published_rows_on_gpu(gpu, split_kat, graph):
    """Every distinct (L, bucket) of the published node-split runs on this graph:
    the GPU split reproduces MAX_EDGES_ALLOWED and ns_nnode (result.vi)."""
    g = synth_from_dd(split_kat["dd"][graph], 1)
    rows = [r for r in split_kat["rows"] if r["graph"] == graph]
    seen = {}
    for r in rows:
        seen[(r["level"], r["bucket"])] = (r["mda"], r["ns_nnode"], r["maxdeg"])
    with gga.DeviceGraph(g) as dg:
        for (level, bucket), (mda, ns, maxdeg) in sorted(seen.items()):
            announce(f"kat {graph} L={level} bucket={bucket}")
            st = dg.split(level=level, buckets=bucket if level < 0 else 0)
            assert (st["mda"], st["ns_nnode"], st["maxdeg"]) == (mda, ns, maxdeg), (level, bucket)


@pytest.mark.parametrize("buckets", [2, 5, 10, 15, 37])
def test_histogram_threshold_equals_oracle(gpu, split_kat, buckets):
    for graph in ("lj", "webB"):
        g = synth_from_dd(split_kat["dd"][graph], 2)
        announce(f"histogram {graph} buckets={buckets}")
        with gga.DeviceGraph(g) as dg:
            st = dg.split(level=-1, buckets=buckets)
            assert st["mda"] == oracle.mda_histogram(g.degrees(), buckets)
            assert st["ns_nnode"] == oracle.split_count(g.degrees(), st["mda"])


@pytest.mark.parametrize("T", [32, 64, 128, 256, 1024])
def test_rmat16_sssp_threshold_sweep(gpu, T):
    g = gga.gen_rmat(16, 16, seed=7, wlo=1, whi=100)
    src = g.maxdeg()[1]
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
    want_bfs = oracle.bfs_queue(g.row_ptr, g.dst, src)
    serial, _, _ = oracle.serial_sp(g.row_ptr, g.dst, g.w, src)
    np.testing.assert_array_equal(serial, want)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=T)
        for mode in MODES:
            for rep in (DENSE if mode != "topo" else ["auto"]):
                announce(f"rmat16 sssp T={T} mode={mode} frontier={rep}")
                d, st = dg.sssp(src, mode=mode, count=True, dense_min=DENSE[rep])
                np.testing.assert_array_equal(d, want)
                assert st["m_reach"] == oracle.m_reach(g.row_ptr, want)
                assert st["edges_inspected"] >= st["m_reach"]
                d, st = dg.bfs(src, mode=mode, dense_min=DENSE[rep])
                np.testing.assert_array_equal(d, want_bfs)


def test_grid_bfs_sssp(gpu):
    g = gga.gen_grid(200, 300, 0.005, seed=11, wlo=1, whi=1000)
    want_s = oracle.dijkstra(g.row_ptr, g.dst, g.w, 0)
    want_b = oracle.bfs_queue(g.row_ptr, g.dst, 0)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=4)
        for mode in MODES:
            for rep in (DENSE if mode != "topo" else ["auto"]):
                announce(f"grid200x300 mode={mode} frontier={rep}")
                d, _ = dg.sssp(0, mode=mode, dense_min=DENSE[rep])
                np.testing.assert_array_equal(d, want_s)
                d, _ = dg.bfs(0, mode=mode, dense_min=DENSE[rep])
                np.testing.assert_array_equal(d, want_b)


def test_result_same_alone_and_after_other_queries(gpu):
    g = gga.gen_rmat(14, 16, seed=3, wlo=1, whi=100)
    rng = np.random.default_rng(0)
    srcs = rng.integers(0, g.n, 12)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=64)
        alone = {}
        for s in srcs[:4]:
            alone[int(s)] = dg.sssp(int(s), mode="worklist")[0]
        for i, s in enumerate(srcs):
            announce(f"interleaved query {i} src={s}")
            dg.bfs(int(s), mode=MODES[i % 3])
            dg.sssp(int(srcs[-1 - i]), mode=MODES[(i + 1) % 3])
        for s, want in alone.items():
            for mode in MODES:
                np.testing.assert_array_equal(dg.sssp(s, mode=mode)[0], want)


def test_device_verify_and_errors(gpu):
    g = gga.gen_rmat(12, 8, seed=9, wlo=1, whi=100)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=16)
        d, _ = dg.sssp(0)
        r = dg.verify(0, bfs=False)
        assert r["nerr"] == 0 and r["untight"] == 0 and r["src_ok"]
        assert r["reached"] == int((d != INF).sum())
        np.testing.assert_array_equal(dg.copy_dist(), d)
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            dg.sssp(g.n)
        with pytest.raises(_lib.GgaError, match="GGA_ERANGE"):
            dg.split(level=-1, buckets=1 << 20)  # more buckets than the kernel holds


def test_instrumented_worklist_timeline_and_convergence(gpu):
    """The instrumented relax gives the same distances; it records one entry
    per workgroup per step with plausible placement, and the convergence rows
    end at zero difference with a non-increasing difference count (distances
    only decrease toward the final ones)."""
    g = gga.gen_rmat(14, 16, seed=3, wlo=1, whi=100)
    src = g.maxdeg()[1]
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=32)
        announce("instrumented sssp rmat14")
        d, st, tl, conv = dg.trace(src, timeline_cap=1 << 20, conv_final=want)
        np.testing.assert_array_equal(d, want)
        grid = st["timeline_grid"]
        assert grid > 0 and tl.size == st["timeline_records"] and tl.size % grid == 0
        assert tl.size // grid >= st["iterations"]
        assert (tl["ticks"][: grid * st["iterations"]] > 0).all()
        assert ((tl["place"] >> 8) < 8).all()                  # XCC ids of MI355X
        assert len(set((tl["place"] >> 8).tolist())) > 1        # workgroups spread over XCDs
        assert conv.size >= st["iterations"] and conv["diff_count"][-1] == 0 and conv["norm2"][-1] == 0
        assert (np.diff(conv["diff_count"].astype(np.int64)) <= 0).all()
        assert conv["diff_count"][0] > 0
        ms = gga.timeline_makespan(tl, grid, 256)
        assert 0 < ms < 1e3


@pytest.mark.parametrize("delta", [1, 7, 50, 300, 100000])
def test_near_far_sssp_equals_oracle(gpu, golden, delta):
    """Near-far buckets of width delta: distances equal the oracle on the golden
    vectors, R-MAT s16 (several thresholds and both list representations) and
    a grid (whose shortest paths cross many buckets)."""
    index, arrays, _ = golden
    for e in (e for e in index if e["algo"] == "sssp"):
        rp, dst, w = golden_graph(arrays, e["graph"])
        g = gga.csr_from_arrays(rp, dst, w)
        with gga.DeviceGraph(g) as dg:
            for mda in (1, 3):
                dg.split(mda=mda)
                for rep in DENSE:
                    d, _ = dg.sssp(e["src"], mode="worklist", inf=e["inf"], delta=delta, dense_min=DENSE[rep])
                    np.testing.assert_array_equal(d, arrays[e["key"]], err_msg=f"{e['key']} mda={mda} {rep}")
    g = gga.gen_rmat(16, 16, seed=7, wlo=1, whi=100)
    src = g.maxdeg()[1]
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
    with gga.DeviceGraph(g) as dg:
        for T in (32, 1024):
            dg.split(mda=T)
            for rep in DENSE:
                announce(f"near-far rmat16 delta={delta} T={T} lists={rep}")
                d, st = dg.sssp(src, mode="worklist", count=True, delta=delta, dense_min=DENSE[rep])
                np.testing.assert_array_equal(d, want)
                assert st["m_reach"] == oracle.m_reach(g.row_ptr, want)
    g = gga.gen_grid(120, 150, 0.0, seed=11, wlo=1, whi=1000)
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, 0)
    with gga.DeviceGraph(g) as dg:
        dg.split()
        announce(f"near-far grid120x150 delta={delta}")
        np.testing.assert_array_equal(dg.sssp(0, mode="worklist", delta=delta)[0], want)
        np.testing.assert_array_equal(dg.sssp(7, mode="worklist", delta=delta, sync_every=1)[0],
                                      oracle.dijkstra(g.row_ptr, g.dst, g.w, 7))


@pytest.mark.slow
def test_full_size_config3_grid_sssp_bfs_equal_oracle(gpu):
    """Config 3 at full size (4096^2 grid + 0.5% shortcuts): the default
    (AUTO), worklist and persistent (one cooperative launch, near-far SSSP)
    modes against the oracle's Dijkstra / queue BFS."""
    g = gga.gen_grid(4096, 4096, 0.005, seed=11, wlo=1, whi=1000)
    want_s = oracle.dijkstra(g.row_ptr, g.dst, g.w, 0)
    want_b = oracle.bfs_queue(g.row_ptr, g.dst, 0)
    with gga.DeviceGraph(g) as dg:
        dg.split()
        for mode in ("auto", "worklist", "persistent"):
            announce(f"config3 full size sssp mode={mode}")
            np.testing.assert_array_equal(dg.sssp(0, mode=mode)[0], want_s)
            announce(f"config3 full size bfs mode={mode}")
            np.testing.assert_array_equal(dg.bfs(0, mode=mode)[0], want_b)


@pytest.mark.slow
def test_full_size_config2_rmat22_sssp_equals_oracle(gpu):
    g = gga.gen_rmat(22, 16, seed=7, wlo=1, whi=100)
    src = g.maxdeg()[1]
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
    with gga.DeviceGraph(g) as dg:
        for T in (32, 64, 128, 256, 1024):
            dg.split(mda=T)
            announce(f"config2 full size sssp T={T}")
            np.testing.assert_array_equal(dg.sssp(src)[0], want)
        np.testing.assert_array_equal(dg.bfs(src)[0], oracle.bfs_queue(g.row_ptr, g.dst, src))
