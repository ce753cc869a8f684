"""GPU regression tests: state that outlives one query, and the device verifier.

1. Queue-step stamps (mark[]) stay in the graph's workspace between queries;
   a near-far query with a narrow bucket runs empty steps at every threshold
   jump, so its step numbers run ahead of its non-empty step count.  Every
   later query on the same handle (plain worklist, persistent, near-far
   again) must still equal the oracle.
2. The device verifier (k_verify_edges / k_verify_vertices behind
   gga_graph_verify, replacing verifysolution, myutils.h:1-18) must count
   exactly what the oracle's restatement counts on distances with planted
   errors: too small (tightness), too large (fixpoint), a wrong source.
"""
import ctypes
import sys

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
import oracle

pytestmark = pytest.mark.gpu
INF = gga.INF


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _h2d(dptr: int, arr: np.ndarray) -> None:
    """Host -> device copy into a library-owned device buffer (test only)."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    arr = np.ascontiguousarray(arr, dtype=np.uint32)
    assert hip.hipMemcpy(dptr, arr.ctypes.data, arr.nbytes, 1) == 0  # hipMemcpyHostToDevice


def test_stamps_do_not_leak_across_queries(gpu):
    g = gga.gen_grid(40, 50, shortcut_frac=0.0, seed=5, wlo=1, whi=1000)
    srcs = [0, 777, 1999, 1234, 3]
    want = {s: oracle.dijkstra(g.row_ptr, g.dst, g.w, s) for s in srcs}
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=4)
        for rnd in range(2):
            for s in srcs:
                # delta = 1 below every weight: a threshold jump (an empty
                # step) after nearly every bucket
                announce(f"round {rnd} src={s} near-far delta=1")
                d, st = dg.sssp(s, mode="worklist", delta=1, dense_min=0xFFFFFFFF)
                np.testing.assert_array_equal(d, want[s], err_msg=f"near-far src={s}")
                assert st["iterations"] > 2
                for t in srcs:
                    for mode in ("worklist", "persistent"):
                        announce(f"  then {mode} src={t}")
                        d, _ = dg.sssp(t, mode=mode, delta=0xFFFFFFFF, dense_min=0xFFFFFFFF)
                        np.testing.assert_array_equal(d, want[t], err_msg=f"{mode} src={t} after near-far")
                    d, _ = dg.bfs(t, mode="worklist", dense_min=0xFFFFFFFF)
                    np.testing.assert_array_equal(d, oracle.bfs_queue(g.row_ptr, g.dst, t))


def test_failed_query_does_not_poison_the_next(gpu):
    """A query stopped by its step cap (GGA_ETIMEDOUT) leaves stamps of
    unknown step numbers; the next query on the handle is still exact.  (A
    mesh large enough that two steps cannot finish it, although its waves
    follow their wavefronts past their batches.)"""
    g = gga.gen_grid(200, 200, shortcut_frac=0.0, seed=2, wlo=1, whi=50)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=4)
        for mode in ("worklist", "persistent"):
            announce(f"capped {mode}")
            with pytest.raises(_lib.GgaError, match="GGA_ETIMEDOUT"):
                dg.sssp(0, mode=mode, max_iters=2, delta=0xFFFFFFFF, dense_min=0xFFFFFFFF)
            for s in (0, g.n // 2 + 100, g.n - 1):
                d, _ = dg.sssp(s, mode="worklist", delta=0xFFFFFFFF, dense_min=0xFFFFFFFF)
                np.testing.assert_array_equal(d, oracle.dijkstra(g.row_ptr, g.dst, g.w, s))


def _plant(rng, dist, inf, k_small, k_large):
    d = dist.copy()
    reached = np.flatnonzero((d != inf) & (d > 1))
    pick = rng.choice(reached, size=k_small + k_large, replace=False)
    d[pick[:k_small]] -= 1                       # too small: no tight predecessor
    d[pick[k_small:]] += 5                       # too large: an edge relaxes it
    return d


@pytest.mark.parametrize("bfs", [False, True])
def test_device_verifier_counts_planted_errors(gpu, bfs):
    g = gga.gen_rmat(12, 8, seed=21, wlo=1, whi=100)
    w = None if bfs else g.w
    src = g.maxdeg()[1]
    rng = np.random.default_rng(4)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=16)
        for case in ("clean", "small", "large", "mixed", "source", "unreached"):
            run = dg.bfs if bfs else dg.sssp
            d, _ = run(src, mode="worklist")
            if case == "small":
                d = _plant(rng, d, INF, 7, 0)
            elif case == "large":
                d = _plant(rng, d, INF, 0, 9)
            elif case == "mixed":
                d = _plant(rng, d, INF, 5, 11)
            elif case == "source":
                d = d.copy()
                d[src] = 3
            elif case == "unreached":
                d = d.copy()
                reached = np.flatnonzero((d != INF) & (np.arange(g.n) != src))
                d[rng.choice(reached, size=6, replace=False)] = INF
            ptr = _lib.load().gga_graph_dist_device(dg._h)
            assert ptr, "no device result"
            _h2d(ptr, d)
            announce(f"verify {'bfs' if bfs else 'sssp'} case={case}")
            r = dg.verify(src, bfs=bfs)
            nerr = oracle.verify(g.row_ptr, g.dst, w, d)
            bad = oracle.tight(g.row_ptr, g.dst, w, src, d)
            assert r["nerr"] == nerr, (case, r, nerr)
            assert r["src_ok"] == bool(d[src] == 0), case
            assert r["untight"] + (0 if r["src_ok"] else 1) == bad, (case, r, bad)
            assert r["reached"] == int((d != INF).sum())
            host = gga.verify(g, src, d, bfs=bfs)
            assert (host["nerr"], host["untight"], host["src_ok"]) == (r["nerr"], r["untight"], r["src_ok"])
            if case == "clean":
                assert nerr == 0 and bad == 0
            else:
                assert nerr + bad > 0, f"planted case {case} went unnoticed by the oracle"


def test_concurrent_persistent_queries_on_two_graphs(gpu):
    """Two threads, each driving its own graph, run cooperative-launch
    queries (persistent single-source, 256-lane bitmask BFS) at the same
    time: every call succeeds and equals the oracle (the library runs one
    cooperative grid at a time)."""
    import threading
    gs = [gga.gen_grid(60, 70, 0.01, seed=31, wlo=1, whi=1000), gga.gen_rmat(13, 16, seed=32, wlo=1, whi=100)]
    srcs = [[0, 1000, 4199], [0, 77, 4000]]
    want = [{s: oracle.dijkstra(g.row_ptr, g.dst, g.w, s) for s in ss} for g, ss in zip(gs, srcs)]
    want_b = [{s: oracle.bfs_queue(g.row_ptr, g.dst, s) for s in ss} for g, ss in zip(gs, srcs)]
    errors = []
    dgs = [gga.DeviceGraph(g) for g in gs]
    try:
        for dg in dgs:
            dg.split(mda=16)

        def worker(i):
            try:
                for rep in range(4):
                    for s in srcs[i]:
                        d, _ = dgs[i].sssp(s, mode="persistent")
                        if not np.array_equal(d, want[i][s]):
                            errors.append(f"graph {i} src {s} rep {rep}: persistent sssp differs")
                    lv, _ = dgs[i].multi_bfs(np.array(srcs[i], dtype=np.uint32))
                    for k, s in enumerate(srcs[i]):
                        if not np.array_equal(lv[k], want_b[i][s]):
                            errors.append(f"graph {i} src {s} rep {rep}: bitmask bfs differs")
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(f"graph {i}: {e!r}")

        announce("two threads, cooperative launches on two graphs")
        th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(timeout=300)
        assert not any(t.is_alive() for t in th), "worker still running"
    finally:
        for dg in dgs:
            dg.close()
    assert not errors, errors


def test_unsplit_hubs_walked_in_pieces(gpu):
    """L = 0 (no split, the reference's baseline and dynamic-parallelism
    case): the split statistics and gga_graph_chunks keep the reference's
    one-chunk-per-vertex table, while the kernels walk hubs cut into pieces
    of at most GGA_PIECE_MAX edges; every mode is still exact."""
    g = gga.gen_rmat(16, 16, seed=7, wlo=1, whi=100)
    maxdeg, src = g.maxdeg()
    assert maxdeg > 1024
    deg = g.degrees().astype(np.int64)
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
    want_b = oracle.bfs_queue(g.row_ptr, g.dst, src)
    with gga.DeviceGraph(g) as dg:
        st = dg.split()
        assert st["ns_nnode"] == g.n and st["mda"] == maxdeg and st["max_chunk"] == maxdeg
        assert st["exec_chunks"] == int(np.maximum(1, -(-deg // 1024)).sum())
        cb, own = dg.chunks()
        assert cb.size == g.n + 1 and own.size == g.n
        ref_cb, ref_own = oracle.node_split(g.row_ptr, maxdeg)[:2]
        np.testing.assert_array_equal(cb, ref_cb)
        np.testing.assert_array_equal(own, ref_own)
        for mode in ("topo", "worklist", "persistent"):
            announce(f"unsplit rmat16 mode={mode}")
            np.testing.assert_array_equal(dg.sssp(src, mode=mode)[0], want)
            np.testing.assert_array_equal(dg.bfs(src, mode=mode)[0], want_b)
        st = dg.split(mda=4096)
        assert st["exec_chunks"] > st["ns_nnode"]
        np.testing.assert_array_equal(dg.sssp(src, mode="worklist")[0], want)


@pytest.mark.parametrize("block", [128, 256, 512])
def test_worklist_workgroup_sizes_equal_oracle(gpu, block):
    """The worklist kernel at 128 / 256 / 512 threads per workgroup: same
    distances in every frontier form; an unsupported size is an error."""
    cases = [(gga.gen_rmat(14, 16, seed=5, wlo=1, whi=100), 64),
             (gga.gen_grid(90, 80, 0.005, seed=6, wlo=1, whi=1000), 4)]
    for g, T in cases:
        src = g.maxdeg()[1]
        want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
        want_b = oracle.bfs_queue(g.row_ptr, g.dst, src)
        with gga.DeviceGraph(g) as dg:
            dg.split(mda=T)
            for dense in (0, 1, 0xFFFFFFFF):
                announce(f"block={block} n={g.n} dense_min={dense}")
                d, _ = dg.sssp(src, mode="worklist", block=block, dense_min=dense, count=True)
                np.testing.assert_array_equal(d, want)
                d, _ = dg.bfs(src, mode="worklist", block=block, dense_min=dense)
                np.testing.assert_array_equal(d, want_b)
            with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
                dg.sssp(src, mode="worklist", block=192)


def test_tuner_picks_a_workgroup_size(gpu):
    g = gga.gen_rmat(15, 16, seed=9, wlo=1, whi=100)
    src = g.maxdeg()[1]
    with gga.DeviceGraph(g) as dg:
        announce("tune rmat15")
        tr = dg.tune(src)
        assert tr["block"] in (128, 256, 512)
        assert all(ms > 0 for ms in tr["block_ms"].values()), tr["block_ms"]
        d, _ = dg.sssp(src)   # the tuned size and grid
        np.testing.assert_array_equal(d, oracle.dijkstra(g.row_ptr, g.dst, g.w, src))


def test_near_far_flag_steps_past_the_byte_wrap(gpu):
    """Near-far SSSP with every step a flag step and one chunk per edge
    (T = 1), on queries of more than 256 steps: the step stamps' byte wraps,
    listing vertices whose stamp matches from 256 steps before -- including
    deferred ones above the threshold, which must not be expanded (a
    concurrent deferred improvement of such a vertex could otherwise be
    lost).  The loss depends on timing, so many queries run."""
    g = gga.gen_rmat(13, 8, seed=77, wlo=1, whi=100000)
    srcs = [int(s) for s in np.flatnonzero(g.degrees() > 8)[:6]]
    wrapped = 0
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=1)
        for s in srcs:
            for inf in (INF, gga.REF_INF):
                want = oracle.dijkstra(g.row_ptr, g.dst, g.w, s, inf=inf)
                for delta in (7, 100):
                    announce(f"near-far flags T=1 src={s} delta={delta} inf={inf}")
                    d, st = dg.sssp(s, mode="worklist", inf=inf, dense_min=1, delta=delta)
                    np.testing.assert_array_equal(d, want, err_msg=f"src={s} delta={delta} inf={inf}")
                    wrapped += st["iterations"] > 256
            d, st = dg.sssp(s, mode="persistent", dense_min=1, delta=7)
            np.testing.assert_array_equal(d, oracle.dijkstra(g.row_ptr, g.dst, g.w, s))
    assert wrapped >= 4, f"only {wrapped} queries ran past 256 steps"


DENSE = {"auto": 0, "flags": 1, "queue": 0xFFFFFFFF}


def test_list_rebuild_vector_path_edges(gpu):
    """The compaction and refill passes read each thread's eight vertices
    with vector loads when all eight exist, one at a time otherwise.  Graphs
    whose vertex count is not a multiple of 8 (15, 221, 1 073, 2 047) take
    both paths in the last tile; every step is a flag step (dense_min = 1)
    so every list goes through them, with and without near-far buckets
    (whose refill reads dist and last), single-source and 32-lane engine,
    and must equal the oracle."""
    graphs = [gga.gen_grid(3, 5, 0.0, seed=41, wlo=1, whi=9),
              gga.gen_grid(13, 17, 0.02, seed=42, wlo=1, whi=1000),
              gga.gen_grid(37, 29, 0.01, seed=43, wlo=1, whi=100),
              gga.gen_grid(23, 89, 0.05, seed=44, wlo=1, whi=1000)]
    for g in graphs:
        assert g.n % 8 != 0
        rng = np.random.default_rng(g.n)
        srcs = [0, g.n - 1, int(rng.integers(0, g.n))]
        with gga.DeviceGraph(g) as dg:
            dg.split()
            for s in srcs:
                want_b = oracle.bfs_queue(g.row_ptr, g.dst, s)
                want_s = oracle.dijkstra(g.row_ptr, g.dst, g.w, s)
                for mode in ("worklist", "persistent"):
                    announce(f"vector path n={g.n} src={s} {mode}")
                    d, _ = dg.bfs(s, mode=mode, dense_min=1)
                    np.testing.assert_array_equal(d, want_b, err_msg=f"bfs n={g.n} {s} {mode}")
                    for delta in (0xFFFFFFFF, 1, 50):
                        d, _ = dg.sssp(s, mode=mode, dense_min=1, delta=delta)
                        np.testing.assert_array_equal(d, want_s, err_msg=f"sssp n={g.n} {s} {mode} {delta}")
            lanes = [int(x) for x in rng.choice(g.n, 32, replace=g.n < 32)]
            announce(f"vector path engine n={g.n}")
            out, _ = dg.multi_sp(lanes, bfs_lanes=0xFFFF, mode="worklist", dense_min=1)
            for i, s in enumerate(lanes):
                want = oracle.bfs_queue(g.row_ptr, g.dst, s) if i < 16 else oracle.dijkstra(g.row_ptr, g.dst, g.w, s)
                np.testing.assert_array_equal(out[i], want, err_msg=f"engine lane {i} n={g.n} src={s}")


def test_mesh_wavefront_continuation_equals_oracle(gpu):
    """On meshes a wave follows the vertices it lowered past its batch within
    the step (single source and the 32-lane engine).  The extra rounds only
    bring improvements forward, so every result must still equal the
    oracle's, in every mode and list representation, BFS (which then takes
    queue stamps, being no longer level-synchronous within a step) and SSSP
    with and without near-far buckets -- and the step count must fall well
    below the BFS eccentricity, showing that the rounds ran."""
    g = gga.gen_grid(160, 190, 0.0, seed=23, wlo=1, whi=1000)
    srcs = [0, g.n // 2 + 77, g.n - 1]
    with gga.DeviceGraph(g) as dg:
        dg.split()
        for s in srcs:
            want_b = oracle.bfs_queue(g.row_ptr, g.dst, s)
            want_s = oracle.dijkstra(g.row_ptr, g.dst, g.w, s)
            ecc = int(want_b.max())
            for mode in ("worklist", "persistent", "auto"):
                for rep in DENSE:
                    announce(f"mesh continuation src={s} mode={mode} frontier={rep}")
                    d, st = dg.bfs(s, mode=mode, dense_min=DENSE[rep])
                    np.testing.assert_array_equal(d, want_b, err_msg=f"bfs src={s} {mode} {rep}")
                    if mode == "worklist":
                        assert st["iterations"] < ecc // 2, (st["iterations"], ecc)
                    for delta in (0xFFFFFFFF, 0, 1, 2000):
                        d, _ = dg.sssp(s, mode=mode, dense_min=DENSE[rep], delta=delta)
                        np.testing.assert_array_equal(d, want_s, err_msg=f"sssp src={s} {mode} {rep} {delta}")
        # the engine: 16 BFS + 16 SSSP lanes on the same mesh
        rng = np.random.default_rng(9)
        lanes = [int(x) for x in rng.choice(g.n, 32, replace=False)]
        eng = gga.MultiSourceEngine(dg)
        try:
            announce("mesh continuation engine 32 lanes")
            eng.start(dict(enumerate(lanes)), bfs=set(range(16)))
            live, steps = 0xFFFFFFFF, 0
            while live and steps < 100000:
                live = eng.run(max_steps=12, mode="worklist", sync_every=12)
                steps += 12
            assert live == 0
            out = eng.fetch(range(32))
            eccs = []
            for i, s in enumerate(lanes):
                wb = oracle.bfs_queue(g.row_ptr, g.dst, s)
                want = wb if i < 16 else oracle.dijkstra(g.row_ptr, g.dst, g.w, s)
                np.testing.assert_array_equal(out[i], want, err_msg=f"engine lane {i} src {s}")
                eccs.append(int(wb.max()))
            assert steps < max(eccs), (steps, max(eccs))
        finally:
            eng.close()
