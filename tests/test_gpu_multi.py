"""GPU parity of the 32-lane multi-source engine (gga_ms_*, gga_multi_sp).

Every lane is checked against the oracle's single-source result from the same
source: batched runs in both launch modes, hipGraph-replayed steps, duplicate
and edge-less sources, and lanes recycled while other lanes keep running
(a lane's result alone equals its result in the mix).
"""
import sys

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
import oracle
from conftest import golden_graph

pytestmark = pytest.mark.gpu

INF = gga.INF


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def want_for(g, src, bfs):
    if bfs:
        return oracle.bfs_queue(g.row_ptr, g.dst, int(src))
    return oracle.dijkstra(g.row_ptr, g.dst, g.w, int(src))


def nonisolated(g, k, seed):
    deg = g.degrees()
    cand = np.flatnonzero(deg > 0)
    return np.random.default_rng(seed).choice(cand, size=k, replace=False).astype(np.uint32)


# next-list building: default switch, compaction every step, pushes every step
DENSE = {"auto": 0, "flags": 1, "queue": 0xFFFFFFFF}


LAYOUT = {"vertex": 1, "lane": 2}


@pytest.mark.parametrize("layout", list(LAYOUT))
@pytest.mark.parametrize("rep", list(DENSE))
@pytest.mark.parametrize("mode", ["worklist", "persistent"])
@pytest.mark.parametrize("T", [1, 16, 256])
def test_multi_sp_rmat_each_lane_equals_oracle(gpu, mode, T, rep, layout):
    g = gga.gen_rmat(14, 16, seed=13, wlo=1, whi=100)
    srcs = nonisolated(g, 32, T)
    srcs[5] = srcs[4]                    # duplicate source in two lanes
    bfs_lanes = 0xA5A5A5A5
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=T)
        announce(f"multi_sp rmat14 T={T} mode={mode} lists={rep} layout={layout}")
        out, st = dg.multi_sp(srcs, bfs_lanes=bfs_lanes, mode=mode, count=True, dense_min=DENSE[rep],
                              ms_layout=LAYOUT[layout])
        reached = 0
        mreach = 0
        for i, s in enumerate(srcs):
            want = want_for(g, s, (bfs_lanes >> i) & 1)
            np.testing.assert_array_equal(out[i], want, err_msg=f"lane {i} src {s}")
            reached += int((want != INF).sum())
            mreach += oracle.m_reach(g.row_ptr, want)
        assert st["reached"] == reached
        assert st["m_reach"] == mreach
        assert st["edges_inspected"] >= mreach


def test_multi_sp_golden_graphs_and_small_k(gpu, golden):
    index, arrays, _ = golden
    for name in sorted({e["graph"] for e in index}):
        rp, dst, w = golden_graph(arrays, name)
        g = gga.csr_from_arrays(rp, dst, w)
        n = g.n
        srcs = np.arange(min(n, 32), dtype=np.uint32)
        with gga.DeviceGraph(g) as dg:
            for mda in (1, 3):
                dg.split(mda=mda)
                for mode in ("worklist", "persistent"):
                    for k in sorted({1, srcs.size}):
                        announce(f"multi_sp golden {name} mda={mda} mode={mode} k={k}")
                        out, _ = dg.multi_sp(srcs[:k], bfs_lanes=0x3, mode=mode)
                        for i in range(k):
                            want = want_for(g, srcs[i], i < 2)
                            np.testing.assert_array_equal(out[i], want, err_msg=f"{name} lane {i}")


def test_multi_sp_hipgraph_replay_and_ref_inf(gpu):
    g = gga.gen_grid(60, 70, 0.01, seed=11, wlo=1, whi=1000)
    srcs = np.array([0, 4199, 77, 2100] * 8, dtype=np.uint32)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=4)
        for rep in range(3):
            announce(f"multi_sp grid60x70 hipgraph replay {rep}")
            out, st = dg.multi_sp(srcs, bfs_lanes=0xFFFF0000, mode="worklist", use_graph=1,
                                  dense_min=list(DENSE.values())[rep], ms_layout=1 + rep % 2)
            for i, s in enumerate(srcs):
                np.testing.assert_array_equal(out[i], want_for(g, s, i >= 16), err_msg=f"lane {i}")
        announce("multi_sp grid60x70 inf=1e9")
        out, _ = dg.multi_sp(srcs[:3], mode="persistent", inf=gga.REF_INF)
        for i in range(3):
            want = want_for(g, srcs[i], False)
            want = np.where(want == INF, gga.REF_INF, want).astype(np.uint32)
            np.testing.assert_array_equal(out[i], want)


@pytest.mark.parametrize("layout", list(LAYOUT))
def test_engine_lanes_recycled_while_others_run(gpu, layout):
    """Continuous batching at the engine level: lanes finish at different
    steps, are harvested and restarted with new sources (BFS or SSSP) while the
    remaining lanes keep running; every result equals the single-source one."""
    g = gga.gen_rmat(13, 8, seed=5, wlo=1, whi=100)
    rng = np.random.default_rng(1)
    queue = [(int(s), bool(rng.integers(0, 2))) for s in nonisolated(g, 120, 3)]
    queue += [(int(np.flatnonzero(g.degrees() == 0)[0]), False)] if (g.degrees() == 0).any() else []
    results = {}
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=32)
        with gga.MultiSourceEngine(dg, layout=layout) as eng:
            lane_q = {}
            pending = list(enumerate(queue))
            steps = 0
            while pending or lane_q:
                free = [l for l in range(32) if l not in lane_q]
                start = {}
                bfs = set()
                for l in free:
                    if not pending:
                        break
                    qi, (s, b) = pending.pop(0)
                    lane_q[l] = qi
                    start[l] = s
                    if b:
                        bfs.add(l)
                if start:
                    eng.start(start, bfs)
                live = eng.run(max_steps=2, mode="worklist" if steps % 2 else "persistent",
                               dense_min=list(DENSE.values())[steps % 3])
                steps += 1
                done = [l for l in lane_q if not (live >> l) & 1]
                if done:
                    assert all((eng.done >> l) & 1 for l in done)
                    got = eng.fetch(done)
                    for l in done:
                        results[lane_q.pop(l)] = got[l]
                assert steps < 10000
            announce(f"engine recycling: {len(results)} queries in {steps} rounds")
            with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
                eng.start({0: g.n})          # source out of range
    for qi, (s, b) in enumerate(queue):
        np.testing.assert_array_equal(results[qi], want_for(g, s, b), err_msg=f"query {qi} src {s}")


def test_engine_errors(gpu):
    g = gga.gen_rmat(10, 8, seed=2, wlo=1, whi=10)
    with gga.DeviceGraph(g) as dg:
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            dg.multi_sp(np.arange(33, dtype=np.uint32))
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            dg.multi_sp(np.array([g.n], dtype=np.uint32))
        with gga.MultiSourceEngine(dg) as eng:
            eng.start({0: 1})
            with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
                eng.start({0: 2})            # lane 0 is running
            with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
                eng.fetch([0])               # not finished
            assert eng.run() == 0
            np.testing.assert_array_equal(eng.fetch([0])[0], want_for(g, 1, False))
        dg.split(mda=4)                      # engines of the old chunk table are stale


@pytest.mark.parametrize("alpha", [0, 14, 1])
def test_multi_bfs_256_lanes_equal_oracle(gpu, alpha):
    """256 bitmask lanes, push-only (alpha 0xFFFFFFFF), the default switch, and
    pull from the first large level on; duplicate sources, k not a multiple of 64."""
    g = gga.gen_rmat(13, 16, seed=13, wlo=1, whi=100)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=32)
        for k in (256, 77, 1):
            srcs = nonisolated(g, k, k)
            if k > 3:
                srcs[3] = srcs[2]
            pa = 0xFFFFFFFF if alpha == 0 else alpha
            announce(f"multi_bfs rmat13 k={k} pull_alpha={pa}")
            out, st = dg.multi_bfs(srcs, pull_alpha=pa)
            reached = 0
            mreach = 0
            for i, s in enumerate(srcs):
                want = want_for(g, s, True)
                np.testing.assert_array_equal(out[i], want, err_msg=f"lane {i} src {s}")
                reached += int((want != INF).sum())
                mreach += oracle.m_reach(g.row_ptr, want)
            assert st["reached"] == reached and st["m_reach"] == mreach
            if alpha == 0:
                assert st["pull_levels"] == 0


def test_multi_bfs_golden_graphs(gpu, golden):
    """The bitmask BFS on every golden graph (a single vertex, no edges,
    self-loops and duplicate edges, the reference's testgraph, a grid): every
    vertex as a source (up to 256 lanes), push-only and pull-eager."""
    index, arrays, _ = golden
    for name in sorted({e["graph"] for e in index}):
        rp, dst, w = golden_graph(arrays, name)
        g = gga.csr_from_arrays(rp, dst, w)
        srcs = np.arange(min(g.n, 256), dtype=np.uint32)
        with gga.DeviceGraph(g) as dg:
            for mda in (1, 3):
                dg.split(mda=mda)
                for pa in (0xFFFFFFFF, 1):
                    announce(f"multi_bfs golden {name} mda={mda} k={srcs.size} pull_alpha={pa}")
                    out, _ = dg.multi_bfs(srcs, pull_alpha=pa)
                    for i, s in enumerate(srcs):
                        np.testing.assert_array_equal(out[i], want_for(g, s, True),
                                                      err_msg=f"{name} mda={mda} lane {i}")


def test_multi_bfs_directed_grid_and_isolated(gpu):
    """Directed graph (pull must use in-edges), an isolated source, REF_INF."""
    rp = np.array([0, 2, 3, 3, 4, 4], dtype=np.uint32)     # 0->1,0->2  1->3  3->1  4 isolated
    dst = np.array([1, 2, 3, 1], dtype=np.uint32)
    g = gga.csr_from_arrays(rp, dst, None)
    with gga.DeviceGraph(g) as dg:
        for pa in (0xFFFFFFFF, 1):
            out, _ = dg.multi_bfs(np.array([0, 4, 3, 2], dtype=np.uint32), inf=gga.REF_INF, pull_alpha=pa)
            for i, s in enumerate([0, 4, 3, 2]):
                want = want_for(g, s, True)
                want = np.where(want == INF, gga.REF_INF, want).astype(np.uint32)
                np.testing.assert_array_equal(out[i], want)
    g = gga.gen_grid(40, 50, 0.01, seed=11, wlo=1, whi=1000)
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=2)
        srcs = np.arange(0, 2000, 8, dtype=np.uint32)
        announce("multi_bfs grid40x50 k=250")
        out, st = dg.multi_bfs(srcs)
        for i, s in enumerate(srcs):
            np.testing.assert_array_equal(out[i], want_for(g, s, True), err_msg=f"lane {i}")
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            dg.multi_bfs(np.zeros(257, dtype=np.uint32))


def lollipop(path, leaves):
    """A path 0 .. path-1 (both directions) whose vertex 0 is also a hub with
    `leaves` leaf neighbours: deep levels, and a degree spread that does not
    read as a mesh."""
    src, dst = [], []
    for v in range(path - 1):
        src += [v, v + 1]
        dst += [v + 1, v]
    for i in range(leaves):
        src += [0, path + i]
        dst += [path + i, 0]
    n = path + leaves
    order = np.argsort(np.array(src), kind="stable")
    d = np.array(dst, dtype=np.uint32)[order]
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(rp, np.array(src) + 1, 1)
    return gga.csr_from_arrays(np.cumsum(rp).astype(np.uint32), d, None)


def test_multi_bfs_kept_fronts_and_deep_levels(gpu):
    """Levels below the kept-front count are written at the end from the front
    slots, later ones at discovery: a 300-level lollipop and a sparse R-MAT
    (levels past the kept slots, unreached vertices) mix the two in every
    lane, push-only and pull-eager, against the oracle."""
    cases = [("lollipop300+60", lollipop(300, 60), 2),
             ("rmat12 ef2", gga.gen_rmat(12, 2, seed=21, wlo=1, whi=1), 1)]
    for name, g, mda in cases:
        with gga.DeviceGraph(g) as dg:
            dg.split(mda=mda)
            for k in (200, 64):
                srcs = nonisolated(g, k, 3 + k)
                for pa in (0xFFFFFFFF, 1):
                    announce(f"multi_bfs {name} k={k} pull_alpha={pa}")
                    out, st = dg.multi_bfs(srcs, pull_alpha=pa)
                    depth = 0
                    for i, s in enumerate(srcs):
                        want = want_for(g, s, True)
                        np.testing.assert_array_equal(out[i], want, err_msg=f"{name} lane {i} src {s}")
                        depth = max(depth, int(want[want != INF].max()))
                    assert st["iterations"] >= depth >= 8   # level 8 is the first written at discovery


def test_autotuner_keeps_results_and_seeds_with_histogram(gpu):
    """gga_tune: the histogram heuristic (10 buckets) is the seed, the chosen
    threshold is one of the timed candidates, and every query after tuning
    (single source and multi-source in the chosen layout) equals the oracle."""
    for g, src in ((gga.gen_rmat(14, 16, seed=13, wlo=1, whi=100), 0),
                   (gga.gen_grid(80, 90, 0.005, seed=11, wlo=1, whi=1000), 0)):
        with gga.DeviceGraph(g) as dg:
            announce(f"tune n={g.n}")
            r = dg.tune(src, bfs=False, layout=True)
            assert r["seed_mda"] == oracle.mda_histogram(g.degrees(), 10)
            assert r["mda"] in r["tried"] and r["best_ms"] <= r["seed_ms"]
            assert r["layout"] in (1, 2) and dg.tuned_layout() == {1: "vertex", 2: "lane"}[r["layout"]]
            want = want_for(g, src, False)
            np.testing.assert_array_equal(dg.sssp(src)[0], want)
            np.testing.assert_array_equal(dg.bfs(src)[0], want_for(g, src, True))
            srcs = nonisolated(g, 32, 1)
            out, _ = dg.multi_sp(srcs, bfs_lanes=0xF0F0F0F0)
            for i, s in enumerate(srcs):
                np.testing.assert_array_equal(out[i], want_for(g, s, (0xF0F0F0F0 >> i) & 1))


@pytest.mark.parametrize("delta", [7, 500, 4000])
def test_near_far_engine_lanes_equal_oracle(gpu, delta):
    """Near-far SSSP lanes (per-lane thresholds, refills of deferred pairs)
    beside BFS lanes: every lane equals the oracle on a pure grid and R-MAT,
    both layouts, batched and with lanes recycled mid-run."""
    for g in (gga.gen_grid(60, 80, 0.0, seed=17, wlo=1, whi=1000), gga.gen_rmat(13, 8, seed=5, wlo=1, whi=100)):
        srcs = nonisolated(g, 32, delta)
        with gga.DeviceGraph(g) as dg:
            dg.split(mda=8)
            for layout in LAYOUT:
                announce(f"near-far engine n={g.n} delta={delta} layout={layout}")
                out, st = dg.multi_sp(srcs, bfs_lanes=0x0000FF00, delta=delta, ms_layout=LAYOUT[layout],
                                      use_graph=int(layout == "vertex"))
                for i, s in enumerate(srcs):
                    np.testing.assert_array_equal(out[i], want_for(g, s, 8 <= i < 16), err_msg=f"lane {i}")
            with gga.MultiSourceEngine(dg) as eng:
                eng.set_delta(delta)
                queue = [(int(s), i % 3 == 0) for i, s in enumerate(nonisolated(g, 70, delta + 1))]
                lane_q, pending, results, rounds = {}, list(enumerate(queue)), {}, 0
                while pending or lane_q:
                    start, bfs = {}, set()
                    for l in range(32):
                        if l in lane_q or not pending:
                            continue
                        qi, (s, b) = pending.pop(0)
                        lane_q[l] = qi
                        start[l] = s
                        if b:
                            bfs.add(l)
                    if start:
                        eng.start(start, bfs)
                    live = eng.run(max_steps=3, mode="worklist")
                    rounds += 1
                    done = [l for l in lane_q if not (live >> l) & 1]
                    if done:
                        got = eng.fetch(done)
                        for l in done:
                            results[lane_q.pop(l)] = got[l]
                    assert rounds < 20000
                for qi, (s, b) in enumerate(queue):
                    np.testing.assert_array_equal(results[qi], want_for(g, s, b), err_msg=f"query {qi}")
                with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
                    eng.start({0: int(srcs[0])})
                    eng.run(mode="persistent")
