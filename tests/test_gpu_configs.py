"""GPU: the batched and served workloads at their own sizes, against the oracle.

Config 4: R-MAT s20 ef16 seed 13, 256 BFS sources as bitmask lanes in one
cooperative launch and 32 SSSP sources in one 32-lane engine (hipGraph-replayed
steps, threshold and engine layout from the autotuner), sources uniform among
vertices with out-edges.  The reference answers one source per process
(bfs_sssp.cu:203); every checked lane must equal the oracle's single-source
result (queue BFS / Dijkstra, pinned against the restatement of
findSerialBFS/SSSP, graph.h:115-184).

Config 5: the resident set R-MAT s16-s23 ef16 + grids 512^2 .. 4096^2, seed
17, a fixed trace of 48 mixed BFS/SSSP queries over every graph arriving as
a Poisson stream, served with continuous batching; every result fingerprint must equal the fingerprint of
the oracle's distances.

The oracle runs in a thread pool (its C functions release the GIL).
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd.serve import Server, checksum, poisson_trace
import oracle

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
POOL = 8


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def config4_sources(g):
    rng = np.random.default_rng(13)
    cand = np.flatnonzero(g.degrees() > 0)
    return (rng.choice(cand, 256, replace=False).astype(np.uint32),
            rng.choice(cand, 32, replace=False).astype(np.uint32))


def test_config4_batched_lanes_equal_oracle(gpu):
    g = gga.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=13, wlo=1, whi=100)
    bsrc, ssrc = config4_sources(g)
    bfs_check = list(range(0, 256, 6)) + [255]          # 44 lanes spread over all four words
    with ThreadPoolExecutor(POOL) as ex:
        want_b = ex.map(lambda i: oracle.bfs_queue(g.row_ptr, g.dst, int(bsrc[i])), bfs_check)
        want_s = ex.map(lambda s: oracle.dijkstra(g.row_ptr, g.dst, g.w, int(s)), ssrc)
        with gga.DeviceGraph(g) as dg:
            announce("config4 tune (multi-source objective)")
            tr = dg.tune(int(ssrc[0]), bfs=False, layout=True, multi=True)
            announce(f"config4 256-lane BFS (T={tr['mda']})")
            lv, sb = dg.multi_bfs(bsrc)
            announce("config4 32-lane SSSP, hipGraph steps")
            ds, ss = dg.multi_sp(ssrc, use_graph=1)
        for i, w in zip(bfs_check, want_b):
            np.testing.assert_array_equal(lv[i], w, err_msg=f"BFS lane {i} src {bsrc[i]}")
        for i, w in enumerate(want_s):
            np.testing.assert_array_equal(ds[i], w, err_msg=f"SSSP lane {i} src {ssrc[i]}")
    assert sb["iterations"] > 0 and ss["iterations"] > 0


def config5_set():
    gs = [("rmat%d" % s, gga.gen_rmat(s, 16, 0.57, 0.19, 0.19, seed=17, wlo=1, whi=100)) for s in range(16, 24)]
    gs += [("grid%d" % a, gga.gen_grid(a, a, 0.0, seed=17, wlo=1, whi=1000)) for a in (512, 1024, 2048, 4096)]
    return gs


def test_config5_served_fingerprints_equal_oracle(gpu):
    announce("config5 generating the resident set")
    gs = config5_set()
    sizes = [g.n for _, g in gs]
    degs = [np.flatnonzero(g.degrees() > 0).astype(np.uint32) for _, g in gs]
    # 48 queries arriving at 2000 q/s, every graph at least twice, half BFS
    trace = poisson_trace(sizes, 48, rate=2000.0, seed=41, degrees=degs)
    for k in range(len(gs)):
        trace["graph"][2 * k] = k
        trace["src"][2 * k] = degs[k][(7919 * k) % degs[k].size]
        trace["graph"][2 * k + 1] = k
        trace["src"][2 * k + 1] = degs[k][(104729 * k + 1) % degs[k].size]
    trace["bfs"] = np.arange(48) % 2

    def want(r):
        _, g = gs[int(r["graph"])]
        d = (oracle.bfs_queue(g.row_ptr, g.dst, int(r["src"])) if r["bfs"]
             else oracle.dijkstra(g.row_ptr, g.dst, g.w, int(r["src"])))
        return checksum(d)

    with ThreadPoolExecutor(POOL) as ex:
        fut = [ex.submit(want, r) for r in trace]
        dgs = []
        try:
            for name, g in gs:
                dg = gga.DeviceGraph(g)
                if name.startswith("rmat"):
                    dg.split(mda=64)
                else:
                    dg.split()
                dgs.append(dg)
            with Server(hbm_budget=16 << 30, steps_per_round=12, use_graph=True) as srv:
                for dg in dgs:
                    srv.add_graph(dg)
                announce("config5 serving 48 mixed queries")
                res = srv.replay(trace)
                st = srv.stats()
        finally:
            for dg in dgs:
                dg.close()
        assert (res["status"] == 0).all(), res["status"]
        assert st["joined"] > 0 and st["peak_running"] > 1, st
        for i, f in enumerate(fut):
            r = trace[i]
            assert int(res["checksum"][i]) == f.result(), (
                f"query {i}: {gs[int(r['graph'])][0]} src {int(r['src'])} {'bfs' if r['bfs'] else 'sssp'}")
