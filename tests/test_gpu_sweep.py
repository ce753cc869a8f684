"""GPU: a seeded sweep of single-source SSSP/BFS configurations against the
oracle, bit-exact -- graph family and size, weight range (unit, small, wide,
near-saturating), split threshold, launch mode, frontier representation
(stamps + compaction, queue pushes, the default switch), near-far bucket
width and source.  It crosses the step-level paths (re-listed owners left to
the next step, refills, list order) with each other at small sizes.
"""
import sys

import numpy as np
import pytest

import gpugraphalgo_amd as gga
import oracle

pytestmark = pytest.mark.gpu

INF = gga.INF
DENSE = {"auto": 0, "flags": 1, "queue": 0xFFFFFFFF}


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def make_graph(rng, kind):
    wide = [(1, 1), (1, 10), (1, 100), (1, 100000), (1 << 28, 1 << 29)][int(rng.integers(0, 5))]
    seed = int(rng.integers(1, 1 << 30))
    if kind == "rmat":
        return gga.gen_rmat(int(rng.integers(9, 14)), int(rng.integers(2, 17)), seed=seed,
                            wlo=wide[0], whi=wide[1]), wide
    side = int(rng.integers(8, 70))
    return gga.gen_grid(side, side + int(rng.integers(0, 9)), float(rng.choice([0.0, 0.01, 0.05])),
                        seed=seed, wlo=wide[0], whi=wide[1]), wide


@pytest.mark.parametrize("case", range(24))
def test_seeded_configuration_sweep(gpu, case):
    rng = np.random.default_rng(1000 + case)
    kind = "rmat" if case % 3 else "grid"
    g, wide = make_graph(rng, kind)
    srcs = np.flatnonzero(g.degrees() > 0)
    with gga.DeviceGraph(g) as dg:
        for T in rng.choice([1, 2, 7, 16, 64, 1024], size=2, replace=False):
            dg.split(mda=int(T))
            for _ in range(3):
                src = int(rng.choice(srcs)) if srcs.size else 0
                mode = str(rng.choice(["worklist", "persistent", "auto"]))
                rep = str(rng.choice(list(DENSE)))
                delta = int(rng.choice([0xFFFFFFFF, 0, 1, 7, 100, 5000]))
                inf = int(rng.choice([INF, gga.REF_INF]))
                announce(f"case {case} {kind} n={g.n} m={g.m} w={wide} T={T} src={src} "
                         f"mode={mode} frontier={rep} delta={delta} inf={inf}")
                want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src, inf=inf)
                d, st = dg.sssp(src, mode=mode, inf=inf, count=True, dense_min=DENSE[rep], delta=delta)
                np.testing.assert_array_equal(d, want, err_msg=f"sssp case {case}")
                assert st["reached"] == int((want != inf).sum())
                d, _ = dg.bfs(src, mode=mode, inf=inf, dense_min=DENSE[rep])
                np.testing.assert_array_equal(d, oracle.bfs_queue(g.row_ptr, g.dst, src, inf=inf),
                                              err_msg=f"bfs case {case}")
