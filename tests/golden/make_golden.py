"""Writes golden_sp.npz: BFS/SSSP golden vectors from textbook library routines.

Expected outputs come from scipy.sparse.csgraph.dijkstra / shortest_path
(unweighted=True) and are cross-checked against networkx before writing, so a
case is written only when two independent libraries agree.  Inputs are built
here with numpy (fixed seeds) except the R-MAT scale-14 case, whose CSR comes
from the library's own generator and is pinned by its sha256 (the test
regenerates it and checks the hash first).

Distances are exact integers (float64 holds them exactly); a distance >= inf
is stored as inf (unreached), matching the saturating uint32 convention.
Run from the repo root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra, shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF
REF_INF = 1000000000


def csr_from_edges(n, src, dst, w):
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_ptr, src + 1, 1)
    return np.cumsum(row_ptr).astype(np.uint32), dst.astype(np.uint32), w.astype(np.uint32)


def expected(row_ptr, dst, w, src, inf, bfs):
    n = row_ptr.size - 1
    s = np.repeat(np.arange(n), np.diff(row_ptr.astype(np.int64)))
    ww = np.ones(dst.size) if bfs else w.astype(np.float64)
    keep = s != dst  # self-loops never shorten a path with positive weights
    s, d, ww = s[keep], dst[keep].astype(np.int64), ww[keep]
    # duplicates: keep the minimum weight (csr_matrix would sum them)
    if s.size:
        key = s * n + d
        order = np.lexsort((ww, key))
        key, ww2 = key[order], ww[order]
        first = np.ones(key.size, dtype=bool)
        first[1:] = key[1:] != key[:-1]
        s, d, ww = (key[first] // n), (key[first] % n), ww2[first]
    A = sp.csr_matrix((ww, (s, d)), shape=(n, n))
    if bfs:
        dd = shortest_path(A, directed=True, unweighted=True, indices=src)
    else:
        dd = dijkstra(A, directed=True, indices=src)
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    G.add_weighted_edges_from(zip(s.tolist(), d.tolist(), ww.tolist()))
    nxd = nx.single_source_dijkstra_path_length(G, src)
    out = np.full(n, inf, dtype=np.uint64)
    for v in range(n):
        x = dd[v]
        if np.isfinite(x):
            assert nxd[v] == x, (v, nxd[v], x)
            if x < inf:
                out[v] = int(x)
        else:
            assert v not in nxd
    return out.astype(np.uint32)


def main():
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    rng = np.random.default_rng(20240607)
    arrays, index = {}, []

    def add(name, row_ptr, dst, w, sources, bfs_too=True, infs=(INF,), weighted=True):
        arrays[f"{name}.row_ptr"] = row_ptr
        arrays[f"{name}.dst"] = dst
        arrays[f"{name}.w"] = w
        for src in sources:
            for inf in infs:
                for bfs in ([True, False] if bfs_too and weighted else [not weighted]):
                    key = f"{name}.{'bfs' if bfs else 'sssp'}.{src}.{inf}"
                    arrays[key] = expected(row_ptr, dst, w, src, inf, bfs)
                    index.append({"graph": name, "algo": "bfs" if bfs else "sssp", "src": int(src),
                                  "inf": int(inf), "key": key})

    # graph.h:557-598 Graph::testgraph, 11 nodes, 0-based: 0->1,2,3; 2->4..8; 3->9,10
    src = np.array([0, 0, 0, 2, 2, 2, 2, 2, 3, 3])
    dst = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
    rp, d, w = csr_from_edges(11, src, dst, np.ones(10, dtype=np.int64))
    add("testgraph", rp, d, w, [0, 2, 10], weighted=False)

    # single vertex, and a graph with no edges
    add("single", np.zeros(2, dtype=np.uint32), np.zeros(0, dtype=np.uint32),
        np.zeros(0, dtype=np.uint32), [0])
    add("noedges", np.zeros(6, dtype=np.uint32), np.zeros(0, dtype=np.uint32),
        np.zeros(0, dtype=np.uint32), [0, 4])

    # random multigraphs with self-loops and duplicate arcs
    for i, (n, ef) in enumerate([(2, 3), (17, 4), (100, 5), (1000, 8), (3000, 3)]):
        m = n * ef
        s = rng.integers(0, n, m)
        t = rng.integers(0, n, m)
        hub = rng.integers(0, n)  # one heavy vertex so splitting has work
        s[: m // 5] = hub
        ww = rng.integers(1, 101, m)
        rp, d, w = csr_from_edges(n, s, t, ww)
        add(f"rand{i}", rp, d, w, sorted({0, int(hub), int(rng.integers(0, n))}),
            infs=(INF, REF_INF))

    # 20 x 20 grid (both arc directions) + 0.5% shortcuts, weights [1, 1000]
    R = C = 20
    es = []
    for r in range(R):
        for c in range(C):
            u = r * C + c
            if c + 1 < C:
                es.append((u, u + 1))
            if r + 1 < R:
                es.append((u, u + C))
    k = max(1, round(0.005 * len(es)))
    while k:
        u, v = rng.integers(0, R * C, 2)
        if u != v and (min(u, v), max(u, v)) not in es:
            es.append((int(min(u, v)), int(max(u, v))))
            k -= 1
    uw = rng.integers(1, 1001, len(es))
    s = np.array([a for a, _ in es] + [b for _, b in es])
    t = np.array([b for _, b in es] + [a for a, _ in es])
    rp, d, w = csr_from_edges(R * C, s, t, np.concatenate([uw, uw]))
    add("grid20", rp, d, w, [0, R * C - 1])

    # saturation: 0 -> 1 (3e9), 1 -> 2 (2e9), 0 -> 3 (999999999), 3 -> 4 (5)
    s = np.array([0, 1, 0, 3])
    t = np.array([1, 2, 3, 4])
    ww = np.array([3000000000, 2000000000, 999999999, 5], dtype=np.int64)
    rp, d, w = csr_from_edges(5, s, t, ww)
    add("saturate", rp, d, w, [0], bfs_too=False, infs=(INF, REF_INF))

    # R-MAT scale 14, edge factor 8, seed 1, dedup, BFS from 0 (the CPU reference config)
    import gpugraphalgo_amd as gga
    g = gga.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1, wlo=1, whi=1, dedup=True)
    h = hashlib.sha256(g.row_ptr.tobytes() + g.dst.tobytes()).hexdigest()
    key = "rmat14.bfs.0"
    arrays[key] = expected(g.row_ptr, g.dst, g.w, 0, INF, True)
    meta = {"index": index, "rmat14": {"sha256": h, "n": g.n, "m": g.m, "key": key}}
    np.savez_compressed(os.path.join(HERE, "golden_sp.npz"), **arrays)
    with open(os.path.join(HERE, "golden_sp.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(len(index), "vectors; rmat14 m =", g.m)


if __name__ == "__main__":
    main()
