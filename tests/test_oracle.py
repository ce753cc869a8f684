"""Pins the CPU oracle (oracle/gga_oracle.c) before anything is compared to it.

Each restated reference function is checked against something that does not
share its code: scipy/networkx golden vectors, brute force, the reference's
own published result rows, or closed forms.
"""
import itertools

import numpy as np
import pytest

import oracle
from conftest import golden_graph

INF = oracle.INF


def test_serial_sp_matches_golden_vectors(golden):
    index, arrays, _ = golden
    for e in index:
        rp, dst, w = golden_graph(arrays, e["graph"])
        bfs = e["algo"] == "bfs"
        d, sweeps, _ = oracle.serial_sp(rp, dst, None if bfs else w, e["src"], e["inf"])
        np.testing.assert_array_equal(d, arrays[e["key"]], err_msg=e["key"])
        assert 1 <= sweeps <= max(1, rp.size - 1)


def test_textbook_routines_match_golden_vectors(golden):
    index, arrays, _ = golden
    for e in index:
        rp, dst, w = golden_graph(arrays, e["graph"])
        if e["algo"] == "bfs":
            d = oracle.bfs_queue(rp, dst, e["src"], e["inf"])
        else:
            d = oracle.dijkstra(rp, dst, w, e["src"], e["inf"])
        np.testing.assert_array_equal(d, arrays[e["key"]], err_msg=e["key"])


@pytest.mark.parametrize("mda", [1, 2, 3, 7, 64])
def test_split_relax_matches_golden_vectors(golden, mda):
    """The reference's split-graph algorithm (pseudo-children, remapped dst) must
    give the unsplit distances: pins orc_node_split's renumbering end to end."""
    index, arrays, _ = golden
    for e in index:
        rp, dst, w = golden_graph(arrays, e["graph"])
        bfs = e["algo"] == "bfs"
        d, _ = oracle.split_relax(rp, dst, None if bfs else w, mda, e["src"], e["inf"])
        np.testing.assert_array_equal(d, arrays[e["key"]], err_msg=f"{e['key']} mda={mda}")


def test_rmat14_bfs_config(golden):
    """R-MAT s14 ef8 seed 1, BFS from 0: generator pinned by hash, levels by scipy."""
    import hashlib
    import gpugraphalgo_amd as gga
    _, arrays, meta = golden
    r = meta["rmat14"]
    g = gga.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1, wlo=1, whi=1, dedup=True)
    assert (g.n, g.m) == (r["n"], r["m"])
    assert hashlib.sha256(g.row_ptr.tobytes() + g.dst.tobytes()).hexdigest() == r["sha256"]
    d, sweeps, mx = oracle.serial_sp(g.row_ptr, g.dst, None, 0)
    np.testing.assert_array_equal(d, arrays[r["key"]])
    assert mx == int(d[d != INF].max())


def _floyd(n, edges, inf):
    D = [[None] * n for _ in range(n)]
    for i in range(n):
        D[i][i] = 0
    for u, v, w in edges:
        if u != v and (D[u][v] is None or w < D[u][v]):
            D[u][v] = w
    for k, i, j in itertools.product(range(n), repeat=3):
        if D[i][k] is not None and D[k][j] is not None:
            s = D[i][k] + D[k][j]
            if D[i][j] is None or s < D[i][j]:
                D[i][j] = s
    return [[inf if (x is None or x >= inf) else x for x in row] for row in D]


def test_serial_sp_brute_force_tiny():
    rng = np.random.default_rng(7)
    for trial in range(60):
        n = int(rng.integers(1, 8))
        m = int(rng.integers(0, 3 * n + 1))
        us, vs = rng.integers(0, n, m), rng.integers(0, n, m)
        ws = rng.integers(1, 6 if trial % 2 else 2**31, m)
        order = np.lexsort((vs, us))
        us, vs, ws = us[order], vs[order], ws[order]
        rp = np.zeros(n + 1, dtype=np.uint32)
        np.add.at(rp, us + 1, 1)
        rp = np.cumsum(rp).astype(np.uint32)
        for inf in (INF, oracle.REF_INF):
            D = _floyd(n, list(zip(us.tolist(), vs.tolist(), ws.tolist())), inf)
            for s in range(n):
                d, _, _ = oracle.serial_sp(rp, vs, ws, s, inf)
                assert d.tolist() == D[s], (trial, s, inf)


def test_verify_and_tight_catch_planted_errors(golden):
    _, arrays, _ = golden
    rp, dst, w = golden_graph(arrays, "rand3")
    good = arrays["rand3.sssp.0.4294967295"]
    assert oracle.verify(rp, dst, w, good) == 0
    assert oracle.tight(rp, dst, w, 0, good) == 0
    reached = np.flatnonzero((good != INF) & (np.arange(good.size) != 0))
    v = int(reached[len(reached) // 2])
    too_big = good.copy()
    too_big[v] += 5            # a relaxable edge now exists into v
    assert oracle.verify(rp, dst, w, too_big) > 0
    too_small = good.copy()
    too_small[v] -= 1          # still a fixpoint locally, but v has no tight predecessor
    assert oracle.tight(rp, dst, w, 0, too_small) > 0
    bad_src = good.copy()
    bad_src[0] = 1
    assert oracle.tight(rp, dst, w, 0, bad_src) > 0


def test_maxdeg_first_index():
    rp = np.array([0, 3, 3, 6, 8, 11], dtype=np.uint32)   # degrees 3,0,3,2,3
    assert oracle.maxdeg(rp) == (3, 0)
    rp = np.array([0, 0, 2, 5, 8], dtype=np.uint32)        # degrees 0,2,3,3
    assert oracle.maxdeg(rp) == (3, 2)


def test_split_kat_published_rows(split_kat):
    """MAX_EDGES_ALLOWED and ns_nnode of all 704 published node-split runs
    (result.vi), recomputed from the published degree histograms."""
    bad = []
    for r in split_kat["rows"]:
        dd = np.array(split_kat["dd"][r["graph"]], dtype=np.int64)
        deg, cnt = dd[:, 0].astype(np.uint32), dd[:, 1].astype(np.uint32)
        maxdeg = int(deg[cnt > 0].max())
        if r["level"] > 0:
            mda = oracle.mda_level(maxdeg, r["level"])
        else:
            mda = oracle.mda_histogram_dd(deg, cnt, r["bucket"])
        degs = np.repeat(deg, cnt)
        ns = oracle.split_count(degs, mda)
        got = (maxdeg, mda, ns, int(cnt.sum()), int(degs.astype(np.int64).sum()))
        want = (r["maxdeg"], r["mda"], r["ns_nnode"], r["nnode"], r["nedge"])
        if got != want:
            bad.append((r["file"], r["line"], got, want))
    assert len(split_kat["rows"]) == 704
    assert not bad, bad[:5]


def test_paper_histogram_kat(split_kat):
    """Report.pdf VI.B.2: 10 buckets give MDA 118 on rmat20 and 25 on webs."""
    for g, want in (("rmat20", 118), ("webs", 25)):
        dd = np.array(split_kat["dd"][g], dtype=np.int64)
        assert oracle.mda_histogram_dd(dd[:, 0], dd[:, 1], 10) == want
        assert oracle.mda_histogram(np.repeat(dd[:, 0], dd[:, 1]).astype(np.uint32), 10) == want


def test_node_split_closed_form():
    rng = np.random.default_rng(3)
    deg = rng.integers(0, 300, 400).astype(np.uint32)
    deg[5] = 10_000
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    for mda in (1, 2, 5, 32, 97, 10_000, 20_000):
        cb, own, cc, ps = oracle.node_split(rp, mda)
        C = own.size
        assert C == oracle.split_count(deg, mda)
        assert cb[0] == 0 and cb[-1] == rp[-1]
        sizes = np.diff(cb.astype(np.int64))
        assert (sizes >= 0).all() and (sizes <= mda).all()
        for v in (0, 5, 17, 399):
            ids = np.flatnonzero(own == v)
            d = int(deg[v])
            k = 1 if d == 0 else -(-d // mda)
            assert ids.size == k and ids[0] == v + ps[v]
            assert (np.diff(ids) == 1).all()
            assert cc[ids[0]] == k - 1
            if d:
                local = -(-d // k)
                assert (sizes[ids[:-1]] == local).all()
                assert sizes[ids[-1]] == d - (k - 1) * local
                assert cb[ids[0]] == rp[v]
        # owners are non-decreasing and every vertex owns a chunk
        assert (np.diff(own.astype(np.int64)) >= 0).all()
        assert np.unique(own).size == deg.size


def test_m_reach():
    rp = np.array([0, 2, 3, 3, 5], dtype=np.uint32)
    d = np.array([0, INF, 4, 1], dtype=np.uint32)
    assert oracle.m_reach(rp, d) == 2 + 0 + 2


def test_oracle_routines_agree_property():
    """Property check across the oracle's independent routines on random small
    multigraphs (self-loops, duplicate edges, isolated vertices, weights up to
    near-saturating): the restated serial Bellman-Ford, Dijkstra, the
    reference's split relaxation at several thresholds and, for unit weights,
    the queue BFS all give one answer, and verify / tightness accept it."""
    hyp = pytest.importorskip("hypothesis")
    st = hyp.strategies

    @hyp.settings(max_examples=150, deadline=None, derandomize=True)
    @hyp.given(n=st.integers(1, 12), data=st.data())
    def check(n, data):
        m = data.draw(st.integers(0, 4 * n))
        us = np.array(data.draw(st.lists(st.integers(0, n - 1), min_size=m, max_size=m)), dtype=np.int64)
        vs = np.array(data.draw(st.lists(st.integers(0, n - 1), min_size=m, max_size=m)), dtype=np.uint32)
        wmax = data.draw(st.sampled_from([1, 7, 1000, 2**31]))
        ws = np.array(data.draw(st.lists(st.integers(1, wmax), min_size=m, max_size=m)), dtype=np.uint32)
        order = np.argsort(us, kind="stable")
        us, vs, ws = us[order], vs[order], ws[order]
        rp = np.zeros(n + 1, dtype=np.uint32)
        np.add.at(rp, us + 1, 1)
        rp = np.cumsum(rp).astype(np.uint32)
        src = data.draw(st.integers(0, n - 1))
        inf = data.draw(st.sampled_from([INF, oracle.REF_INF]))
        want = oracle.dijkstra(rp, vs, ws, src, inf)
        d, _, _ = oracle.serial_sp(rp, vs, ws, src, inf)
        assert d.tolist() == want.tolist()
        for T in (1, 2, 5):
            ds, _ = oracle.split_relax(rp, vs, ws, T, src, inf)
            assert ds.tolist() == want.tolist(), T
        assert oracle.verify(rp, vs, ws, want, inf) == 0
        assert oracle.tight(rp, vs, ws, src, want, inf) == 0
        bq = oracle.bfs_queue(rp, vs, src, inf)
        db, _, _ = oracle.serial_sp(rp, vs, None, src, inf)
        assert bq.tolist() == db.tolist()

    check()
