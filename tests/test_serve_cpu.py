"""Serving loop policy on the CPU (dry-run servers: engines are only accounted,
a query completes a fixed number of rounds after admission): admission by the
device-memory budget, eviction of idle engines, rejection of queries whose
engine exceeds the whole budget, continuous refilling of freed lanes, no
head-of-line blocking across graphs, and the fingerprint the GPU results use.
"""
import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
from gpugraphalgo_amd.serve import Server, TRACE_DTYPE, checksum, poisson_trace


def engine_bytes(n, chunks):
    return int(_lib.load().gga_ms_bytes(n, chunks))


def trace_of(rows):
    t = np.zeros(len(rows), dtype=TRACE_DTYPE)
    for i, (g, s, b, a) in enumerate(rows):
        t[i] = (g, s, b, a)
    return t


def test_engine_bytes_formula():
    n, c = 1000, 1500
    ntiles = (n + 2047) // 2048
    assert engine_bytes(n, c) == 128 * n + 8 * n + 8 * c + 4 * (ntiles + 1) + 4 * 44 * 32 + 32 * 1024 + 8 * 32


def test_budget_never_exceeded_and_all_complete():
    sizes = [(1000, 1200), (5000, 6000), (20000, 21000), (3000, 3100)]
    budget = engine_bytes(20000, 21000) + engine_bytes(1000, 1200)   # fits the big one + one small
    with Server(hbm_budget=budget, dry_run=True, dry_steps=2) as s:
        ids = [s.add_graph_dry(n, c) for n, c in sizes]
        rng = np.random.default_rng(0)
        rows = [(int(g), int(rng.integers(0, sizes[g][0])), int(rng.integers(0, 2)), 0.0)
                for g in rng.integers(0, len(ids), 200)]
        res = s.replay(trace_of(rows))
        st = s.stats()
    assert (res["status"] == 0).all()
    assert st["peak_bytes"] <= budget
    assert st["evictions"] > 0
    assert st["completed"] == 200
    # a query is admitted no earlier than it arrived and completes after admission
    assert (res["admitted"] >= res["arrival"]).all() and (res["done"] >= res["admitted"]).all()


def test_engine_larger_than_budget_is_rejected():
    with Server(hbm_budget=engine_bytes(1000, 1000), dry_run=True) as s:
        small = s.add_graph_dry(1000, 1000)
        big = s.add_graph_dry(100000, 100000)
        res = s.replay(trace_of([(big, 5, 1, 0.0), (small, 3, 0, 0.0), (big, 7, 0, 0.0)]))
    assert list(res["status"]) == [_lib.GGA_EBUDGET, 0, _lib.GGA_EBUDGET]


def test_freed_lanes_are_refilled_continuously():
    """4 lanes, 10 queries at once on one graph, 3 collected step batches
    each: admission in waves of 4, each wave admitted in the round that frees
    the lanes, FIFO order."""
    with Server(lanes=4, dry_run=True, dry_steps=3, engines_per_graph=1) as s:
        g = s.add_graph_dry(100, 100)
        for i in range(10):
            s.submit(g, i, i % 2 == 0)
        rounds_seen = []
        done = []
        while True:
            act = s.pump()
            r = s.poll()
            done += list(r["id"])
            rounds_seen.append((act, len(done)))
            if len(done) == 10:
                break
            assert len(rounds_seen) < 50
        st = s.stats()
    assert done == list(range(10))                 # FIFO within one graph
    assert max(a for a, _ in rounds_seen) <= 4     # never more lanes than allowed
    # pump 0 launches; pumps 1-3 collect; the lanes freed in pump 3 restart in pump 3
    assert [d for _, d in rounds_seen[:10]] == [0, 0, 0, 4, 4, 4, 8, 8, 8, 10]
    assert st["engines_created"] == 1


def test_no_head_of_line_blocking_across_graphs():
    """Graph A's engine is full; a later query of graph B is admitted in the
    same round instead of waiting behind A's queue."""
    with Server(lanes=1, dry_run=True, dry_steps=5, engines_per_graph=1) as s:
        a = s.add_graph_dry(100, 100)
        b = s.add_graph_dry(100, 100)
        s.submit(a, 1, True)
        s.submit(a, 2, True)     # waits for A's only lane
        s.submit(b, 3, False)    # must not wait behind it
        s.pump()
        assert s.stats()["engines_created"] == 2
        s.drain()
        r = s.poll()
    by_id = {int(i): k for k, i in enumerate(r["id"])}
    adm = {i: r["admitted"][k] for i, k in by_id.items()}
    done = {i: r["done"][k] for i, k in by_id.items()}
    assert adm[1] >= done[0]          # A's second query waited for A's lane
    assert adm[2] == adm[0]           # B's query was admitted in the first round


def test_second_engine_takes_a_full_engines_queries():
    """Graph A's only lane is busy: its next query starts at once on a second
    engine of A; a third waits (two engines per graph).  With the budget for
    one engine, the second query waits for the first lane instead."""
    with Server(lanes=1, dry_run=True, dry_steps=4, engines_per_graph=2) as s:
        a = s.add_graph_dry(100, 100)
        res = s.replay(trace_of([(a, 1, 1, 0.0), (a, 2, 0, 0.0), (a, 3, 1, 0.0)]))
        st = s.stats()
    assert (res["status"] == 0).all() and st["engines_created"] == 2
    assert res["admitted"][1] == res["admitted"][0] and res["engine"][1] != res["engine"][0]
    assert res["admitted"][2] >= min(res["done"][0], res["done"][1])
    with Server(hbm_budget=engine_bytes(100, 100), lanes=1, dry_run=True, dry_steps=4,
                engines_per_graph=2) as s:
        a = s.add_graph_dry(100, 100)
        res = s.replay(trace_of([(a, 1, 1, 0.0), (a, 2, 0, 0.0)]))
        st = s.stats()
    assert st["engines_created"] == 1 and res["admitted"][1] >= res["done"][0]


def test_queries_wait_for_busy_engines_then_run():
    """Budget for one engine: B's query waits until A's lanes drain, then A is evicted."""
    with Server(hbm_budget=engine_bytes(100, 100), lanes=2, dry_run=True, dry_steps=2) as s:
        a = s.add_graph_dry(100, 100)
        b = s.add_graph_dry(100, 100)
        res = s.replay(trace_of([(a, 0, 1, 0.0), (a, 1, 1, 0.0), (b, 2, 0, 0.0), (a, 3, 0, 0.0)]))
        st = s.stats()
    assert (res["status"] == 0).all()
    assert res["admitted"][2] >= max(res["done"][0], res["done"][1])
    assert st["evictions"] >= 1 and st["peak_bytes"] <= engine_bytes(100, 100)


def test_submit_validates_graph_and_source():
    with Server(dry_run=True) as s:
        g = s.add_graph_dry(10, 10)
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            s.submit(g + 1, 0, True)
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            s.submit(g, 10, True)


def test_poisson_trace_shape():
    t = poisson_trace([100, 200, 300], 1000, rate=500.0, seed=3)
    assert t["arrival"][0] == 0.0 and (np.diff(t["arrival"]) >= 0).all()
    assert 1.5 < t["arrival"][-1] < 2.5                     # ~1000 / 500 s
    assert set(np.unique(t["graph"])) == {0, 1, 2}
    assert 0.4 < t["bfs"].mean() < 0.6
    assert (t["src"] < np.array([100, 200, 300])[t["graph"]]).all()


def test_checksum_is_order_independent_fingerprint():
    rng = np.random.default_rng(5)
    d = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)

    def mix(x):
        x = (x + np.uint64(0x9E3779B97F4A7C15))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))

    with np.errstate(over="ignore"):
        v = np.arange(d.size, dtype=np.uint64)
        want = int(mix((v << np.uint64(32)) | d.astype(np.uint64)).sum(dtype=np.uint64))
    assert checksum(d) == want
    e = d.copy()
    e[17] ^= 1
    assert checksum(e) != checksum(d)


def test_submit_and_poll_from_other_threads_while_pumping():
    """Submitters on two threads (thousands of queries, so the query table
    grows under the pumping thread) and a poller on a third: every query
    completes and is polled at most once, each replay result matches its own
    query."""
    import threading
    with Server(dry_run=True, dry_steps=2, lanes=8) as s:
        g = [s.add_graph_dry(1000, 1100), s.add_graph_dry(2000, 2100)]
        n_each = 3000
        ids = {0: [], 1: []}
        polled = []
        done = threading.Event()

        def submitter(k):
            for i in range(n_each):
                ids[k].append(s.submit(g[k], i % 1000, bool(i & 1)))

        def poller():
            while not done.is_set():
                polled.extend(int(x) for x in s.poll()["id"])
            polled.extend(int(x) for x in s.poll()["id"])

        th = [threading.Thread(target=submitter, args=(k,)) for k in (0, 1)]
        pt = threading.Thread(target=poller)
        pt.start()
        for t in th:
            t.start()
        rows = [(g[0], 7, 1, 0.0), (g[1], 9, 0, 0.0)] * 50
        res = s.replay(trace_of(rows))   # pumps until idle, meanwhile others submit
        for t in th:
            t.join()
        s.drain()
        done.set()
        pt.join()
        st = s.stats()
    assert (res["status"] == 0).all()
    assert list(res["graph"]) == [r[0] for r in rows] and list(res["src"]) == [r[1] for r in rows]
    submitted = ids[0] + ids[1]
    assert len(set(submitted)) == 2 * n_each
    # every result is polled at most once; the submitted ones all arrive, and
    # the replay's own may reach a concurrent poller before the replay ends
    assert len(polled) == len(set(polled))
    assert set(submitted) <= set(polled)
    assert set(polled) - set(submitted) <= {int(x) for x in res["id"]}
    assert st["completed"] == 2 * n_each + len(rows)


def test_joined_counts_queries_admitted_beside_running_lanes():
    """Two queries start; two more arrive while their lanes run and join the
    same engine (continuous batching), so four run at once."""
    with Server(lanes=8, dry_run=True, dry_steps=3, engines_per_graph=1) as s:
        g = s.add_graph_dry(100, 100)
        s.submit(g, 1, True)
        s.submit(g, 2, False)
        s.pump()                       # admitted and launched
        st = s.stats()
        assert st["joined"] == 0 and st["peak_running"] == 2
        s.submit(g, 3, True)
        s.submit(g, 4, False)
        s.pump()                       # the first two are still running
        st = s.stats()
        assert st["joined"] == 2 and st["peak_running"] == 4
        s.drain()
        assert s.stats()["completed"] == 4


def test_mesh_reserve_range():
    """The workgroup-slot reserve for mesh engines accepts -1 (none) and 0..6
    eighths; anything else is an argument error."""
    for r in (-1, 0, 1, 6):
        Server(dry_run=True, mesh_reserve=r).close()
    for r in (-2, 7, 8):
        with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
            Server(dry_run=True, mesh_reserve=r)
