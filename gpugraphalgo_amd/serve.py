"""Python view of the serving loop (gga_server_*, csrc/host/serve.c).

The scheduler, admission and batching run in the C library; this module only
builds query traces and marshals them.  A trace is a numpy structured array
with fields graph, src, bfs, arrival (seconds from the replay's start).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, load
from .graph import DeviceGraph

TRACE_DTYPE = np.dtype([("graph", np.uint32), ("src", np.uint32), ("bfs", np.uint32),
                        ("arrival", np.float64)])
RESULT_FIELDS = ("id", "graph", "src", "bfs", "lane", "engine", "status", "arrival", "admitted", "done",
                 "checksum")


def checksum(dist: np.ndarray) -> int:
    """gga_checksum: sum over v of splitmix64(v << 32 | dist[v]) mod 2^64."""
    d = np.ascontiguousarray(dist, dtype=np.uint32)
    return int(load().gga_checksum(d.ctypes.data_as(ctypes.c_void_p), d.size))


def poisson_trace(sizes, n_queries: int, rate: float, seed: int, bfs_frac: float = 0.5,
                  degrees=None) -> np.ndarray:
    """Poisson arrivals at `rate` q/s, uniform graph, uniform source (among
    vertices with out-edges when `degrees` is given), BFS with probability bfs_frac."""
    rng = np.random.default_rng(seed)
    t = np.zeros(n_queries, dtype=TRACE_DTYPE)
    gaps = rng.exponential(1.0 / rate, n_queries) if rate > 0 else np.zeros(n_queries)
    t["arrival"] = np.cumsum(gaps) - (gaps[0] if n_queries else 0.0)
    t["graph"] = rng.integers(0, len(sizes), n_queries)
    for i in range(n_queries):
        g = int(t["graph"][i])
        if degrees is not None:
            cand = degrees[g]
            t["src"][i] = cand[rng.integers(0, cand.size)]
        else:
            t["src"][i] = rng.integers(0, sizes[g])
    t["bfs"] = rng.random(n_queries) < bfs_frac
    return t


def _queries(trace, outs=None):
    q = (_lib.Query * len(trace))()
    for i, r in enumerate(trace):
        q[i].graph = int(r["graph"])
        q[i].src = int(r["src"])
        q[i].bfs = int(r["bfs"])
        q[i].arrival = float(r["arrival"])
        q[i].out_on_device = 0
        q[i].out = None
        if outs is not None and outs[i] is not None:
            o = outs[i]
            if isinstance(o, np.ndarray):
                q[i].out = o.ctypes.data
            else:                      # (device pointer,)
                q[i].out = int(o[0])
                q[i].out_on_device = 1
    return q


def _results(res) -> dict:
    out = {f: np.array([getattr(r, f) for r in res]) for f in RESULT_FIELDS}
    out["checksum"] = np.array([r.checksum for r in res], dtype=np.uint64)
    return out


class Server:
    """gga_server: resident graphs, budgeted engines, continuous batching."""

    def __init__(self, hbm_budget: int = 0, steps_per_round: int = 0, lanes: int = 0,
                 use_graph: bool = False, dense_min: int = 0, host_results: bool = False,
                 dry_run: bool = False, dry_steps: int = 0, layout: str = "auto",
                 engines_per_graph: int = 0, mesh_reserve: int = 0):
        o = _lib.ServerOpts(hbm_budget, steps_per_round, lanes, 1 if use_graph else 0, dense_min,
                            1 if host_results else 0, 1 if dry_run else 0, dry_steps,
                            {"auto": 0, "vertex": 1, "lane": 2}[layout], engines_per_graph, mesh_reserve)
        self._h = ctypes.c_void_p()
        check(load().gga_server_create(ctypes.byref(o), ctypes.byref(self._h)), "gga_server_create")
        self.graphs: list = []

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().gga_server_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_graph(self, dg: DeviceGraph) -> int:
        gid = ctypes.c_uint32()
        check(load().gga_server_add_graph(self._h, dg._h, ctypes.byref(gid)), "gga_server_add_graph")
        self.graphs.append(dg)
        return int(gid.value)

    def add_graph_dry(self, n: int, chunks: int) -> int:
        gid = ctypes.c_uint32()
        check(load().gga_server_add_graph_dry(self._h, n, chunks, ctypes.byref(gid)),
              "gga_server_add_graph_dry")
        self.graphs.append(None)
        return int(gid.value)

    def submit(self, graph: int, src: int, bfs: bool, out=None) -> int:
        q = _queries(np.array([(graph, src, int(bfs), 0.0)], dtype=TRACE_DTYPE),
                     None if out is None else [out])
        qid = ctypes.c_uint64()
        check(load().gga_server_submit(self._h, q, ctypes.byref(qid)), "gga_server_submit")
        return int(qid.value)

    def pump(self) -> int:
        act = ctypes.c_uint32()
        check(load().gga_server_pump(self._h, ctypes.byref(act)), "gga_server_pump")
        return int(act.value)

    def drain(self) -> None:
        check(load().gga_server_drain(self._h), "gga_server_drain")

    def poll(self, cap: int = 4096) -> dict:
        res = (_lib.QueryResult * cap)()
        n = ctypes.c_uint32()
        check(load().gga_server_poll(self._h, res, cap, ctypes.byref(n)), "gga_server_poll")
        return _results(res[: n.value])

    def replay(self, trace: np.ndarray, outs=None) -> dict:
        q = _queries(trace, outs)
        res = (_lib.QueryResult * len(trace))()
        check(load().gga_server_replay(self._h, q, len(trace), res), "gga_server_replay")
        return _results(res)

    def replay_plain(self, trace: np.ndarray, mode: str = "persistent", outs=None) -> dict:
        q = _queries(trace, outs)
        res = (_lib.QueryResult * len(trace))()
        check(load().gga_server_replay_plain(self._h, q, len(trace), _lib.MODES[mode], res),
              "gga_server_replay_plain")
        return _results(res)

    def stats(self) -> dict:
        st = _lib.ServerCounters()
        check(load().gga_server_stats(self._h, ctypes.byref(st)), "gga_server_stats")
        return {f: getattr(st, f) for f, _ in st._fields_}


def latency_summary(res: dict) -> dict:
    ok = res["status"] == 0
    lat = (res["done"] - res["arrival"])[ok] * 1e3
    span = float(res["done"][ok].max() - res["arrival"].min()) if ok.any() else 0.0
    return {"completed": int(ok.sum()), "qps": (int(ok.sum()) / span) if span > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50)) if lat.size else 0.0,
            "p99_ms": float(np.percentile(lat, 99)) if lat.size else 0.0,
            "makespan_s": span}
