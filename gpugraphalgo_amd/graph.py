"""Host-side mirror of the reference's graph program over the C-ABI.

The reference drives one graph through (bfs_sssp.cu:103-539):
    readFromGR -> NodeSplit(L, bucket) -> cudaCopy -> fixpoint loop -> Ary2File / writeResult
Here the same steps are
    read_gr / gen_rmat / gen_grid  ->  DeviceGraph(csr)  ->  .split(level=L, buckets=b)
    ->  .bfs(src) / .sssp(src)  ->  write_dist / write_result_row
all executed by libgga.so (C host + HIP kernels); this module only marshals
numpy arrays through ctypes.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import INF, REF_INF, MODES, check, load

_u32 = np.uint32


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class CsrGraph:
    """Host CSR: row_ptr[n + 1], dst[m], w[m] (uint32, 0-based)."""

    row_ptr: np.ndarray
    dst: np.ndarray
    w: np.ndarray
    has_weights: bool = True

    @property
    def n(self) -> int:
        return int(self.row_ptr.size - 1)

    @property
    def m(self) -> int:
        return int(self.dst.size)

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr.astype(np.int64)).astype(_u32)

    def maxdeg(self) -> tuple[int, int]:
        """(max out-degree, first vertex with it) -- getMaxdeg/getMaxdegIndex."""
        d = self.degrees()
        if d.size == 0:
            return 0, 0
        i = int(np.argmax(d))
        return int(d[i]), i

    def _as_c(self) -> _lib.Csr:
        c = _lib.Csr()
        c.n, c.m = self.n, self.m
        c.row_ptr = self.row_ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        c.dst = self.dst.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        c.w = self.w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        c.has_weights = 1 if self.has_weights else 0
        return c


def _take_csr(c: _lib.Csr) -> CsrGraph:
    """Copies a library-owned gga_csr into numpy and frees it."""
    L = load()
    try:
        n, m = c.n, c.m
        rp = np.ctypeslib.as_array(c.row_ptr, shape=(n + 1,)).copy()
        dst = np.ctypeslib.as_array(c.dst, shape=(max(m, 1),))[:m].copy()
        w = np.ctypeslib.as_array(c.w, shape=(max(m, 1),))[:m].copy()
        return CsrGraph(rp, dst, w, bool(c.has_weights))
    finally:
        L.gga_csr_free(ctypes.byref(c))


def csr_from_arrays(row_ptr, dst, w=None) -> CsrGraph:
    rp = np.ascontiguousarray(row_ptr, dtype=_u32)
    d = np.ascontiguousarray(dst, dtype=_u32)
    if w is None:
        return CsrGraph(rp, d, np.ones(d.size, dtype=_u32), False)
    return CsrGraph(rp, d, np.ascontiguousarray(w, dtype=_u32), True)


def read_gr(path: str) -> CsrGraph:
    """Galois .gr v1 reader (Graph::readFromGR, graph.h:599-682)."""
    c = _lib.Csr()
    check(load().gga_csr_read_gr(os.fsencode(path), ctypes.byref(c)), f"read_gr({path})")
    return _take_csr(c)


def write_gr(path: str, g: CsrGraph) -> None:
    c = g._as_c()
    check(load().gga_csr_write_gr(os.fsencode(path), ctypes.byref(c)), f"write_gr({path})")


def gen_rmat(scale: int, edge_factor: int, a=0.57, b=0.19, c=0.19, seed=1, wlo=1, whi=1,
             dedup=True) -> CsrGraph:
    out = _lib.Csr()
    check(load().gga_gen_rmat(scale, edge_factor, a, b, c, seed, wlo, whi, 1 if dedup else 0,
                              ctypes.byref(out)), "gen_rmat")
    g = _take_csr(out)
    g.has_weights = not (wlo == 1 and whi == 1)
    return g


def gen_grid(rows: int, cols: int, shortcut_frac=0.005, seed=11, wlo=1, whi=1000) -> CsrGraph:
    out = _lib.Csr()
    check(load().gga_gen_grid(rows, cols, shortcut_frac, seed, wlo, whi, ctypes.byref(out)),
          "gen_grid")
    return _take_csr(out)


def verify(g: CsrGraph, src: int, dist: np.ndarray, bfs: bool, inf: int = INF) -> dict:
    """Host fixpoint + tightness check (verifysolution, myutils.h:1-18, plus tightness)."""
    r = _lib.VerifyResult()
    dist = np.ascontiguousarray(dist, dtype=_u32)
    check(load().gga_verify(g.n, _ptr(g.row_ptr), _ptr(g.dst), None if bfs else _ptr(g.w), src,
                            _ptr(dist), inf, ctypes.byref(r)), "verify")
    return {"nerr": r.nerr, "untight": r.untight, "src_ok": bool(r.src_ok), "reached": r.reached}


def write_dist(path: str, dist: np.ndarray) -> None:
    """Ary2File (myutils.h:51-62): one "%u\\n" per original vertex."""
    dist = np.ascontiguousarray(dist, dtype=_u32)
    check(load().gga_write_dist(os.fsencode(path), _ptr(dist), dist.size), "write_dist")


def write_degree_dist(path: str, g: CsrGraph) -> None:
    """DD_<graph>.txt: "deg,count" per occurring out-degree (printfDegreeDistribution,
    graph.h:98-114)."""
    check(load().gga_write_degree_dist(path.encode(), g.n, _ptr(g.row_ptr)), "write_degree_dist")


def write_result_row(path: str, **fields) -> None:
    """writeResult (myutils.h:86-113)."""
    r = _lib.ResultRow()
    for k, v in fields.items():
        if isinstance(v, str):
            v = v.encode()
        setattr(r, k, v)
    check(load().gga_write_result_row(os.fsencode(path), ctypes.byref(r)), "write_result_row")


def device_count() -> int:
    return int(load().gga_device_count())


def _stats_dict(s: _lib.Stats) -> dict:
    return {f: getattr(s, f) for f, _ in _lib.Stats._fields_}


def _split_dict(s: _lib.SplitStats) -> dict:
    return {f: getattr(s, f) for f, _ in _lib.SplitStats._fields_}


class DeviceGraph:
    """A CSR resident in HBM on one device (gga_graph)."""

    def __init__(self, g: CsrGraph, device: int = 0):
        L = load()
        self._h = ctypes.c_void_p()
        self.n, self.m = g.n, g.m
        w = _ptr(g.w) if g.has_weights else None
        check(L.gga_graph_from_csr32(g.n, g.m, _ptr(g.row_ptr), _ptr(g.dst), w, device,
                                     ctypes.byref(self._h)), "gga_graph_from_csr32")
        self.split_stats: dict | None = None

    @classmethod
    def load_gr(cls, path: str, device: int = 0) -> "DeviceGraph":
        self = cls.__new__(cls)
        L = load()
        self._h = ctypes.c_void_p()
        check(L.gga_graph_load_gr(os.fsencode(path), device, ctypes.byref(self._h)),
              f"gga_graph_load_gr({path})")
        self.n = int(L.gga_graph_n(self._h))
        self.m = int(L.gga_graph_m(self._h))
        self.split_stats = None
        return self

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            load().gga_graph_free(self._h)
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

    @property
    def device_bytes(self) -> int:
        return int(load().gga_graph_device_bytes(self._h))

    def split(self, mda: int = 0, level: int = 0, buckets: int = 0) -> dict:
        """gga_graph_split: mda > 0 explicit T; level > 0 T = ceil(maxdeg / level);
        level < 0 histogram heuristic; all zero = no split."""
        st = _lib.SplitStats()
        check(load().gga_graph_split(self._h, mda, level, buckets, ctypes.byref(st)), "split")
        self.split_stats = _split_dict(st)
        return self.split_stats

    def chunks(self) -> tuple[np.ndarray, np.ndarray]:
        C = self.split_stats["ns_nnode"]
        cb = np.empty(C + 1, dtype=_u32)
        own = np.empty(C, dtype=_u32)
        check(load().gga_graph_chunks(self._h, _ptr(cb), _ptr(own)), "chunks")
        return cb, own

    def _run(self, fn, src: int, mode: str, inf: int, count: bool, keep_on_device: bool,
             **kw) -> tuple[np.ndarray | None, dict]:
        o = _opts(mode, inf, count, kw)
        st = _lib.Stats()
        out = None if keep_on_device else np.empty(self.n, dtype=_u32)
        check(fn(self._h, src, _ptr(out), ctypes.byref(o), ctypes.byref(st)), fn.__name__)
        return out, _stats_dict(st)

    def bfs(self, src: int, mode: str = "auto", inf: int = INF, count: bool = False,
            keep_on_device: bool = False, **kw):
        return self._run(load().gga_bfs, src, mode, inf, count, keep_on_device, **kw)

    def sssp(self, src: int, mode: str = "auto", inf: int = INF, count: bool = False,
             keep_on_device: bool = False, **kw):
        return self._run(load().gga_sssp, src, mode, inf, count, keep_on_device, **kw)

    def trace(self, src: int, bfs: bool = False, timeline_cap: int = 0, conv_final=None,
              conv_cap: int = 4096, inf: int = INF, **kw):
        """A worklist-mode query with the instrumentation on: per-workgroup
        timeline records (start, ticks, placement) for the first
        timeline_cap // grid steps, and, given the final distances, the
        per-step convergence rows.  Returns (dist, stats, timeline, conv)."""
        tl = np.zeros(timeline_cap, dtype=[("start", np.uint64), ("ticks", np.uint32), ("place", np.uint32)])
        rows = np.zeros(conv_cap if conv_final is not None else 0,
                        dtype=[("step", np.uint32), ("diff_count", np.uint32), ("norm2", np.float64)])
        fin = None if conv_final is None else np.ascontiguousarray(conv_final, dtype=_u32)
        kw = dict(kw)
        if timeline_cap:
            kw["timeline"] = tl.ctypes.data
            kw["timeline_cap"] = timeline_cap
        if fin is not None:
            kw["conv_final"] = fin.ctypes.data
            kw["conv_rows"] = rows.ctypes.data
            kw["conv_cap"] = conv_cap
        fn = load().gga_bfs if bfs else load().gga_sssp
        d, st = self._run(fn, src, "worklist", inf, True, False, **kw)
        return d, st, tl[: st["timeline_records"]], rows[: st["conv_steps"]]

    def copy_dist(self) -> np.ndarray:
        out = np.empty(self.n, dtype=_u32)
        check(load().gga_graph_copy_dist(self._h, _ptr(out)), "copy_dist")
        return out

    def verify(self, src: int, bfs: bool, inf: int = INF) -> dict:
        r = _lib.VerifyResult()
        check(load().gga_graph_verify(self._h, src, 1 if bfs else 0, inf, ctypes.byref(r)),
              "graph_verify")
        return {"nerr": r.nerr, "untight": r.untight, "src_ok": bool(r.src_ok),
                "reached": r.reached}

    def multi_sp(self, srcs, bfs_lanes=0, mode: str = "auto", inf: int = INF,
                 count: bool = False, keep_on_device: bool = False, **kw):
        """gga_multi_sp: up to 32 sources in one traversal.  bfs_lanes: bit i set
        -> source i is a BFS (hop counts).  Returns (k x n array or None, stats)."""
        srcs = np.ascontiguousarray(srcs, dtype=_u32)
        k = int(srcs.size)
        o = _opts(mode, inf, count, kw)
        st = _lib.Stats()
        out = None if keep_on_device else np.empty((k, self.n), dtype=_u32)
        check(load().gga_multi_sp(self._h, _ptr(srcs), k, int(bfs_lanes), _ptr(out), ctypes.byref(o),
                                  ctypes.byref(st)), "gga_multi_sp")
        return out, _stats_dict(st)

    def multi_bfs(self, srcs, inf: int = INF, keep_on_device: bool = False, **kw):
        """gga_multi_bfs: up to 256 BFS in one launch (bitmask lanes).
        Returns (k x n levels or None, stats)."""
        srcs = np.ascontiguousarray(srcs, dtype=_u32)
        k = int(srcs.size)
        o = _opts("persistent", inf, False, kw)
        st = _lib.Stats()
        out = None if keep_on_device else np.empty((k, self.n), dtype=_u32)
        check(load().gga_multi_bfs(self._h, _ptr(srcs), k, _ptr(out), ctypes.byref(o), ctypes.byref(st)),
              "gga_multi_bfs")
        return out, _stats_dict(st)


    def tune(self, src: int, bfs: bool = False, layout: bool = False, multi: bool = False) -> dict:
        """gga_tune: times split thresholds (seeded by the histogram heuristic),
        workgroup size, grid and push/compaction switch on queries from src
        (multi: 32-lane multi-source runs instead); keeps the best."""
        r = _lib.TuneResult()
        flags = (1 if bfs else 0) | (2 if layout else 0) | (4 if multi else 0)
        check(load().gga_tune(self._h, src, flags, ctypes.byref(r)), "gga_tune")
        out = {f: getattr(r, f) for f, _ in r._fields_
               if f not in ("tried_mda", "tried_ms", "layout_ms", "block_ms")}
        out["block_ms"] = {128: float(r.block_ms[0]), 256: float(r.block_ms[1]), 512: float(r.block_ms[2])}
        k = min(r.ntried, 16)
        out["tried"] = {int(r.tried_mda[i]): float(r.tried_ms[i]) for i in range(k)}
        out["layout_ms"] = (float(r.layout_ms[0]), float(r.layout_ms[1]))
        self.split_stats = None
        return out

    def tuned_layout(self) -> str:
        return {1: "vertex", 2: "lane"}[int(load().gga_tune_layout(self._h))]


def _opts(mode, inf: int, count: bool, kw: dict) -> _lib.Opts:
    o = _lib.Opts()
    o.mode = MODES[mode] if isinstance(mode, str) else int(mode)
    o.inf = inf
    o.count = 1 if count else 0
    names = {f for f, _ in _lib.Opts._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown gga_opts field {k!r}")
        setattr(o, k, v)
    return o


class MultiSourceEngine:
    """gga_ms: 32 traversal lanes on one DeviceGraph; lanes start, converge and
    restart independently (the serving loop's unit of continuous batching)."""

    LANES = 32

    LAYOUTS = {"auto": 0, "vertex": 1, "lane": 2}

    def __init__(self, dg: DeviceGraph, inf: int = INF, layout: str = "auto"):
        self.dg = dg
        self.n = dg.n
        self._h = ctypes.c_void_p()
        check(load().gga_ms_create(dg._h, inf, self.LAYOUTS[layout], ctypes.byref(self._h)), "gga_ms_create")

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            load().gga_ms_free(self._h)
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

    def start(self, lanes: dict, bfs: set | frozenset = frozenset()) -> None:
        """lanes: {lane: source}; bfs: lanes that count hops."""
        src = np.zeros(self.LANES, dtype=_u32)
        mask = unit = 0
        for lane, s in lanes.items():
            src[lane] = s
            mask |= 1 << lane
            if lane in bfs:
                unit |= 1 << lane
        check(load().gga_ms_start(self._h, mask, _ptr(src), unit), "gga_ms_start")

    def set_delta(self, delta: int) -> None:
        """gga_ms_set_delta: near-far bucket width of the SSSP lanes (0: off)."""
        check(load().gga_ms_set_delta(self._h, delta), "gga_ms_set_delta")

    def run(self, max_steps: int = 0, mode: str = "worklist", count: bool = False, **kw) -> int:
        """Runs steps; returns the mask of lanes with work left."""
        o = _opts(mode, 0, count, kw)
        live = ctypes.c_uint32()
        check(load().gga_ms_run(self._h, max_steps, ctypes.byref(o), ctypes.byref(live)), "gga_ms_run")
        return int(live.value)

    @property
    def done(self) -> int:
        return int(load().gga_ms_done(self._h))

    def fetch(self, lanes) -> dict:
        """{lane: distances} for finished lanes."""
        arrs = {i: np.empty(self.n, dtype=_u32) for i in lanes}
        ptrs = (ctypes.c_void_p * self.LANES)()
        mask = 0
        for i, a in arrs.items():
            ptrs[i] = a.ctypes.data
            mask |= 1 << i
        check(load().gga_ms_fetch(self._h, mask, ptrs, 0), "gga_ms_fetch")
        return arrs

    def stats(self) -> dict:
        st = _lib.Stats()
        check(load().gga_ms_stats(self._h, ctypes.byref(st)), "gga_ms_stats")
        return _stats_dict(st)


def timeline_makespan(tl: np.ndarray, grid: int, slots: int) -> float:
    """gga_timeline_makespan: list-scheduling estimate (ms) from timeline records."""
    tl = np.ascontiguousarray(tl)
    return float(load().gga_timeline_makespan(tl.ctypes.data_as(ctypes.c_void_p), tl.size, grid, slots))


def write_conv(path: str, rows: np.ndarray, level: int, maxdeg: int, graph: str, algo: str,
               max_dist: int) -> None:
    """gga_write_conv: the conv.vi block of printConv_rtime_init/_rtime/_final."""
    rows = np.ascontiguousarray(rows)
    check(load().gga_write_conv(os.fsencode(path), level, maxdeg, graph.encode(), algo.encode(),
                                rows.ctypes.data_as(ctypes.c_void_p), rows.size, max_dist), "write_conv")


__all__ = ["timeline_makespan", "write_conv", "CsrGraph", "DeviceGraph", "MultiSourceEngine", "csr_from_arrays", "read_gr", "write_gr", "gen_rmat",
           "gen_grid", "verify", "write_dist", "write_result_row", "device_count", "INF",
           "REF_INF"]
