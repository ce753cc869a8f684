"""CPU oracle for the BFS/SSSP node-splitting path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or the timed CPU
baseline.  The product (``gpugraphalgo_amd``) never imports it.

Thin ctypes wrapper over ``liboracle.so`` (built from ``gga_oracle.c`` by
``oracle/Makefile``).  Every function restates a reference function; the
citations and the way each one is pinned are in ``gga_oracle.h``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compiles liboracle.so in place (gcc)."""
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
        vp = ctypes.c_void_p
        L.orc_serial_sp.restype = u32
        L.orc_serial_sp.argtypes = [u32, _u32p, _u32p, vp, u32, u32, _u32p, ctypes.POINTER(u32)]
        L.orc_verify.restype = u64
        L.orc_verify.argtypes = [u32, _u32p, _u32p, vp, _u32p, u32]
        L.orc_tight.restype = u64
        L.orc_tight.argtypes = [u32, _u32p, _u32p, vp, u32, _u32p, u32]
        L.orc_maxdeg.restype = u32
        L.orc_maxdeg.argtypes = [u32, _u32p, ctypes.POINTER(u32)]
        L.orc_mda_level.restype = u32
        L.orc_mda_level.argtypes = [u32, i32]
        L.orc_mda_histogram.restype = u32
        L.orc_mda_histogram.argtypes = [u32, _u32p, i32]
        L.orc_mda_histogram_dd.restype = u32
        L.orc_mda_histogram_dd.argtypes = [u32, _u32p, _u32p, i32]
        L.orc_split_count.restype = u64
        L.orc_split_count.argtypes = [u32, _u32p, u32]
        L.orc_node_split.restype = u32
        L.orc_node_split.argtypes = [u32, _u32p, u32, vp, vp, vp, vp]
        L.orc_split_relax.restype = u32
        L.orc_split_relax.argtypes = [u32, _u32p, _u32p, vp, u32, u32, u32, _u32p]
        L.orc_dijkstra.restype = None
        L.orc_dijkstra.argtypes = [u32, _u32p, _u32p, vp, u32, u32, _u32p]
        L.orc_bfs_queue.restype = None
        L.orc_bfs_queue.argtypes = [u32, _u32p, _u32p, u32, u32, _u32p]
        L.orc_m_reach.restype = u64
        L.orc_m_reach.argtypes = [u32, _u32p, _u32p, u32]
        _LIB = L
    return _LIB


INF = 0xFFFFFFFF
REF_INF = 1000000000


def _c(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _wptr(w):
    return None if w is None else w.ctypes.data_as(ctypes.c_void_p)


def serial_sp(row_ptr, dst, w, src, inf=INF):
    """findSerialSSSP (w given) / findSerialBFS (w None).  Returns (dist, sweeps, max_dist)."""
    row_ptr, dst = _c(row_ptr), _c(dst)
    w = None if w is None else _c(w)
    n = row_ptr.size - 1
    dist = np.empty(n, dtype=np.uint32)
    mx = ctypes.c_uint32(0)
    sweeps = lib().orc_serial_sp(n, row_ptr, dst, _wptr(w), src, inf, dist, ctypes.byref(mx))
    return dist, int(sweeps), int(mx.value)


def verify(row_ptr, dst, w, dist, inf=INF) -> int:
    row_ptr, dst, dist = _c(row_ptr), _c(dst), _c(dist)
    w = None if w is None else _c(w)
    return int(lib().orc_verify(row_ptr.size - 1, row_ptr, dst, _wptr(w), dist, inf))


def tight(row_ptr, dst, w, src, dist, inf=INF) -> int:
    row_ptr, dst, dist = _c(row_ptr), _c(dst), _c(dist)
    w = None if w is None else _c(w)
    return int(lib().orc_tight(row_ptr.size - 1, row_ptr, dst, _wptr(w), src, dist, inf))


def maxdeg(row_ptr):
    row_ptr = _c(row_ptr)
    idx = ctypes.c_uint32(0)
    d = lib().orc_maxdeg(row_ptr.size - 1, row_ptr, ctypes.byref(idx))
    return int(d), int(idx.value)


def mda_level(maxdeg_, level) -> int:
    return int(lib().orc_mda_level(maxdeg_, level))


def mda_histogram(deg, buckets) -> int:
    deg = _c(deg)
    return int(lib().orc_mda_histogram(deg.size, deg, buckets))


def mda_histogram_dd(deg, cnt, buckets) -> int:
    deg, cnt = _c(deg), _c(cnt)
    return int(lib().orc_mda_histogram_dd(deg.size, deg, cnt, buckets))


def split_count(deg, mda) -> int:
    deg = _c(deg)
    return int(lib().orc_split_count(deg.size, deg, mda))


def node_split(row_ptr, mda):
    """Returns (chunk_begin[C+1], owner[C], child_count[C], presum[n+1])."""
    row_ptr = _c(row_ptr)
    n = row_ptr.size - 1
    deg = np.diff(row_ptr.astype(np.int64)).astype(np.uint32)
    C = split_count(deg, mda)
    cb = np.empty(C + 1, dtype=np.uint32)
    own = np.empty(C, dtype=np.uint32)
    cc = np.empty(C, dtype=np.uint32)
    ps = np.empty(n + 1, dtype=np.uint32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    got = lib().orc_node_split(n, row_ptr, mda, p(cb), p(own), p(cc), p(ps))
    assert got == C
    return cb, own, cc, ps


def split_relax(row_ptr, dst, w, mda, src, inf=INF):
    row_ptr, dst = _c(row_ptr), _c(dst)
    w = None if w is None else _c(w)
    n = row_ptr.size - 1
    dist = np.empty(n, dtype=np.uint32)
    sweeps = lib().orc_split_relax(n, row_ptr, dst, _wptr(w), mda, src, inf, dist)
    return dist, int(sweeps)


def dijkstra(row_ptr, dst, w, src, inf=INF):
    row_ptr, dst = _c(row_ptr), _c(dst)
    w = None if w is None else _c(w)
    n = row_ptr.size - 1
    dist = np.empty(n, dtype=np.uint32)
    lib().orc_dijkstra(n, row_ptr, dst, _wptr(w), src, inf, dist)
    return dist


def bfs_queue(row_ptr, dst, src, inf=INF):
    row_ptr, dst = _c(row_ptr), _c(dst)
    n = row_ptr.size - 1
    dist = np.empty(n, dtype=np.uint32)
    lib().orc_bfs_queue(n, row_ptr, dst, src, inf, dist)
    return dist


def m_reach(row_ptr, dist, inf=INF) -> int:
    row_ptr, dist = _c(row_ptr), _c(dist)
    return int(lib().orc_m_reach(row_ptr.size - 1, row_ptr, dist, inf))
