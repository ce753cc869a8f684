"""gga_sssp_noalloc (static device scratch, no allocation) against gga_sssp.

Calls go through the C-ABI; the caller's output buffer is plain device
memory from hipMalloc.  Each call must equal gga_sssp on the same graph and
options bit for bit (and the oracle's Dijkstra), also when the previous call
left a larger graph's workspace in the scratch array.
"""
import ctypes
import sys

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
import oracle

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


class DeviceBuffer:
    """n uint32 of device memory owned by the test (the caller's output)."""

    def __init__(self, n: int):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), 4 * max(n, 1)) == 0

    def fill(self, value: int) -> None:
        host = np.full(self.n, value, dtype=np.uint32)
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1) == 0

    def read(self) -> np.ndarray:
        host = np.empty(self.n, dtype=np.uint32)
        assert self.hip.hipMemcpy(host.ctypes.data, self.ptr, host.nbytes, 2) == 0
        return host

    def free(self) -> None:
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


class OffsetBuffer(DeviceBuffer):
    """Device output that starts 4 bytes into its allocation: a caller's
    buffer need not be 16-byte aligned."""

    def __init__(self, n: int):
        super().__init__(n + 1)
        self.base = self.ptr
        self.ptr = ctypes.c_void_p(self.base.value + 4)
        self.n = n

    def free(self) -> None:
        if self.base:
            self.hip.hipFree(self.base)
            self.base = ctypes.c_void_p()
            self.ptr = ctypes.c_void_p()


def noalloc(dg, src, buf, **kw):
    o = gga.graph._opts(kw.pop("mode", "auto"), kw.pop("inf", gga.INF), kw.pop("count", False), kw)
    st = _lib.Stats()
    rc = _lib.load().gga_sssp_noalloc(dg._h, src, buf.ptr, ctypes.byref(o), ctypes.byref(st))
    _lib.check(rc, "gga_sssp_noalloc")
    return buf.read(), gga.graph._stats_dict(st)


CASES = [
    # (graph, split threshold, runs): the second graph is the smaller one
    ("rmat13", lambda: gga.gen_rmat(13, 16, seed=3, wlo=1, whi=100), 32,
     [dict(mode="worklist"), dict(mode="persistent"), dict(mode="topo"),
      dict(mode="worklist", dense_min=1), dict(mode="worklist", delta=40), dict(mode="auto", count=True)]),
    ("grid30x40", lambda: gga.gen_grid(30, 40, shortcut_frac=0.01, seed=8, wlo=1, whi=1000), 4,
     [dict(mode="worklist"), dict(mode="persistent"), dict(mode="worklist", delta=0xFFFFFFFF)]),
]


def test_noalloc_equals_gga_sssp_two_calls_in_a_row(gpu):
    graphs = []
    try:
        for name, make, T, runs in CASES:
            g = make()
            dg = gga.DeviceGraph(g)
            dg.split(mda=T)
            graphs.append((name, g, dg, runs))
        for rnd in range(2):
            for name, g, dg, runs in graphs:
                buf = DeviceBuffer(g.n)
                try:
                    for src in (g.maxdeg()[1], g.n // 3):
                        want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
                        for kw in runs:
                            announce(f"round {rnd} {name} n={g.n} src={src} {kw}")
                            ref, rst = dg.sssp(src, **dict(kw))
                            buf.fill(0xA5A5A5A5)
                            got, st = noalloc(dg, src, buf, **dict(kw))
                            np.testing.assert_array_equal(got, ref, err_msg=f"{name} {kw} vs gga_sssp")
                            np.testing.assert_array_equal(got, want, err_msg=f"{name} {kw} vs oracle")
                            assert st["reached"] == rst["reached"] and st["m_reach"] == rst["m_reach"]
                            if kw.get("count"):
                                assert st["edges_inspected"] > 0
                finally:
                    buf.free()
    finally:
        for _, _, dg, _ in graphs:
            dg.close()


def test_noalloc_misaligned_output(gpu):
    """The output buffer 4 bytes past a 16-byte boundary: the near-far
    refill, which reads dist with 16-byte loads when it can, must take the
    word-by-word path and still equal gga_sssp and the oracle."""
    for name, g, T in (("grid30x41", gga.gen_grid(30, 41, shortcut_frac=0.01, seed=9, wlo=1, whi=1000), 4),
                       ("rmat12", gga.gen_rmat(12, 8, seed=5, wlo=1, whi=100), 16)):
        with gga.DeviceGraph(g) as dg:
            dg.split(mda=T)
            buf = OffsetBuffer(g.n)
            try:
                for src in (0, g.n // 2):
                    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src)
                    for kw in (dict(mode="worklist", delta=40, dense_min=1), dict(mode="worklist", delta=1),
                               dict(mode="worklist")):
                        announce(f"misaligned output {name} src={src} {kw}")
                        ref, _ = dg.sssp(src, **dict(kw))
                        buf.fill(0x5A5A5A5A)
                        got, _ = noalloc(dg, src, buf, **dict(kw))
                        np.testing.assert_array_equal(got, ref, err_msg=f"{name} {kw} vs gga_sssp")
                        np.testing.assert_array_equal(got, want, err_msg=f"{name} {kw} vs oracle")
            finally:
                buf.free()


def test_noalloc_errors(gpu):
    g = gga.gen_rmat(10, 8, seed=1, wlo=1, whi=10)
    with gga.DeviceGraph(g) as dg:
        buf = DeviceBuffer(g.n)
        try:
            with pytest.raises(_lib.GgaError, match="GGA_ENOSPLIT"):
                noalloc(dg, 0, buf)
            dg.split(mda=16)
            with pytest.raises(_lib.GgaError, match="GGA_EINVAL"):
                noalloc(dg, g.n, buf)
            st = _lib.Stats()
            o = gga.graph._opts("auto", gga.INF, False, {})
            assert _lib.load().gga_sssp_noalloc(dg._h, 0, None, ctypes.byref(o), ctypes.byref(st)) == -1
            got, _ = noalloc(dg, 0, buf)
            np.testing.assert_array_equal(got, oracle.dijkstra(g.row_ptr, g.dst, g.w, 0))
        finally:
            buf.free()
    # 12 M vertices, no edges: 21 bytes per vertex of workspace > 128 MiB
    n = 12 << 20
    big = gga.csr_from_arrays(np.zeros(n + 1, dtype=np.uint32), np.zeros(0, dtype=np.uint32),
                              np.zeros(0, dtype=np.uint32))
    with gga.DeviceGraph(big) as dg:
        dg.split(mda=1)
        buf = DeviceBuffer(n)
        try:
            announce("oversized workspace")
            with pytest.raises(_lib.GgaError, match="GGA_ERANGE"):
                noalloc(dg, 0, buf)
        finally:
            buf.free()
