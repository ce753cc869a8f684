"""CPU-side tests of libgga.so: C-ABI exports, .gr I/O, generators, writers,
host verify and the error paths that need no device."""
import ctypes
import hashlib
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
import oracle
from conftest import ROOT, golden_graph

INF = gga.INF


def header_symbols():
    with open(os.path.join(ROOT, "include", "gga.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gga_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(syms) == bound, set(syms) ^ bound


def test_strerror_and_version():
    lib = _lib.load()
    assert lib.gga_strerror(0) == b"ok"
    assert lib.gga_strerror(-3) == b"malformed .gr file"
    assert b"gfx950" in lib.gga_version()


def random_csr(rng, n, m, wmax=100):
    s = rng.integers(0, n, m)
    t = rng.integers(0, n, m)
    order = np.lexsort((t, s))
    s, t = s[order], t[order]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, s + 1, 1)
    return gga.csr_from_arrays(np.cumsum(rp), t, rng.integers(1, wmax + 1, m))


@pytest.mark.parametrize("m", [0, 1, 7, 1000])
def test_gr_round_trip(tmp_path, m):
    rng = np.random.default_rng(m)
    g = random_csr(rng, 50, m)
    p = str(tmp_path / "g.gr")
    gga.write_gr(p, g)
    size = os.path.getsize(p)
    assert size == 32 + 8 * 50 + 4 * m + (4 if m % 2 else 0) + 4 * m
    h = gga.read_gr(p)
    np.testing.assert_array_equal(h.row_ptr, g.row_ptr)
    np.testing.assert_array_equal(h.dst, g.dst)
    np.testing.assert_array_equal(h.w, g.w)


def write_raw_gr(path, version, esz, n, row_end, dst, w=None, pad=True, truncate=0):
    m = len(dst)
    b = struct.pack("<4Q", version, esz, n, m)
    b += struct.pack(f"<{n}Q", *row_end)
    b += struct.pack(f"<{m}I", *dst)
    if m % 2 and pad:
        b += b"\0\0\0\0"
    if w is not None:
        b += struct.pack(f"<{m}I", *w)
    with open(path, "wb") as f:
        f.write(b[: len(b) - truncate])


def test_gr_edge_size_zero_means_unit_weights(tmp_path):
    p = str(tmp_path / "u.gr")
    write_raw_gr(p, 1, 0, 3, [2, 3, 3], [1, 2, 2])
    g = gga.read_gr(p)
    assert not g.has_weights
    np.testing.assert_array_equal(g.w, [1, 1, 1])
    np.testing.assert_array_equal(g.row_ptr, [0, 2, 3, 3])


@pytest.mark.parametrize("case,err", [
    ("version", "GGA_EFORMAT"), ("edgesize", "GGA_EFORMAT"), ("truncated", "GGA_EFORMAT"),
    ("dst", "GGA_EFORMAT"), ("rowend", "GGA_EFORMAT"), ("missing", "GGA_EIO"),
])
def test_gr_rejects_malformed(tmp_path, case, err):
    p = str(tmp_path / "bad.gr")
    args = dict(version=1, esz=4, n=3, row_end=[2, 3, 3], dst=[1, 2, 2], w=[5, 6, 7])
    if case == "version":
        args["version"] = 2
    elif case == "edgesize":
        args["esz"] = 8
    elif case == "dst":
        args["dst"] = [1, 3, 2]
    elif case == "rowend":
        args["row_end"] = [2, 1, 3]
    if case != "missing":
        write_raw_gr(p, args["version"], args["esz"], args["n"], args["row_end"], args["dst"],
                     args["w"], truncate=5 if case == "truncated" else 0)
    with pytest.raises(_lib.GgaError, match=err):
        gga.read_gr(p)


def test_rmat_generator_invariants():
    a = gga.gen_rmat(12, 8, seed=5, wlo=1, whi=100)
    b = gga.gen_rmat(12, 8, seed=5, wlo=1, whi=100)
    c = gga.gen_rmat(12, 8, seed=6, wlo=1, whi=100)
    for x in ("row_ptr", "dst", "w"):
        np.testing.assert_array_equal(getattr(a, x), getattr(b, x))
    assert not np.array_equal(a.dst, c.dst)
    src = np.repeat(np.arange(a.n), a.degrees())
    keys = src.astype(np.int64) * a.n + a.dst
    assert (np.diff(keys) > 0).all()          # sorted, no duplicates
    assert (src != a.dst).all()               # no self-loops
    assert a.w.min() >= 1 and a.w.max() <= 100
    assert a.maxdeg()[1] == 0                 # R-MAT hub at the lowest id
    raw = gga.gen_rmat(12, 8, seed=5, wlo=1, whi=100, dedup=False)
    assert raw.m == 8 << 12 and raw.m > a.m


def test_grid_generator_structure():
    R, C = 33, 17
    g = gga.gen_grid(R, C, 0.005, seed=11, wlo=1, whi=1000)
    n = R * C
    grid_edges = R * (C - 1) + (R - 1) * C
    k = round(0.005 * grid_edges)
    assert g.n == n and g.m == 2 * (grid_edges + k)
    src = np.repeat(np.arange(n), g.degrees())
    fwd = set(zip(src.tolist(), g.dst.tolist()))
    wmap = dict(zip(zip(src.tolist(), g.dst.tolist()), g.w.tolist()))
    assert all((v, u) in fwd for u, v in fwd)                     # symmetric
    assert all(wmap[(u, v)] == wmap[(v, u)] for u, v in fwd)      # one weight per edge
    nonlocal_ = [(u, v) for u, v in fwd if u < v and not (v - u == C or (v - u == 1 and v % C))]
    assert len(nonlocal_) == k
    for r in range(R):
        for c in range(C):
            u = r * C + c
            if c + 1 < C:
                assert (u, u + 1) in fwd
            if r + 1 < R:
                assert (u, u + C) in fwd


def test_host_verify_matches_oracle(golden):
    _, arrays, _ = golden
    rp, dst, w = golden_graph(arrays, "rand3")
    g = gga.csr_from_arrays(rp, dst, w)
    good = arrays["rand3.sssp.0.4294967295"]
    r = gga.verify(g, 0, good, bfs=False)
    assert r == {"nerr": 0, "untight": 0, "src_ok": True, "reached": int((good != INF).sum())}
    v = int(np.flatnonzero(good != INF)[5])
    bad = good.copy()
    bad[v] += 3
    assert gga.verify(g, 0, bad, bfs=False)["nerr"] == oracle.verify(rp, dst, w, bad) > 0
    bad = good.copy()
    bad[v] -= 1
    assert gga.verify(g, 0, bad, bfs=False)["untight"] == oracle.tight(rp, dst, w, 0, bad) > 0


def test_write_dist_format(tmp_path):
    p = str(tmp_path / "d.txt")
    d = np.array([0, 5, INF, gga.REF_INF], dtype=np.uint32)
    gga.write_dist(p, d)
    assert open(p).read() == "0\n5\n4294967295\n1000000000\n"


def test_write_result_row_format(tmp_path):
    p = str(tmp_path / "result.vi")
    gga.write_result_row(p, exename="./na_sssp.exe", algo="SSSP", graphfile="rmat20.gr",
                         bucket=10, nnode=1048576, nedge=8259994, ns_nnode=1053682,
                         ns_nedge=8259994, extranode=5106, maxdeg=1181, level=-1, mda=118,
                         iterations=17, runtime_ms=160, time_ms=163.592606, ktime_ms=162.917404,
                         sktime_ms=0.0)
    line = open(p).read()
    # field order and formats of the reference row (K/result.vi:699)
    assert line == ("./na_sssp.exe,SSSP,rmat20.gr,10,1048576,8259994,1053682,8259994,5106,1181,"
                    "4294967295,118,17,160,163.592606,162.917404,0.000000\n")


def test_graph_name():
    lib = _lib.load()
    buf = ctypes.create_string_buffer(64)
    lib.gga_graph_name(b"/home/vivek/rmat20.gr", buf, 64)
    assert buf.value == b"rmat20.gr"


def test_no_device_is_an_error_not_a_fallback():
    if gga.device_count() > 0:
        pytest.skip("a device is visible")
    g = gga.gen_rmat(8, 4, seed=1)
    with pytest.raises(_lib.GgaError, match="GGA_ENODEV"):
        gga.DeviceGraph(g)


def test_cli_usage_and_errors(tmp_path):
    exe = _lib.CLI_PATH
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stdout
    r = subprocess.run([exe, str(tmp_path / "nope.gr"), "1"], capture_output=True, text=True)
    assert r.returncode == 2
    # a file too short for the .gr header is rejected before any device work
    bad = tmp_path / "bad.gr"
    bad.write_bytes(b"not a graph")
    r = subprocess.run([exe, str(bad), "1"], capture_output=True, text=True)
    assert r.returncode == 2 and "malformed .gr" in r.stderr
    # --algo selects the algorithm (its name line comes first, as in bfs_sssp.cu)
    r = subprocess.run([exe, "--algo", "bfs"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stdout.startswith("BFS\n")


def test_unknown_option_names_raise():
    """A misspelt option must not be dropped silently (ctypes would keep it as
    a plain Python attribute)."""
    from gpugraphalgo_amd.graph import _opts
    o = _opts("worklist", INF, False, {"dense_min": 7, "delta": 3})
    assert (o.dense_min, o.delta, o.mode) == (7, 3, 1)
    with pytest.raises(TypeError, match="dens_min"):
        _opts("worklist", INF, False, {"dens_min": 7})


@pytest.mark.parametrize("graph", ["webs", "rmat20"])
def test_degree_distribution_matches_published_file(tmp_path, split_kat, graph):
    """A graph with the published degree sequence writes back the published
    DD_<graph>.gr.txt histogram (printfDegreeDistribution, graph.h:98-114):
    one "deg,count" line per occurring degree, ascending."""
    dd = np.asarray(split_kat["dd"][graph], dtype=np.int64)
    deg = np.repeat(dd[:, 0], dd[:, 1]).astype(np.uint32)
    np.random.default_rng(3).shuffle(deg)
    rp = np.concatenate([[0], np.cumsum(deg, dtype=np.int64)]).astype(np.uint32)
    g = gga.csr_from_arrays(rp, np.zeros(int(rp[-1]), dtype=np.uint32), None)
    p = str(tmp_path / f"DD_{graph}.gr.txt")
    gga.write_degree_dist(p, g)
    with open(p) as f:
        text = f.read()
    assert text == "".join(f"{d},{c}\n" for d, c in dd.tolist())


def test_degree_distribution_small_cases(tmp_path):
    p = str(tmp_path / "dd.txt")
    # degrees 2, 0, 2, 5: zero-degree vertices are listed too, gaps are skipped
    rp = np.array([0, 2, 2, 4, 9], dtype=np.uint32)
    gga.write_degree_dist(p, gga.csr_from_arrays(rp, np.zeros(9, dtype=np.uint32), None))
    assert open(p).read() == "0,1\n2,2\n5,1\n"
    rp = np.zeros(4, dtype=np.uint32)   # no edges at all
    gga.write_degree_dist(p, gga.csr_from_arrays(rp, np.zeros(0, dtype=np.uint32), None))
    assert open(p).read() == "0,3\n"
