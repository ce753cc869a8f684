"""GPU: the gga_run executable end to end, the drop-in for the reference's
{n,na,d,da}_{bfs,sssp}.exe (bfs_sssp.cu:103-539).

Each case writes a Galois .gr file, runs the executable as a child process,
and checks its distance file (Ary2File) against the oracle, its own
verification line, and the writeResult row's split fields against the
oracle's NodeSplit restatement.  This file sorts before the other GPU tests,
so the child processes start before this process touches the device.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import gpugraphalgo_amd as gga
from gpugraphalgo_amd import _lib
import oracle

pytestmark = pytest.mark.gpu

REF_INF = 1000000000


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def run_cli(exe, args, cwd):
    announce(f"cli {os.path.basename(exe)} {' '.join(args)}")
    r = subprocess.run([exe] + args, capture_output=True, text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout}\n{r.stderr}"
    assert "verifying: no of errors = 0." in r.stdout, r.stdout
    return r.stdout


def read_dist(path):
    with open(path) as f:
        return np.array([int(x) for x in f.read().split()], dtype=np.uint64)


def read_row(path):
    with open(path) as f:
        rows = [ln.rstrip("\n").split(",") for ln in f if ln.strip()]
    assert rows and all(len(r) == 17 for r in rows), rows
    keys = ["exe", "algo", "graph", "bucket", "nnode", "nedge", "ns_nnode", "ns_nedge",
            "extranode", "maxdeg", "level", "mda", "iterations", "runtime", "time", "ktime",
            "sktime"]
    return [dict(zip(keys, r)) for r in rows]


@pytest.fixture(scope="module")
def grfile(tmp_path_factory):
    g = gga.gen_rmat(12, 8, seed=5, wlo=1, whi=60)
    p = str(tmp_path_factory.mktemp("cli") / "rmat12.gr")
    gga.write_gr(p, g)
    return p, g


def test_cli_sssp_level_split_outfile(grfile, tmp_path):
    path, g = grfile
    deg = g.degrees()
    md = oracle.maxdeg(g.row_ptr)[0]
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, 0)
    for mode in ("auto", "topo", "worklist", "persistent"):
        out = tmp_path / f"d_{mode}.txt"
        res = tmp_path / "result.vi"
        stdout = run_cli(_lib.CLI_PATH, [path, "4", str(out), "--algo", "sssp", "--mode", mode,
                                         "--result", str(res)], str(tmp_path))
        assert stdout.startswith("SSSP\n") and "rmat12.gr" in stdout
        np.testing.assert_array_equal(read_dist(out), want.astype(np.uint64), err_msg=mode)
    rows = read_row(tmp_path / "result.vi")
    assert len(rows) == 4
    mda = oracle.mda_level(md, 4)
    for r in rows:
        assert r["algo"] == "SSSP" and r["graph"] == "rmat12.gr"
        assert (int(r["nnode"]), int(r["nedge"]), int(r["maxdeg"])) == (g.n, g.m, md)
        assert int(r["level"]) == 4 and int(r["bucket"]) == 0 and int(r["mda"]) == mda
        assert int(r["ns_nnode"]) == oracle.split_count(deg, mda)
        assert int(r["extranode"]) == int(r["ns_nnode"]) - g.n
        assert int(r["ns_nedge"]) == g.m


def test_cli_bfs_by_program_name_histogram_split(grfile, tmp_path):
    """L = -1 with a bucket count picks MDA by the degree histogram; the
    program name selects BFS as the reference's n_bfs.exe does."""
    path, g = grfile
    exe = str(tmp_path / "n_bfs.exe")
    os.symlink(_lib.CLI_PATH, exe)
    out = tmp_path / "d.txt"
    stdout = run_cli(exe, [path, "-1", "7", "--out", str(out), "--result", "r.vi"], str(tmp_path))
    assert stdout.startswith("BFS\n")
    np.testing.assert_array_equal(read_dist(out),
                                  oracle.bfs_queue(g.row_ptr, g.dst, 0).astype(np.uint64))
    (r,) = read_row(tmp_path / "r.vi")
    mda = oracle.mda_histogram(g.degrees(), 7)
    assert r["algo"] == "BFS" and int(r["bucket"]) == 7 and int(r["mda"]) == mda
    assert int(r["level"]) == 0xFFFFFFFF   # splitlevel -1 printed with %u
    assert int(r["ns_nnode"]) == oracle.split_count(g.degrees(), mda)


def test_cli_ref_inf_source_and_explicit_mda(grfile, tmp_path):
    """--ref-inf writes MYINFINITY for unreached vertices; --src moves the
    source (the reference hard-codes 0); --mda overrides L."""
    path, g = grfile
    src = int(np.flatnonzero(g.degrees() > 0)[3])
    out = tmp_path / "d.txt"
    run_cli(_lib.CLI_PATH, [path, "--algo", "sssp", "--src", str(src), "--ref-inf", "--mda", "16",
                            "--out", str(out), "--result", "r.vi"], str(tmp_path))
    want = oracle.dijkstra(g.row_ptr, g.dst, g.w, src, inf=REF_INF)
    got = read_dist(out)
    assert (got == REF_INF).any() and got[src] == 0
    np.testing.assert_array_equal(got, want.astype(np.uint64))
    (r,) = read_row(tmp_path / "r.vi")
    assert int(r["mda"]) == 16 and int(r["ns_nnode"]) == oracle.split_count(g.degrees(), 16)


def test_cli_no_split_row(grfile, tmp_path):
    """Two arguments: no split (EnableNodeSplit = 0), mda field 0."""
    path, g = grfile
    run_cli(_lib.CLI_PATH, [path, "--algo", "bfs", "--result", "r.vi"], str(tmp_path))
    (r,) = read_row(tmp_path / "r.vi")
    assert int(r["mda"]) == 0 and int(r["ns_nnode"]) == g.n and int(r["extranode"]) == 0


def test_mysc_sweep_reproduces_published_rows(tmp_path, split_kat):
    """The reference's sweep script (mysc:20-51): L = 1..40, then L = -1 with
    buckets {2, 5, 10, 15}, each a separate run of the executable.  Run
    through gga_run (linked as na_bfs.exe) on a .gr with webs' published
    degree sequence, the 44 result.vi rows must carry the published split
    fields of na_bfs.exe on webs (Kepler result.vi) -- MDA, node count after
    the split, maxdeg, n, m -- and the first run's --dd file the published
    DD_webs.gr.txt."""
    dd = np.asarray(split_kat["dd"]["webs"], dtype=np.int64)
    deg = np.repeat(dd[:, 0], dd[:, 1]).astype(np.uint32)
    rng = np.random.default_rng(11)
    rng.shuffle(deg)
    rp = np.concatenate([[0], np.cumsum(deg, dtype=np.int64)]).astype(np.uint32)
    dst = rng.integers(0, deg.size, int(rp[-1]), dtype=np.int64).astype(np.uint32)
    g = gga.csr_from_arrays(rp, dst, None)
    gga.write_gr(str(tmp_path / "webs.gr"), g)
    exe = str(tmp_path / "na_bfs.exe")
    os.symlink(_lib.CLI_PATH, exe)
    published = {(r["level"], r["bucket"]): r for r in split_kat["rows"]
                 if r["graph"] == "webs" and r["exe"] == "na_bfs.exe" and r["file"].startswith("Kepler")}
    runs = [[str(L)] for L in range(1, 41)] + [["-1", str(b)] for b in (2, 5, 10, 15)]
    assert len(runs) == len(published) == 44
    for k, args in enumerate(runs):
        extra = ["--dd", "--mode", "worklist"] if k == 0 else ["--mode", "worklist"]
        out = run_cli(exe, ["webs.gr"] + args + extra, str(tmp_path))
        assert "BFS" in out.splitlines()[0]
    with open(tmp_path / "DD_webs.gr.txt") as f:
        assert f.read() == "".join(f"{d},{c}\n" for d, c in dd.tolist())
    rows = read_row(str(tmp_path / "result.vi"))
    assert len(rows) == 44
    degs = g.degrees()
    for args, r in zip(runs, rows):
        L = int(args[0])
        b = int(args[1]) if L == -1 else 0
        want = published[(L, b)]
        assert r["exe"].endswith("na_bfs.exe") and r["algo"] == "BFS" and r["graph"] == "webs.gr"
        got = {k: int(r[k]) for k in ("bucket", "nnode", "nedge", "maxdeg", "mda", "ns_nnode")}
        got["level"] = int(np.uint32(int(r["level"])).astype(np.int32))
        exp = {k: want[k] for k in ("bucket", "nnode", "nedge", "maxdeg", "mda", "ns_nnode", "level")}
        assert got == exp, (args, got, exp)
        assert int(r["ns_nnode"]) == oracle.split_count(degs, int(r["mda"]))
        assert int(r["extranode"]) == int(r["ns_nnode"]) - int(r["nnode"])
