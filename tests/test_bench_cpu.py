"""bench.py pieces that run without a GPU: the multi-GPU refusal, the
algorithmic-bytes model of the relax kernel (12 B per SSSP edge inspection,
8 B per BFS one, 16 B per chunk expanded and 8 B per relaxation in worklist
mode; 12 / 4 B in the topology-driven schedule), and the committed rocprofv3
traffic summary the roofline reads."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from gpugraphalgo_amd.profiling import read_summary  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_algorithmic_bytes_per_unit():
    st = {"edges_inspected": 1000, "chunks_expanded": 10, "relaxations": 7}
    assert bench.algorithmic_bytes(st, "worklist", False) == 12 * 1000 + 16 * 10 + 8 * 7
    assert bench.algorithmic_bytes(st, "worklist", True) == 8 * 1000 + 16 * 10 + 8 * 7
    assert bench.algorithmic_bytes(st, "topo", False) == 12 * 1000 + 12 * 10 + 4 * 7


def test_committed_traffic_summary_is_readable():
    """The committed L2-counter summary: its byte totals add up from request
    counts times sizes, and its copy check counted exactly the bytes copied."""
    s = read_summary(bench.TRAFFIC)
    assert "k_wl_step" in s["kernel"]
    f = {k: float(v) for k, v in s.items() if k not in ("command", "kernel")}
    reads = 32 * f["read_requests_32B"] + 64 * f["read_requests_64B"] + 128 * f["read_requests_128B"]
    # the counts are per-dispatch means printed rounded: compare to 1e-6
    assert abs(f["read_bytes_per_dispatch"] - reads) <= 1e-6 * reads
    writes = 64 * f["write_requests_64B"] + 32 * (f["write_requests"] - f["write_requests_64B"] - f["atomics_past_l2"])
    assert abs(f["write_bytes_per_dispatch"] - writes) <= 1e-6 * writes
    assert abs(f["atomic_bytes_per_dispatch"] - 32 * f["atomics_past_l2"]) <= 1e-6 * f["atomic_bytes_per_dispatch"]
    total = f["read_bytes_per_dispatch"] + f["write_bytes_per_dispatch"] + f["atomic_bytes_per_dispatch"]
    assert abs(f["traffic_bytes_per_dispatch"] - total) <= 1
    assert 0.0 < f["l2_hit_rate"] < 1.0
    assert abs(f["l2_hit_rate"] - f["l2_hits"] / (f["l2_hits"] + f["l2_misses"])) < 1e-3
    assert abs(f["copy_read_bytes_counted"] / f["copy_bytes"] - 1.0) < 1e-3
    assert abs(f["copy_write_bytes_counted"] / f["copy_bytes"] - 1.0) < 1e-3
