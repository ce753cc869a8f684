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
    s = read_summary(bench.TRAFFIC)
    assert s["kernel"] == "k_wl_step"
    fetch, write = float(s["fetch_bytes_per_dispatch"]), float(s["write_bytes_per_dispatch"])
    assert float(s["traffic_bytes_per_dispatch"]) == round(fetch + write) or \
        abs(float(s["traffic_bytes_per_dispatch"]) - (fetch + write)) <= 1
    assert 1.9 < float(s["fetch_factor"]) < 2.1 and 0.9 < float(s["write_factor"]) < 1.1
