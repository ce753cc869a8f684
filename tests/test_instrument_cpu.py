"""Instrumentation host logic on the CPU: the list-scheduling estimate of the
reference's combineClocks (metrics.h:70-92) and the conv.vi text block of
printConv_rtime_init / _rtime / _final (convergence.h:81-140)."""
import numpy as np

import gpugraphalgo_amd as gga

REC = [("start", np.uint64), ("ticks", np.uint32), ("place", np.uint32)]


def test_makespan_is_greedy_list_schedule():
    # 2 steps x 5 workgroups onto 2 slots, launch order, least-loaded first (first on ties)
    ticks = [5, 3, 4, 2, 6,    # step 1: slots -> [5], [3], [3+4=7]? no: 5->s0, 3->s1, 4->s1(3<5) =7, 2->s0=7, 6->s0(7==7, first)=13
             1, 1, 1, 1, 1]    # step 2: [3, 2] -> 3
    tl = np.zeros(10, dtype=REC)
    tl["ticks"] = ticks
    ms = gga.timeline_makespan(tl, grid=5, slots=2)
    assert ms == (13 + 3) / 1e5


def test_makespan_one_slot_is_the_sum_and_many_slots_the_max():
    rng = np.random.default_rng(0)
    tl = np.zeros(64, dtype=REC)
    tl["ticks"] = rng.integers(1, 1000, 64)
    assert gga.timeline_makespan(tl, grid=64, slots=1) == tl["ticks"].sum() / 1e5
    assert gga.timeline_makespan(tl, grid=64, slots=64) == tl["ticks"].max() / 1e5


def test_write_conv_matches_reference_text(tmp_path):
    rows = np.zeros(3, dtype=[("step", np.uint32), ("diff_count", np.uint32), ("norm2", np.float64)])
    rows["step"] = [1, 2, 3]
    rows["diff_count"] = [120, 7, 0]
    rows["norm2"] = [12345.678901, 2.5, 0.0]
    p = tmp_path / "conv.vi"
    gga.write_conv(str(p), rows, level=4, maxdeg=1181, graph="rmat20.gr", algo="SSSP", max_dist=97)
    bar = "-" * 88
    want = ("\n\n\n" + bar + "\nDist_Level_4_MaxDeg_1181_rmat20.gr_SSSP\n"
            "Iteration:1   = 12345.7(120)\nIteration:2   = 2.5(7)\nIteration:3   = 0(0)\n"
            "Total Iteration : 3     Max Dist : 97\n" + bar + "\n")
    assert p.read_text() == want
