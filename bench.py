#!/usr/bin/env python3
"""Headline benchmark: node-split SSSP on R-MAT scale 22, edge factor 16, one MI355X.

Workload (the single-GPU configuration the metric is quoted on): R-MAT
2^22 vertices, 16 * 2^22 generated arcs (a, b, c, d = .57, .19, .19, .05,
seed 7, duplicates and self-loops removed), uint32 weights uniform in
[1, 100], source = the highest-degree vertex (lowest id on ties).  Split
threshold T: every point of {32, 64, 128, 256, 1024} is timed and the best
is the headline (--no-sweep: only --mda, default 128).  Beside it: a cold
first query on a freshly uploaded graph, and 20 queries from distinct sources.

A step is one full SSSP query (distance init -> convergence) with the graph
already resident in HBM and the result left in HBM.  value = GTEPS =
steps * m_reach / wall time of the K timed steps.  roofline.achieved =
algorithmic bytes of the relax kernel (per launch) / its average launch
duration, both from instrumented queries run right after the timed ones:
bytes from the device work counters (12 B per edge inspected, 16 B per chunk
expanded, 8 B per successful relaxation), durations from HIP events the
library records around each relax launch on the stream the kernel runs on
(left out of the timed steps: each event adds a dispatch gap).  cpu_baseline = the oracle's restatement of the
reference's serial SSSP (graph.h:115-148), one thread, full run on the same
graph and source.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode auto|worklist|persistent|topo]
                  [--mda T --no-sweep] [--no-cpu] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "GTEPS + % HBM roofline (1 GPU); served BFS/SSSP queries/s with p50/p99 ms"
HBM_PEAK_GBS = 8000.0
SWEEP = (32, 64, 128, 256, 1024)
DEFAULT_T = 128  # best of SWEEP on MI355X (profiles/sweep_rmat22_sssp.txt)
TRAFFIC = os.path.join(ROOT, "profiles", "k_wl_step_rmat22_sssp_counters.txt")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(st: dict, mode: str, bfs: bool) -> int:
    """Bytes the relax kernel must move for the work it did (see module doc)."""
    per_edge = 8 if bfs else 12
    per_chunk = 12 if mode == "topo" else 16
    per_relax = 4 if mode == "topo" else 8
    return (per_edge * st["edges_inspected"] + per_chunk * st["chunks_expanded"]
            + per_relax * st["relaxations"])


def run_config(dg, src, T, mode, steps, warmup, bfs=False):
    """W warmup + K timed queries, uninstrumented; then instrumented ones
    (work counters + a HIP event pair around every relax launch, on the
    graph's stream) for the byte model and the relax kernel's durations --
    the events add dispatch gaps, so they stay out of the timed K."""
    dg.split(mda=T)
    run = dg.bfs if bfs else dg.sssp
    for _ in range(warmup):
        run(src, mode=mode, keep_on_device=True)
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = run(src, mode=mode, keep_on_device=True)
        stats.append(st)
    wall = time.perf_counter() - t0
    counted = [run(src, mode=mode, keep_on_device=True, count=True)[1] for _ in range(max(3, steps // 4))]
    return wall, stats, counted


def summarize(wall, stats, counted, mode, bfs):
    k, kc = len(stats), len(counted)
    m_reach = stats[0]["m_reach"]
    ktime_ms = sum(s["ktime_ms"] for s in counted) / kc
    launches = sum(s["launches"] for s in counted) / kc
    bytes_q = sum(algorithmic_bytes(s, mode, bfs) for s in counted) / kc
    return {
        "ms_per_step": wall / k * 1e3,
        "gteps": k * m_reach / wall / 1e9,
        "m_reach": m_reach,
        "device_ms": sum(s["time_ms"] for s in stats) / k,
        "counted_queries": kc,
        "counted_device_ms": sum(s["time_ms"] for s in counted) / kc,
        "ktime_ms": ktime_ms,
        "launches": launches,
        "bytes_per_query": bytes_q,
        "achieved_gbs": bytes_q / (ktime_ms * 1e-3) / 1e9 if ktime_ms > 0 else 0.0,
        "iterations": stats[0]["iterations"],
        "edges_inspected": sum(s["edges_inspected"] for s in counted) / kc,
        "chunks_expanded": sum(s["chunks_expanded"] for s in counted) / kc,
        "relaxations": sum(s["relaxations"] for s in counted) / kc,
        "mode": stats[0]["mode"],
    }


def cold_first_query(gga, g, src, T, mode):
    """One query on a freshly uploaded and split graph: no step-count history,
    so the host checks fall back to their default batch (8 steps per check)."""
    with gga.DeviceGraph(g) as dg:
        dg.split(mda=T)
        t0 = time.perf_counter()
        _, st = dg.sssp(src, mode=mode, keep_on_device=True)
        wall = time.perf_counter() - t0
    return {"ms": round(wall * 1e3, 3), "device_ms": round(st["time_ms"], 3),
            "gteps": round(st["m_reach"] / wall / 1e9, 3), "steps": st["iterations"]}


def unsplit_query(dg, src, mode, split_ms, reps=10):
    """L = 0, the reference's no-split baseline (one chunk per vertex, the
    97 670-edge hub included): the kernels walk its hubs in pieces of at most
    GGA_PIECE_MAX edges, spread over many waves."""
    st = dg.split()
    for _ in range(2):
        dg.sssp(src, mode=mode, keep_on_device=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        dg.sssp(src, mode=mode, keep_on_device=True)
    ms = (time.perf_counter() - t0) / reps * 1e3
    return {"ms": round(ms, 3), "vs_split": round(ms / split_ms, 3), "chunks": st["ns_nnode"],
            "walked_chunks": st["exec_chunks"], "max_chunk": st["max_chunk"]}


def distinct_sources(dg, g, mode, k=20):
    """k queries from k different sources (uniform among vertices with out-edges,
    seed 20), one after another on the warm graph: each query's host-check batch
    comes from the previous query's step count, not its own."""
    import numpy as np
    rng = np.random.default_rng(20)
    srcs = rng.choice(np.flatnonzero(g.degrees() > 0), k, replace=False)
    walls, reach = [], 0
    for s in srcs:
        t0 = time.perf_counter()
        _, st = dg.sssp(int(s), mode=mode, keep_on_device=True)
        walls.append(time.perf_counter() - t0)
        reach += st["m_reach"]
    return {"queries": k, "mean_ms": round(1e3 * sum(walls) / k, 3), "max_ms": round(1e3 * max(walls), 3),
            "gteps": round(reach / sum(walls) / 1e9, 3)}


RATES = os.path.join(ROOT, "gpugraphalgo_amd", "bin", "gga_rates")


def random_access_ceiling(scale: int, pass_pct: int) -> dict | None:
    """Runs the memory-rate probe (gpugraphalgo_amd/csrc/probes/rates.hip) with
    the relax kernel's access pattern at this workload's scale and guard-pass
    share: streamed col/w, a gather of dist[v] per edge from an R-MAT column
    distribution, a device-scope atomicMin where the guard passes."""
    import subprocess
    if not os.path.exists(RATES):
        return None
    cmd = [RATES, "--scale", str(scale), "--pass", str(pass_pct), "--medges", "32"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        log(f"gga_rates failed: {r.stderr.strip()}")
        return None
    out = json.loads(r.stdout)
    out["command"] = " ".join(["gpugraphalgo_amd/bin/gga_rates"] + cmd[1:])
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(g, src, m_reach):
    import oracle
    import numpy as np
    t0 = time.perf_counter()
    d, sweeps, _ = oracle.serial_sp(g.row_ptr, g.dst, g.w, src)
    t = time.perf_counter() - t0
    mr = oracle.m_reach(g.row_ptr, d)
    assert mr == m_reach, (mr, m_reach)
    return {"value": mr / t / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": (f"one full run of the restated serial Gauss-Seidel SSSP (graph.h:115-148) "
                       f"on the same graph and source: {t:.2f} s, {sweeps} sweeps, 1 of "
                       f"{os.cpu_count()} host cores ({cpu_model()})")}, d


def batched_leg(gga, np, quick=False, coop=True):
    """Config 4: 256 BFS (bitmask lanes, one cooperative launch) + 32 SSSP
    (32-lane engine, hipGraph-replayed steps) on R-MAT s20 ef16 seed 13,
    sources uniform among vertices with out-edges; a sample of lanes is
    checked bit-exact against single-source runs of the same sources.
    coop=False leaves out the cooperative 256-lane BFS."""
    g = gga.gen_rmat(20, 16, 0.57, 0.19, 0.19, seed=13, wlo=1, whi=100)
    rng = np.random.default_rng(13)
    cand = np.flatnonzero(g.degrees() > 0)
    bsrc = rng.choice(cand, 256, replace=False).astype(np.uint32)
    ssrc = rng.choice(cand, 32, replace=False).astype(np.uint32)
    out = {"workload": "R-MAT s20 ef16 seed 13, 256 BFS + 32 SSSP sources", "n": g.n, "m": g.m}
    with gga.DeviceGraph(g) as dg:
        tr = dg.tune(int(ssrc[0]), bfs=False, layout=not quick, multi=True)
        out["tuned"] = {"split_threshold": tr["mda"], "seed_threshold": tr["seed_mda"], "grid": tr["grid"],
                        "tried_ms": {str(t): round(ms, 3) for t, ms in tr["tried"].items()},
                        "dense_min": tr["dense_min"], "layout": {1: "vertex", 2: "lane"}[tr["layout"]]}
        reps = 3 if quick else 5
        wb, sb = None, None
        for _ in range(2 if coop else 0):
            dg.multi_bfs(bsrc, keep_on_device=True)
        if coop:
            t0 = time.perf_counter()
            for _ in range(reps):
                _, sb = dg.multi_bfs(bsrc, keep_on_device=True)
            wb = (time.perf_counter() - t0) / reps
        for _ in range(2):
            dg.multi_sp(ssrc, use_graph=1, keep_on_device=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            _, ss = dg.multi_sp(ssrc, use_graph=1, keep_on_device=True)
        ws = (time.perf_counter() - t0) / reps
        # single-source baselines on the same sources (worklist mode, result kept on the device)
        t0 = time.perf_counter()
        mr1 = sum(dg.bfs(int(s), keep_on_device=True)[1]["m_reach"] for s in bsrc[:32])
        w1b = (time.perf_counter() - t0) / 32
        t0 = time.perf_counter()
        mr2 = sum(dg.sssp(int(s), keep_on_device=True)[1]["m_reach"] for s in ssrc)
        w1s = (time.perf_counter() - t0) / 32
        ds, _ = dg.multi_sp(ssrc, use_graph=1)
        if coop:
            lv, _ = dg.multi_bfs(bsrc)
            for i in range(0, 256, 37):
                assert np.array_equal(lv[i], dg.bfs(int(bsrc[i]))[0]), f"BFS lane {i}"
        for i in range(0, 32, 5):
            assert np.array_equal(ds[i], dg.sssp(int(ssrc[i]))[0]), f"SSSP lane {i}"
    if coop:
        out["bfs256"] = {"ms": round(wb * 1e3, 3), "gteps": round(sb["m_reach"] / wb / 1e9, 2),
                         "levels": sb["iterations"], "pull_levels": sb["pull_levels"],
                         "single_source_ms_per_query": round(w1b * 1e3, 3),
                         "speedup_vs_single": round(w1b * 256 / wb, 2)}
    out["sssp32"] = {"ms": round(ws * 1e3, 3), "gteps": round(ss["m_reach"] / ws / 1e9, 2),
                     "steps": ss["iterations"], "single_source_ms_per_query": round(w1s * 1e3, 3),
                     "speedup_vs_single": round(w1s * 32 / ws, 2)}
    out["checked_lanes_bit_exact"] = (len(range(0, 256, 37)) if coop else 0) + len(range(0, 32, 5))
    del mr1, mr2
    return out


def grid_leg(gga, np, quick=False, coop=True):
    """Config 3: 4096 x 4096 grid, 4-neighbour undirected, + 0.5% random
    shortcut edges (seed 11), weights [1, 1000], source 0 (corner); BFS and
    SSSP in each launch mode.  Checked on the device (fixpoint + tightness +
    dist[src] = 0, which fixes the shortest-path distances).  coop=False
    leaves out the persistent (cooperative) launch mode."""
    side = 1024 if quick else 4096
    g = gga.gen_grid(side, side, 0.005, seed=11, wlo=1, whi=1000)
    out = {"workload": f"{side}x{side} grid + 0.5% shortcuts, weights [1,1000], source 0",
           "n": g.n, "m": g.m}
    with gga.DeviceGraph(g) as dg:
        dg.split()
        for algo in ("bfs", "sssp"):
            run = dg.bfs if algo == "bfs" else dg.sssp
            res = {}
            for mode in ("auto", "worklist", "persistent") if coop else ("worklist",):
                run(0, mode=mode, keep_on_device=True)
                reps = 3
                t0 = time.perf_counter()
                for _ in range(reps):
                    _, st = run(0, mode=mode, keep_on_device=True)
                wall = (time.perf_counter() - t0) / reps
                ver = dg.verify(0, bfs=(algo == "bfs"))
                assert ver["nerr"] == 0 and ver["untight"] == 0 and ver["src_ok"], (algo, mode, ver)
                res[mode] = {"ms": round(wall * 1e3, 3), "gteps": round(st["m_reach"] / wall / 1e9, 3),
                             "steps": st["iterations"],
                             "mode_used": {0: "topo", 1: "worklist", 2: "persistent"}.get(st["mode"], st["mode"])}
            out[algo] = res
    return out


def engine_leg(gga, np, coop=True, side=4096):
    """One 32-lane engine (16 BFS + 16 SSSP lanes from random sources, seed 5)
    on the side x side grid of the serving set (weights [1, 1000], seed 17),
    run until every lane converges, checked every 12 steps: the time that
    sets the serving latency of long-diameter graphs.  Beside it, one of its
    BFS and one of its SSSP queries alone (worklist mode)."""
    g = gga.gen_grid(side, side, 0.0, seed=17, wlo=1, whi=1000)
    rng = np.random.default_rng(5)
    srcs = rng.choice(g.n, 32, replace=False)
    with gga.DeviceGraph(g) as dg:
        dg.split()
        solo = {}
        for name, fn, s in (("bfs", dg.bfs, int(srcs[0])), ("sssp", dg.sssp, int(srcs[16]))):
            fn(s, mode="worklist")
            t0 = time.perf_counter()
            _, st = fn(s, mode="worklist")
            solo[name] = {"ms": round((time.perf_counter() - t0) * 1e3, 2), "steps": int(st["iterations"])}
        eng = gga.MultiSourceEngine(dg)
        try:
            eng.start({i: int(s) for i, s in enumerate(srcs)}, bfs=set(range(16)))
            fin, steps = {}, 0
            t0 = time.perf_counter()
            while steps < 100000:
                live = eng.run(max_steps=12, mode="worklist", sync_every=12)
                steps += 12
                for i in range(32):
                    if i not in fin and not (live >> i) & 1:
                        fin[i] = steps
                if live == 0:
                    break
            wall = time.perf_counter() - t0
        finally:
            eng.close()
    b = sorted(fin.get(i, steps) for i in range(16))
    q = sorted(fin.get(i, steps) for i in range(16, 32))
    return {"workload": f"32-lane engine on the {side}x{side} grid (16 BFS + 16 SSSP lanes) to convergence",
            "ms": round(wall * 1e3, 3), "steps": steps, "us_per_step": round(wall / steps * 1e6, 2),
            "bfs_lane_steps_median": b[8], "sssp_lane_steps_median": q[8],
            "queries_per_s": round(32 / wall, 1), "alone_worklist": solo}


def serving_leg(gga, np, full=True, quick=False):
    """Config 5 serving: resident R-MAT graphs (s16.. ef16, weights [1,100])
    and 4-neighbour grids (weights [1,1000]), seed 17; BFS/SSSP 50/50, uniform
    graph and source (among vertices with out-edges).  The server (32-lane
    engines, HBM-budget admission, continuous batching) against the same
    queries answered one persistent launch per query."""
    from gpugraphalgo_amd.serve import Server, latency_summary, poisson_trace
    scales = range(16, 24) if full else range(16, 21)
    sides = (512, 1024, 2048, 4096) if full else (512, 1024)
    gs = [("rmat%d" % s, gga.gen_rmat(s, 16, 0.57, 0.19, 0.19, seed=17, wlo=1, whi=100)) for s in scales]
    gs += [("grid%d" % a, gga.gen_grid(a, a, 0.0, seed=17, wlo=1, whi=1000)) for a in sides]
    dgs = []
    for name, g in gs:
        dg = gga.DeviceGraph(g)
        if name.startswith("rmat"):
            dg.split(mda=64)
        else:
            dg.split()
        dgs.append(dg)
    sizes = [g.n for _, g in gs]
    degs = [np.flatnonzero(g.degrees() > 0).astype(np.uint32) for _, g in gs]
    nq = 120 if quick else 400
    nplain = 40 if quick else 80
    out = {"workload": ("resident set: R-MAT s16-s23 + grids 512^2, 1024^2, 2048^2, 4096^2" if full else
                        "reduced resident set: R-MAT s16-s20 + grids 512^2, 1024^2") +
                       ", BFS/SSSP 50/50, uniform graph and source",
           "graphs": [n for n, _ in gs], "resident_gb": round(sum(d.device_bytes for d in dgs) / 1e9, 3)}
    with Server(hbm_budget=16 << 30, steps_per_round=12, use_graph=True) as srv:
        for dg in dgs:
            srv.add_graph(dg)
        sat = poisson_trace(sizes, nq, rate=0, seed=21, degrees=degs)
        srv.replay(sat[: min(48, nq)])                    # engines created, graphs captured
        r = srv.replay(sat)
        s = latency_summary(r)
        p = srv.replay_plain(sat[:nplain])
        sp = latency_summary(p)
        assert np.array_equal(p["checksum"], r["checksum"][:nplain]), "served != one-query-per-launch"
        out["saturating"] = {"queries": nq, "qps": round(s["qps"], 1), "p50_ms": round(s["p50_ms"], 2),
                             "p99_ms": round(s["p99_ms"], 2)}
        out["plain_one_launch_per_query"] = {"queries": nplain, "qps": round(sp["qps"], 1),
                                             "p50_ms": round(sp["p50_ms"], 2), "p99_ms": round(sp["p99_ms"], 2)}
        out["speedup_qps"] = round(s["qps"] / sp["qps"], 2) if sp["qps"] else None
        out["results_bit_exact_vs_plain"] = nplain
        # the same queries one at a time in AUTO mode (worklist launches on the
        # R-MAT graphs, the persistent kernel on the grids once a query there
        # ran long), after one BFS and one SSSP query per graph have set AUTO's
        # evidence
        warm = np.zeros(2 * len(gs), dtype=sat.dtype)
        for gi in range(len(gs)):
            warm[2 * gi] = (gi, int(degs[gi][0]), 0, 0.0)
            warm[2 * gi + 1] = (gi, int(degs[gi][0]), 1, 0.0)
        srv.replay_plain(warm, mode="auto")
        pa = latency_summary(srv.replay_plain(sat[:nplain], mode="auto"))
        out["plain_auto_mode_per_query"] = {"queries": nplain, "qps": round(pa["qps"], 1),
                                            "p50_ms": round(pa["p50_ms"], 2), "p99_ms": round(pa["p99_ms"], 2)}
        out["speedup_qps_vs_auto"] = round(s["qps"] / pa["qps"], 2) if pa["qps"] else None
        # Poisson arrivals over the configured range (100 to 20 000 q/s offered)
        # and at half the saturating rate
        rates = [100.0, 1000.0, 5000.0, 20000.0, round(0.5 * s["qps"], 1)] if not quick else [100.0, 1000.0]
        out["poisson"] = []
        for rate in sorted(rates):
            tr = poisson_trace(sizes, nq, rate=max(1.0, rate), seed=31, degrees=degs)
            rr = latency_summary(srv.replay(tr))
            out["poisson"].append({"offered_qps": round(rate, 1), "achieved_qps": round(rr["qps"], 1),
                                   "p50_ms": round(rr["p50_ms"], 2), "p99_ms": round(rr["p99_ms"], 2),
                                   "queries": nq})
        out["server"] = {k: v for k, v in srv.stats().items() if k in ("rounds", "engines_created", "peak_bytes")}
    for dg in dgs:
        dg.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="auto")
    ap.add_argument("--mda", type=int, default=DEFAULT_T, help="split threshold T")
    ap.add_argument("--no-sweep", action="store_true",
                    help="time only --mda instead of every T in {32, 64, 128, 256, 1024} (the default "
                         "sweep reports the best as the headline)")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="profiling run: one T, no CPU leg")
    ap.add_argument("--no-grid", action="store_true", help="skip the grid leg (4096^2 grid + shortcuts)")
    ap.add_argument("--no-batched", action="store_true", help="skip the multi-source leg (R-MAT s20, 256 BFS + 32 SSSP)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg (12 resident graphs)")
    ap.add_argument("--serve-reduced", action="store_true",
                    help="serving leg on a reduced resident set (R-MAT s16-s20, grids 512^2 and 1024^2) "
                         "instead of the full one (R-MAT s16-s23, grids 512^2 to 4096^2)")
    ap.add_argument("--leg", choices=("grid", "batched", "serve", "engine"),
                    help="profiling run: only this leg (grid, batched, serving or engine), its JSON object on stdout")
    ap.add_argument("--no-coop", action="store_true",
                    help="with --leg: leave out cooperative launches (rocprofv3 can lose its output at "
                         "process exit after them)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times), to scale rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE against a known byte count")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch  # plumbing only: a plain copy kernel of known size
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.int32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 4 * n, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; multi-GPU sharding is a design "
                                   "note only (queries shard with no exchange)", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import gpugraphalgo_amd as gga

    if args.leg:
        leg = {"grid": grid_leg, "batched": batched_leg, "engine": engine_leg}.get(args.leg)
        res = leg(gga, np, coop=not args.no_coop) if leg else serving_leg(gga, np, full=not args.serve_reduced)
        print(json.dumps({"leg": args.leg, **res}))
        return 0

    log(f"generating R-MAT s{args.scale} ef{args.edge_factor} seed {args.seed}")
    g = gga.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed, wlo=1, whi=100)
    maxdeg, src = g.maxdeg()
    log(f"n={g.n} m={g.m} maxdeg={maxdeg} src={src}")
    sweep = (args.mda,) if (args.no_sweep or args.quick) else SWEEP
    results = {}
    with gga.DeviceGraph(g) as dg:
        for T in sweep:
            log(f"T={T} mode={args.mode}: {args.warmup} warmup + {args.steps} timed steps")
            wall, stats, counted = run_config(dg, src, T, args.mode, args.steps, args.warmup)
            results[T] = summarize(wall, stats, counted, "topo" if stats[0]["mode"] == 0 else "worklist", False)
            results[T]["split"] = dict(dg.split_stats)
            log(f"T={T}: {results[T]['ms_per_step']:.3f} ms/step, {results[T]['gteps']:.2f} GTEPS, "
                f"{results[T]['achieved_gbs']:.0f} GB/s algorithmic, {results[T]['iterations']} steps")
        best = min(results, key=lambda t: results[t]["ms_per_step"])
        r = results[best]
        dg.split(mda=best)
        d_gpu, _ = dg.sssp(src, mode=args.mode)
        ver = dg.verify(src, bfs=False)
        assert ver["nerr"] == 0 and ver["untight"] == 0 and ver["src_ok"], ver
        distinct = None if args.quick else distinct_sources(dg, g, args.mode)
        unsplit = None if args.quick else unsplit_query(dg, src, args.mode, r["ms_per_step"])
    cold = None if args.quick else cold_first_query(gga, g, src, best, args.mode)

    cpu = None
    if not (args.no_cpu or args.quick):
        log("cpu baseline: serial SSSP restatement, 1 thread")
        cpu, d_cpu = cpu_baseline(g, src, r["m_reach"])
        assert np.array_equal(d_cpu, d_gpu), "GPU distances differ from the CPU reference"
        log(f"cpu: {cpu['value']:.4f} GTEPS; GPU result bit-exact")

    batched = serving = grid = None
    if not (args.no_grid or args.quick):
        log("grid leg (4096^2 grid + shortcuts)")
        grid = grid_leg(gga, np)
        log(f"grid: {json.dumps(grid)}")
    if not (args.no_batched or args.quick):
        log("batched multi-source leg (R-MAT s20)")
        batched = batched_leg(gga, np)
        log(f"batched: {json.dumps(batched)}")
    if not (args.no_serve or args.quick):
        log("serving leg (%s)" % ("reduced resident set" if args.serve_reduced else "12 resident graphs"))
        serving = serving_leg(gga, np, full=not args.serve_reduced)
        if not args.serve_reduced:
            serving["mesh_engine"] = engine_leg(gga, np)
        log(f"serving: {json.dumps(serving)}")

    ceiling = None
    if not args.quick and r["edges_inspected"] > 0:
        pass_pct = max(1, min(100, round(100.0 * r["relaxations"] / r["edges_inspected"])))
        ceiling = random_access_ceiling(args.scale, pass_pct)
    random_access = None
    if ceiling:
        relax = {(x["layout"], x["atomics"]): x for x in ceiling["relax"]}
        achieved = r["edges_inspected"] / (r["ktime_ms"] * 1e-3) / 1e9 if r["ktime_ms"] > 0 else 0.0
        cap = relax[("unsliced", True)]["g_edges_per_s"]
        random_access = {
            "achieved_g_edges_per_s": round(achieved, 1),
            "probe_g_edges_per_s": cap,
            "frac": round(achieved / cap, 3) if cap else None,
            "probe": ceiling,
        }
    traffic = None
    if os.path.exists(TRAFFIC) and r["mode"] == 1:
        from gpugraphalgo_amd.profiling import read_summary
        traffic = float(read_summary(TRAFFIC)["traffic_bytes_per_dispatch"])
    mode_name = {0: "topo", 1: "worklist", 2: "persistent"}.get(r["mode"], str(r["mode"]))
    kname = {0: "k_topo_step", 1: "k_wl_step", 2: "k_persistent"}.get(r["mode"], "?")
    line = {
        "metric": METRIC,
        "value": round(r["gteps"], 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT (built-in seeded generator), weights uniform u32 [1,100]",
        "config": {
            "workload": "node-split SSSP, R-MAT scale 22 edge factor 16",
            "scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed,
            "n": g.n, "m": g.m, "source": src, "maxdeg": maxdeg, "m_reach": r["m_reach"],
            "split_threshold": best, "chunks": r["split"]["ns_nnode"], "mode": mode_name,
            "relax_steps": r["iterations"],
            "sweep_ms_per_step": {str(t): round(results[t]["ms_per_step"], 4) for t in results},
            "sweep_gteps": {str(t): round(results[t]["gteps"], 3) for t in results},
            "cold_first_query": cold,
            "distinct_sources": distinct,
            "unsplit": unsplit,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(r["achieved_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": (os.path.relpath(TRAFFIC, ROOT) + " (rocprofv3 --pmc L2 request counters of "
                               "k_wl_step on this workload: reads by request size, writes, atomics; a "
                               "committed profile, not measured in this run)")
                              if traffic is not None else None,
            "l2_hit_rate": float(read_summary(TRAFFIC)["l2_hit_rate"]) if traffic is not None else None,
            "bytes_per_launch": round(r["bytes_per_query"] / max(r["launches"], 1)),
            "launches_per_step": r["launches"],
            "avg_launch_ms": round(r["ktime_ms"] / max(r["launches"], 1), 5),
            "measured_over": (f"{r['counted_queries']} instrumented queries after the timed ones (work "
                              f"counters; HIP events around each relax launch on the graph's stream), "
                              f"{r['counted_device_ms']:.3f} ms each"),
            "work": {"edges_inspected": r["edges_inspected"], "chunks_expanded": r["chunks_expanded"],
                     "relaxations": r["relaxations"]},
            "random_access": random_access,
        },
        "cpu_baseline": cpu,
        "grid": grid,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
