"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m gpugraphalgo_amd.profiling stats DIR --cmd "..."
        DIR/*kernel_stats.csv (rocprofv3 --kernel-trace --stats --output-format csv)
        -> per-kernel calls / average / total table
    python -m gpugraphalgo_amd.profiling traffic --kernel NAME --fetch D1 --write D2 \
            --copy-fetch D3 --copy-write D4 --copy-bytes B --cmd "..."
        D1..D4/*counter_collection.csv (rocprofv3 --pmc FETCH_SIZE, --pmc WRITE_SIZE, for
        the workload and for a plain copy of B bytes) -> HBM bytes per dispatch of NAME,
        each counter scaled by what it reads for the copy (bytes / (KiB * 1024)).
        --last N: only the last N dispatches of NAME (one run at the end of the workload).

Summaries are `key = value` text; read_summary() parses them back.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys


def _one(pattern: str) -> str:
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise FileNotFoundError(pattern)
    return hits[0]


def summarize_stats(d: str, cmd: str, top: int = 12) -> str:
    path = _one(os.path.join(d, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = [f"# rocprofv3 kernel statistics", f"command = {cmd}", f"total_kernel_ms = {total / 1e6:.3f}",
           "", f"{'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}"]
    for r in rows[:top]:
        name = r["Name"].replace("(anonymous namespace)::", "")
        name = name.split("(")[0][:70]
        out.append(f"{name:70s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:10.2f} "
                   f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['Percentage']):6.2f}")
    return "\n".join(out) + "\n"


def _counter(d: str, name: str, kernel_substr: str):
    path = _one(os.path.join(d, "**", "*counter_collection.csv"))
    vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(path))
            if kernel_substr in r["Kernel_Name"] and r["Counter_Name"].startswith(name)]
    return vals


def summarize_traffic(kernel: str, fetch: str, write: str, copy_fetch: str, copy_write: str,
                      copy_bytes: int, cmd: str, last: int = 0) -> str:
    cf = _counter(copy_fetch, "FETCH_SIZE", "copyBuffer")
    cw = _counter(copy_write, "WRITE_SIZE", "copyBuffer")
    f_fetch = copy_bytes / (sum(cf) / len(cf) * 1024.0)
    f_write = copy_bytes / (sum(cw) / len(cw) * 1024.0)
    kf = _counter(fetch, "FETCH_SIZE", kernel)
    kw = _counter(write, "WRITE_SIZE", kernel)
    if last:
        kf, kw = kf[-last:], kw[-last:]
    fetch_b = sum(kf) / len(kf) * 1024.0 * f_fetch
    write_b = sum(kw) / len(kw) * 1024.0 * f_write
    lines = [
        "# HBM traffic per dispatch (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)",
        f"command = {cmd}",
        f"kernel = {kernel}",
        f"dispatches = {len(kf)}" + (f" (the last {last})" if last else ""),
        f"copy_bytes = {copy_bytes}",
        f"fetch_factor = {f_fetch:.4f}",
        f"write_factor = {f_write:.4f}",
        f"fetch_kib_per_dispatch_raw = {sum(kf) / len(kf):.1f}",
        f"write_kib_per_dispatch_raw = {sum(kw) / len(kw):.1f}",
        f"fetch_bytes_per_dispatch = {fetch_b:.0f}",
        f"write_bytes_per_dispatch = {write_b:.0f}",
        f"traffic_bytes_per_dispatch = {fetch_b + write_b:.0f}",
    ]
    return "\n".join(lines) + "\n"


def read_summary(path: str) -> dict:
    out = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--cmd", required=True)
    t = sub.add_parser("traffic")
    for k in ("kernel", "fetch", "write", "copy-fetch", "copy-write", "cmd"):
        t.add_argument("--" + k, required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--last", type=int, default=0)
    a = ap.parse_args(argv)
    if a.what == "stats":
        sys.stdout.write(summarize_stats(a.dir, a.cmd))
    else:
        sys.stdout.write(summarize_traffic(a.kernel, a.fetch, a.write, a.copy_fetch, a.copy_write,
                                           a.copy_bytes, a.cmd, a.last))
    return 0


if __name__ == "__main__":
    sys.exit(main())
