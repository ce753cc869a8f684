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

    python -m gpugraphalgo_amd.profiling counters --kernel NAME --hitmiss D1 --rdsize D2 \
            --wratom D3 --dram D4 --copy-rd D5 --copy-wr D6 --copy-bytes B --cmd "..."
        Di/*counter_collection.csv of rocprofv3 --pmc passes (at most four TCC
        counters each): TCC_HIT/MISS/REQ; TCC_EA0_RDREQ by request size (32,
        64, 128 B); TCC_EA0_WRREQ (32 / 64 B) and TCC_ATOMIC / TCC_EA0_ATOMIC;
        DRAM-destined requests -> L2 hit rate, bytes read and written past L2
        per dispatch of exactly NAME (request counts times sizes: no scale
        factor), atomics per dispatch; the same size passes over a copy of B
        bytes check the byte accounting.

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


def _per_dispatch(d: str, kernel: str, exact: bool = True, last: int = 0) -> dict:
    """Mean over the dispatches of `kernel` (the last `last` of them, when
    given) of each counter in d's collection."""
    path = _one(os.path.join(d, "**", "*counter_collection.csv"))
    per: dict = {}
    for r in csv.DictReader(open(path)):
        name = r["Kernel_Name"].replace("(anonymous namespace)::", "")
        if (name.split("(")[0].strip() == kernel) if exact else (kernel in name):
            per.setdefault(r["Dispatch_Id"], {}).setdefault(r["Counter_Name"], 0.0)
            per[r["Dispatch_Id"]][r["Counter_Name"]] += float(r["Counter_Value"])
    if not per:
        raise ValueError(f"no dispatch of {kernel} in {path}")
    if last:
        per = {k: per[k] for k in sorted(per, key=int)[-last:]}
    keys = {k for v in per.values() for k in v}
    return {"_dispatches": len(per), **{k: sum(v.get(k, 0.0) for v in per.values()) / len(per) for k in keys}}


def _read_bytes(c: dict) -> float:
    return (32.0 * c["TCC_EA0_RDREQ_32B_sum"] + 64.0 * c["TCC_EA0_RDREQ_64B_sum"]
            + 128.0 * c["TCC_EA0_RDREQ_128B_sum"])


def _write32(c: dict) -> float:
    # WRREQ counts 32- and 64-byte transactions; atomics travel on the same
    # interface as 32-byte ones.  When WRREQ falls short of the atomics plus
    # the 64-byte writes (the counters do not add up for some atomic mixes),
    # no 32-byte plain write is inferred.
    return max(0.0, c["TCC_EA0_WRREQ_sum"] - c["TCC_EA0_WRREQ_64B_sum"] - c.get("TCC_EA0_ATOMIC_sum", 0.0))


def _write_bytes(c: dict) -> float:
    return 64.0 * c["TCC_EA0_WRREQ_64B_sum"] + 32.0 * _write32(c)


def summarize_counters(kernel: str, hitmiss: str, rdsize: str, wratom: str, dram: str, copy_rd: str,
                       copy_wr: str, copy_bytes: int, cmd: str, last: int = 0) -> str:
    hm = _per_dispatch(hitmiss, kernel, last=last)
    rd = _per_dispatch(rdsize, kernel, last=last)
    wa = _per_dispatch(wratom, kernel, last=last)
    dr = _per_dispatch(dram, kernel, last=last)
    cr = _per_dispatch(copy_rd, "copyBuffer", exact=False)
    cw = _per_dispatch(copy_wr, "copyBuffer", exact=False)
    # the copy: its largest dispatches are the timed 1 GiB copies
    rb, wb = _read_bytes(rd), _write_bytes(wa)
    ab = 32.0 * wa["TCC_EA0_ATOMIC_sum"]   # 32-byte requests on the write interface
    lines = [
        "# L2 (TCC) counters per dispatch: hit rate, bytes past L2 by request size, atomics",
        f"command = {cmd}",
        f"kernel = {kernel}",
        f"dispatches = {rd['_dispatches']}" + (f" (the last {last})" if last else ""),
        f"l2_requests = {hm['TCC_REQ_sum']:.0f}",
        f"l2_hits = {hm['TCC_HIT_sum']:.0f}",
        f"l2_misses = {hm['TCC_MISS_sum']:.0f}",
        f"l2_hit_rate = {hm['TCC_HIT_sum'] / max(hm['TCC_HIT_sum'] + hm['TCC_MISS_sum'], 1.0):.4f}",
        f"read_requests_32B = {rd['TCC_EA0_RDREQ_32B_sum']:.0f}",
        f"read_requests_64B = {rd['TCC_EA0_RDREQ_64B_sum']:.0f}",
        f"read_requests_128B = {rd['TCC_EA0_RDREQ_128B_sum']:.0f}",
        f"read_requests_dram = {dr['TCC_EA0_RDREQ_DRAM_sum']:.0f}",
        f"write_requests = {wa['TCC_EA0_WRREQ_sum']:.0f}",
        f"write_requests_64B = {wa['TCC_EA0_WRREQ_64B_sum']:.0f}",
        f"write_requests_32B_inferred = {_write32(wa):.0f}",
        f"write_requests_dram = {dr['TCC_EA0_WRREQ_DRAM_sum']:.0f}",
        f"atomics = {wa['TCC_ATOMIC_sum']:.0f}",
        f"atomics_past_l2 = {wa['TCC_EA0_ATOMIC_sum']:.0f}",
        f"read_bytes_per_dispatch = {rb:.0f}",
        f"write_bytes_per_dispatch = {wb:.0f}",
        f"atomic_bytes_per_dispatch = {ab:.0f}",
        f"traffic_bytes_per_dispatch = {rb + wb + ab:.0f}",
        f"copy_bytes = {copy_bytes}",
        f"copy_read_bytes_counted = {_read_bytes(cr):.0f}",
        f"copy_write_bytes_counted = {64.0 * cw['TCC_EA0_WRREQ_64B_sum'] + 32.0 * (cw['TCC_EA0_WRREQ_sum'] - cw['TCC_EA0_WRREQ_64B_sum']):.0f}",
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
    c = sub.add_parser("counters")
    for k in ("kernel", "hitmiss", "rdsize", "wratom", "dram", "copy-rd", "copy-wr", "cmd"):
        c.add_argument("--" + k, required=True)
    c.add_argument("--copy-bytes", type=int, required=True)
    c.add_argument("--last", type=int, default=0)
    a = ap.parse_args(argv)
    if a.what == "stats":
        sys.stdout.write(summarize_stats(a.dir, a.cmd))
    elif a.what == "counters":
        sys.stdout.write(summarize_counters(a.kernel, a.hitmiss, a.rdsize, a.wratom, a.dram, a.copy_rd,
                                            a.copy_wr, a.copy_bytes, a.cmd, a.last))
    else:
        sys.stdout.write(summarize_traffic(a.kernel, a.fetch, a.write, a.copy_fetch, a.copy_write,
                                           a.copy_bytes, a.cmd, a.last))
    return 0


if __name__ == "__main__":
    sys.exit(main())
