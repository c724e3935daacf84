"""Summarises rocprofv3 databases of bench.py runs into a markdown table.

Inputs are the rocpd SQLite files that `rocprofv3 -d DIR -o NAME` writes:
a kernel trace (--kernel-trace --stats), a FETCH_SIZE pass, a WRITE_SIZE
pass, and the same two counters over the copy-calibration kernel
(bench.py --copy-calibration MIB).  Dispatches are grouped by kernel name and
grid size, so the 4,096-world batch (4,096 workgroups) and the plain world
(1 workgroup) are reported apart.

usage: python beat-box_amd/rocprof_summary.py --trace T.db --fetch F.db
       --write W.db --cal-fetch CF.db --cal-write CW.db --cal-mib 256 [--out X.md]
       python beat-box_amd/rocprof_summary.py --trace T.db [--bench-log L] [--out X.md]
(without the counter databases the HBM columns stay empty; --bench-log copies
the "[bench]" lines of the traced run's stderr under the table)
"""
from __future__ import annotations

import argparse
import sqlite3
import statistics
from collections import defaultdict


def short(name: str) -> str:
    name = name.replace("void ", "").replace("(bb::Dev)", "")
    if name.startswith("bb::k_copy_calibrate"):
        return "bb::k_copy_calibrate"
    return name


def trace_rows(path: str):
    c = sqlite3.connect(path)
    rows = c.execute("select name, grid_x, workgroup_x, duration, vgpr_count, accum_vgpr_count, "
                     "sgpr_count, lds_size, scratch_size from kernels").fetchall()
    out = defaultdict(list)
    meta = {}
    for name, gx, wx, dur, vg, ag, sg, lds, scr in rows:
        key = (short(name), gx // max(wx, 1))
        out[key].append(dur)
        meta[key] = (vg, ag, sg, lds, scr)
    return out, meta


def counter_rows(path: str, counter: str):
    c = sqlite3.connect(path)
    rows = c.execute("select kernel_name, grid_size, workgroup_size, value from counters_collection "
                     "where counter_name = ?", (counter,)).fetchall()
    out = defaultdict(list)
    for name, gs, ws, val in rows:
        out[(short(name), gs // max(ws, 1))].append(val)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--cal-fetch", default=None)
    ap.add_argument("--cal-write", default=None)
    ap.add_argument("--cal-mib", type=int, default=256)
    ap.add_argument("--out", default=None)
    ap.add_argument("--json-out", default=None, help="per-kernel figures as JSON (read by bench.py)")
    ap.add_argument("--command", default="", help="command lines that produced the databases (recorded)")
    ap.add_argument("--bench-log", default=None,
                    help="stderr of the traced bench.py run: its [bench] lines are copied under the table")
    ap.add_argument("--bench-json", default=None,
                    help="the JSON line bench.py printed in the counter runs: its step window and the "
                         "algorithmic bytes per launch of each stage are recorded with the counters")
    a = ap.parse_args()
    dur, meta = trace_rows(a.trace)
    counters = all((a.fetch, a.write, a.cal_fetch, a.cal_write))
    fetch = counter_rows(a.fetch, "FETCH_SIZE") if counters else {}
    write = counter_rows(a.write, "WRITE_SIZE") if counters else {}
    cf = counter_rows(a.cal_fetch, "FETCH_SIZE") if counters else {}
    cw = counter_rows(a.cal_write, "WRITE_SIZE") if counters else {}
    kib = a.cal_mib * 1024
    cal_f = [v for k, vs in cf.items() if k[0] == "bb::k_copy_calibrate" for v in vs]
    cal_w = [v for k, vs in cw.items() if k[0] == "bb::k_copy_calibrate" for v in vs]
    ff = statistics.mean(cal_f) / kib if cal_f else float("nan")
    fw = statistics.mean(cal_w) / kib if cal_w else float("nan")
    lines = []
    if a.command:
        lines.append("Commands:")
        lines.append("")
        for c in a.command.split(";;"):
            lines.append(f"    {c.strip()}")
        lines.append("")
    if counters:
        lines.append(f"Copy-kernel calibration ({a.cal_mib} MiB read + {a.cal_mib} MiB written per dispatch): "
                     f"FETCH_SIZE reads {ff:.3f}x, WRITE_SIZE reads {fw:.3f}x the true bytes; the HBM "
                     f"columns below are divided by these factors.")
    else:
        lines.append("Kernel trace only (no counter passes): the HBM columns are empty.")
    lines.append("")
    lines.append("| kernel | workgroups | calls | avg us | min us | max us | VGPR | AGPR | SGPR | LDS B | "
                 "scratch B | HBM read MB/dispatch | HBM write MB/dispatch | GB/s (read+write / avg) |")
    lines.append("|---|---|---|---|---|---|---|---|---|---|---|---|---|---|")
    table = {}
    for key in sorted(dur, key=lambda k: -sum(dur[k])):
        ds = dur[key]
        avg = statistics.mean(ds)
        vg, ag, sg, lds, scr = meta[key]
        fr = statistics.mean(fetch[key]) / ff / 1024 if fetch.get(key) else None
        wr = statistics.mean(write[key]) / fw / 1024 if write.get(key) else None
        gbs = ((fr or 0) + (wr or 0)) * 1e6 / avg if (fr is not None and avg > 0) else None
        table[f"{key[0]}@{key[1]}"] = dict(kernel=key[0], workgroups=key[1], calls=len(ds), avg_us=avg / 1e3,
                                           read_bytes=None if fr is None else fr * 1e6,
                                           write_bytes=None if wr is None else wr * 1e6)
        lines.append(f"| {key[0]} | {key[1]} | {len(ds)} | {avg / 1e3:.1f} | {min(ds) / 1e3:.1f} | "
                     f"{max(ds) / 1e3:.1f} | {vg} | {ag} | {sg} | {lds} | {scr} | "
                     f"{'' if fr is None else f'{fr:.2f}'} | {'' if wr is None else f'{wr:.2f}'} | "
                     f"{'' if gbs is None else f'{gbs:.0f}'} |")
    if a.bench_log:
        with open(a.bench_log) as f:
            logged = [x.rstrip() for x in f if x.startswith("[bench]")]
        lines += ["", "bench.py output of the traced run:", ""] + [f"    {x}" for x in logged]
    text = "\n".join(lines) + "\n"
    if a.json_out:
        import json
        extra = {}
        if a.bench_json:
            with open(a.bench_json) as f:
                line = [x for x in f.read().splitlines() if x.startswith("{")][-1]
            bench = json.loads(line)
            extra = dict(window=dict(steps=bench["steps"], warmup=bench["warmup"]),
                         launch_params=bench["config"].get("launch_params"),
                         model_bytes_per_launch={k: v["model_bytes"] for k, v in bench["config"]["stages"].items()})
        with open(a.json_out, "w") as f:
            json.dump(dict(fetch_factor=ff, write_factor=fw, **extra, kernels=table), f, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
