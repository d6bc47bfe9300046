"""Summaries of rocprofv3 output for the particle step.

  python -m spearhed_amd.profiling stats <kernel_stats.csv> [--title T] [--command C]
      markdown table of per-kernel average durations (rocprofv3 --kernel-trace --stats);
  python -m spearhed_amd.profiling traffic <fetch_counter_collection.csv> <write_counter_collection.csv>
      <probe_fetch.csv> <probe_write.csv> --probe-bytes B --kernel NAME --particles N --tuning JSON
      HBM bytes per launch of one kernel from a FETCH_SIZE pass and a WRITE_SIZE pass, each scaled by
      what the same counter reads for the calibration copy (sph_probe_copy of B bytes reads B and
      writes B); prints a markdown block and the JSON record the benchmark reads.

FETCH_SIZE and WRITE_SIZE are reported in KiB per dispatch.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import sys


def short_name(name: str) -> str:
    base = name.split("(")[0].replace("void ", "").replace("sph::", "")
    if "(anonymous namespace)::" in name:
        base = name.split("::")[1].split("(")[0]
    return base.strip()


def kernel_stats(path: str):
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows:
        out.append({"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
                    "avg_us": float(r["AverageNs"]) / 1e3, "min_us": float(r["MinNs"]) / 1e3,
                    "max_us": float(r["MaxNs"]) / 1e3, "pct": float(r["Percentage"])})
    return out


def stats_markdown(path: str, title: str = "", command: str = "") -> str:
    lines = []
    if title:
        lines += [f"# {title}", ""]
    if command:
        lines += ["Command:", "", "```", command, "```", ""]
    lines += ["| kernel | calls | avg us | min us | max us | % time |", "|---|---:|---:|---:|---:|---:|"]
    for k in kernel_stats(path):
        lines.append(f"| {k['kernel']} | {k['calls']} | {k['avg_us']:.1f} | {k['min_us']:.1f} | "
                     f"{k['max_us']:.1f} | {k['pct']:.2f} |")
    return "\n".join(lines) + "\n"


def counter_means(path: str, counter: str):
    """Mean per-dispatch counter value by kernel (short name)."""
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def traffic(fetch_csv, write_csv, probe_fetch_csv, probe_write_csv, probe_bytes, kernel):
    pf = counter_means(probe_fetch_csv, "FETCH_SIZE")["k_probe_copy"][0]
    pw = counter_means(probe_write_csv, "WRITE_SIZE")["k_probe_copy"][0]
    kib = probe_bytes / 1024.0
    f_scale, w_scale = kib / pf, kib / pw  # counter KiB -> true KiB
    fetch = counter_means(fetch_csv, "FETCH_SIZE")
    write = counter_means(write_csv, "WRITE_SIZE")
    table = {}
    for k in sorted(set(fetch) | set(write)):
        fk = fetch.get(k, (0.0, 0))[0] * f_scale * 1024
        wk = write.get(k, (0.0, 0))[0] * w_scale * 1024
        table[k] = {"fetch_bytes": fk, "write_bytes": wk}
    main = [k for k in table if k.startswith(kernel)]
    if not main:
        raise SystemExit(f"kernel {kernel} not in the counter files")
    rec = table[main[0]]
    return {"fetch_scale": f_scale, "write_scale": w_scale, "kernels": table, "kernel": main[0],
            "bytes_per_launch": rec["fetch_bytes"] + rec["write_bytes"]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title", default="")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("probe_fetch")
    t.add_argument("probe_write")
    t.add_argument("--probe-bytes", type=int, required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--particles", type=int, required=True)
    t.add_argument("--tuning", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.title, a.command))
        return
    res = traffic(a.fetch, a.write, a.probe_fetch, a.probe_write, a.probe_bytes, a.kernel)
    lines = [f"Calibration: FETCH_SIZE x {res['fetch_scale']:.3f}, WRITE_SIZE x {res['write_scale']:.3f} "
             f"(copy of {a.probe_bytes} bytes)", "",
             "| kernel | HBM read MB/launch | HBM write MB/launch |", "|---|---:|---:|"]
    for k, v in res["kernels"].items():
        lines.append(f"| {k} | {v['fetch_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} |")
    print("\n".join(lines))
    rec = {"particles": a.particles, "tuning": json.loads(a.tuning), "kernel": res["kernel"],
           "bytes_per_launch": res["bytes_per_launch"],
           "fetch_bytes": res["kernels"][res["kernel"]]["fetch_bytes"],
           "write_bytes": res["kernels"][res["kernel"]]["write_bytes"],
           "fetch_scale": res["fetch_scale"], "write_scale": res["write_scale"]}
    print("JSON " + json.dumps(rec))


if __name__ == "__main__":
    main()
