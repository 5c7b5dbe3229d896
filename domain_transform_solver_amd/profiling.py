"""Summaries of rocprofv3 output for the solver kernels.

  python -m domain_transform_solver_amd.profiling stats <results.db | kernel_stats.csv>
      per-kernel dispatch count, total / average / min / max duration (us);
  python -m domain_transform_solver_amd.profiling traffic <fetch_dir> <write_dir>
          --stream-elems N [--json out.json]
      HBM bytes per dispatch from two `rocprofv3 --pmc FETCH_SIZE` /
      `--pmc WRITE_SIZE` runs (KiB per dispatch, averaged per kernel), each
      scaled by the factor that maps the same counter onto the known traffic
      of the k_stream5 calibration kernel over N floats (reads 16 N bytes,
      writes 4 N bytes).

The JSON written by `traffic` is what bench.py reports as roofline.traffic.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sqlite3
import sys


def short_name(name: str) -> str:
    """Kernel name without its argument list."""
    return re.sub(r"\(.*", "", name).strip()


def kernel_stats(path: str) -> list[tuple[str, int, float, float, float, float]]:
    """(name, count, total_us, avg_us, min_us, max_us), largest total first,
    from a rocprofv3 SQLite results database or a *_kernel_stats.csv."""
    agg: dict[str, list[float]] = collections.defaultdict(list)
    if path.endswith(".db"):
        con = sqlite3.connect(path)
        for name, start, end in con.execute("select name, start, end from kernels"):
            agg[short_name(name)].append((end - start) / 1e3)
        rows = [(n, len(v), sum(v), sum(v) / len(v), min(v), max(v)) for n, v in agg.items()]
    else:
        rows = []
        for r in csv.DictReader(open(path)):
            n = int(r["Calls"])
            rows.append((short_name(r["Name"]), n, float(r["TotalDurationNs"]) / 1e3,
                         float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                         float(r["MaxNs"]) / 1e3))
    return sorted(rows, key=lambda r: -r[2])


def format_stats(rows, limit: int = 40) -> str:
    tot = sum(r[2] for r in rows) or 1.0
    out = [f"{'total_us':>11} {'share':>6} {'calls':>6} {'avg_us':>9} {'min_us':>9} "
           f"{'max_us':>9}  kernel"]
    for name, n, t, a, lo, hi in rows[:limit]:
        out.append(f"{t:11.1f} {100 * t / tot:5.1f}% {n:6d} {a:9.2f} {lo:9.2f} {hi:9.2f}  "
                   f"{name[:100]}")
    return "\n".join(out)


def pmc_means(directory: str) -> dict[str, dict[str, float]]:
    """Mean value per dispatch of every counter for every kernel, from the
    *_counter_collection.csv files under a rocprofv3 --pmc output folder."""
    acc: dict[str, dict[str, list[float]]] = collections.defaultdict(
        lambda: collections.defaultdict(list))
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *counter_collection.csv under {directory}")
    for f in files:
        for r in csv.DictReader(open(f)):
            acc[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(
                float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


def traffic(fetch_dir: str, write_dir: str, stream_elems: int) -> dict:
    fetch = pmc_means(fetch_dir)
    write = pmc_means(write_dir)
    cal = next((k for k in fetch if "k_stream5" in k), None)
    if cal is None or cal not in write:
        raise ValueError("calibration kernel k_stream5 missing from the counter runs")
    f_kib = fetch[cal]["FETCH_SIZE"]
    w_kib = write[cal]["WRITE_SIZE"]
    f_scale = 16.0 * stream_elems / (f_kib * 1024.0)
    w_scale = 4.0 * stream_elems / (w_kib * 1024.0)
    kernels = {}
    for k in fetch:
        if k == cal or k not in write:
            continue
        fb = fetch[k].get("FETCH_SIZE", 0.0) * 1024.0 * f_scale
        wb = write[k].get("WRITE_SIZE", 0.0) * 1024.0 * w_scale
        kernels[k] = {"fetch_bytes": fb, "write_bytes": wb, "bytes": fb + wb,
                      "fetch_kib_raw": fetch[k].get("FETCH_SIZE", 0.0),
                      "write_kib_raw": write[k].get("WRITE_SIZE", 0.0)}
    return {"calibration": {"kernel": cal, "elements": stream_elems,
                            "fetch_kib_raw": f_kib, "write_kib_raw": w_kib,
                            "fetch_scale": f_scale, "write_scale": w_scale},
            "kernels": kernels}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("path")
    a.add_argument("--limit", type=int, default=40)
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--stream-elems", type=int, required=True)
    b.add_argument("--json")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(format_stats(kernel_stats(args.path), args.limit))
        return 0
    t = traffic(args.fetch_dir, args.write_dir, args.stream_elems)
    c = t["calibration"]
    print(f"calibration {c['kernel']} over {c['elements']} floats: FETCH_SIZE "
          f"{c['fetch_kib_raw']:.1f} KiB -> x{c['fetch_scale']:.3f}, WRITE_SIZE "
          f"{c['write_kib_raw']:.1f} KiB -> x{c['write_scale']:.3f}")
    for k, v in sorted(t["kernels"].items(), key=lambda kv: -kv[1]["bytes"]):
        print(f"{v['bytes'] / 1e6:10.2f} MB/dispatch (fetch {v['fetch_bytes'] / 1e6:8.2f}, "
              f"write {v['write_bytes'] / 1e6:8.2f})  {k[:90]}")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(t, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
