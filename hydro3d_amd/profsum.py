"""Summaries of rocprofv3 CSV output, per kernel and launch grid.

    python -m hydro3d_amd.profsum trace <run_kernel_trace.csv> [--match REGEX]
    python -m hydro3d_amd.profsum pmc <run_counter_collection.csv> [--match REGEX]
    python -m hydro3d_amd.profsum stats <run_kernel_stats.csv> [--match REGEX]

`trace` prints, for every (kernel, workgroups per launch) pair, the number of
launches and the total / average / min / max duration in microseconds: one
kernel serves every multigrid level, and the grid size tells the levels
apart.  `pmc` prints the average value of each counter per dispatch for the
same pairs.  `stats` reprints rocprofv3's own --stats table with short names.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys


def short_name(name: str) -> str:
    m = re.search(r"(k_\w+(<[^>]*>)?)", name)
    if m:
        return m.group(1)
    return name.split("(")[0][:60]


def _workgroups(row) -> int:
    if "Grid_Size_X" in row:
        g = int(row["Grid_Size_X"]) * int(row["Grid_Size_Y"]) * int(row["Grid_Size_Z"])
        w = int(row["Workgroup_Size_X"]) * int(row["Workgroup_Size_Y"]) * int(row["Workgroup_Size_Z"])
    else:
        g, w = int(row["Grid_Size"]), int(row["Workgroup_Size"])
    return g // max(w, 1)


def trace_summary(path: str, match: str | None = None):
    """{(kernel, workgroups): [count, total_us, min_us, max_us]}"""
    agg = collections.OrderedDict()
    with open(path) as fh:
        for r in csv.DictReader(fh):
            name = short_name(r["Kernel_Name"])
            if match and not re.search(match, name):
                continue
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            key = (name, _workgroups(r))
            a = agg.setdefault(key, [0, 0.0, float("inf"), 0.0])
            a[0] += 1
            a[1] += d
            a[2] = min(a[2], d)
            a[3] = max(a[3], d)
    return agg


def pmc_summary(path: str, match: str | None = None):
    """{(kernel, workgroups, counter): [count, sum]}"""
    agg = collections.OrderedDict()
    with open(path) as fh:
        for r in csv.DictReader(fh):
            name = short_name(r["Kernel_Name"])
            if match and not re.search(match, name):
                continue
            key = (name, _workgroups(r), r["Counter_Name"])
            a = agg.setdefault(key, [0, 0.0])
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    return agg


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("mode", choices=["trace", "pmc", "stats"])
    ap.add_argument("csv")
    ap.add_argument("--match", default=None, help="regex on the short kernel name")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        print(f"{'kernel':40s} {'calls':>6s} {'total_us':>10s} {'avg_us':>9s} {'pct':>6s} "
              f"{'min_us':>9s} {'max_us':>9s}")
        with open(a.csv) as fh:
            for r in csv.DictReader(fh):
                k = short_name(r["Name"])
                if a.match and not re.search(a.match, k):
                    continue
                print(f"{k:40s} {int(r['Calls']):6d} {float(r['TotalDurationNs']) / 1e3:10.1f} "
                      f"{float(r['AverageNs']) / 1e3:9.2f} {float(r['Percentage']):6.2f} "
                      f"{float(r['MinNs']) / 1e3:9.2f} {float(r['MaxNs']) / 1e3:9.2f}")
    elif a.mode == "trace":
        agg = trace_summary(a.csv, a.match)
        print(f"{'kernel':40s} {'wgs':>7s} {'calls':>6s} {'total_us':>10s} {'avg_us':>9s} "
              f"{'min_us':>9s} {'max_us':>9s}")
        for (k, wg), (n, tot, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            print(f"{k:40s} {wg:7d} {n:6d} {tot:10.1f} {tot / n:9.2f} {mn:9.2f} {mx:9.2f}")
    else:
        agg = pmc_summary(a.csv, a.match)
        print(f"{'kernel':40s} {'wgs':>7s} {'counter':>12s} {'dispatches':>10s} {'avg_value':>14s}")
        for (k, wg, c), (n, tot) in agg.items():
            print(f"{k:40s} {wg:7d} {c:>12s} {n:10d} {tot / n:14.1f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
