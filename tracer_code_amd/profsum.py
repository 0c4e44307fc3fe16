"""Summaries of rocprofv3 CSV output, grouped by kernel and launch size.

    python -m tracer_code_amd.profsum trace <run_kernel_trace.csv> [top]
    python -m tracer_code_amd.profsum pmc <run_counter_collection.csv> [top]

`trace` prints, per (kernel, grid size), the launch count and the total and
average kernel durations; `pmc` prints the average counter value per
dispatch.  The multigrid levels of one kernel differ only in grid size, so
grouping by it separates e.g. the level-1 sweep from the coarser ones.
"""
import csv
import re
import sys
from collections import defaultdict


def short(name):
    name = re.sub(r"^void ", "", name)
    head = name.split("(")[0]
    return head


def trace(path, top=40):
    agg = defaultdict(lambda: [0, 0.0])
    total = 0.0
    for r in csv.DictReader(open(path)):
        d = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        key = (short(r["Kernel_Name"]), int(r["Grid_Size"]) if "Grid_Size" in r else int(r.get("Grid_Size_X", 0)))
        agg[key][0] += 1
        agg[key][1] += d
        total += d
    rows = sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]
    print(f"| kernel | grid size | calls | total ms | avg us | % of kernel time |")
    print(f"|---|---:|---:|---:|---:|---:|")
    for (k, g), (n, t) in rows:
        print(f"| {k} | {g} | {n} | {t / 1e6:.2f} | {t / n / 1e3:.2f} | {100 * t / total:.1f} |")
    print(f"\ntotal kernel time: {total / 1e6:.1f} ms over {sum(v[0] for v in agg.values())} launches")


def pmc(path, top=40):
    agg = defaultdict(list)
    for r in csv.DictReader(open(path)):
        agg[(short(r["Kernel_Name"]), int(r["Grid_Size"]), r["Counter_Name"])].append(float(r["Counter_Value"]))
    rows = sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]
    print("| kernel | grid size | counter | dispatches | avg value | min | max |")
    print("|---|---:|---|---:|---:|---:|---:|")
    for (k, g, c), v in rows:
        print(f"| {k} | {g} | {c} | {len(v)} | {sum(v) / len(v):.2f} | {min(v):.2f} | {max(v):.2f} |")


if __name__ == "__main__":
    kind, path = sys.argv[1], sys.argv[2]
    n = int(sys.argv[3]) if len(sys.argv) > 3 else 40
    (trace if kind == "trace" else pmc)(path, n)
