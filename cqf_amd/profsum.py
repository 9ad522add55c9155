"""Summaries of rocprofv3 CSV output (kernel stats and counter collection).

    python -m cqf_amd.profsum stats  <run>_kernel_stats.csv
    python -m cqf_amd.profsum counters <run>_counter_collection.csv

`stats` prints calls / total / average duration per kernel; `counters`
prints, per kernel and counter, the number of dispatches and the mean value
per dispatch (FETCH_SIZE and WRITE_SIZE are in KiB).
"""
import csv
import sys
from collections import defaultdict


def _short(name, width=60):
    name = name.strip()
    return name if len(name) <= width else name[:width - 3] + "..."


def stats(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | total ms | avg us | % |")
    print("|---|---|---|---|---|")
    for r in rows:
        print(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
              f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")


def counters(path):
    acc = defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(path)):
        k = (r["Kernel_Name"].strip(), r["Counter_Name"].strip())
        acc[k][0] += 1
        acc[k][1] += float(r["Counter_Value"])
    print("| kernel | counter | dispatches | mean per dispatch |")
    print("|---|---|---|---|")
    for (kern, ctr), (n, tot) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"| {_short(kern)} | {ctr} | {n} | {tot / n:.1f} |")


if __name__ == "__main__":
    {"stats": stats, "counters": counters}[sys.argv[1]](sys.argv[2])
