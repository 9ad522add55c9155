"""Summaries of rocprofv3 CSV output (kernel stats and counter collection).

    python -m cqf_amd.profsum stats  <run>_kernel_stats.csv
    python -m cqf_amd.profsum counters <run>_counter_collection.csv [--by-grid] [--first N]
    python -m cqf_amd.profsum trace <run>_kernel_trace.csv

`stats` prints calls / total / average duration per kernel; `counters`
prints, per kernel and counter, the number of dispatches and the mean value
per dispatch (FETCH_SIZE and WRITE_SIZE are in KiB); `trace` and
`counters --by-grid` split each kernel by grid size, which separates
launches of one kernel over different amounts of work.
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


def counters(path, by_grid=False, first=0):
    """`first` > 0: only each kernel's first `first` dispatches (in dispatch
    order), e.g. the insert launches of a run whose deletes follow."""
    acc = defaultdict(lambda: [0, 0.0])
    seen = defaultdict(list)  # kernel -> dispatch ids in order
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    for r in rows:
        kern = r["Kernel_Name"].strip()
        d = int(r["Dispatch_Id"])
        if d not in seen[kern]:
            seen[kern].append(d)
        if first and len(seen[kern]) > first:
            continue
        k = (kern, r["Counter_Name"].strip(), r["Grid_Size"] if by_grid else "")
        acc[k][0] += 1
        acc[k][1] += float(r["Counter_Value"])
    print("| kernel | grid | counter | dispatches | mean per dispatch |")
    print("|---|---|---|---|---|")
    for (kern, ctr, grid), (n, tot) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"| {_short(kern)} | {grid or '-'} | {ctr} | {n} | {tot / n:.1f} |")


def trace(path):
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        acc[(r["Kernel_Name"].strip(), r["Grid_Size_X"])].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    print("| kernel | grid | calls | avg us | min us | max us |")
    print("|---|---|---|---|---|---|")
    for (kern, grid), d in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        print(f"| {_short(kern)} | {grid} | {len(d)} | {sum(d) / len(d):.2f} | {min(d):.2f} | {max(d):.2f} |")


if __name__ == "__main__":
    cmd, path = sys.argv[1], sys.argv[2]
    if cmd == "counters":
        rest = sys.argv[3:]
        first = int(rest[rest.index("--first") + 1]) if "--first" in rest else 0
        counters(path, by_grid="--by-grid" in rest, first=first)
    else:
        {"stats": stats, "trace": trace}[cmd](path)
