"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m seplib_amd.rocprof_summary stats DIR        # --kernel-trace --stats run
    python -m seplib_amd.rocprof_summary pmc DIR COUNTER  # one --pmc run

Reads the CSV files rocprofv3 writes with ``--output-format csv`` under DIR
(searched recursively) and prints a Markdown table: per kernel, the number
of dispatches, the mean duration (kernel trace) or the mean counter value
per dispatch (counter collection)."""
import csv
import glob
import os
import sys
from collections import defaultdict


def _rows(d, suffix):
    for path in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(path, newline="") as f:
            yield from csv.DictReader(f)


def stats(d, top=25):
    agg = defaultdict(lambda: [0, 0.0])
    for r in _rows(d, "kernel_trace.csv"):
        a = agg[r["Kernel_Name"]]
        a[0] += 1
        a[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    tot = sum(v[1] for v in agg.values()) or 1.0
    print("| kernel | dispatches | mean us | total us | share |")
    print("|---|---:|---:|---:|---:|")
    for name, (cnt, us) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        print("| `%s` | %d | %.2f | %.0f | %.1f%% |" % (name[:90], cnt, us / cnt, us,
                                                     100.0 * us / tot))


def pmc(d, counter):
    agg = defaultdict(lambda: [0, 0.0])
    for r in _rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        a = agg[r["Kernel_Name"]]
        a[0] += 1
        a[1] += float(r["Counter_Value"])
    print("| kernel | dispatches | mean %s per dispatch |" % counter)
    print("|---|---:|---:|")
    for name, (cnt, v) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print("| `%s` | %d | %.1f |" % (name[:90], cnt, v / cnt))


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3])
    else:
        sys.exit(__doc__)
