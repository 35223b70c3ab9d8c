"""Summaries of rocprofv3 output for the engine's kernels.

    python -m cuda_a_star_amd.profiles DIR [--bench BENCH_JSON]

DIR is the output directory of one or more rocprofv3 runs made with
``-f csv``.  The summary lists, per kernel (template arguments kept, so the
step kernel's per-lane instantiations stay apart):

  * from ``*kernel_trace.csv``: dispatches, mean / median / total duration;
  * from ``*counter_collection.csv``: the mean of each counter per dispatch
    (FETCH_SIZE and WRITE_SIZE are in KiB).

With --bench (a file holding the JSON line bench.py printed in the same run)
it adds the bench's own per-lane HIP-event means and algorithmic bytes per
launch, for comparison with the trace and the counters.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import statistics
import sys
from collections import defaultdict


def _rows(directory: str, suffix: str):
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            yield from csv.DictReader(f)


def kernel_durations(directory: str):
    """name -> list of dispatch durations (ns) from the kernel traces."""
    out = defaultdict(list)
    for r in _rows(directory, "kernel_trace.csv"):
        out[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return out


def counter_means(directory: str):
    """name -> {counter: mean value per dispatch} from the counter collections."""
    acc = defaultdict(lambda: defaultdict(list))
    for r in _rows(directory, "counter_collection.csv"):
        acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: statistics.fmean(v) for c, v in d.items()} for k, d in acc.items()}


def short(name: str, width: int = 70) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def summarize(directory: str, bench: dict | None = None, top: int = 12) -> str:
    lines = []
    dur = kernel_durations(directory)
    if dur:
        total = sum(sum(v) for v in dur.values())
        lines.append("| kernel | dispatches | mean us | median us | total ms | share |")
        lines.append("|---|---|---|---|---|---|")
        for name, v in sorted(dur.items(), key=lambda kv: -sum(kv[1]))[:top]:
            lines.append(f"| `{short(name)}` | {len(v)} | {statistics.fmean(v) / 1e3:.2f} | "
                         f"{statistics.median(v) / 1e3:.2f} | {sum(v) / 1e6:.1f} | {sum(v) / total:.1%} |")
    cm = counter_means(directory)
    if cm:
        names = sorted({c for d in cm.values() for c in d})
        lines.append("")
        lines.append("| kernel | " + " | ".join(f"{c} per dispatch" for c in names) + " |")
        lines.append("|---|" + "---|" * len(names))
        for name, d in sorted(cm.items(), key=lambda kv: -max(kv[1].values()))[:top]:
            lines.append(f"| `{short(name)}` | " + " | ".join(f"{d.get(c, float('nan')):.1f}" for c in names) + " |")
    if bench and bench.get("roofline"):
        lines.append("")
        lines.append("bench.py's HIP-event means for the same run (sampled chunks of the timed stream):")
        lines.append("")
        lines.append("| lane | launches | timed | event mean us | algorithmic bytes per launch |")
        lines.append("|---|---|---|---|---|")
        for ln in bench["roofline"]["lanes"]:
            lines.append(f"| {ln['lane']} | {ln['launches']} | {ln['timed_launches']} | {ln['avg_launch_us']} | "
                         f"{ln['algo_bytes_per_launch']:.0f} |")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("directory")
    ap.add_argument("--bench", help="bench.py's JSON line of the same run")
    ap.add_argument("--top", type=int, default=12)
    args = ap.parse_args(argv)
    bench = None
    if args.bench:
        with open(args.bench) as f:
            lines = [l for l in f if l.startswith("{")]
        bench = json.loads(lines[-1]) if lines else None
    print(summarize(args.directory, bench, args.top))
    return 0


if __name__ == "__main__":
    sys.exit(main())
