"""Summaries of rocprofv3 output for the engine's kernels.

    python -m cuda_a_star_amd.profiles DIR [--bench BENCH_JSON]
    python -m cuda_a_star_amd.profiles --write OUT PROFILES

The second form writes the committed summaries (PROFILES/*.md and
step_kernel_traffic.json) from one measurement pass whose rocprofv3 runs and
bench lines were written under OUT (the layout is listed in write_summaries).

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
import gzip
import json
import os
import statistics
import sys
from collections import defaultdict


def _rows(directory: str, suffix: str):
    """Rows of the CSVs (plain or gzip-compressed) under `directory` ending in `suffix`."""
    paths = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    plain = set(paths)
    paths += [p for p in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}.gz"), recursive=True))
              if p[:-3] not in plain]
    for path in paths:
        opener = gzip.open if path.endswith(".gz") else open
        with opener(path, "rt", newline="") as f:
            yield from csv.DictReader(f)


def kernel_durations(directory: str):
    """name -> list of dispatch durations (ns) from the kernel traces."""
    out = defaultdict(list)
    for r in _rows(directory, "kernel_trace.csv"):
        out[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return out


def lane_timed_means(directory: str, bench: dict) -> dict:
    """Per launch lane, the mean duration (us) of the step dispatches of the
    timed stream in a kernel trace: each lane launches on a queue of its own,
    identified by the step kernel instantiation most of its dispatches use, and
    the lane's first launches_before_timed dispatches (warm-up) are skipped."""
    queues = defaultdict(list)
    for r in _rows(directory, "kernel_trace.csv"):
        if "step_kernel" in r["Kernel_Name"]:
            queues[(r["Queue_Id"], r["Stream_Id"])].append(
                (int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    out = {}
    for ln in bench["roofline"]["lanes"]:
        sub = LANE_KERNELS.get(ln["lane"])
        if sub is None:
            continue
        best = max(queues.values(), key=lambda v: sum(sub in n for _, _, n in v), default=None)
        if not best or not any(sub in n for _, _, n in best):
            continue
        best.sort()
        timed = best[ln.get("launches_before_timed", 0):][:ln["launches"]]
        if timed:
            out[ln["lane"]] = {"dispatches": len(timed), "mean_us": sum(b - a for a, b, _ in timed) / len(timed) / 1e3}
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


def bench_line(path: str) -> dict:
    """The last JSON line bench.py printed into `path`."""
    with open(path) as f:
        return json.loads([l for l in f if l.startswith("{")][-1])


FLAGS = "--skip-batch --skip-p24 --plain-sample 0 --dropin-sample 0 --skip-cpu --latency-load 0"
# each lane's step kernel: the name prefix that finds it in rocprofv3's output, and the full instantiation
# under the served tuning (arity 4, the kinds it runs, 128 threads, pops per round: 2 on the grid lane)
LANE_KERNELS = {0: "step_kernel<4, 30u", 1: "step_kernel<4, 8u"}
LANE_FULL = {0: "step_kernel<4, 30u, 128, 1>", 1: "step_kernel<4, 8u, 128, 2>"}
LANE_NAMES = {0: f"lane 0 (15-puzzles, grids below 2 M cells): `{LANE_FULL[0]}`",
              1: f"lane 1 (grids of at least 2 M cells, two pops per round): `{LANE_FULL[1]}`"}
WORKLOAD = ("the serving stream (seed 5, 4000 QPS offered): one warm-up segment of 512 requests, then 1024 timed "
            "requests on two launch lanes")


def _pick(cm: dict, sub: str, counter: str) -> float:
    vals = [d[counter] for k, d in cm.items() if sub in k and counter in d]
    if not vals:
        raise KeyError(f"no {counter} for a kernel matching {sub!r}")
    return vals[0]


def write_summaries(out: str, profiles: str) -> dict:
    """Writes the committed summaries from one measurement pass under `out`:

    prof_trace/, bench_trace.json    rocprofv3 --kernel-trace --stats -- bench.py FLAGS
    prof_fetch/, bench_fetch.json    rocprofv3 --pmc FETCH_SIZE -- bench.py FLAGS
    prof_write/, bench_write.json    rocprofv3 --pmc WRITE_SIZE -- bench.py FLAGS
    prof_l2/, bench_l2.json          rocprofv3 --pmc TCC_HIT TCC_MISS -- bench.py FLAGS
    prof_copy_fetch/, prof_copy_write/, prof_copy_l2/
                                     the same counters over bench.py --copy-calibration
    prof_sq/ (optional)              rocprofv3 --pmc SQ_... -- bench.py FLAGS
    bench_default.json               python3 bench.py (no profiler)
    bench_long.json (optional)       python3 bench.py --steps 20 --warmup 5

    Returns the traffic record (also written as step_kernel_traffic.json)."""
    J = lambda *a: os.path.join(out, *a)
    bt, bd = bench_line(J("bench_trace.json")), bench_line(J("bench_default.json"))
    dur = kernel_durations(J("prof_trace"))

    def mean_us(sub):
        v = [x for k, xs in dur.items() if sub in k for x in xs]
        return sum(v) / len(v) / 1e3

    lanes = (0, 1)
    timed = lane_timed_means(J("prof_trace"), bt)
    trace = [f"# Kernel trace of the serving stream (step, path, fill kernels)", "",
             "Commands (graph replay is off under rocprofv3: the engine sees ROCPROF_OUTPUT_PATH; OUT is any directory):",
             "", "```",
             f"rocprofv3 --kernel-trace --stats -f csv -d $OUT/prof_trace -o trace -- python3 bench.py {FLAGS} > $OUT/bench_trace.json",
             "python3 bench.py > $OUT/bench_default.json",
             "python3 -m cuda_a_star_amd.profiles $OUT/prof_trace --bench $OUT/bench_trace.json", "```", "",
             f"Workload: {WORKLOAD}. Every dispatch of the run is in the table.", "",
             summarize(J("prof_trace"), bt), "",
             "Agreement of the HIP-event chunk means with the trace (the timed stream's dispatches: each lane's "
             "own queue, its warm-up dispatches skipped):", "",
             "| lane | trace mean, timed stream (us) | trace mean, every dispatch of the kernel (us) | "
             "event mean, this profiled run (us) | event mean, unprofiled default run (us) |", "|---|---|---|---|---|"]
    for l in lanes:
        tm = timed.get(l)
        trace.append(f"| {LANE_NAMES[l]} | {tm['mean_us']:.2f} ({tm['dispatches']} dispatches) | "
                     f"{mean_us(LANE_KERNELS[l]):.2f} | {bt['roofline']['lanes'][l]['avg_launch_us']} | "
                     f"{bd['roofline']['lanes'][l]['avg_launch_us']} |" if tm else
                     f"| {LANE_NAMES[l]} | - | {mean_us(LANE_KERNELS[l]):.2f} | "
                     f"{bt['roofline']['lanes'][l]['avg_launch_us']} | {bd['roofline']['lanes'][l]['avg_launch_us']} |")
    trace += ["", "The event pair around a timed chunk also spans the gaps between its dispatches, and under the "
              "profiler its own work there, so the profiled run's event means lie above the trace means; the "
              "unprofiled default run's event means are what bench.py reports.", ""]
    with open(os.path.join(profiles, "step_kernel_serving_stream_kernel_trace.md"), "w") as f:
        f.write("\n".join(trace))

    cf, cw = counter_means(J("prof_copy_fetch")), counter_means(J("prof_copy_write"))
    copy_kib = (1 << 30) / 1024
    fs = copy_kib / _pick(cf, "copyBuffer", "FETCH_SIZE")
    ws = copy_kib / _pick(cw, "copyBuffer", "WRITE_SIZE")
    mf, mw = counter_means(J("prof_fetch")), counter_means(J("prof_write"))
    bf, bw = bench_line(J("bench_fetch.json")), bench_line(J("bench_write.json"))
    ratio, rows = {}, []
    for l in lanes:
        fb = _pick(mf, LANE_KERNELS[l], "FETCH_SIZE") * 1024 * fs
        wb = _pick(mw, LANE_KERNELS[l], "WRITE_SIZE") * 1024 * ws
        af = bf["roofline"]["lanes"][l]["algo_bytes_per_launch"]
        aw = bw["roofline"]["lanes"][l]["algo_bytes_per_launch"]
        ratio[str(l)] = fb / af + wb / aw
        rows.append(f"| {LANE_NAMES[l]} | {fb / 1e6:.2f} | {af / 1e6:.2f} | {fb / af:.3f} | {wb / 1e6:.2f} | "
                    f"{aw / 1e6:.2f} | {wb / aw:.3f} | {ratio[str(l)]:.3f} |")
    traffic = ["# HBM traffic of the step kernel on the serving stream", "",
               "Commands (one counter per run; device counters and timing as in the default bench):", "", "```",
               f"rocprofv3 --pmc FETCH_SIZE -f csv -d $OUT/prof_fetch -o fetch -- python3 bench.py {FLAGS} > $OUT/bench_fetch.json",
               f"rocprofv3 --pmc WRITE_SIZE -f csv -d $OUT/prof_write -o write -- python3 bench.py {FLAGS} > $OUT/bench_write.json",
               "rocprofv3 --pmc FETCH_SIZE -f csv -d $OUT/prof_copy_fetch -o copy -- python3 bench.py --copy-calibration",
               "rocprofv3 --pmc WRITE_SIZE -f csv -d $OUT/prof_copy_write -o copy -- python3 bench.py --copy-calibration",
               "python3 -m cuda_a_star_amd.profiles --write $OUT profiles", "```", "",
               f"Copy calibration: a 1 GiB device copy (`__amd_rocclr_copyBuffer`) reads "
               f"{_pick(cf, 'copyBuffer', 'FETCH_SIZE'):.0f} KiB as FETCH_SIZE and "
               f"{_pick(cw, 'copyBuffer', 'WRITE_SIZE'):.0f} KiB as WRITE_SIZE per dispatch, so the counters are "
               f"scaled by x{fs:.3f} (fetch) and x{ws:.3f} (write).", "",
               "Per dispatch of each lane's step kernel (counter runs serialise dispatches, so they hold more "
               "requests per launch than an unprofiled run; the ratio carries over, the per-launch bytes do not):", "",
               "| lane | fetched MB | algorithmic MB (fetch run) | fetch ratio | written MB | algorithmic MB (write run) "
               "| write ratio | HBM bytes per algorithmic byte |", "|---|---|---|---|---|---|---|---|"] + rows + [
               "", "The excess over the algorithmic bytes is whole lines moved for 8- and 16-byte scattered accesses: "
               "each grid expansion touches the popped cell's closed-set line, its neighbours' weight and closed-set "
               "lines (one line per 4 x 4 tile of cells: a neighbour across the tile's edge is another line) and heap "
               "entries; the LDS copy of each working heap's top levels is read and written back every launch; idle "
               "heaps clear their candidate slices. None of that is counted as algorithmic.", ""]
    with open(os.path.join(profiles, "step_kernel_serving_stream_hbm_traffic.md"), "w") as f:
        f.write("\n".join(traffic))

    rec = {"kernels": {str(l): LANE_FULL[l] for l in lanes}, "workload": WORKLOAD,
           "command": f"rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE|TCC_HIT TCC_MISS -f csv -- python3 bench.py {FLAGS}",
           "fetch_scale": fs, "write_scale": ws, "hbm_bytes_per_algo_byte": ratio}
    if os.path.isdir(J("prof_l2")):
        ml, cl = counter_means(J("prof_l2")), counter_means(J("prof_copy_l2"))
        bl = bench_line(J("bench_l2.json"))
        per_load = _pick(cl, "atomic_rate_kernel", "TCC_MISS") / (1 << 26)
        mis, rows = {}, []
        for l in lanes:
            m = _pick(ml, LANE_KERNELS[l], "TCC_MISS")
            h = _pick(ml, LANE_KERNELS[l], "TCC_HIT")
            a = bl["roofline"]["lanes"][l]["algo_bytes_per_launch"]
            mis[str(l)] = m / a
            rows.append(f"| {LANE_NAMES[l]} | {h / 1e6:.3f} | {m / 1e6:.3f} | {h / (h + m):.1%} | {a / 1e6:.2f} | "
                        f"{m / a * 1e3:.3f} |")
        rec["l2_misses_per_algo_byte"] = mis
        rec["l2_misses_per_scattered_load"] = per_load
        l2 = ["# L2 (TCC) hits and misses of the step kernel on the serving stream", "", "Commands:", "", "```",
              f"rocprofv3 --pmc TCC_HIT TCC_MISS -f csv -d $OUT/prof_l2 -o l2 -- python3 bench.py {FLAGS} > $OUT/bench_l2.json",
              "rocprofv3 --pmc TCC_HIT TCC_MISS -f csv -d $OUT/prof_copy_l2 -o copy -- python3 bench.py --copy-calibration",
              "python3 -m cuda_a_star_amd.profiles --write $OUT profiles", "```", "",
              f"Calibration: the scattered-load probe of `bench.py --copy-calibration` (`atomic_rate_kernel`, 2^26 "
              f"random 64-bit accesses over 1 GiB per dispatch) counts {per_load:.3f} TCC_MISS per access, so a "
              "TCC miss stands for one scattered line served by the memory side.", "",
              "Per dispatch of each lane's step kernel:", "",
              "| lane | TCC_HIT (M) | TCC_MISS (M) | hit rate | algorithmic MB (this run) | misses per algorithmic KB |",
              "|---|---|---|---|---|---|"] + rows + [
              "", "bench.py scales the misses per algorithmic byte by its own run's algorithmic bytes per launch and "
              "divides by the launch time (`roofline.scattered`), against the device's rate for scattered 64-bit "
              "loads measured in the same run (`ga_atomic_rate` kind 2).", ""]
        with open(os.path.join(profiles, "step_kernel_serving_stream_l2_misses.md"), "w") as f:
            f.write("\n".join(l2))
    if os.path.isdir(J("prof_sq")):
        sq = ["# Shader-sequencer counters of the step kernel on the serving stream", "", "Command:", "", "```",
              "rocprofv3 --pmc SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_INST_ANY SQ_INSTS_VALU SQ_INSTS_SALU "
              f"SQ_INSTS_VMEM_RD SQ_INSTS_LDS -f csv -d $OUT/prof_sq -o sq -- python3 bench.py {FLAGS}",
              "python3 -m cuda_a_star_amd.profiles $OUT/prof_sq", "```", "", summarize(J("prof_sq"), None, 6), "",
              "Per wave of the step kernels:", "", "| kernel | SALU | VALU | LDS | VMEM reads | wave cycles |",
              "|---|---|---|---|---|---|"]
        for name, d in sorted(counter_means(J("prof_sq")).items()):
            if "step_kernel" not in name or not d.get("SQ_WAVES"):
                continue
            w = d["SQ_WAVES"]
            sq.append(f"| `{short(name)}` | {d['SQ_INSTS_SALU'] / w:.0f} | {d['SQ_INSTS_VALU'] / w:.0f} | "
                      f"{d['SQ_INSTS_LDS'] / w:.0f} | {d['SQ_INSTS_VMEM_RD'] / w:.1f} | {d['SQ_WAVE_CYCLES'] / w:.0f} |")
        sq.append("")
        with open(os.path.join(profiles, "step_kernel_serving_stream_sq.md"), "w") as f:
            f.write("\n".join(sq))
    with open(os.path.join(profiles, "step_kernel_traffic.json"), "w") as f:
        json.dump(rec, f, indent=1)
    with open(os.path.join(profiles, "bench_serving_default.json"), "w") as f:
        json.dump({"command": "python3 bench.py", "line": bd}, f, indent=1)
    if os.path.exists(J("bench_long.json")):
        with open(os.path.join(profiles, "bench_serving_20steps.json"), "w") as f:
            json.dump({"command": "python3 bench.py --steps 20 --warmup 5", "line": bench_line(J("bench_long.json"))},
                      f, indent=1)
    return rec


def main(argv=None) -> int:
    if argv is None:
        argv = sys.argv[1:]
    if argv[:1] == ["--write"]:
        if len(argv) != 3:
            print("usage: python -m cuda_a_star_amd.profiles --write OUT PROFILES", file=sys.stderr)
            return 2
        print(json.dumps(write_summaries(argv[1], argv[2]), indent=1))
        return 0
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
