#!/usr/bin/env python3
"""Served GA* vs one-at-a-time plain GA* on one MI355X.

Workload: the serving stream of the benchmark configurations -- a seeded
Poisson arrival stream at a fixed offered QPS; 60 % 15-puzzles (walk depth
U{30..70}), 40 % grids (side U{512..4096}, weights 1..9, 20 % walls); every
grid resident in HBM before the timed region; admission budget 200 GB of heap
and closed-set bytes.

A "step" is one stream segment of --requests requests.  The W warm-up
segments are served first (untimed); then the K timed segments are served as
one continuous arrival stream with continuous batching (ga_serve_run), so no
segment drains before the next one arrives.  The offered load (--qps) lies
above what the engine serves, so the timed stream measures the saturation
throughput:  value = completed solves / wall seconds of the timed stream.

Measured outside the timed region:
  latency    the same requests served again with their arrivals spread to
             --latency-load x the measured saturation rate (below capacity):
             p50/p99 of arrival -> result on the host at that offered load
  batched    the 1024 mixed instances of the batch configuration, one call
             (median of --batch-reps calls), on the launch lanes and in the
             one-launch form (ga_config.persist)
  24-puzzle  256 random-walk 5x5 boards, one call
  one at a time  --plain-sample requests of the timed stream, plain GA*
             (ga_solve) back to back on the one-at-a-time tuning
  CPU        the oracle (sequential A*) built with -march=native on this
             host: throughput on pinned threads, latency on one pinned core
With --host-cells the grids stay in host memory and are copied to the device
at admission, inside the timed region (the PCIe-inclusive rate; not the
default).
The step kernel's roofline comes from sampled chunks of each launch lane:
every --time-every-th chunk is timed (its launches issued one by one between
one pair of HIP events on the lane's stream, device counters off), and the
chunks half-way between are counted (device counters on: the algorithmic
bytes, scaled to every launch).  The dominant kernel is the lane whose
launches take the most kernel time.
Its HBM traffic is this run's algorithmic bytes per launch times the
HBM-bytes-per-algorithmic-byte ratio of the committed rocprofv3 counter runs
(profiles/step_kernel_traffic.json), and its scattered accesses (L2 misses
per launch, by the same scaling from the committed TCC counter run) are set
against the device's rate for scattered 64-bit loads, measured in this run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MAX_ITERATIONS = 1_000_000    # per request on the GPU
MAX_ORACLE_STATES = 12_000_000  # per CPU request; larger puzzles are skipped (and counted)
METRIC = "optimal solves/sec + p50/p99 latency: served batch vs one-at-a-time GA*"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return 0.0
    k = min(len(xs) - 1, max(0, int(-(-p * len(xs) // 1)) - 1))
    return xs[k]


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it (a virtual machine may give only the vendor's line), with its
    family / model numbers, clock, vector width and the cores this process may use."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # the first processor's block
                k, _, v = line.partition(":")
                info[k.strip()] = v.strip()
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    extra = []
    if "cpu family" in info and "model" in info:
        extra.append(f"family {info['cpu family']} model {info['model']}")
    if "cpu MHz" in info:
        extra.append(f"{float(info['cpu MHz']) / 1000:.2f} GHz")
    flags = info.get("flags", "").split()
    extra.append("AVX-512" if "avx512f" in flags else "AVX2" if "avx2" in flags else "no AVX2")
    extra.append(f"{len(host_cores(1 << 20))} of {os.cpu_count()} logical CPUs usable")
    return f"{name} ({', '.join(extra)})"


def host_cores(limit: int):
    """The cores this process may run on (at most `limit`), for pinned workers."""
    try:
        cores = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cores = list(range(os.cpu_count() or 1))
    return cores[:max(1, limit)]


def run_pinned(cores, n_items: int, seconds: float, solve):
    """Workers pinned one per core take items 0, 1, ... in order until `seconds`
    have passed; solve(i) -> (done: bool, result).  Returns (per-item wall
    seconds of the items done, results by item, elapsed seconds)."""
    lock = threading.Lock()
    state = {"next": 0}
    times, results = {}, {}
    deadline = time.perf_counter() + seconds
    t0 = time.perf_counter()

    def worker(core):
        try:
            os.sched_setaffinity(0, {core})  # Linux: pins the calling thread
        except (AttributeError, OSError):
            pass
        while True:
            with lock:
                i = state["next"]
                if i >= n_items or time.perf_counter() > deadline:
                    return
                state["next"] += 1
            t = time.perf_counter()
            done, res = solve(i)
            dt = time.perf_counter() - t
            with lock:
                results[i] = res
                if done:
                    times[i] = dt

    ts = [threading.Thread(target=worker, args=(c,)) for c in cores]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    return times, results, time.perf_counter() - t0


P24_N, P24_DEPTH = 256, (30, 60)


def puzzle24_leg(e, heaps: int, cores, cpu_seconds: float):
    """256 seeded 24-puzzles (walks of depth U{30..60} from the ordered board, no
    immediate backtrack) in one ga_solve_batch; the CPU oracle (sequential A*
    on boards with the same pattern databases) takes the same instances in
    order on pinned workers for `cpu_seconds`, and every instance it finishes
    checks the GPU cost.  Then 24 walks of depth U{80..120} in one batch."""
    import cuda_a_star_amd as ga
    from cuda_a_star_amd import workloads as wl

    st = wl.Stream(24)
    boards = [ga.gen_board(5, st.uniform_int(*P24_DEPTH), wl.instance_seed(24, i)) for i in range(P24_N)]
    reqs = [ga.puzzle5_request(b, num_queues=heaps) for b in boards]
    e.solve_batch(reqs[:16])  # warm-up
    t = time.perf_counter()
    sols = e.solve_batch(reqs)
    gpu_t = time.perf_counter() - t
    ok = sum(s.status == ga.OK for s in sols)
    out = {"solves_per_sec": round(ok / gpu_t, 3), "instances": P24_N, "ok": ok, "seconds": round(gpu_t, 3),
           "workload": f"24-puzzle walks of depth U{{{P24_DEPTH[0]}..{P24_DEPTH[1]}}} from the ordered board, "
                       f"{heaps} heaps each, one batch"}
    # walks of depth U{80..120}: out of reach of Manhattan; the pattern databases (ga_config.pdb24,
    # built on the device at the first such request) bring them within reach
    deep = wl.puzzle24_deep(25, 24)
    d_reqs = [ga.puzzle5_request(b, num_queues=heaps) for _, _, b in deep]
    t = time.perf_counter()
    d_sols = e.solve_batch(d_reqs)
    d_t = time.perf_counter() - t
    d_ok = sum(s.status == ga.OK for s in d_sols)
    out["deep_pdb"] = {"solves_per_sec": round(d_ok / d_t, 3), "instances": len(deep), "ok": d_ok,
                       "seconds": round(d_t, 3),
                       "expanded_mean": round(sum(s.expanded for s in d_sols) / len(d_sols), 1),
                       "workload": "24-puzzle walks of depth U{80..120}, additive pattern databases (5-5-5-5-4), "
                                   f"{heaps} heaps each, one batch; costs checked in the GPU tests against "
                                   "the pattern-database oracle"}
    d_status = {}
    for s in d_sols:
        d_status[ga.status_str(s.status)] = d_status.get(ga.status_str(s.status), 0) + 1
    out["deep_pdb"]["statuses"] = d_status
    out["deep_pdb"]["grows_max"] = max(s.grows for s in d_sols)
    if cpu_seconds <= 0:
        return out
    import oracle

    oracle.pdb24()  # the oracle's own tables (built once, outside the timed window)

    def solve(i):  # the same heuristic as the GPU: A* with the pattern databases
        rc, res, _ = oracle.astar_board_pdb(boards[i], max_states=MAX_ORACLE_STATES, want_path=False)
        if rc == oracle.E_LIMIT:
            return False, None
        return True, res.cost if rc == oracle.OK else -1

    times, results, el = run_pinned(cores, P24_N, cpu_seconds, solve)
    match = sum(1 for i in times if sols[i].status == ga.OK and sols[i].cost == results[i])
    capped = sum(1 for i in results if i not in times)
    out["cpu_oracle"] = {"solves_per_sec": round(len(times) / el, 3), "cores": len(cores), "kind": "port",
                         "heuristic": "additive pattern databases (the GPU's)",
                         "sample": f"first {len(results)} instances in order on {len(cores)} pinned threads, "
                                   f"{el:.1f}s; {capped} above {MAX_ORACLE_STATES} stored states skipped"}
    out["gpu_cost_matches"] = f"{match}/{len(times)}"
    return out


def cpu_baseline(insts, seconds: float, latency_seconds: float, cores, gpu_costs: dict, build: str):
    """The oracle (sequential A*; `build`: "native" when oracle.use_native()
    compiled it for this host) over a bounded sample of the timed stream, in
    stream order:
      throughput  one worker pinned per core on `cores` for `seconds`
      latency     one worker pinned to the first core for `latency_seconds`:
                  wall time of each request's search
    Every request solved checks the GPU's cost for it."""
    import oracle

    import cuda_a_star_amd as ga

    ordered = ga.puzzle_goal(4)
    if any(x.kind == "puzzle" and x.goal == ordered for x in insts):
        oracle.pdb15()  # the tables, built before the timed region as the GPU builds its own before the stream

    def solve(i):
        x = insts[i]
        if x.kind == "puzzle":
            # the same heuristic as the GPU: the 15-puzzle pattern databases towards the ordered goal
            if x.goal == ordered:
                rc, res, _ = oracle.astar_puzzle_pdb(x.start, max_states=MAX_ORACLE_STATES, want_path=False)
            else:
                rc, res, _ = oracle.astar_puzzle(4, x.start, x.goal, max_states=MAX_ORACLE_STATES, want_path=False)
            if rc == oracle.E_LIMIT:
                return False, None
        else:
            rc, res, _ = oracle.astar_grid(x.cells, x.side, x.side, x.start, x.goal, 4, want_path=False)
        return True, res.cost if rc == oracle.OK else -1

    times, results, el = run_pinned(cores, len(insts), seconds, solve)
    checked = [i for i in times if i in gpu_costs]
    match = sum(1 for i in checked if gpu_costs[i] == results[i])
    capped = sum(1 for i in results if i not in times)
    lat_times, _, lat_el = run_pinned(cores[:1], len(insts), latency_seconds, solve)
    lat_ms = [v * 1e3 for v in lat_times.values()]
    return {
        "value": round(len(times) / el, 3), "unit": "solves/s", "cores": len(cores), "kind": "port",
        "cpu_model": cpu_model(), "build": f"oracle/oracle.c, gcc -O3 {'-march=native' if build == 'native' else '(portable)'}",
        "sample": f"first {len(results)} requests of the timed stream in stream order, sequential A* (binary heap, "
                  f"lazy deletion) per request, 15-puzzles with the GPU's pattern databases, grids with "
                  f"Manhattan x w_min, one thread pinned per core on {len(cores)} cores, {el:.1f}s; "
                  f"{capped} puzzles above {MAX_ORACLE_STATES} stored states skipped",
        "latency_mode": {"cores": 1, "requests": len(lat_ms), "seconds": round(lat_el, 1),
                         "p50_ms": round(pct(lat_ms, 0.5), 3), "p99_ms": round(pct(lat_ms, 0.99), 3),
                         "solves_per_sec": round(len(lat_ms) / lat_el, 3) if lat_el else None},
        "gpu_cost_matches": f"{match}/{len(checked)}",
    }


def lane_roofline(st, cfg, elapsed_s):
    """Roofline line of the dominant step-kernel lane (most kernel time) from
    the sampled launches of the timed stream."""
    n = cfg.lanes + 1
    lanes = []
    for l in range(n):
        if not st.lane_timed_launches[l] or not st.lane_counted_launches[l]:
            continue
        avg_ms = st.lane_step_ms[l] / st.lane_timed_launches[l]
        lanes.append({
            "lane": l, "launches": st.lane_launches[l], "timed_launches": st.lane_timed_launches[l],
            "counted_launches": st.lane_counted_launches[l],
            "avg_launch_us": round(avg_ms * 1e3, 3),
            "kernel_ms_est": round(avg_ms * st.lane_launches[l], 1),
            "algo_bytes_per_launch": round(st.lane_algo_bytes[l] / st.lane_launches[l], 1),
            "achieved_GBps": round(st.lane_algo_bytes[l] / st.lane_launches[l] / (avg_ms * 1e-3) / 1e9, 2),
            "atomics_per_launch": round(st.lane_atomics[l] / st.lane_counted_launches[l], 1),
            # the lane's chunks on its stream (one event pair around every chunk) over the timed stream's wall time
            "busy_frac": round(st.lane_busy_ms[l] / (elapsed_s * 1e3), 4),

        })
    if not lanes:
        return None
    dom = max(lanes, key=lambda x: x["kernel_ms_est"])
    kernel = {0: "step_kernel<4, all kinds> (lane 0: 15-puzzles and grids below deep_cells)",
              1: "step_kernel<4, 4-connected grids" + (", two pops per round" if cfg.step_flags & 4 else "")
                 + "> (lane 1: grids of at least deep_cells cells)"}
    rf = {"bound": "hbm", "achieved": dom["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
          "frac": round(dom["achieved_GBps"] / HBM_PEAK_GBS, 5), "traffic": None,
          "kernel": kernel.get(dom["lane"], f"step kernel of lane {dom['lane']}"), "lane": dom["lane"],
          "avg_launch_us": dom["avg_launch_us"], "algo_bytes_per_launch": dom["algo_bytes_per_launch"],
          "event_timing": "one HIP event pair around each timed chunk of step launches (issued one by one, "
                          "device counters off), on the lane's stream; time / launches",
          "lanes": lanes,
          "expansions_per_launch": round(st.expanded / st.step_launches, 1),
          # heap-group iterations with work / launched (every mapped heap, every sampled launch)
          "working_group_fraction": round(st.work_groups / st.heap_groups, 4) if st.heap_groups else None,
          "sift_down_hbm_fraction": round(st.sift_down_hbm / st.sift_down, 4) if st.sift_down else None,
          # all algorithmic bytes of the timed stream over its wall time (lanes run concurrently)
          "achieved_device_GBps": round(st.algo_bytes / elapsed_s / 1e9, 2),
          # the run loop's host time by phase, as fractions of the timed stream's wall time
          "host_frac": {k: round(getattr(st, f"host_{k}_ms") / (elapsed_s * 1e3), 4)
                        for k in ("admit", "snapshot", "launch", "path")}}
    tpath = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")
    if os.path.exists(tpath):  # HBM bytes per algorithmic byte from the committed rocprofv3 --pmc runs
        with open(tpath) as f:
            tj = json.load(f)
        ratio = tj.get("hbm_bytes_per_algo_byte", {})
        r = ratio.get(str(dom["lane"])) if isinstance(ratio, dict) else None
        if r:
            rf["traffic"] = round(r * dom["algo_bytes_per_launch"], 1)
            rf["traffic_source"] = (f"profiles/step_kernel_traffic.json: {r:.3f} HBM bytes per algorithmic byte "
                                    f"(lane {dom['lane']}) x this run's algorithmic bytes per launch")
        m = tj.get("l2_misses_per_algo_byte", {}).get(str(dom["lane"]))
        if m:  # scattered lines served by the memory side per launch (L2 misses, from the TCC counter run)
            misses = m * dom["algo_bytes_per_launch"]
            rf["scattered"] = {"l2_misses_per_launch": round(misses, 1),
                               "achieved_G_per_s": round(misses / (dom["avg_launch_us"] * 1e-6) / 1e9, 3),
                               "source": f"profiles/step_kernel_traffic.json: {m * 1e3:.4f} L2 misses per algorithmic "
                                         f"KB (lane {dom['lane']}) x this run's algorithmic bytes per launch"}
    # the dominant lane's global atomics (counted launches): one CAS per closed-set probe (hashed sets),
    # one atomicMin per lookup
    rf["atomics_per_launch"] = dom["atomics_per_launch"]
    rf["atomic_rate_G_per_s"] = round(dom["atomics_per_launch"] / (dom["avg_launch_us"] * 1e-6) / 1e9, 3)
    return rf


def dropin_leg(idx, insts, served_costs):
    """The header drop-in's own loop (include/gastar/a_star.h: one instance per
    a_star<>() call, as the reference's a_star.h:223-282): gastar_cli solves
    the 15-puzzles idx one after another (--each-pair 1) with the reference's
    4096 queues and reports each search's time as the reference's CLI does
    (main.cu:26-29), in microseconds.  Rate = puzzles / summed search time;
    costs are checked against the served ones."""
    import subprocess
    import tempfile

    import cuda_a_star_amd as ga

    cli = os.path.join(ROOT, "cuda_a_star_amd", "gastar_cli")

    def tiles(state):
        return ",".join(str(t) if t else "_" for t in ga.state_to_board(state, 4))

    with tempfile.TemporaryDirectory() as tmp:
        src, dst = os.path.join(tmp, "in.txt"), os.path.join(tmp, "out.txt")
        with open(src, "w") as f:
            for i in idx:
                f.write(f"{tiles(insts[i].start)} {tiles(insts[i].goal)}\n")
        t = time.perf_counter()
        r = subprocess.run([cli, "--version", "sliding", "--size", "4", "--num-queues", "4096", "--each-pair", "1",
                            "--time-unit", "us", "--input-data", src, "--output-data", dst],
                           capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t
        if r.returncode != 0:
            raise RuntimeError(f"gastar_cli failed ({r.returncode}): {r.stdout[-500:]} {r.stderr[-500:]}")
        times, costs = [], []
        for line in open(dst).read().split():
            if line.isdigit():  # a result's elapsed time starts it; its states follow, one per line
                times.append(int(line))
                costs.append(-1)
            else:
                costs[-1] += 1
    if len(times) != len(idx):
        raise RuntimeError(f"gastar_cli answered {len(times)} of {len(idx)} puzzles")
    mism = sum(1 for i, c in zip(idx, costs) if served_costs.get(i) != c)
    search_s = sum(times) / 1e6
    return {"solves_per_sec": round(len(idx) / search_s, 3), "puzzles": len(idx), "search_s": round(search_s, 3),
            "process_wall_s": round(wall, 3), "cost_mismatches_vs_served": mism,
            "workload": "the first 15-puzzles of the timed stream (walk depth 30..70), 4096 heaps (the reference's "
                        "default), one a_star<>() call each through gastar_cli; rate from the summed search times"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=512, help="requests per step (stream segment)")
    ap.add_argument("--qps", type=float, default=4000.0,
                    help="offered load of the timed stream (Poisson arrivals); above capacity, so that the timed "
                         "stream measures the saturation throughput")
    ap.add_argument("--latency-load", type=float, default=0.8,
                    help="offered load of the latency leg, as a fraction of the measured saturation rate (0 = skip)")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--plain-sample", type=int, default=200, help="requests solved one at a time")
    ap.add_argument("--dropin-sample", type=int, default=100,
                    help="15-puzzles of the stream solved one at a time through the header drop-in (gastar_cli)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-latency-seconds", type=float, default=5.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores available)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--no-profile", action="store_true", help="skip per-step HIP events and counters")
    ap.add_argument("--time-every", type=int, default=8,
                    help="sample every n-th chunk: HIP events and device counters in its step launches")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true", help="skip the batched (1024-instance) leg")
    ap.add_argument("--batch-reps", type=int, default=5, help="timed calls per batched form (the median is reported)")
    ap.add_argument("--skip-p24", action="store_true", help="skip the 24-puzzle leg")
    ap.add_argument("--p24-heaps", type=int, default=8192,
                    help="heaps per 24-puzzle (their own launch lane; 0 = the served 15-puzzle count)")
    ap.add_argument("--override", default="", help="served tuning overrides, e.g. lanes=3,rounds=2")
    ap.add_argument("--host-cells", action="store_true",
                    help="pass grids in host memory (copied over PCIe at admission) instead of resident in HBM")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy and the scattered-load probe (for scaling counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        import cuda_a_star_amd as ga

        with ga.Engine(0) as e:  # 2^26 scattered 64-bit loads over 1 GiB per dispatch
            rate = e.atomic_rate(1 << 30, 1 << 26, 2)
        print(json.dumps({"copy_bytes": 1 << 30, "copies": 5, "scattered_loads": 1 << 26,
                          "scattered_load_G_per_s": round(rate / 1e9, 2)}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances are independent: a multi-GPU run would shard the request stream)", file=sys.stderr)
        return 2

    import cuda_a_star_amd as ga
    from cuda_a_star_amd import tuning
    from cuda_a_star_amd import workloads as wl

    M, K, W = args.requests, args.steps, args.warmup
    n_total = (W + K) * M
    t = time.perf_counter()
    insts = wl.config5(args.seed, n_total, args.qps, threads=16)
    log(f"[bench] generated {n_total} requests in {time.perf_counter() - t:.1f}s")

    served_cfg = tuning.load()
    for kv in filter(None, args.override.split(",")):
        k, v = kv.split("=")
        setattr(served_cfg, k, int(v))
    under_profiler = bool(os.environ.get("ROCPROF_OUTPUT_PATH"))  # the engine then launches without graphs
    # The iteration limit only bounds a pathological request (it would count as failed); the longest
    # searches of this stream take about 6,000 iterations.
    e = ga.Engine(0, budget_bytes=int(args.budget_gb * (1 << 30)), chunk_bytes=8 << 30,
                  profile=0 if args.no_profile else max(1, args.time_every), max_iterations=MAX_ITERATIONS,
                  **served_cfg.engine)
    dev = {}
    t = time.perf_counter()
    for i, x in enumerate(insts):
        if x.kind == "grid" and not args.host_cells:
            dev[i] = e.upload(x.cells)
    log(f"[bench] uploaded {len(dev)} grids in {time.perf_counter() - t:.1f}s"
        + (" (--host-cells: grids are copied from host memory at admission, inside the timed region)"
           if args.host_cells else ""))

    def served_request(i):
        x = insts[i]
        nq = served_cfg.heaps(x.kind, x.side)
        return x.request(cells_device=dev.get(i), num_queues=nq)

    def segment(first, count):
        idx = list(range(first, first + count))
        t0 = insts[idx[0]].arrival_ns
        return idx, [served_request(i) for i in idx], [insts[i].arrival_ns - t0 for i in idx]

    # ---- warm-up (untimed): the first W*M requests of the stream ----
    if W:
        idx, reqs, arr = segment(0, W * M)
        sols, ss = e.serve(reqs, arr)
        log(f"[bench] warmup: {ss.completed}/{len(reqs)} in {ss.wall_ns / 1e9:.2f}s "
            f"({ss.solves_per_sec:.1f}/s, p50 {ss.p50_ms:.1f} ms, p99 {ss.p99_ms:.1f} ms)")

    # ---- timed: K steps of M requests, served as one continuous arrival stream ----
    st_before = e.stats()
    launches_before = st_before.step_launches  # step dispatches of this process before the timed segment
    lane_before = list(st_before.lane_launches)
    e.reset_stats()
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for the barrier
        sync = lambda: None
    idx, reqs, arr = segment(W * M, K * M)
    sync()
    t_start = time.perf_counter()
    sols, ss = e.serve(reqs, arr)  # raises GAStarError if the run fails
    sync()
    t_total = time.perf_counter() - t_start
    st = e.stats()
    ok = sum(1 for s in sols if s.status == ga.OK)
    failed = len(sols) - ok
    served_costs = {i: s.cost for i, s in zip(idx, sols) if s.status == ga.OK}
    sat_lat = [s.latency_ns / 1e6 for s in sols if s.status == ga.OK]
    served_sps = ok / t_total
    log(f"[bench] timed (saturation, {args.qps:.0f} QPS offered): {ok}/{len(reqs)} in {t_total:.2f}s = "
        f"{served_sps:.1f} solves/s, p50 {pct(sat_lat, 0.5):.1f} ms, p99 {pct(sat_lat, 0.99):.1f} ms, "
        f"inflight max {ss.max_inflight}")
    roofline = None
    if not args.no_profile and st.step_launches and st.timed_launches and st.counted_launches:
        roofline = lane_roofline(st, served_cfg, t_total)
        roofline["launches_before_timed"] = launches_before
        for ln in roofline["lanes"]:  # per lane, so that a kernel trace can pick out the timed dispatches
            ln["launches_before_timed"] = lane_before[ln["lane"]]
        peak_min = e.atomic_rate(1 << 30, 1 << 26, 0)
        peak_cas = e.atomic_rate(1 << 30, 1 << 26, 1)
        roofline["atomic_peak_G_per_s"] = {"fetch_min": round(peak_min / 1e9, 2), "cas": round(peak_cas / 1e9, 2)}
        # the dominant lane's atomics at the lower of the two random-atomic rates, per launch time
        roofline["atomic_frac"] = round(roofline["atomic_rate_G_per_s"] * 1e9 / min(peak_min, peak_cas), 4)
        # scattered 64-bit loads over 1 GiB: the rate at which the memory system serves random lines
        peak_load = e.atomic_rate(1 << 30, 1 << 26, 2)
        roofline["scattered_load_peak_G_per_s"] = round(peak_load / 1e9, 2)
        if "scattered" in roofline:
            sc = roofline["scattered"]
            sc["peak_G_per_s"] = round(peak_load / 1e9, 2)
            # (a lower bound on the memory side's load: some of the misses are atomics, whose rate is lower;
            # atomics on neighbouring cells share lines, so they are not priced one by one)
            sc["frac"] = round(sc["achieved_G_per_s"] * 1e9 / peak_load, 4)
    log(f"[bench] roofline {roofline}")

    # ---- latency leg: the same requests at a stated offered load below saturation ----
    latency = None
    if args.latency_load > 0 and served_sps > 0:
        lam = args.latency_load * served_sps
        scale = args.qps / lam  # arrival gaps of the stream drawn at --qps, stretched to lam
        arr2 = [int(a * scale) for a in arr]
        sols2, ss2 = e.serve(reqs, arr2)
        lat2 = [s.latency_ns / 1e6 for s in sols2 if s.status == ga.OK]
        by_kind = {}
        for kind in ("puzzle", "grid"):
            kl = [s.latency_ns / 1e6 for i, s in zip(idx, sols2) if insts[i].kind == kind and s.status == ga.OK]
            if kl:
                by_kind[kind] = {"n": len(kl), "p50_ms": round(pct(kl, 0.5), 2), "p99_ms": round(pct(kl, 0.99), 2)}
        mism = sum(1 for i, s in zip(idx, sols2) if s.status != ga.OK or served_costs.get(i) != s.cost)
        latency = {"offered_qps": round(lam, 1), "load_vs_saturation": args.latency_load,
                   "completed": len(lat2), "solves_per_sec": round(ss2.solves_per_sec, 3),
                   "p50_ms": round(pct(lat2, 0.5), 3), "p99_ms": round(pct(lat2, 0.99), 3),
                   "latency_by_kind": by_kind, "max_inflight": ss2.max_inflight,
                   "cost_mismatches_vs_saturation_leg": mism}
        log(f"[bench] latency leg at {lam:.0f} QPS offered: {latency}")

    # ---- batched: the 1024 mixed instances of the batch configuration in one ga_solve_batch ----
    # twice over: on the launch lanes (one launch per iteration, chunks replayed from graphs) and in the
    # one-launch form (every iteration of an instance inside one launch, one workgroup per instance);
    # each the median of --batch-reps calls
    batch = None
    if not args.skip_batch:
        b_insts = wl.config4(4, 1024, threads=16)
        b_work = ("1024 mixed: even = 15-puzzle walk depth 20..60, odd = grid side {256,512,1024,2048}, "
                  "walls 10..25%")

        def batch_leg(eng, label):
            dev = {i: eng.upload(x.cells) for i, x in enumerate(b_insts) if x.kind == "grid"}
            reqs = [x.request(cells_device=dev.get(i), num_queues=served_cfg.heaps(x.kind, x.side))
                    for i, x in enumerate(b_insts)]
            eng.solve_batch(reqs[:64])  # warm-up
            eng.reset_stats()
            times, sols = [], None
            for _ in range(max(1, args.batch_reps)):
                t = time.perf_counter()
                sols = eng.solve_batch(reqs)
                times.append(time.perf_counter() - t)
            for a in dev.values():
                eng.free_device(a)
            times.sort()
            b_t = times[len(times) // 2]
            ok = sum(1 for s_ in sols if s_.status == ga.OK)
            log(f"[bench] batched ({label}): {ok}/{len(reqs)} in {b_t:.3f}s (median of {len(times)}) = "
                f"{ok / b_t:.1f} solves/s; all: {[round(x, 3) for x in times]}")
            return {"solves_per_sec": round(ok / b_t, 3), "instances": len(reqs), "ok": ok,
                    "seconds": round(b_t, 3), "runs_s": [round(x, 3) for x in times]}, sols

        batch, lane_sols = batch_leg(e, "launch lanes")
        batch["workload"] = b_work
        with ga.Engine(0, budget_bytes=100 << 30, chunk_bytes=8 << 30, max_iterations=MAX_ITERATIONS, persist=1,
                       persist_block=1024, persist_cells=2048 * 2048, **served_cfg.engine) as e1:
            one, one_sols = batch_leg(e1, "one launch")
            one["persist_instances_per_call"] = e1.stats().persist_instances // max(1, args.batch_reps)
        one["form"] = ("ga_config.persist: every puzzle and grid runs all of its iterations inside one launch, "
                       "one 1024-thread workgroup (256 heaps) per instance; those that outgrow their structures "
                       "restart on the launch lanes")
        one["cost_mismatches_vs_lanes"] = sum(1 for a_, b_ in zip(lane_sols, one_sols)
                                              if a_.status != b_.status or a_.cost != b_.cost)
        batch["one_launch"] = one

    cores = host_cores(args.cpu_threads or 16)
    oracle_build = "portable"
    if not args.skip_cpu:  # the timed CPU legs use the oracle compiled for this host
        import oracle

        oracle_build = oracle.use_native()

    # ---- 24-puzzle: 5x5 boards on the wide-key kernel, one batch, vs the CPU oracle ----
    p24 = None
    if not args.skip_p24:
        p24 = puzzle24_leg(e, args.p24_heaps or served_cfg.heaps("puzzle", 5), cores, 0.0 if args.skip_cpu else 8.0)
        log(f"[bench] 24-puzzle: {p24}")

    # ---- one at a time: plain GA*, same requests, sequential ----
    timed_idx = list(range(W * M, (W + K) * M))
    plain_sps, mismatch, plain_n = None, 0, 0
    if args.plain_sample > 0:
        step = max(1, len(timed_idx) // args.plain_sample)
        sample = timed_idx[::step][: args.plain_sample]
        plain_cfg = tuning.load(mode="plain")
        e1 = ga.Engine(0, budget_bytes=int(args.budget_gb * (1 << 30)), chunk_bytes=8 << 30,
                       max_iterations=MAX_ITERATIONS, **plain_cfg.engine)
        t = time.perf_counter()
        for i in sample:
            x = insts[i]
            sol = e1.solve(x.request(cells_device=dev.get(i), num_queues=plain_cfg.heaps(x.kind, x.side)))
            if sol.status != ga.OK or served_costs.get(i, sol.cost) != sol.cost:
                mismatch += 1
        plain_t = time.perf_counter() - t
        plain_n = len(sample)
        plain_sps = plain_n / plain_t
        e1.close()
        log(f"[bench] one-at-a-time: {plain_n} in {plain_t:.2f}s = {plain_sps:.2f} solves/s, "
            f"served-vs-alone cost mismatches {mismatch}")

    # ---- the header drop-in: a_star<>() through gastar_cli, the same 15-puzzles one at a time ----
    dropin = None
    if args.dropin_sample > 0:
        dropin = dropin_leg([i for i in timed_idx if insts[i].kind == "puzzle"][: args.dropin_sample], insts,
                            served_costs)
        log(f"[bench] drop-in a_star<>: {dropin}")

    # ---- CPU oracle on the host cores ----
    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline([insts[i] for i in timed_idx], args.cpu_seconds, args.cpu_latency_seconds, cores,
                           {k - W * M: v for k, v in served_costs.items()}, oracle_build)
        log(f"[bench] cpu oracle: {cpu}")

    for a in dev.values():
        e.free_device(a)
    e.close()

    out = {
        "metric": METRIC,
        "value": round(served_sps, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(t_total / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "serving stream: seeded Poisson arrivals; 60% 15-puzzles (walk depth 30..70), "
                        "40% grids (side 512..4096, weights 1..9, 20% walls); 200 GB heap+closed-set budget",
            "requests_per_step": M,
            "offered_qps": args.qps,
            "seed": args.seed,
            "heaps_per_instance": served_cfg.describe(),
            "graph_replay": bool(served_cfg.use_graph) and not under_profiler,
            "grids": "copied from host memory at admission (PCIe inside the timed region)" if args.host_cells
                     else "resident in HBM before the timed region",
        },
        # latency at the stated offered load below saturation (the latency leg)
        "p50_ms": latency["p50_ms"] if latency else None,
        "p99_ms": latency["p99_ms"] if latency else None,
        "latency": latency,
        "saturation_latency": {"offered_qps": args.qps, "p50_ms": round(pct(sat_lat, 0.5), 3),
                               "p99_ms": round(pct(sat_lat, 0.99), 3), "max_inflight": ss.max_inflight,
                               "arena_peak_GB": round(st.peak_bytes / 2**30, 1), "budget_GB": args.budget_gb,
                               "grows": sum(s.grows for s in sols), "retries": sum(s.retries for s in sols),
                               "note": "offered load above capacity: a backlog builds over the stream"},
        "completed": ok,
        "failed": failed,
        "one_at_a_time_solves_per_sec": round(plain_sps, 3) if plain_sps else None,
        "dropin_one_at_a_time": dropin,
        "one_at_a_time_sample": plain_n,
        "batched": batch,
        "puzzle24": p24,
        "speedup_vs_one_at_a_time": round(served_sps / plain_sps, 2) if plain_sps else None,
        "served_vs_alone_cost_mismatches": mismatch,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    bad = failed or mismatch or (latency and latency["cost_mismatches_vs_saturation_leg"])
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
