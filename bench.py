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
segment drains before the next one arrives.
value = completed solves / wall seconds of the timed stream.

Also measured (outside the timed region): the same requests solved one at a
time by plain GA* (ga_solve), and the CPU oracle on the host cores.
The step kernel's roofline comes from the sampled chunks (every
--time-every-th chunk of each launch stream): HIP events around their step
launches (with graph replay one event pair brackets a chunk) and the device
counters, which count in those launches only.  Its HBM traffic is this run's algorithmic bytes per launch times
the HBM-bytes-per-algorithmic-byte ratio of the committed rocprofv3 counter
runs (profiles/step_kernel_traffic.json).
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


P24_N, P24_DEPTH = 256, (30, 60)


def puzzle24_leg(e, heaps: int, threads: int, cpu_seconds: float):
    """256 seeded 24-puzzles (walks of depth U{30..60} from the ordered board, no
    immediate backtrack) in one ga_solve_batch; the CPU oracle (sequential A*
    on boards) takes the same instances in order on `threads` workers for
    `cpu_seconds`, and every instance it finishes checks the GPU cost."""
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
    if cpu_seconds <= 0:
        return out
    import oracle

    lock = threading.Lock()
    state = {"next": 0, "done": 0, "match": 0, "capped": 0}
    deadline = time.perf_counter() + cpu_seconds
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                i = state["next"]
                if i >= P24_N or time.perf_counter() > deadline:
                    return
                state["next"] += 1
            rc, res, _ = oracle.astar_board(5, boards[i], ga.board_goal(5), max_states=MAX_ORACLE_STATES,
                                            want_path=False)
            with lock:
                if rc == oracle.E_LIMIT:
                    state["capped"] += 1
                    continue
                state["done"] += 1
                state["match"] += int(rc == oracle.OK and sols[i].status == ga.OK and sols[i].cost == res.cost)

    ts = [threading.Thread(target=worker) for _ in range(threads)]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    el = time.perf_counter() - t0
    out["cpu_oracle"] = {"solves_per_sec": round(state["done"] / el, 3), "cores": threads, "kind": "port",
                         "sample": f"first {state['done']} instances in order, {el:.1f}s; {state['capped']} above "
                                   f"{MAX_ORACLE_STATES} stored states skipped"}
    out["gpu_cost_matches"] = f"{state['match']}/{state['done']}"
    return out


def cpu_baseline(insts, seconds: float, threads: int, gpu_costs: dict):
    """CPU oracle (sequential A*) on the host cores over a bounded sample of
    the same stream: requests are taken in stream order by `threads` workers
    until `seconds` have passed."""
    import oracle

    lock = threading.Lock()
    state = {"next": 0, "done": 0, "match": 0, "checked": 0, "capped": 0}
    deadline = time.perf_counter() + seconds
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                i = state["next"]
                if i >= len(insts) or time.perf_counter() > deadline:
                    return
                state["next"] += 1
            x = insts[i]
            if x.kind == "puzzle":
                rc, res, _ = oracle.astar_puzzle(4, x.start, x.goal, max_states=MAX_ORACLE_STATES, want_path=False)
                if rc == oracle.E_LIMIT:
                    with lock:
                        state["capped"] += 1
                    continue
                cost = res.cost if rc == oracle.OK else -1
            else:
                rc, res, _ = oracle.astar_grid(x.cells, x.side, x.side, x.start, x.goal, 4, want_path=False)
                cost = res.cost if rc == oracle.OK else -1
            with lock:
                state["done"] += 1
                if i in gpu_costs:
                    state["checked"] += 1
                    state["match"] += int(gpu_costs[i] == cost)

    ts = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    return state["done"] / el, state["done"], el, state["checked"], state["match"], state["capped"]


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=512, help="requests per step (stream segment)")
    ap.add_argument("--qps", type=float, default=1500.0, help="offered load (Poisson arrivals)")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--plain-sample", type=int, default=40, help="requests solved one at a time")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cpu count)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--no-profile", action="store_true", help="skip per-step HIP events and counters")
    ap.add_argument("--time-every", type=int, default=8,
                    help="sample every n-th chunk: HIP events and device counters in its step launches")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batch", action="store_true", help="skip the batched (1024-instance) leg")
    ap.add_argument("--skip-p24", action="store_true", help="skip the 24-puzzle leg")
    ap.add_argument("--p24-heaps", type=int, default=0, help="heaps per 24-puzzle (0 = the served puzzle count)")
    ap.add_argument("--override", default="", help="served tuning overrides, e.g. lanes=3,rounds=2")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (for scaling HBM byte counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 1 << 30, "copies": 5}), flush=True)
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
    # The iteration limit only bounds a pathological request (it would count as failed); the longest
    # searches of this stream take about 6,000 iterations.
    e = ga.Engine(0, budget_bytes=int(args.budget_gb * (1 << 30)), chunk_bytes=8 << 30,
                  profile=0 if args.no_profile else max(1, args.time_every), max_iterations=MAX_ITERATIONS,
                  **served_cfg.engine)
    dev = {}
    t = time.perf_counter()
    for i, x in enumerate(insts):
        if x.kind == "grid":
            dev[i] = e.upload(x.cells)
    log(f"[bench] uploaded {len(dev)} grids in {time.perf_counter() - t:.1f}s")

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
    launches_before = e.stats().step_launches  # step dispatches of this process before the timed segment
    e.reset_stats()
    lat, ok, failed, served_costs = [], 0, 0, {}
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is only used for the barrier
        sync = lambda: None
    idx, reqs, arr = segment(W * M, K * M)
    sync()
    t_start = time.perf_counter()
    sols, ss = e.serve(reqs, arr)
    sync()
    t_total = time.perf_counter() - t_start
    for i, sol in zip(idx, sols):
        if sol.status == ga.OK:
            ok += 1
            lat.append(sol.latency_ns / 1e6)
            served_costs[i] = sol.cost
        else:
            failed += 1
    log(f"[bench] timed: {ss.completed}/{len(reqs)} in {t_total:.2f}s, p50 {ss.p50_ms:.1f} ms, "
        f"p99 {ss.p99_ms:.1f} ms, inflight max {ss.max_inflight}")
    by_kind = {}
    for kind in ("puzzle", "grid"):
        kl = [sol.latency_ns / 1e6 for i, sol in zip(idx, sols) if insts[i].kind == kind and sol.status == ga.OK]
        if kl:
            by_kind[kind] = {"n": len(kl), "p50_ms": round(pct(kl, 0.5), 2), "p99_ms": round(pct(kl, 0.99), 2)}
    slow = sorted(zip(idx, sols), key=lambda p: -p[1].latency_ns)[:5]
    slowest = [{"kind": insts[i].kind, "side": insts[i].side, "depth": insts[i].depth,
                "latency_ms": round(sol.latency_ns / 1e6, 1), "iterations": sol.iterations,
                "expanded": sol.expanded, "retries": sol.retries} for i, sol in slow]
    log(f"[bench] latency by kind {by_kind}; slowest {slowest}")
    st = e.stats()
    served_sps = ok / t_total

    # ---- kernel roofline (timed segments) ----
    roofline = None
    if not args.no_profile and st.step_launches and st.timed_launches:
        avg_ms = st.step_ms_total / st.timed_launches  # launches of the sampled chunks
        bytes_per_launch = st.algo_bytes / st.step_launches
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        traffic, traffic_src = None, None
        tpath = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")
        if os.path.exists(tpath):  # HBM bytes per algorithmic byte from the committed rocprofv3 --pmc runs
            with open(tpath) as f:
                tj = json.load(f)
            if "hbm_bytes_per_algo_byte" in tj:
                traffic = round(tj["hbm_bytes_per_algo_byte"] * bytes_per_launch, 1)
                traffic_src = (f"profiles/{os.path.basename(tpath)}: {tj['hbm_bytes_per_algo_byte']:.3f} HBM bytes "
                               f"per algorithmic byte x this run's algorithmic bytes per launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                    "kernel": f"step_kernel<{served_cfg.heap_arity}>", "avg_launch_us": round(avg_ms * 1e3, 3),
                    "algo_bytes_per_launch": round(bytes_per_launch, 1),
                    "expansions_per_launch": round(st.expanded / st.step_launches, 1),
                    # heap-group iterations with work / launched (every mapped heap, every launch)
                    "working_group_fraction": round(st.work_groups / st.heap_groups, 4) if st.heap_groups else None,
                    "launches": st.step_launches,
                    "timed_launches": st.timed_launches,
                    "launches_before_timed": launches_before,
                    "event_timing": ("per chunk of step launches (graph replay): chunk time / launches"
                                     if served_cfg.use_graph else "per step launch"),
                    # launch lanes run concurrently, so the per-launch figure above shares the device;
                    # this one is all algorithmic bytes of the timed stream over its wall time
                    "achieved_device_GBps": round(st.algo_bytes / t_total / 1e9, 2),
                    # global atomics issued by the step kernel: one CAS per closed-set probe (hashed sets)
                    # plus one atomicMin per lookup; counters and reductions excluded
                    # (probes and lookups are counted in the sampled launches only)
                    "atomics_per_launch": round((st.probes + st.lookups) / st.timed_launches, 1),
                    "atomic_rate_G_per_s": round((st.probes + st.lookups) / st.timed_launches / (avg_ms * 1e-3) / 1e9, 3)}
        # The device's rate for those atomics (random 64-bit, agent scope, over a 1 GiB table: no cache
        # reuse) and the share of the average launch that the launch's CAS probes and fetch-mins would
        # take at those rates.
        peak_min = e.atomic_rate(1 << 30, 1 << 26, 0)
        peak_cas = e.atomic_rate(1 << 30, 1 << 26, 1)
        t_atomics = (st.probes / peak_cas + st.lookups / peak_min) / st.timed_launches
        roofline["atomic_peak_G_per_s"] = {"fetch_min": round(peak_min / 1e9, 2), "cas": round(peak_cas / 1e9, 2)}
        roofline["atomic_frac"] = round(t_atomics / (avg_ms * 1e-3), 4)
    log(f"[bench] served {ok} ok / {failed} failed in {t_total:.2f}s = {served_sps:.1f} solves/s; roofline {roofline}")

    # ---- batched: the 1024 mixed instances of the batch configuration in one ga_solve_batch ----
    batch = None
    if not args.skip_batch:
        b_insts = wl.config4(4, 1024, threads=16)
        b_dev = {i: e.upload(x.cells) for i, x in enumerate(b_insts) if x.kind == "grid"}
        b_reqs = [x.request(cells_device=b_dev.get(i), num_queues=served_cfg.heaps(x.kind, x.side))
                  for i, x in enumerate(b_insts)]
        e.solve_batch(b_reqs[:64])  # warm-up
        t = time.perf_counter()
        b_sols = e.solve_batch(b_reqs)
        b_t = time.perf_counter() - t
        b_ok = sum(1 for s_ in b_sols if s_.status == ga.OK)
        batch = {"solves_per_sec": round(b_ok / b_t, 3), "instances": len(b_reqs), "ok": b_ok, "seconds": round(b_t, 3),
                 "workload": "1024 mixed: even = 15-puzzle walk depth 20..60, odd = grid side {256,512,1024,2048}, "
                             "walls 10..25%"}
        for a in b_dev.values():
            e.free_device(a)
        log(f"[bench] batched: {b_ok}/{len(b_reqs)} in {b_t:.2f}s = {b_ok / b_t:.1f} solves/s")

    # ---- 24-puzzle: 5x5 boards on the wide-key kernel, one batch, vs the CPU oracle ----
    p24 = None
    if not args.skip_p24:
        p24 = puzzle24_leg(e, args.p24_heaps or served_cfg.heaps("puzzle", 5), args.cpu_threads or min(16, os.cpu_count() or 1),
                           0.0 if args.skip_cpu else 8.0)
        log(f"[bench] 24-puzzle: {p24}")

    # ---- one at a time: plain GA*, same requests, sequential ----
    timed_idx = list(range(W * M, (W + K) * M))
    plain_sps, mismatch = None, 0
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
        plain_sps = len(sample) / plain_t
        e1.close()
        log(f"[bench] one-at-a-time: {len(sample)} in {plain_t:.2f}s = {plain_sps:.2f} solves/s, "
            f"served-vs-alone cost mismatches {mismatch}")

    # ---- CPU oracle on the host cores ----
    cpu = None
    if not args.skip_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sps, n, el, checked, match, capped = cpu_baseline([insts[i] for i in timed_idx], args.cpu_seconds, threads,
                                                  {k - W * M: v for k, v in served_costs.items()})
        cpu = {"value": round(sps, 3), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"first {n} requests of the timed stream in stream order, sequential A* "
                         f"(binary heap, lazy deletion) per request, {threads} threads, {el:.1f}s; "
                         f"{capped} puzzles above {MAX_ORACLE_STATES} stored states skipped",
               "gpu_cost_matches": f"{match}/{checked}"}
        log(f"[bench] cpu oracle: {n} in {el:.1f}s = {sps:.2f} solves/s ({threads} threads), "
            f"GPU costs equal on {match}/{checked}")

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
        },
        "p50_ms": round(pct(lat, 0.50), 3),
        "p99_ms": round(pct(lat, 0.99), 3),
        "latency_by_kind": by_kind,
        "slowest_requests": slowest,
        "completed": ok,
        "failed": failed,
        "one_at_a_time_solves_per_sec": round(plain_sps, 3) if plain_sps else None,
        "batched": batch,
        "puzzle24": p24,
        "speedup_vs_one_at_a_time": round(served_sps / plain_sps, 2) if plain_sps else None,
        "served_vs_alone_cost_mismatches": mismatch,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0 if failed == 0 and mismatch == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
