"""Autotuner for the engine's launch and sizing parameters.

Tuned per mode ("served": many instances per launch; "plain": one instance
per launch):
  heaps per instance   by problem kind and grid size class
  heap arity           lanes per heap group = heap arity (4, 8)
  block size           threads per step-kernel workgroup (128, 256, 512)
  chunk iterations     GA* iterations launched between host checks
  pops per iteration   heap pops tried while entries are stale
  hash load factor     closed-set load limit before a capacity retry
  expansion rounds     heap pops per launch with local pushes in between
  adaptive width       heaps in use at admission, doubled on the device while the
                       open list exceeds grow_thresh entries per heap
  launch lanes         streams with their own iteration clock; the largest grids
                       (>= deep_cells cells) get their own lane
  f slack              a heap expands its top only within this much of the last
                       iteration's minimum f (less speculative expansion)
  graph replay         chunks replayed from a captured hipGraph or launched one by one
  kernel variants      ga_config.step_flags: paired pops on the 4-connected grid lane
                       (4; with 8, also on a lane that holds several kinds), write-back
                       of changed heap-top entries only (1), grids' settled-cell check
                       before the atomic (2)

`python -m cuda_a_star_amd.tuning [--mode served|plain]` sweeps the
candidates on a seeded sample of the serving workload on device 0 and writes
the winners to tuning.json next to this file; `load()` reads them (or the
built-in defaults when no file exists).  Served candidates are scored on a
Poisson arrival stream (ga_serve_run, solves per second of the stream's wall
time); plain candidates on the same requests solved one at a time.  Every
candidate is measured in a process of its own (engines created one after
another in one process run slower).
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import subprocess
import sys
import time
from dataclasses import asdict, dataclass, field
from typing import Dict

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_FILE = os.path.join(HERE, "tuning.json")


@dataclass
class TuneConfig:
    puzzle_heaps: int = 256
    grid_heaps_small: int = 512     # side <= 768
    grid_heaps_mid: int = 1024      # side <= 1536
    grid_heaps_big: int = 2048      # larger
    heap_arity: int = 4
    chunk_iters: int = 8
    pops_per_iter: int = 2
    max_load_pct: int = 50
    rounds: int = 1
    use_graph: int = 0
    adaptive_start: int = 0     # 0 = every request uses all its heaps from the start
    grow_thresh: int = 32
    rounds_grid: int = 0        # expansion rounds for grids; 0 = rounds
    lanes: int = 1              # launch lanes (streams with their own iteration clock)
    deep_cells: int = 4 << 20   # grids with at least this many cells take the last lane
    slack_puzzle: int = 0       # 0 = off; k: expand a heap top only if f <= previous min f + k - 1
    slack_grid: int = 0
    block: int = 256            # threads per step-kernel workgroup
    step_flags: int = 0         # step-kernel variants (ga_config.step_flags)
    max_instances: int = 4096   # instance slots, split evenly over the launch lanes and the 24-puzzle lane

    def heaps(self, kind: str, side: int) -> int:
        if kind == "puzzle":
            return self.puzzle_heaps
        if side <= 768:
            return self.grid_heaps_small
        if side <= 1536:
            return self.grid_heaps_mid
        return self.grid_heaps_big

    @property
    def engine(self) -> Dict[str, int]:
        return {"heap_arity": self.heap_arity, "chunk_iters": self.chunk_iters,
                "pops_per_iter": self.pops_per_iter, "max_load_pct": self.max_load_pct,
                "rounds": self.rounds, "use_graph": self.use_graph,
                "adaptive_start": self.adaptive_start, "grow_thresh": self.grow_thresh,
                "rounds_grid": self.rounds_grid, "lanes": self.lanes, "deep_cells": self.deep_cells,
                "slack_puzzle": self.slack_puzzle, "slack_grid": self.slack_grid, "block": self.block,
                "step_flags": self.step_flags, "max_instances": self.max_instances}

    def describe(self) -> Dict[str, int]:
        return asdict(self)


DEFAULTS = {
    "served": TuneConfig(),
    "plain": TuneConfig(puzzle_heaps=4096, grid_heaps_small=4096, grid_heaps_mid=4096, grid_heaps_big=4096),
}


def load(mode: str = "served") -> TuneConfig:
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            data = json.load(f)
        if mode in data:
            return TuneConfig(**data[mode])
    return TuneConfig(**asdict(DEFAULTS[mode]))


# ---------------------------------------------------------------------------
# sweep
# ---------------------------------------------------------------------------
def _measure_served(insts, cfg: TuneConfig) -> float:
    import cuda_a_star_amd as ga

    with ga.Engine(0, budget_bytes=200 << 30, chunk_bytes=8 << 30, **cfg.engine) as e:
        dev = {i: e.upload(x.cells) for i, x in enumerate(insts) if x.kind == "grid"}
        reqs = [x.request(cells_device=dev.get(i), num_queues=cfg.heaps(x.kind, x.side))
                for i, x in enumerate(insts)]
        t0 = insts[0].arrival_ns
        sols, ss = e.serve(reqs, [x.arrival_ns - t0 for x in insts])
        for a in dev.values():
            e.free_device(a)
    if any(s.status != ga.OK for s in sols) or not ss.wall_ns:
        return 0.0
    return len(sols) / (ss.wall_ns / 1e9)


def _measure_plain(insts, cfg: TuneConfig) -> Dict[str, float]:
    """Solves per second one at a time: over the whole sample ("value", wall
    time) and per kind ("puzzle", "grid": requests of the kind over the sum of
    their search times), so that a parameter of one kind is judged on that
    kind's solves, not drowned in the other's."""
    import cuda_a_star_amd as ga

    zero = {"value": 0.0, "puzzle": 0.0, "grid": 0.0}
    with ga.Engine(0, budget_bytes=200 << 30, chunk_bytes=8 << 30, **cfg.engine) as e:
        dev = {i: e.upload(x.cells) for i, x in enumerate(insts) if x.kind == "grid"}
        search = {"puzzle": [0, 0.0], "grid": [0, 0.0]}
        t = time.perf_counter()
        for i, x in enumerate(insts):
            s = e.solve(x.request(cells_device=dev.get(i), num_queues=cfg.heaps(x.kind, x.side)))
            if s.status != ga.OK:
                return zero
            search[x.kind][0] += 1
            search[x.kind][1] += s.search_ns / 1e9
            if i % 10 == 9:  # a line every ten requests: a long point is visibly alive
                print(f"[tune] plain {i + 1}/{len(insts)} in {time.perf_counter() - t:.1f}s", file=sys.stderr,
                      flush=True)
        dt = time.perf_counter() - t
        for a in dev.values():
            e.free_device(a)
    out = {"value": len(insts) / dt}
    for k, (n, sec) in search.items():
        out[k] = n / sec if sec > 0 else 0.0
    return out


# the kind whose solves decide a parameter in the plain sweep (others: the whole sample)
PLAIN_AXIS_KIND = {"puzzle_heaps": "puzzle", "slack_puzzle": "puzzle", "rounds": "puzzle", "max_load_pct": "puzzle",
                   "grid_heaps_small": "grid", "grid_heaps_mid": "grid", "grid_heaps_big": "grid",
                   "slack_grid": "grid", "rounds_grid": "grid", "step_flags": "grid"}


def _measure(mode: str, n: int, seed: int, qps: float, cfg: TuneConfig) -> Dict[str, float]:
    """One candidate, measured in a process of its own: engines created one
    after another in one process run slower (their streams share the process's
    few hardware queues with the earlier engines'), which would bias a sweep
    against every candidate after the first.  The child regenerates the sample
    from its seed and uses one engine."""
    cmd = [sys.executable, "-m", "cuda_a_star_amd.tuning", "--measure", mode, "--n", str(n), "--seed", str(seed),
           "--qps", str(qps), "--config", json.dumps(asdict(cfg))]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    if r.returncode != 0:
        print(f"[tune] measurement failed ({r.returncode}): {r.stderr[-400:]}", flush=True)
        return {"value": 0.0, "puzzle": 0.0, "grid": 0.0}
    return json.loads(r.stdout.strip().splitlines()[-1])


def sweep(mode: str, n: int, seed: int, qps: float, passes: int = 2) -> TuneConfig:
    best = load(mode)
    best_r = _measure(mode, n, seed, qps, best)
    print(f"[tune] {mode} start {best_r} {asdict(best)}", flush=True)
    if mode == "served":
        axes = {
            "puzzle_heaps": [256, 512, 1024, 2048, 4096, 8192, 16384],
            "grid_heaps_small": [128, 256, 512, 1024, 2048],
            "grid_heaps_mid": [256, 512, 1024, 2048, 4096],
            "grid_heaps_big": [256, 512, 1024, 2048, 4096, 8192],
            "heap_arity": [4, 8],
            "chunk_iters": [8, 16, 32, 64, 128],
            "pops_per_iter": [1, 2, 4],
            "max_load_pct": [50, 70],
            "rounds": [1, 2, 3, 4],
            "use_graph": [0, 1],
            "adaptive_start": [0, 32, 128, 512],
            "grow_thresh": [8, 32, 128],
            "rounds_grid": [0, 1, 2, 3, 4, 6, 8],
            "lanes": [1, 2, 3, 4],
            "deep_cells": [1 << 20, 2 << 20, 4 << 20, 8 << 20],
            "slack_puzzle": [0, 1, 3, 5, 9],
            "slack_grid": [0, 8, 16, 24, 32, 48],
            "block": [128, 256, 512],
            "step_flags": [0, 1, 2, 3, 4, 5, 6, 12],
            "max_instances": [4096, 6144, 8192],
        }
    else:
        axes = {
            "puzzle_heaps": [1024, 2048, 4096, 8192, 16384],
            "slack_puzzle": [0, 1, 3, 5],
            "rounds": [1, 2, 4],
            "grid_heaps_small": [1024, 2048, 4096, 8192],
            "grid_heaps_mid": [1024, 2048, 4096, 8192],
            "grid_heaps_big": [2048, 4096, 8192, 16384],
            "slack_grid": [0, 8, 16, 32, 48],
            "rounds_grid": [0, 1, 2, 3, 4],
            "step_flags": [0, 4],
            "chunk_iters": [4, 8, 16, 32, 64],
            "use_graph": [0, 1],
            "heap_arity": [4, 8],
            "pops_per_iter": [1, 2, 4],
            "max_load_pct": [50, 70],
            "block": [128, 256, 512],
        }
    # coordinate descent: one axis at a time; a plain parameter of one kind is
    # judged on that kind's solves per second of search time
    for _ in range(passes):
        for name, values in axes.items():
            key = PLAIN_AXIS_KIND.get(name, "value") if mode == "plain" else "value"
            if name == "rounds" and not best.rounds_grid:
                key = "value" if mode == "plain" else key  # rounds_grid 0: the grids take `rounds` too
            for v in values:
                if getattr(best, name) == v:
                    continue
                cand = TuneConfig(**{**asdict(best), name: v})
                r = _measure(mode, n, seed, qps, cand)
                print(f"[tune] {mode} {name}={v}: {r[key]:.2f}/s {key} (best {best_r[key]:.2f})", flush=True)
                if r[key] > best_r[key] * 1.02:
                    best, best_r = cand, r
    print(f"[tune] {mode} best {best_r} {asdict(best)}", flush=True)
    return best


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["served", "plain", "both"], default="both")
    ap.add_argument("--n", type=int, default=128, help="requests in the tuning sample")
    ap.add_argument("--seed", type=int, default=77)
    ap.add_argument("--qps", type=float, default=1500.0, help="offered load of the served sample")
    ap.add_argument("--out", default=TUNING_FILE, help="where to write the winners")
    ap.add_argument("--passes", type=int, default=2, help="coordinate-descent passes over the axes")
    ap.add_argument("--measure", choices=["served", "plain"], help="(internal) measure --config once and print it")
    ap.add_argument("--config", help="(internal) the TuneConfig to measure, as JSON")
    args = ap.parse_args(argv)
    if args.measure:
        from cuda_a_star_amd import workloads as wl

        insts = wl.config5(args.seed, args.n, args.qps, threads=16)
        cfg = TuneConfig(**json.loads(args.config))
        if args.measure == "served":
            print(json.dumps({"value": _measure_served(insts, cfg)}), flush=True)
        else:
            print(json.dumps(_measure_plain(insts, cfg)), flush=True)
        return 0
    data = {}
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            data = json.load(f)
    modes = ["served", "plain"] if args.mode == "both" else [args.mode]
    for m in modes:
        n = args.n if m == "served" else max(8, args.n * 3 // 8)
        data[m] = asdict(sweep(m, n, args.seed, args.qps, args.passes))
        with open(args.out, "w") as f:
            json.dump(data, f, indent=2, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
