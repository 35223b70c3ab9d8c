#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the seeded generators and
the CPU oracle (both built from this repository).

  generators.json  digests of generator outputs (pins the instance streams)
  config1.json     8-puzzle walk-30 instances and their optimal costs
  config2.json     15-puzzle walk-60 instances, costs, oracle expansions
  config3.json     the 4096x4096 grid instance: cell digest and cost
  config4.json     the 1024 mixed batch instances and their costs
  config5.json     the first requests of the serving stream and their costs
  puzzle24_pdb.json 24-puzzles of walk depth 80..120 that the pattern-database
                   A* solves, with their costs (Manhattan A* stops at its cap)

Run: python tests/golden/make_golden.py [--threads 8]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import cuda_a_star_amd as ga  # noqa: E402
import oracle  # noqa: E402
from cuda_a_star_amd import workloads as wl  # noqa: E402


def digest(cells) -> str:
    return hashlib.sha256(cells.tobytes()).hexdigest()[:32]


def solve(x) -> dict:
    if x.kind == "puzzle":
        rc, res, _ = oracle.astar_puzzle(x.side, x.start, x.goal, want_path=False)
        assert rc == oracle.OK
        return {"kind": "puzzle", "side": x.side, "seed": x.seed, "depth": x.depth, "start": f"{x.start:#018x}",
                "cost": int(res.cost), "oracle_expanded": int(res.expanded)}
    rc, res, _ = oracle.astar_grid(x.cells, x.side, x.side, x.start, x.goal, 4, want_path=False)
    assert rc == oracle.OK
    return {"kind": "grid", "side": x.side, "seed": x.seed, "wall_p": x.wall_p, "cells_sha256": digest(x.cells),
            "cost": int(res.cost), "oracle_expanded": int(res.expanded)}


def write(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1)
    print("wrote", name, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    only = set(args.only.split(",")) if args.only else None
    pool = ThreadPoolExecutor(max_workers=args.threads)

    if not only or "generators" in only:
        gens = {"puzzle": [], "grid": []}
        for side, depth, nb in [(3, 30, False), (4, 60, True), (4, 70, True)]:
            for seed in range(8):
                gens["puzzle"].append({"side": side, "depth": depth, "no_backtrack": nb, "seed": seed,
                                       "state": f"{ga.gen_puzzle(side, depth, seed, nb):#018x}"})
        for w, h, p, seed in [(64, 48, 0.2, 0), (200, 150, 0.2, 42), (512, 512, 0.1, 3), (1000, 700, 0.25, 9)]:
            cells, draws = ga.gen_grid(w, h, p, seed)
            gens["grid"].append({"width": w, "height": h, "wall_p": p, "seed": seed, "draws": draws,
                                 "cells_sha256": digest(cells)})
        write("generators.json", gens)

    if not only or "config1" in only:
        insts = [wl.config1(s) for s in range(64)]
        write("config1.json", {"instances": list(pool.map(solve, insts))})

    if not only or "config2" in only:
        insts = [wl.config2(s) for s in range(16)]
        write("config2.json", {"instances": list(pool.map(solve, insts))})

    if not only or "config3" in only:
        x = wl.config3(0)
        write("config3.json", {"instances": [solve(x)]})

    if not only or "config4" in only:
        insts = wl.config4(4, 1024, threads=args.threads)
        out = list(pool.map(solve, insts))
        write("config4.json", {"master": 4, "n": 1024, "instances": out})

    if not only or "config5" in only:
        insts = wl.config5(5, 256, 1500.0, threads=args.threads)
        out = list(pool.map(solve, insts))
        for o, x in zip(out, insts):
            o["arrival_ns_at_1500qps"] = x.arrival_ns
        write("config5.json", {"master": 5, "n": 256, "instances": out})
    if not only or "puzzle24_pdb" in only:
        oracle.pdb24()
        cand = wl.puzzle24_deep(25, 24)

        def solve24(c):
            d, seed, b = c
            rc, res, _ = oracle.astar_board_pdb(b, max_states=15_000_000, want_path=False)
            if rc != oracle.OK:
                return None
            rm, rman, _ = oracle.astar_board(5, b, ga.board_goal(5), max_states=3_000_000, want_path=False)
            return {"index": cand.index(c), "depth": d, "seed": seed, "board": list(b), "cost": int(res.cost),
                    "oracle_pdb_expanded": int(res.expanded),
                    "manhattan_within_3M_states": rm == oracle.OK}

        out = [o for o in pool.map(solve24, cand) if o]
        write("puzzle24_pdb.json", {"master": 25, "depth": [80, 120], "tried": len(cand), "instances": out})
    return 0


if __name__ == "__main__":
    sys.exit(main())
