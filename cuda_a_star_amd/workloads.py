"""Seeded workload definitions for the five benchmark configurations.

Every random choice comes from splitmix64 streams (``ga_splitmix64`` in
libgastar.so), so a (config, seed) pair names the same instances everywhere.

  config 1  8-puzzle, 30-move walk with backtracks, N = 256 heaps
  config 2  15-puzzle, 60-move walk without immediate backtracks,
            N = 16384 heaps, closed set of 2^26 slots (load <= 0.5)
  config 3  4096 x 4096 grid, weights 1..9, 20 % walls, 4-connected,
            (0,0) -> (4095,4095), N = 16384 heaps
  config 4  1024 mixed instances: even i -> 15-puzzle with walk depth
            U{20..60}; odd i -> grid with side in {256,512,1024,2048},
            weights 1..9, wall probability U[0.10, 0.25]
  config 5  Poisson arrival stream: 60 % 15-puzzles (depth U{30..70}),
            40 % grids (side U{512..4096}, weights 1..9, 20 % walls)

  deep 24-puzzles  5x5 boards from walks of depth U{80..120} (no immediate
            backtrack) from the ordered board: out of reach of Manhattan,
            solved with the additive pattern databases
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import gen_grid, gen_puzzle, grid_request, puzzle_goal, puzzle_request, splitmix64, Request


@dataclass
class Instance:
    kind: str                   # "puzzle" or "grid"
    side: int                   # puzzle side or grid side
    seed: int
    start: int
    goal: int
    depth: int = 0              # puzzle walk length
    wall_p: float = 0.0         # grid wall probability
    cells: Optional[np.ndarray] = None
    num_queues: int = 0
    table_slots: int = 0
    connectivity: int = 4
    arrival_ns: int = 0

    def request(self, cells_device: Optional[int] = None, num_queues: Optional[int] = None, **kw) -> Request:
        nq = self.num_queues if num_queues is None else num_queues
        if self.kind == "puzzle":
            r = puzzle_request(self.side, self.start, self.goal, nq, **kw)
            if self.table_slots and "table_slots" not in kw:
                r.table_slots = self.table_slots
            return r
        return grid_request(self.cells, self.side, self.side, self.start, self.goal, self.connectivity,
                            nq, cells_device=cells_device, **kw)


class Stream:
    """splitmix64 stream of uniform draws."""

    def __init__(self, seed: int):
        self.state = seed & (2**64 - 1)

    def next(self) -> int:
        v, self.state = splitmix64(self.state)
        return v

    def uniform_int(self, lo: int, hi: int) -> int:   # inclusive
        return lo + self.next() % (hi - lo + 1)

    def uniform(self) -> float:                        # [0, 1)
        return (self.next() >> 11) * (1.0 / (1 << 53))


def instance_seed(master: int, i: int) -> int:
    v, _ = splitmix64((master ^ (i * 0x9E3779B97F4A7C15)) & (2**64 - 1))
    return v


def puzzle8(seed: int = 0) -> Instance:
    s = gen_puzzle(3, 30, seed, no_backtrack=False)
    return Instance("puzzle", 3, seed, s, puzzle_goal(3), depth=30, num_queues=256)


def puzzle15(seed: int = 0, depth: int = 60, num_queues: int = 16384, table_slots: int = 1 << 26) -> Instance:
    s = gen_puzzle(4, depth, seed, no_backtrack=True)
    return Instance("puzzle", 4, seed, s, puzzle_goal(4), depth=depth, num_queues=num_queues,
                    table_slots=table_slots)


def grid(side: int, wall_p: float, seed: int, num_queues: int = 0) -> Instance:
    cells, _ = gen_grid(side, side, wall_p, seed)
    return Instance("grid", side, seed, 0, side * side - 1, wall_p=wall_p, cells=cells, num_queues=num_queues)


def config1(seed: int = 0) -> Instance:
    return puzzle8(seed)


def config2(seed: int = 0) -> Instance:
    return puzzle15(seed)


def config3(seed: int = 0) -> Instance:
    return grid(4096, 0.20, seed, num_queues=16384)


def _materialise(insts: List[Instance], threads: int) -> List[Instance]:
    todo = [x for x in insts if x.kind == "grid" and x.cells is None]

    def make(x: Instance):
        x.cells, _ = gen_grid(x.side, x.side, x.wall_p, x.seed)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(make, todo))
    return insts


def config4(master: int = 4, n: int = 1024, threads: int = 16) -> List[Instance]:
    st = Stream(master)
    out = []
    for i in range(n):
        seed = instance_seed(master, i)
        if i % 2 == 0:
            depth = st.uniform_int(20, 60)
            s = gen_puzzle(4, depth, seed, no_backtrack=True)
            out.append(Instance("puzzle", 4, seed, s, puzzle_goal(4), depth=depth))
        else:
            side = (256, 512, 1024, 2048)[st.uniform_int(0, 3)]
            p = 0.10 + 0.15 * st.uniform()
            out.append(Instance("grid", side, seed, 0, side * side - 1, wall_p=p))
    return _materialise(out, threads)


def config5(master: int = 5, n: int = 1024, qps: float = 200.0, threads: int = 16) -> List[Instance]:
    """Requests with Poisson arrivals at `qps` (exponential gaps from the stream)."""
    st = Stream(master)
    out = []
    t = 0.0
    for i in range(n):
        seed = instance_seed(master, i)
        if st.uniform() < 0.6:
            depth = st.uniform_int(30, 70)
            s = gen_puzzle(4, depth, seed, no_backtrack=True)
            inst = Instance("puzzle", 4, seed, s, puzzle_goal(4), depth=depth)
        else:
            side = st.uniform_int(512, 4096)
            inst = Instance("grid", side, seed, 0, side * side - 1, wall_p=0.20)
        u = st.uniform()
        t += -math.log(1.0 - u) / qps
        inst.arrival_ns = int(t * 1e9)
        out.append(inst)
    return _materialise(out, threads)


def puzzle24_deep(master: int = 25, n: int = 16, depth=(80, 120)):
    """(depth, seed, board) of n 24-puzzles: walks of depth U{80..120} from the
    ordered board, no immediate backtrack."""
    from . import gen_board

    st = Stream(master)
    out = []
    for i in range(n):
        d = st.uniform_int(*depth)
        seed = instance_seed(master, i)
        out.append((d, seed, gen_board(5, d, seed)))
    return out
