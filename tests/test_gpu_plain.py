"""Plain GA* on the GPU (one instance per call) against the CPU oracle.

Costs must equal the oracle's exactly (integer, tolerance 0) and every path
must pass a move-by-move replay.  Instances come from the seeded generators.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu


def check_puzzle(engine, side, start, goal=None, **kw):
    goal = ga.puzzle_goal(side) if goal is None else goal
    sol = engine.solve_puzzle(side, start, goal, **kw)
    rc, ref, _ = oracle.astar_puzzle(side, start, goal, want_path=False)
    if rc == oracle.NO_PATH:
        assert sol.status == ga.NO_PATH
        return sol
    assert rc == oracle.OK
    assert sol.status == ga.OK, ga.status_str(sol.status)
    assert sol.cost == ref.cost
    assert oracle.replay_puzzle(side, sol.path, start, goal) == sol.cost
    return sol


def check_grid(engine, cells, w, h, start=0, goal=None, conn=4, **kw):
    goal = w * h - 1 if goal is None else goal
    sol = engine.solve_grid(cells, w, h, start, goal, conn, **kw)
    ref = oracle.dial_grid(cells, w, h, start, goal, conn)
    if ref < 0:
        assert sol.status == ga.NO_PATH
        return sol
    assert sol.status == ga.OK, ga.status_str(sol.status)
    assert sol.cost == ref
    assert oracle.replay_grid(cells, w, h, conn, sol.path, start, goal) == sol.cost
    return sol


def test_puzzle8_start_is_goal(engine):
    g = ga.puzzle_goal(3)
    sol = engine.solve_puzzle(3, g, g, num_queues=256)
    assert sol.status == ga.OK and sol.cost == 0 and sol.path == [g]


@pytest.mark.parametrize("seed", range(8))
def test_puzzle8_walk30(engine, seed):
    # 8-puzzle, 30-move seeded walk with backtracks allowed, N = 256 heaps
    start = ga.gen_puzzle(3, 30, seed, no_backtrack=False)
    check_puzzle(engine, 3, start, num_queues=256)


def test_puzzle8_unsolvable(engine):
    g = ga.puzzle_goal(3)
    tiles = ga.state_to_board(g, 3)
    tiles[0], tiles[1] = tiles[1], tiles[0]  # one transposition: other parity class
    sol = engine.solve_puzzle(3, ga.board_to_state(tiles), g, num_queues=256)
    assert sol.status == ga.NO_PATH


def test_puzzle8_hardest(engine):
    # the two 31-move states of the 8-puzzle (BFS depth 31)
    n, depth, hist = oracle.bfs_puzzle8(ga.puzzle_goal(3))
    assert hist[31] == 2
    found = 0
    import itertools
    for perm in itertools.permutations(range(9)):
        s = ga.board_to_state(perm)
        if depth[oracle.perm_rank9(s)] == 31:
            sol = check_puzzle(engine, 3, s, num_queues=256)
            assert sol.cost == 31
            found += 1
            if found == 2:
                break
    assert found == 2


@pytest.mark.parametrize("seed", range(4))
def test_puzzle15_walk60(engine, seed):
    start = ga.gen_puzzle(4, 60, seed)
    check_puzzle(engine, 4, start, num_queues=16384)


@pytest.mark.parametrize("seed", range(3))
def test_grid_small(engine, seed):
    cells, _ = ga.gen_grid(64, 48, 0.2, seed)
    check_grid(engine, cells, 64, 48, num_queues=256)


def test_grid_8conn(engine):
    cells, _ = ga.gen_grid(80, 60, 0.25, 11, connectivity=8)
    check_grid(engine, cells, 80, 60, conn=8, num_queues=512)


def test_grid_closed_form(engine):
    w, h, wt = 100, 70, 3
    cells = np.full(w * h, wt, dtype=np.uint8)
    sol = check_grid(engine, cells, w, h, num_queues=1024)
    assert sol.cost == wt * (w - 1 + h - 1)


def test_grid_unreachable(engine):
    cells, _ = ga.gen_grid(40, 40, 0.1, 3)
    cells = cells.reshape(40, 40).copy()
    cells[:, 20] = 0  # wall off the right half
    cells = cells.reshape(-1)
    sol = engine.solve_grid(cells, 40, 40, 0, 40 * 40 - 1, 4, num_queues=256)
    assert sol.status == ga.NO_PATH


@pytest.mark.parametrize("w,h,conn", [(1, 64, 4), (2, 50, 4), (2, 50, 8), (64, 1, 4), (3, 40, 8), (33, 2, 4),
                                        (40, 33, 8), (128, 3, 4), (129, 5, 8), (2, 300, 8), (131, 130, 4),
                                        (256, 129, 8)])
def test_grid_thin_and_window_edges(engine, w, h, conn):
    # widths and heights below, at and just above the path walk's window
    # limits (rows of 1 or 2 cells, the 128-cell window): cost, then replay
    rng = np.random.default_rng(1000 * w + h + conn)
    cells = rng.integers(1, 10, size=w * h).astype(np.uint8)
    check_grid(engine, cells, w, h, conn=conn, num_queues=64)
    # the same grid walked from the other corner (goal at cell 0)
    sol = engine.solve_grid(cells, w, h, w * h - 1, 0, conn, num_queues=64)
    ref = oracle.dial_grid(cells, w, h, w * h - 1, 0, conn)
    assert sol.status == ga.OK and sol.cost == ref
    assert oracle.replay_grid(cells, w, h, conn, sol.path, w * h - 1, 0) == sol.cost


@pytest.mark.parametrize("conn", [4, 8])
def test_grid_serpentine_walk(engine, conn):
    # rows joined at alternating ends: the path runs away from the start along
    # x on every other row, so the walk leaves its windows on every side (and
    # diagonally when 8-connected), about 20,000 cells long
    w, h = 300, 263
    rng = np.random.default_rng(7 + conn)
    cells = rng.integers(1, 10, size=(h, w)).astype(np.uint8)
    for i, r in enumerate(range(3, h, 4)):
        cells[r, :] = 0
        cells[r, w - 1 if i % 2 == 0 else 0] = 1  # the gap: right end, then left end
    sol = check_grid(engine, cells.reshape(-1), w, h, conn=conn, num_queues=1024)
    assert len(sol.path) > 19_000


def test_grid_1024(engine):
    cells, _ = ga.gen_grid(1024, 1024, 0.2, 5)
    check_grid(engine, cells, 1024, 1024, num_queues=4096)


def test_capacity_growth(engine):
    # tiny heaps and closed set: grown in place (paused, grown, resumed), never restarted
    start = ga.gen_puzzle(4, 60, 1)
    sol = check_puzzle(engine, 4, start, num_queues=1024, heap_capacity=16, table_slots=1 << 12)
    assert sol.grows > 0 and sol.retries == 0


def test_capacity_growth_by_restart():
    # the same with growth in place off: the overflow -> restart path
    start = ga.gen_puzzle(4, 60, 1)
    with ga.Engine(0, grow_in_place=0, max_iterations=2_000_000) as e:
        sol = check_puzzle(e, 4, start, num_queues=1024, heap_capacity=16, table_slots=1 << 12)
    assert sol.retries > 0 and sol.grows == 0


def _golden(name):
    with open(os.path.join(os.path.dirname(__file__), "golden", name)) as f:
        return json.load(f)["instances"]


@pytest.mark.parametrize("seed", range(16))
def test_config2_puzzle15_matches_golden(engine, seed):
    # 15-puzzle, 60-move walk without immediate backtracks, N = 16384 heaps,
    # 2^26 closed-set slots (load <= 0.5): every committed instance
    x, g = wl.config2(seed), _golden("config2.json")[seed]
    assert x.start == int(g["start"], 16)
    sol = engine.solve(x.request())
    assert sol.status == ga.OK, ga.status_str(sol.status)
    assert sol.cost == g["cost"]
    assert oracle.replay_puzzle(4, sol.path, x.start, x.goal) == sol.cost


def test_config3_grid4096_matches_golden(engine):
    # 4096 x 4096 grid, weights 1..9, 20 % walls, (0,0) -> (4095,4095), N = 16384 heaps
    x, g = wl.config3(0), _golden("config3.json")[0]
    assert hashlib.sha256(x.cells.tobytes()).hexdigest()[:32] == g["cells_sha256"]
    sol = engine.solve(x.request())
    assert sol.status == ga.OK, ga.status_str(sol.status)
    assert sol.cost == g["cost"]
    assert oracle.replay_grid(x.cells, 4096, 4096, 4, sol.path, x.start, x.goal) == sol.cost


def test_c_abi_solve_entry_points_directly(engine):
    # ga_solve / ga_solve_puzzle / ga_solve_grid through ctypes (the Python
    # wrapper runs through ga_solve_batch): the return code is the request's
    # status, the result holds the cost and a path that replays
    import ctypes

    L = ga.lib()
    start = ga.gen_puzzle(4, 40, 11)
    res = ga.Result()
    rc = L.ga_solve_puzzle(engine._h, 4, start, ga.puzzle_goal(4), 1024, ctypes.byref(res))
    _, ref, _ = oracle.astar_puzzle(4, start, ga.puzzle_goal(4), want_path=False)
    assert rc == ga.OK == res.status and res.cost == ref.cost
    path = [res.path[i] for i in range(res.path_len)]
    L.ga_result_free(ctypes.byref(res))
    assert oracle.replay_puzzle(4, path, start, ga.puzzle_goal(4)) == ref.cost
    cells, _ = ga.gen_grid(300, 200, 0.2, 5)
    rc = L.ga_solve_grid(engine._h, cells.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 300, 200, 0,
                         300 * 200 - 1, 4, 512, ctypes.byref(res))
    assert rc == ga.OK == res.status and res.cost == oracle.dial_grid(cells, 300, 200, 0, 300 * 200 - 1, 4)
    path = [res.path[i] for i in range(res.path_len)]
    L.ga_result_free(ctypes.byref(res))
    assert oracle.replay_grid(cells, 300, 200, 4, path, 0, 300 * 200 - 1) == res.cost
    # a per-request failure comes back as the request's status, not as a failed run
    bad = ga.puzzle_request(4, 0x1111111111111111)
    assert L.ga_solve(engine._h, ctypes.byref(bad), ctypes.byref(res)) == ga.E_INVALID == res.status
    t = ga.state_to_board(ga.puzzle_goal(4), 4)
    t[0], t[1] = t[1], t[0]
    assert L.ga_solve_puzzle(engine._h, 4, ga.board_to_state(t), ga.puzzle_goal(4), 64,
                             ctypes.byref(res)) == ga.NO_PATH
