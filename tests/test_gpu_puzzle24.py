"""The 24-puzzle (5x5 boards, the reference CLI's default size, main.cu:72)
through the C ABI: 128-bit keys on their own kernel and launch lane.  Every
cost equals the CPU oracle's (oracle_astar_board, pinned to BFS and to the
64-bit puzzle search), every path replays move by move."""
import json
import os

import numpy as np
import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def check(start, sol, goal=None):
    goal = ga.board_goal(5) if goal is None else goal
    rc, ref, _ = oracle.astar_board(5, start, goal, max_states=30_000_000, want_path=False)
    assert rc == oracle.OK
    assert sol.status == ga.OK, ga.status_str(sol.status)
    assert sol.cost == ref.cost
    assert len(sol.path) == sol.cost + 1
    assert oracle.replay_board(5, sol.path, start, goal) == sol.cost


@pytest.mark.parametrize("depth,seed", [(0, 1), (1, 2), (12, 3), (30, 4), (45, 5)])
def test_plain_matches_oracle(engine, depth, seed):
    start = ga.gen_board(5, depth, seed)
    sol = engine.solve(ga.puzzle5_request(start, num_queues=1024))
    check(start, sol)


def test_nonstandard_goal(engine):
    goal = ga.gen_board(5, 20, 99)
    start = ga.gen_board(5, 25, 98, goal=goal)
    sol = engine.solve(ga.puzzle5_request(start, goal, num_queues=512))
    check(start, sol, goal)


def test_batch_with_other_kinds(engine):
    """24-puzzles on their own lane next to 15-puzzles and grids in one batch."""
    boards = [ga.gen_board(5, 20 + 3 * i, 100 + i) for i in range(10)]
    others = [wl.puzzle15(wl.instance_seed(7, i), depth=40, num_queues=512, table_slots=0) for i in range(6)]
    others += [wl.grid(128 + 64 * i, 0.2, wl.instance_seed(8, i), num_queues=256) for i in range(4)]
    reqs = [ga.puzzle5_request(b, num_queues=(256, 1024, 4096)[i % 3]) for i, b in enumerate(boards)]
    reqs += [x.request() for x in others]
    sols = engine.solve_batch(reqs)
    for b, sol in zip(boards, sols):
        check(b, sol)
    for x, sol in zip(others, sols[len(boards):]):
        assert sol.status == ga.OK
        if x.kind == "puzzle":
            assert sol.cost == oracle.astar_puzzle(4, x.start, x.goal, want_path=False)[1].cost
        else:
            assert sol.cost == oracle.dial_grid(x.cells, x.side, x.side, x.start, x.goal, 4)


def test_invalid_and_unsolvable(engine):
    b = list(ga.gen_board(5, 30, 5))
    bad = b[:]
    bad[0] = bad[1]
    swapped = b[:]
    i, j = swapped.index(7), swapped.index(8)
    swapped[i], swapped[j] = swapped[j], swapped[i]
    sols = engine.solve_batch([ga.puzzle5_request(bad), ga.puzzle5_request(swapped),
                               ga.puzzle5_request(b, num_queues=512)])
    assert sols[0].status == ga.E_INVALID
    assert sols[1].status == ga.NO_PATH
    check(tuple(b), sols[2])


@pytest.mark.parametrize("grow", [1, 0], ids=["in-place", "restart"])
def test_forced_retries(grow):
    # 32-byte board slots rehashed in place, or the restart path
    boards = [ga.gen_board(5, 36, 200 + i) for i in range(4)]
    reqs = []
    for b in boards:
        r = ga.puzzle5_request(b, num_queues=256, heap_capacity=16)
        r.table_slots = 1 << 10
        reqs.append(r)
    with ga.Engine(0, max_iterations=2_000_000, grow_in_place=grow, pdb24=0) as e:
        sols = e.solve_batch(reqs)
    if grow:
        assert any(s.grows > 0 for s in sols) and all(s.retries == 0 for s in sols)
    else:
        assert any(s.retries > 0 for s in sols)
    for b, sol in zip(boards, sols):
        check(b, sol)
        assert oracle.replay_board(5, sol.path, b, ga.board_goal(5)) == sol.cost


@pytest.mark.parametrize("variant", [dict(heap_arity=8, rounds=2, use_graph=1), dict(heap_arity=8, rounds=1, chunk_iters=7),
                                     dict(heap_arity=4, rounds=4, slack_puzzle=3, lanes=2)],
                         ids=["arity8-graph", "arity8-chunk7", "arity4-slack-lanes2"])
def test_variants(variant):
    boards = [ga.gen_board(5, 28 + 4 * i, 300 + i) for i in range(6)]
    with ga.Engine(0, max_iterations=2_000_000, **variant) as e:
        sols = e.solve_batch([ga.puzzle5_request(b, num_queues=1024) for b in boards])
    for b, sol in zip(boards, sols):
        check(b, sol)


def test_serve_stream_mixed():
    boards = [ga.gen_board(5, 24 + (i % 5) * 5, 400 + i) for i in range(16)]
    puzzles = [wl.puzzle15(wl.instance_seed(9, i), depth=35, num_queues=256, table_slots=0) for i in range(16)]
    reqs, arr = [], []
    for i in range(16):
        reqs += [ga.puzzle5_request(boards[i], num_queues=1024), puzzles[i].request()]
        arr += [i * 2_000_000, i * 2_000_000 + 500_000]
    with ga.Engine(0, lanes=2, max_iterations=2_000_000) as e:
        sols, ss = e.serve(reqs, arr)
    assert ss.completed == len(reqs) and ss.failed == 0
    for i in range(16):
        check(boards[i], sols[2 * i])
        assert sols[2 * i + 1].cost == oracle.astar_puzzle(4, puzzles[i].start, puzzles[i].goal,
                                                           want_path=False)[1].cost


# ---- additive pattern databases (ga_config.pdb24) ----
def test_pdb_tables_built_on_device_equal_the_oracle():
    with ga.Engine(0) as e:
        dev = e.pdb24_tables()
    assert np.array_equal(dev, oracle.pdb24())


def test_deep_walks_with_pdb_match_golden():
    # walks of depth 80..120 that Manhattan cannot bring within budget; served heap count
    gold = json.load(open(os.path.join(GOLDEN, "puzzle24_pdb.json")))["instances"]
    with ga.Engine(0, max_iterations=2_000_000) as e:
        sols = e.solve_batch([ga.puzzle5_request(tuple(o["board"]), num_queues=8192) for o in gold])
    for o, sol in zip(gold, sols):
        assert sol.status == ga.OK, (o["index"], ga.status_str(sol.status))
        assert sol.cost == o["cost"], o["index"]
        assert oracle.replay_board(5, sol.path, tuple(o["board"]), ga.board_goal(5)) == sol.cost


def test_pdb_and_manhattan_agree_and_pdb_expands_less():
    gold = json.load(open(os.path.join(GOLDEN, "puzzle24_pdb.json")))["instances"]
    o = next(x for x in gold if x["manhattan_within_3M_states"] and x["oracle_pdb_expanded"] > 100_000)
    req = lambda: ga.puzzle5_request(tuple(o["board"]), num_queues=4096)
    with ga.Engine(0, max_iterations=2_000_000, pdb24=1) as e:
        p = e.solve(req())
    with ga.Engine(0, max_iterations=2_000_000, pdb24=0) as e:
        m = e.solve(req())
    assert p.status == m.status == ga.OK and p.cost == m.cost == o["cost"]
    assert p.expanded < m.expanded
    # a goal other than the ordered board has no tables: Manhattan, still exact
    b = ga.gen_board(5, 20, 77)
    goal = tuple(ga.gen_board(5, 6, 78))
    with ga.Engine(0, max_iterations=2_000_000) as e:
        sol = e.solve(ga.puzzle5_request(b, goal, num_queues=1024))
    rc, ref, _ = oracle.astar_board(5, b, goal, want_path=False)
    assert rc == oracle.OK and sol.status == ga.OK and sol.cost == ref.cost
