"""ga_solve_noalloc (every device buffer in the library's static scratch)
against ga_solve on the same requests: equal status and cost (integer,
tolerance 0), a path of the same length that replays move by move; two calls
in a row per kind, the second on a smaller input (the scratch keeps the first
call's contents).  Also its limits: a request that needs more than the
scratch holds, and a path buffer that is too small."""
import ctypes

import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu


def noalloc(e, req, path_cap):
    words = 4 if (req.kind == ga.PUZZLE and req.side == 5) else 1
    buf = (ctypes.c_uint64 * max(1, path_cap * words))()
    res = ga.Result()
    rc = ga.lib().ga_solve_noalloc(e._h, ctypes.byref(req), ctypes.byref(res), buf, path_cap)
    assert rc == res.status and not res.path
    if words == 4:
        raw = ctypes.string_at(buf, 32 * res.path_len) if rc == ga.OK else b""
        path = [tuple(raw[32 * k:32 * k + 25]) for k in range(res.path_len if rc == ga.OK else 0)]
    else:
        path = list(buf[:res.path_len]) if rc == ga.OK else []
    return res, path


def existing(e, req):
    res = ga.Result()
    rc = ga.lib().ga_solve(e._h, ctypes.byref(req), ctypes.byref(res))
    assert rc == res.status
    sol = ga._take(res, 5 if (req.kind == ga.PUZZLE and req.side == 5) else 0)
    return sol


def same(e, req, replay):
    ref = existing(e, req)
    res, path = noalloc(e, req, 1 << 16)
    assert res.status == ref.status == ga.OK, (ga.status_str(res.status), ga.status_str(ref.status))
    assert res.cost == ref.cost and res.path_len == len(ref.path) == len(path)
    assert replay(path) == res.cost and replay(ref.path) == ref.cost
    return res


def test_noalloc_matches_solve_two_calls_in_a_row():
    with ga.Engine(0, max_iterations=2_000_000) as e:
        # puzzles: a 15-puzzle, then a smaller 8-puzzle
        big = wl.puzzle15(3, depth=50, num_queues=2048, table_slots=1 << 21)
        res = same(e, big.request(), lambda p: oracle.replay_puzzle(4, p, big.start, big.goal))
        _, ref, _ = oracle.astar_puzzle(4, big.start, big.goal, want_path=False)
        assert res.cost == ref.cost
        small = wl.puzzle8(5)
        res = same(e, small.request(), lambda p: oracle.replay_puzzle(3, p, small.start, small.goal))
        _, ref, _ = oracle.astar_puzzle(3, small.start, small.goal, want_path=False)
        assert res.cost == ref.cost
        # grids: 1024 x 1024, then 200 x 200
        for side in (1024, 200):
            g = wl.grid(side, 0.2, 40 + side, num_queues=1024)
            res = same(e, g.request(),
                       lambda p, g=g: oracle.replay_grid(g.cells, g.side, g.side, 4, p, g.start, g.goal))
            assert res.cost == oracle.dial_grid(g.cells, g.side, g.side, g.start, g.goal, 4)
        # a 24-puzzle, then a shorter one
        for depth, seed in ((34, 8), (16, 9)):
            b = ga.gen_board(5, depth, seed)
            req = ga.puzzle5_request(b, num_queues=1024)
            res = same(e, req, lambda p, b=b: oracle.replay_board(5, p, b, ga.board_goal(5)))
            rc, ref, _ = oracle.astar_board(5, b, ga.board_goal(5), want_path=False)
            assert rc == oracle.OK and res.cost == ref.cost


def test_noalloc_limits():
    with ga.Engine(0, max_iterations=2_000_000) as e:
        # 2^26 closed-set slots (1 GiB) do not fit the 256 MiB scratch
        x = wl.puzzle15(1, depth=30, num_queues=1024, table_slots=1 << 26)
        res, _ = noalloc(e, x.request(), 1024)
        assert res.status == ga.E_BUDGET
        assert b"scratch" in ga.lib().ga_engine_last_error(e._h)
        # a path buffer of 3 states for a longer path: the cost and the length needed come back
        x = wl.puzzle8(2)
        _, ref, _ = oracle.astar_puzzle(3, x.start, x.goal, want_path=False)
        res, _ = noalloc(e, x.request(), 3)
        assert res.status == ga.E_INVALID and res.cost == ref.cost and res.path_len == ref.cost + 1
        # and the scratch still serves the next call
        res, path = noalloc(e, x.request(), 64)
        assert res.status == ga.OK and oracle.replay_puzzle(3, path, x.start, x.goal) == ref.cost
        # unsolvable and invalid requests never reach the device
        t = ga.state_to_board(ga.puzzle_goal(4), 4)
        t[0], t[1] = t[1], t[0]
        res, _ = noalloc(e, ga.puzzle_request(4, ga.board_to_state(t)), 8)
        assert res.status == ga.NO_PATH
        res, _ = noalloc(e, ga.puzzle_request(4, 0x1111111111111111), 8)
        assert res.status == ga.E_INVALID
