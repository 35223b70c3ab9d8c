"""The one-launch batch form (ga_config.persist): ga_solve_batch runs every
iteration of an eligible instance inside one launch, one workgroup per
instance.  Costs must equal the CPU oracle's (tolerance 0) and every path must
replay; the requests outside the form (24-puzzles, grids above persist_cells)
and those that outgrow their structures inside the launch finish on the
launch lanes with the same guarantees."""
import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu


def oracle_cost(x):
    if x.kind == "puzzle":
        rc, res, _ = oracle.astar_puzzle(x.side, x.start, x.goal, want_path=False)
        return res.cost if rc == oracle.OK else None
    return oracle.dial_grid(x.cells, x.side, x.side, x.start, x.goal, 4)


def replay_ok(x, sol):
    if x.kind == "puzzle":
        return oracle.replay_puzzle(x.side, sol.path, x.start, x.goal) == sol.cost
    return oracle.replay_grid(x.cells, x.side, x.side, 4, sol.path, x.start, x.goal) == sol.cost


def mix(seed, n_puz=20, n_grid=12, sides=(64, 100, 128, 256)):
    st = wl.Stream(seed)
    out = [wl.puzzle15(wl.instance_seed(seed, i), depth=st.uniform_int(20, 60), num_queues=0, table_slots=0)
           for i in range(n_puz)]
    out += [wl.grid(sides[i % len(sides)], 0.1 + 0.15 * st.uniform(), wl.instance_seed(seed + 1, i))
            for i in range(n_grid)]
    return out


@pytest.mark.parametrize("arity,block", [(4, 1024), (4, 256), (8, 512), (8, 1024)])
def test_one_launch_batch_matches_oracle(arity, block):
    insts = mix(11 + block + arity)
    with ga.Engine(0, heap_arity=arity, persist=1, persist_block=block, persist_cells=256 * 256, rounds_grid=3,
                   max_iterations=2_000_000) as e:
        sols = e.solve_batch([x.request() for x in insts])
        st = e.stats()
    assert st.persist_launches >= 1 and st.persist_instances >= len(insts)  # every instance took the form
    for i, (x, sol) in enumerate(zip(insts, sols)):
        assert sol.status == ga.OK, (i, ga.status_str(sol.status))
        assert sol.cost == oracle_cost(x), i
        assert replay_ok(x, sol), i


def test_one_launch_batch_sends_the_rest_to_the_lanes():
    insts = mix(5, n_puz=8, n_grid=4, sides=(64, 96))
    big = wl.grid(300, 0.2, 77)  # above persist_cells: the lanes
    small_table = wl.puzzle15(wl.instance_seed(9, 1), depth=50, num_queues=0, table_slots=1024)  # fills its set
    board = ga.gen_board(5, 24, 3)  # 24-puzzle: the lanes
    g = ga.puzzle_goal(4)
    t = ga.state_to_board(g, 4)
    t[0], t[1] = t[1], t[0]
    unsolvable = ga.puzzle_request(4, ga.board_to_state(t), g, 256)
    invalid = ga.puzzle_request(4, 0x1111111111111111, g, 256)
    reqs = [x.request() for x in insts] + [big.request(), small_table.request(),
                                           ga.puzzle5_request(board, num_queues=1024), unsolvable, invalid]
    with ga.Engine(0, persist=1, persist_block=512, persist_cells=128 * 128, max_iterations=2_000_000) as e:
        sols = e.solve_batch(reqs)
        st = e.stats()
    n = len(insts)
    assert st.persist_instances == n + 1  # the small grids and puzzles and the small-table puzzle
    for i, (x, sol) in enumerate(zip(insts + [big, small_table], sols)):
        assert sol.status == ga.OK, (i, ga.status_str(sol.status))
        assert sol.cost == oracle_cost(x), i
        assert replay_ok(x, sol), i
    assert sols[n + 1].retries >= 1  # outgrew its closed set in the launch, restarted on the lanes
    # ... at once: its workgroup's completion flag reached the host while the wave still ran
    assert st.persist_early_restarts >= 1
    rc, ref, _ = oracle.astar_board(5, board, ga.board_goal(5), want_path=False)
    assert rc == oracle.OK and sols[n + 2].status == ga.OK and sols[n + 2].cost == ref.cost
    assert oracle.replay_board(5, sols[n + 2].path, board, ga.board_goal(5)) == sols[n + 2].cost
    assert sols[n + 3].status == ga.NO_PATH
    assert sols[n + 4].status == ga.E_INVALID


def test_one_launch_batch_iteration_limit():
    # a limit far below what the searches need: every instance ends GA_E_TIMEOUT inside the launch
    insts = mix(21, n_puz=4, n_grid=2, sides=(256,))
    with ga.Engine(0, persist=1, persist_block=256, persist_cells=256 * 256, max_iterations=3) as e:
        sols = e.solve_batch([x.request() for x in insts])
    assert all(s.status == ga.E_TIMEOUT for s in sols), [ga.status_str(s.status) for s in sols]
