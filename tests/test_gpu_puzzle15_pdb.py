"""15-puzzle pattern databases (ga_config.pdb15): the tables the device builds
equal the CPU oracle's byte for byte, and searches with them return the same
optimal costs as with Manhattan (the oracle's A*), with fewer expansions."""
import numpy as np
import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu

GOAL = ga.puzzle_goal(4)


def test_pdb15_tables_built_on_device_equal_the_oracle():
    with ga.Engine(0) as e:
        dev = e.pdb15_tables()
    assert np.array_equal(dev, oracle.pdb15())


def _puzzles():
    st = wl.Stream(15)
    return [ga.gen_puzzle(4, st.uniform_int(20, 70), wl.instance_seed(15, i)) for i in range(24)]


@pytest.mark.parametrize("form", ["solve", "batch"])
def test_pdb15_costs_equal_manhattan_and_oracle(form):
    starts = _puzzles()
    out = {}
    for pdb in (1, 0):
        with ga.Engine(0, pdb15=pdb, max_iterations=2_000_000) as e:
            reqs = [ga.puzzle_request(4, s, GOAL, 2048) for s in starts]
            sols = [e.solve(r) for r in reqs] if form == "solve" else e.solve_batch(reqs)
        out[pdb] = sols
    exp = {pdb: sum(s.expanded for s in sols) for pdb, sols in out.items()}
    for i, s in enumerate(starts):
        rc, ref, _ = oracle.astar_puzzle(4, s, GOAL, want_path=False)
        assert rc == oracle.OK
        for pdb in (1, 0):
            sol = out[pdb][i]
            assert sol.status == ga.OK and sol.cost == ref.cost, (i, pdb)
            assert oracle.replay_puzzle(4, sol.path, s, GOAL) == sol.cost
    assert exp[1] < exp[0]  # the tables prune


def test_pdb15_not_used_for_other_goals():
    # a goal other than the ordered board: Manhattan, the same cost as the oracle's
    goal = ga.gen_puzzle(4, 30, 9)
    start = ga.gen_puzzle(4, 40, 10)
    rc, ref, _ = oracle.astar_puzzle(4, start, goal, want_path=False)
    with ga.Engine(0, pdb15=1, max_iterations=2_000_000) as e:
        sol = e.solve(ga.puzzle_request(4, start, goal, 1024))
    assert rc == oracle.OK and sol.status == ga.OK and sol.cost == ref.cost
    assert oracle.replay_puzzle(4, sol.path, start, goal) == sol.cost


def test_pdb15_one_launch_form_matches_oracle():
    starts = _puzzles()[:12]
    with ga.Engine(0, persist=1, persist_block=1024, max_iterations=2_000_000) as e:
        sols = e.solve_batch([ga.puzzle_request(4, s, GOAL, 256) for s in starts])
        st = e.stats()
    assert st.persist_instances >= len(starts)
    for s, sol in zip(starts, sols):
        rc, ref, _ = oracle.astar_puzzle(4, s, GOAL, want_path=False)
        assert sol.status == ga.OK and sol.cost == ref.cost
        assert oracle.replay_puzzle(4, sol.path, s, GOAL) == sol.cost
