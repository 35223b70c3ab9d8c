"""Batched GA* (many instances per launch) and serving (continuous batching)
on the GPU.  Every cost must equal the CPU oracle's (tolerance 0), every path
must replay, and every result must be the same alone and in the mix."""
import json
import os

import numpy as np
import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def oracle_cost(x):
    if x.kind == "puzzle":
        rc, res, _ = oracle.astar_puzzle(x.side, x.start, x.goal, want_path=False)
        return res.cost if rc == oracle.OK else None
    return oracle.dial_grid(x.cells, x.side, x.side, x.start, x.goal, 4)


def replay_ok(x, sol):
    if x.kind == "puzzle":
        return oracle.replay_puzzle(x.side, sol.path, x.start, x.goal) == sol.cost
    return oracle.replay_grid(x.cells, x.side, x.side, 4, sol.path, x.start, x.goal) == sol.cost


def small_mix(seed=3, n_puz=24, n_grid=12):
    st = wl.Stream(seed)
    out = []
    for i in range(n_puz):
        out.append(wl.puzzle15(wl.instance_seed(seed, i), depth=st.uniform_int(20, 60), num_queues=0,
                               table_slots=0))
        out[-1].num_queues = (64, 256, 1024)[i % 3]
    for i in range(n_grid):
        side = (64, 128, 200, 256)[i % 4]
        g = wl.grid(side, 0.1 + 0.15 * st.uniform(), wl.instance_seed(seed + 1, i))
        g.num_queues = (64, 512, 2048)[i % 3]
        out.append(g)
    return out


def test_batch_small_mix_matches_oracle_and_alone(engine):
    insts = small_mix()
    sols = engine.solve_batch([x.request() for x in insts])
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK, ga.status_str(sol.status)
        assert sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)
    for x, sol in list(zip(insts, sols))[::5]:  # alone == in the mix
        alone = engine.solve(x.request())
        assert alone.status == ga.OK and alone.cost == sol.cost


def test_batch_config4_full_matches_golden(engine):
    # 1024 mixed instances in one batch; costs against the committed oracle fixture
    gold = json.load(open(os.path.join(GOLDEN, "config4.json")))["instances"]
    insts = wl.config4(4, 1024, threads=16)
    cfg = __import__("cuda_a_star_amd.tuning", fromlist=["load"]).load()
    reqs = [x.request(num_queues=cfg.heaps(x.kind, x.side)) for x in insts]
    sols = engine.solve_batch(reqs)
    bad = [(i, s.status, s.cost, g["cost"]) for i, (s, g) in enumerate(zip(sols, gold))
           if s.status != ga.OK or s.cost != g["cost"]]
    assert not bad, bad[:10]
    for i in range(0, 1024, 97):  # replay a spread of paths
        assert replay_ok(insts[i], sols[i])


@pytest.mark.parametrize("grow", [1, 0], ids=["in-place", "restart"])
def test_batch_forced_retries(grow):
    # tiny heaps and closed sets: with in-place growth every search pauses, grows
    # and resumes with its state (no restart); without, it restarts larger
    insts = small_mix(seed=9, n_puz=8, n_grid=4)
    reqs = []
    for x in insts:
        r = x.request(heap_capacity=16)
        if x.kind == "puzzle":
            r.table_slots = 1 << 12
        reqs.append(r)
    with ga.Engine(0, grow_in_place=grow, max_iterations=2_000_000) as e:
        sols = e.solve_batch(reqs)
    if grow:
        assert any(s.grows > 0 for s in sols) and all(s.retries == 0 for s in sols)
    else:
        assert any(s.retries > 0 for s in sols) and all(s.grows == 0 for s in sols)
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK and sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)


def test_overflow_grows_both_structures():
    """Searches that outgrow both their heaps and their closed set: a retry grows
    the structure that overflowed 4x and the other one 2x when it was more than a
    quarter full, then the heaps at least to the closed set's load limit, so the
    default retry budget still reaches the optimum."""
    insts = [wl.puzzle15(wl.instance_seed(61, i), depth=48, num_queues=256, table_slots=0) for i in range(4)]
    reqs = []
    for x in insts:
        r = x.request(heap_capacity=16)
        r.table_slots = 1 << 10
        reqs.append(r)
    with ga.Engine(0, max_retries=6, max_iterations=2_000_000, grow_in_place=0) as e:
        sols = e.solve_batch(reqs)
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK, ga.status_str(sol.status)
        assert sol.cost == oracle_cost(x)
        assert sol.retries >= 2
        assert replay_ok(x, sol)


def test_in_place_growth_carries_the_search_over():
    """Heaps of 16 entries and a closed set of 1024 slots grow in place several
    times (both structures, hence the closed-set rehash with renumbered heap
    entries, records and parent links): the cost is exact, the path replays,
    and the search does about the work of one with large structures from the
    start, where a restart repeats it."""
    insts = [wl.puzzle15(wl.instance_seed(61, i), depth=48, num_queues=256, table_slots=0) for i in range(4)]
    small, big = [], []
    for x in insts:
        r = x.request(heap_capacity=16)
        r.table_slots = 1 << 10
        small.append(r)
        big.append(x.request(heap_capacity=1 << 14, table_slots=1 << 24))
    with ga.Engine(0, max_iterations=2_000_000) as e:
        grown = e.solve_batch(small)
        ref = e.solve_batch(big)
    with ga.Engine(0, max_iterations=2_000_000, grow_in_place=0) as e:
        restarted = e.solve_batch(small)
    for x, g, r, rs in zip(insts, grown, ref, restarted):
        assert g.status == r.status == rs.status == ga.OK
        assert g.cost == r.cost == rs.cost == oracle_cost(x)
        assert g.grows >= 2 and g.retries == 0 and r.grows == 0 and rs.retries >= 2
        assert replay_ok(x, g)
    # Expansion counts of GA* vary from run to run (successors generated twice in one
    # iteration race for the closed-set slot, and the winner decides which heap holds
    # the state), so the work is compared over the four searches: growing in place
    # does about the work of large structures from the start, a restart repeats it
    # (a result counts the expansions of every attempt).
    g_tot, r_tot, rs_tot = (sum(s.expanded for s in v) for v in (grown, ref, restarted))
    assert g_tot <= 1.3 * r_tot
    assert rs_tot > g_tot


def test_batch_handles_invalid_and_unsolvable(engine):
    g = ga.puzzle_goal(4)
    t = ga.state_to_board(g, 4)
    t[0], t[1] = t[1], t[0]
    bad_parity = ga.puzzle_request(4, ga.board_to_state(t), g, 256)
    invalid = ga.puzzle_request(4, 0x1111111111111111, g, 256)
    good = wl.puzzle15(1, depth=30, num_queues=256, table_slots=0)
    sols = engine.solve_batch([bad_parity, invalid, good.request()])
    assert sols[0].status == ga.NO_PATH
    assert sols[1].status == ga.E_INVALID
    assert sols[2].status == ga.OK and sols[2].cost == oracle_cost(good)


def serving_stream(n=64, qps=400.0, seed=21):
    st = wl.Stream(seed)
    out, t = [], 0.0
    for i in range(n):
        if st.uniform() < 0.6:
            x = wl.puzzle15(wl.instance_seed(seed, i), depth=st.uniform_int(30, 50), num_queues=256, table_slots=0)
        else:
            x = wl.grid(st.uniform_int(128, 512), 0.2, wl.instance_seed(seed, i), num_queues=512)
        t += -np.log(1.0 - st.uniform()) / qps
        x.arrival_ns = int(t * 1e9)
        out.append(x)
    return out


def test_serve_stream_matches_alone(engine):
    insts = serving_stream()
    sols, ss = engine.serve([x.request() for x in insts], [x.arrival_ns for x in insts])
    assert ss.completed == len(insts) and ss.failed == 0
    assert ss.max_inflight > 1          # requests overlapped on the device
    assert 0 < ss.p50_ms <= ss.p99_ms
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK
        assert sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)
        assert sol.latency_ns > 0
    for x, sol in list(zip(insts, sols))[::8]:
        alone = engine.solve(x.request())
        assert alone.cost == sol.cost


@pytest.mark.parametrize("lanes", [2, 3])
def test_serve_stream_on_several_lanes(lanes):
    """Large grids on their own launch lanes: results equal the oracle, and
    requests still overlap."""
    insts = serving_stream(n=48, qps=800.0, seed=40 + lanes)
    with ga.Engine(0, lanes=lanes, deep_cells=1 << 16, rounds=2, max_iterations=2_000_000) as e:
        sols, ss = e.serve([x.request() for x in insts], [x.arrival_ns for x in insts])
    assert ss.completed == len(insts) and ss.max_inflight > 1
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK and sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)


def test_serve_tight_budget_admits_in_turn():
    insts = serving_stream(n=24, qps=5000.0, seed=5)
    with ga.Engine(0, budget_bytes=512 << 20, chunk_bytes=256 << 20, max_iterations=2_000_000) as e:
        sols, ss = e.serve([x.request() for x in insts], [x.arrival_ns for x in insts])
        assert ss.completed == len(insts)
        st = e.stats()
        assert st.peak_bytes <= 768 << 20
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK and sol.cost == oracle_cost(x)


def test_serve_waiting_lane_does_not_hold_up_other_lanes():
    """Admission is first come, first served per launch lane: a grid that
    waits for its lane's only slot does not keep the puzzles behind it off the
    other lane; within a lane the order of arrival holds."""
    grids = [wl.grid(256, 0.2, 300 + i, num_queues=256) for i in range(2)]
    puzzles = [wl.puzzle15(wl.instance_seed(31, i), depth=30, num_queues=256, table_slots=1 << 16) for i in range(4)]
    insts = grids + puzzles  # all arrive at once, the grids first
    # lanes=2 plus the 24-puzzle lane, one slot each; grids of >= 2^16 cells take lane 1
    with ga.Engine(0, lanes=2, deep_cells=1 << 16, max_instances=3, max_iterations=2_000_000) as e:
        sols, ss = e.serve([x.request() for x in insts], [0] * len(insts))
    assert ss.completed == len(insts)
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK and sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)
    admitted = [s.latency_ns - s.search_ns for s in sols]  # arrival 0: admission time since the start
    # the second grid waits for the first to leave lane 1; the first puzzle does not wait for it
    assert admitted[1] >= admitted[0] + sols[0].search_ns
    assert admitted[2] < admitted[1]
    # lane 0 admits its puzzles one at a time, in arrival order
    assert admitted[2] <= admitted[3] <= admitted[4] <= admitted[5]


def test_serve_request_over_budget_is_rejected():
    x = wl.grid(512, 0.2, 1, num_queues=256)
    with ga.Engine(0, budget_bytes=1 << 20, chunk_bytes=1 << 20) as e:
        sols, ss = e.serve([x.request()], [0])
    assert sols[0].status == ga.E_BUDGET


VARIANTS = [
    dict(heap_arity=4, rounds=1, use_graph=1, pops_per_iter=1),
    dict(heap_arity=8, rounds=1, use_graph=0, pops_per_iter=2),
    dict(heap_arity=8, rounds=2, use_graph=0, pops_per_iter=1, chunk_iters=5),
    dict(heap_arity=4, rounds=4, use_graph=0, pops_per_iter=2),
    dict(heap_arity=8, rounds=8, use_graph=1, pops_per_iter=4),
    dict(heap_arity=4, rounds=2, use_graph=1, pops_per_iter=2),
    dict(heap_arity=4, rounds=1, pops_per_iter=2, adaptive_start=16, grow_thresh=2),
    dict(heap_arity=8, rounds=2, use_graph=1, adaptive_start=64, grow_thresh=4),
    dict(heap_arity=8, rounds=1, adaptive_start=1, grow_thresh=1),
    dict(heap_arity=8, rounds=2, adaptive_start=256, grow_thresh=32),
    dict(heap_arity=4, rounds=2, rounds_grid=16),
    dict(heap_arity=8, rounds=1, rounds_grid=8, adaptive_start=32, grow_thresh=8),
    dict(heap_arity=4, rounds=2, lanes=2, deep_cells=1 << 14),
    dict(heap_arity=8, rounds=2, lanes=3, deep_cells=1 << 16, use_graph=1),
    dict(heap_arity=4, rounds=1, lanes=4, deep_cells=1 << 15, chunk_iters=3),
    dict(heap_arity=4, rounds=2, slack_puzzle=1, slack_grid=1),
    dict(heap_arity=8, rounds=1, slack_puzzle=3, slack_grid=5, lanes=2, deep_cells=1 << 14),
    dict(heap_arity=4, rounds=2, slack_grid=2, adaptive_start=16, grow_thresh=4),
    dict(heap_arity=4, rounds=1, block=128, use_graph=1),
    dict(heap_arity=4, rounds=2, block=512, lanes=2, deep_cells=1 << 14),
    dict(heap_arity=8, rounds=1, block=512, use_graph=1, slack_puzzle=3),
    dict(heap_arity=8, rounds=2, block=128, adaptive_start=16, grow_thresh=4),
    # paired pops on the 4-connected grid lane (lane 1 holds grids of >= 2^14 cells)
    dict(heap_arity=4, rounds=1, rounds_grid=2, step_flags=4, lanes=2, deep_cells=1 << 14),
    dict(heap_arity=4, rounds=1, rounds_grid=3, step_flags=6, lanes=2, deep_cells=1 << 14, use_graph=1,
         slack_grid=8),
    # ... and on a lane that holds puzzles and grids together (the all-kinds kernel with two pops)
    dict(heap_arity=4, rounds=1, rounds_grid=3, step_flags=12, use_graph=1, slack_grid=8),
    dict(heap_arity=4, rounds=2, rounds_grid=2, step_flags=13, block=128),
]


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: "-".join(f"{k}{x}" for k, x in v.items()))
def test_engine_variants_match_oracle(variant):
    """Every tunable combination the autotuner may pick keeps costs exact,
    including a frontier width that starts small and doubles on the device."""
    insts = small_mix(seed=31 + variant["heap_arity"] + variant["rounds"] + variant.get("adaptive_start", 0),
                      n_puz=12, n_grid=8)
    with ga.Engine(0, max_iterations=2_000_000, **variant) as e:
        sols = e.solve_batch([x.request() for x in insts])
        alone = [e.solve(insts[i].request()) for i in (0, len(insts) - 1)]
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK, ga.status_str(sol.status)
        assert sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)
    assert alone[0].cost == sols[0].cost and alone[1].cost == sols[-1].cost


@pytest.mark.parametrize("use_graph", [0, 1])
def test_lane_kernel_follows_resident_kinds(use_graph):
    """A lane runs a single-kind step kernel while it holds one kind and the
    all-kinds kernel otherwise; arrivals in phases (8-puzzles, then 15-puzzles
    and grids together, then grids alone) make the lanes switch back and forth,
    with and without graph replay."""
    st = wl.Stream(77)
    phases = []
    for i in range(10):
        p8 = wl.puzzle8(wl.instance_seed(70, i))
        p8.num_queues = 64
        phases.append(p8)
    for i in range(12):
        phases.append(wl.puzzle15(wl.instance_seed(71, i), depth=st.uniform_int(30, 45), num_queues=256,
                                  table_slots=0))
        phases.append(wl.grid(st.uniform_int(96, 320), 0.2, wl.instance_seed(72, i), num_queues=256))
    for i in range(8):
        phases.append(wl.grid(st.uniform_int(96, 320), 0.2, wl.instance_seed(73, i), num_queues=256))
    arrivals = [k * 3_000_000 for k in range(len(phases))]
    with ga.Engine(0, lanes=2, deep_cells=1 << 14, use_graph=use_graph, chunk_iters=4,
                   max_iterations=2_000_000) as e:
        sols, ss = e.serve([x.request() for x in phases], arrivals)
    assert ss.completed == len(phases) and ss.failed == 0
    for x, sol in zip(phases, sols):
        assert sol.status == ga.OK, ga.status_str(sol.status)
        assert sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)
