"""The batched and served benchmark configurations on the GPU, under exactly
the engine configuration that bench.py times: the tuned "served" parameters
(two launch lanes, chunks replayed from captured graphs, f-slack gating, the
tuned heap counts), a 200 GB admission budget, sampled profiling, grids
resident in HBM.  Costs must equal the committed oracle fixtures (integer,
tolerance 0), a spread of paths must replay move by move, and a sample of
requests must give the same cost alone as in the mix."""
import json
import os

import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import tuning
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
BUDGET = 200 << 30


def bench_engine(cfg):
    # the engine of bench.py's served leg (bench.py main: ga.Engine(...))
    return ga.Engine(0, budget_bytes=BUDGET, chunk_bytes=8 << 30, profile=8, max_iterations=1_000_000, **cfg.engine)


def replay_ok(x, sol):
    if x.kind == "puzzle":
        return oracle.replay_puzzle(x.side, sol.path, x.start, x.goal) == sol.cost
    return oracle.replay_grid(x.cells, x.side, x.side, 4, sol.path, x.start, x.goal) == sol.cost


def test_served_config5_stream_matches_golden():
    # the serving stream of the benchmark (60 % 15-puzzles of walk depth 30..70,
    # 40 % grids of side 512..4096 with 20 % walls, Poisson arrivals at 1500 QPS):
    # its first 256 requests, served as bench.py serves them
    gold = json.load(open(os.path.join(GOLDEN, "config5.json")))["instances"]
    insts = wl.config5(5, 256, 1500.0, threads=16)
    cfg = tuning.load("served")
    assert cfg.use_graph == 1 and cfg.lanes == 2  # the configuration under test is the timed one
    with bench_engine(cfg) as e:
        dev = {i: e.upload(x.cells) for i, x in enumerate(insts) if x.kind == "grid"}
        reqs = [x.request(cells_device=dev.get(i), num_queues=cfg.heaps(x.kind, x.side)) for i, x in enumerate(insts)]
        sols, ss = e.serve(reqs, [x.arrival_ns for x in insts])
        for a in dev.values():
            e.free_device(a)
    assert ss.completed == len(insts) and ss.failed == 0
    assert ss.max_inflight > 1  # requests overlapped on the device
    bad = [(i, ga.status_str(s.status), s.cost, g["cost"]) for i, (s, g) in enumerate(zip(sols, gold))
           if s.status != ga.OK or s.cost != g["cost"]]
    assert not bad, bad[:10]
    for i in range(0, len(insts), 17):  # a spread of paths, both kinds
        assert replay_ok(insts[i], sols[i]), i
    # alone == in the mix, on the one-at-a-time configuration
    plain = tuning.load("plain")
    with ga.Engine(0, budget_bytes=BUDGET, chunk_bytes=8 << 30, max_iterations=1_000_000, **plain.engine) as e1:
        for i in range(3, len(insts), 41):
            x = insts[i]
            alone = e1.solve(x.request(num_queues=plain.heaps(x.kind, x.side)))
            assert alone.status == ga.OK and alone.cost == sols[i].cost, i


def test_served_big_grids_expand_near_the_sequential_search():
    # the serving stream's grids of at least deep_cells cells (the grid lane's), under the served
    # tuning: with weights 1..9 and Manhattan x w_min, the oracle's sequential A* expands about 80 % of
    # their cells (f < C* covers nearly the whole grid). GA* with its heaps per big grid and two pops
    # per round must give the same cost within 1.25x the sequential expansions (1.10-1.11 measured)
    cfg = tuning.load("served")
    insts = [x for x in wl.config5(5, 1024, 4000.0, threads=16)
             if x.kind == "grid" and x.side * x.side >= cfg.deep_cells][:4]
    assert len(insts) == 4 and all(x.side >= 1448 for x in insts)
    with bench_engine(cfg) as e:
        sols = e.solve_batch([x.request(num_queues=cfg.heaps(x.kind, x.side)) for x in insts])
        st = e.stats()
    # every request ran on the grid lane, whose closed sets are dense: one atomicMin per lookup, nothing per probe
    assert st.lookups > 0 and st.lane_atomics[cfg.lanes - 1] == st.lookups
    for i, (x, sol) in enumerate(zip(insts, sols)):
        rc, ref, _ = oracle.astar_grid(x.cells, x.side, x.side, x.start, x.goal, 4, want_path=False)
        assert rc == oracle.OK and ref.expanded > 0.7 * x.side * x.side, (i, x.side, ref.expanded)
        assert sol.status == ga.OK and sol.cost == ref.cost, (i, x.side)
        ratio = sol.expanded / ref.expanded
        print(f"grid {x.side}^2: oracle A* {ref.expanded} expansions, GA* {sol.expanded} ({ratio:.3f}x)")
        assert 0.9 <= ratio <= 1.25, (i, x.side, ratio)


def test_batched_config4_under_served_tuning_matches_golden():
    # 1024 mixed instances in one ga_solve_batch on the graph-replayed served engine
    gold = json.load(open(os.path.join(GOLDEN, "config4.json")))["instances"]
    insts = wl.config4(4, 1024, threads=16)
    cfg = tuning.load("served")
    with bench_engine(cfg) as e:
        dev = {i: e.upload(x.cells) for i, x in enumerate(insts) if x.kind == "grid"}
        reqs = [x.request(cells_device=dev.get(i), num_queues=cfg.heaps(x.kind, x.side)) for i, x in enumerate(insts)]
        sols = e.solve_batch(reqs)
        st = e.stats()
        for a in dev.values():
            e.free_device(a)
    bad = [(i, ga.status_str(s.status), s.cost, g["cost"]) for i, (s, g) in enumerate(zip(sols, gold))
           if s.status != ga.OK or s.cost != g["cost"]]
    assert not bad, bad[:10]
    assert st.timed_launches > 0 and st.algo_bytes > 0  # the sampled profiling ran
    for i in range(0, 1024, 61):
        assert replay_ok(insts[i], sols[i]), i


def test_batched_config4_one_launch_matches_golden():
    # the same 1024 instances with the one-launch form: every puzzle and grid (up to 2048^2) runs
    # all of its iterations inside one launch, one workgroup of 1024 threads per instance
    gold = json.load(open(os.path.join(GOLDEN, "config4.json")))["instances"]
    insts = wl.config4(4, 1024, threads=16)
    cfg = tuning.load("served")
    with ga.Engine(0, budget_bytes=BUDGET, chunk_bytes=8 << 30, max_iterations=1_000_000, persist=1,
                   persist_block=1024, persist_cells=2048 * 2048, **cfg.engine) as e:
        dev = {i: e.upload(x.cells) for i, x in enumerate(insts) if x.kind == "grid"}
        reqs = [x.request(cells_device=dev.get(i)) for i, x in enumerate(insts)]
        sols = e.solve_batch(reqs)
        st = e.stats()
        for a in dev.values():
            e.free_device(a)
    assert st.persist_instances >= 1024
    bad = [(i, ga.status_str(s.status), s.cost, g["cost"]) for i, (s, g) in enumerate(zip(sols, gold))
           if s.status != ga.OK or s.cost != g["cost"]]
    assert not bad, bad[:10]
    for i in range(0, 1024, 61):
        assert replay_ok(insts[i], sols[i]), i
