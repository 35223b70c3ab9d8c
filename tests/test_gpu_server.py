"""Online serving on the GPU (ga_serve_start / submit / poll / drain / stop):
requests submitted while others are resident are admitted into the same
launches; every cost must equal the CPU oracle's (tolerance 0), every path
must replay, and each result must be the same alone as in the mix."""
import time

import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import tuning
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


def mixed(seed, n):
    st = wl.Stream(seed)
    out = []
    for i in range(n):
        if st.uniform() < 0.6:
            out.append(wl.puzzle15(wl.instance_seed(seed, i), depth=st.uniform_int(30, 55), num_queues=1024,
                                   table_slots=0))
        else:
            out.append(wl.grid(st.uniform_int(256, 1200), 0.2, wl.instance_seed(seed, i), num_queues=1024))
    return out


def test_server_admits_submissions_while_others_run():
    first, second, third = mixed(81, 12), mixed(82, 12), mixed(83, 8)
    cfg = tuning.load("served")
    with ga.Engine(0, budget_bytes=64 << 30, max_iterations=2_000_000, **cfg.engine) as e:
        with e.server() as srv:
            t1 = [srv.submit(x.request()) for x in first]
            time.sleep(0.01)
            # the second wave arrives while requests of the first are still on the device
            # (a poll that returns a result forgets the ticket: keep what it returns)
            got = {}
            for t in t1:
                s = srv.poll(t)
                if s is not None:
                    got[t] = s
            pending_at_second = len(t1) - len(got)
            t2 = [srv.submit(x.request()) for x in second]
            # poll until some have finished, then submit the third wave
            deadline = time.time() + 60
            n_first = len(got)
            while len(got) == n_first and time.time() < deadline:
                for t in t1 + t2:
                    if t not in got:
                        s = srv.poll(t)
                        if s is not None:
                            got[t] = s
                time.sleep(0.002)
            t3 = [srv.submit(x.request()) for x in third]
            srv.drain()
            for t in t1 + t2 + t3:
                if t not in got:
                    got[t] = srv.poll(t)
            ss = srv.status()
            with pytest.raises(ga.GAStarError):  # a ticket is forgotten once polled
                srv.poll(t1[0])
        # the engine is free again once the server has stopped
        alone = {i: e.solve(x.request()) for i, x in enumerate(first + second + third) if i % 5 == 0}
    insts = first + second + third
    sols = [got[t] for t in t1 + t2 + t3]
    assert pending_at_second > 0  # submissions overlapped with resident requests
    assert ss.completed == len(insts) and ss.failed == 0 and ss.max_inflight > 1
    assert 0 < ss.p50_ms <= ss.p99_ms
    for x, sol in zip(insts, sols):
        assert sol.status == ga.OK, ga.status_str(sol.status)
        assert sol.cost == oracle_cost(x)
        assert replay_ok(x, sol)
        assert sol.latency_ns > 0
    for i, a in alone.items():
        assert a.status == ga.OK and a.cost == sols[i].cost, i


def test_server_checks_requests_and_owns_the_engine():
    with ga.Engine(0, max_iterations=2_000_000) as e:
        srv = e.server()
        with pytest.raises(ga.GAStarError):  # one server per engine
            e.server()
        x = wl.puzzle8(3)
        with pytest.raises(ga.GAStarError):  # the engine's other entry points wait for the server
            e.solve(x.request())
        bad = srv.submit(ga.puzzle_request(4, 0x1111111111111111))
        t = ga.state_to_board(ga.puzzle_goal(4), 4)
        t[0], t[1] = t[1], t[0]
        unsolvable = srv.submit(ga.puzzle_request(4, ga.board_to_state(t)))
        good = srv.submit(x.request())
        b = ga.gen_board(5, 24, 4)
        board = srv.submit(ga.puzzle5_request(b, num_queues=1024))  # its boards are copied at submission
        srv.drain()
        assert srv.poll(bad).status == ga.E_INVALID
        assert srv.poll(unsolvable).status == ga.NO_PATH
        sol = srv.poll(good)
        _, ref8, _ = oracle.astar_puzzle(3, x.start, x.goal, want_path=False)
        assert sol.status == ga.OK and sol.cost == ref8.cost
        sol = srv.poll(board)
        rc, ref, _ = oracle.astar_board(5, b, ga.board_goal(5), want_path=False)
        assert sol.status == ga.OK and sol.cost == ref.cost
        assert oracle.replay_board(5, sol.path, b, ga.board_goal(5)) == sol.cost
        srv.stop()
        assert e.solve(x.request()).cost == ref8.cost  # the engine is usable again after stop
