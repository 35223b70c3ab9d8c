"""Engine profiling on the GPU: profile = 1 counts and times every step,
profile = n > 1 samples every n-th chunk (HIP events and the profile-only
device counters in its steps only), profile = 0 leaves both off.  The search
itself must not depend on it."""
import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _solve(profile, chunk_iters=4):
    g = wl.grid(256, 0.2, wl.instance_seed(11, 0))
    with ga.Engine(0, profile=profile, chunk_iters=chunk_iters, use_graph=1, lanes=1) as e:
        sol = e.solve(g.request(num_queues=256))
        return g, sol, e.stats()


def test_profile_off_counts_nothing():
    g, sol, st = _solve(0)
    assert sol.status == ga.OK
    assert sol.cost == oracle.dial_grid(g.cells, g.side, g.side, g.start, g.goal, 4)
    assert st.step_launches > 0 and st.expanded > 0  # always counted
    assert st.timed_launches == 0 and st.pops == 0 and st.lookups == 0 and st.heap_groups == 0


def test_profile_every_chunk_counts_every_launch():
    g, sol, st = _solve(1)
    assert sol.status == ga.OK
    assert st.timed_launches == st.counted_launches == st.step_launches
    assert st.pops >= st.expanded > 0
    assert 0 < st.work_groups <= st.heap_groups


def test_profile_sampled_chunks_scale_to_every_launch():
    _, full, st1 = _solve(1)
    _, samp, st4 = _solve(4)
    assert samp.status == ga.OK and samp.cost == full.cost
    # a chunk in four is timed (the first chunk of the lane always is), and the
    # chunk half-way between two timed ones is counted
    assert 0 < st4.timed_launches < st4.step_launches
    assert 0 < st4.counted_launches < st4.step_launches
    assert st4.timed_launches + st4.counted_launches <= st4.step_launches
    assert st4.timed_launches % 4 == 0 and st4.counted_launches % 4 == 0  # whole chunks of 4 launches
    assert st4.step_ms_total > 0
    # counters of the counted launches only, yet covering real work
    assert 0 < st4.pops < st1.pops
    assert 0 < st4.work_groups <= st4.heap_groups
    # algorithmic bytes are scaled from the sample to all launches: same order as counting everything
    assert 0.5 * st1.algo_bytes < st4.algo_bytes < 2.0 * st1.algo_bytes


def test_atomic_rate_calibration():
    # random 64-bit atomics over 256 MiB: a positive rate, and bad kinds refused
    with ga.Engine(0) as e:
        for kind in (0, 1):
            r = e.atomic_rate(1 << 28, 1 << 22, kind)
            assert 1e8 < r < 1e12, (kind, r)
        with pytest.raises(ga.GAStarError):
            e.atomic_rate(1 << 28, 1 << 22, 7)


def test_profile_counts_hbm_levels_and_lanes():
    # the algorithmic-byte counters: levels served from the LDS copy of a heap's
    # top are not HBM levels, records from the candidate buffer are a part of all
    # pushes, and the per-lane figures add up to the engine's totals
    x = wl.puzzle15(5, depth=46, num_queues=256, table_slots=1 << 20)
    with ga.Engine(0, profile=1, chunk_iters=4, lanes=1) as e:
        sol = e.solve(x.request())
        st = e.stats()
    assert sol.status == ga.OK
    assert 0 < st.sift_down_hbm <= st.sift_down and st.sift_up_hbm <= st.sift_up
    assert 0 < st.push_in <= st.pushed
    assert st.lane_launches[0] == st.step_launches and st.lane_timed_launches[0] == st.timed_launches
    assert st.lane_counted_launches[0] == st.counted_launches == st.step_launches
    assert sum(st.lane_algo_bytes) == pytest.approx(st.algo_bytes, rel=1e-9)
    assert st.lane_step_ms[0] == pytest.approx(st.step_ms_total, rel=1e-9) and st.step_ms_total > 0
