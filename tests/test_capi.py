"""The C ABI library loads on a CPU-only host, exports every symbol that
include/gastar.h declares, and its struct layouts match the ctypes mirror.
Only host-side entry points are called here (generators, status strings)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cuda_a_star_amd as ga

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gastar.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ga_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_match_python_mirror():
    assert declared_functions() == sorted(ga.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ga.lib()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing


def test_atomic_rate_rejects_bad_arguments_before_any_device_call():
    out = ctypes.c_double()
    L = ga.lib()
    assert L.ga_atomic_rate(None, 1 << 20, 1 << 10, 0, ctypes.byref(out)) == -1  # no engine
    assert L.ga_atomic_rate(None, 1 << 20, 1 << 10, 3, ctypes.byref(out)) == -1  # no such kind


def test_structure_checks_reject_bad_arguments_before_any_device_call():
    from cuda_a_star_amd import checks as ck
    L = ga.lib()
    keys = np.ones((4, 4), dtype=np.uint32)
    g = np.zeros(4, dtype=np.uint32)
    out = np.zeros(4, dtype=np.int32)
    slot = np.zeros(4, dtype=np.uint32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    cs = lambda layout, log2=4, probes=8, serial=0, k=keys, n=4: L.ga_debug_closed_set(
        0, layout, p(k), p(g), p(g), n, log2, probes, serial, p(out), p(slot), None, None, None)
    assert cs(3) == -1  # no such layout
    assert cs(ck.SET_U64, log2=0) == -1 and cs(ck.SET_U64, log2=25) == -1
    assert cs(ck.SET_U64, probes=0) == -1 and cs(ck.SET_U64, serial=2) == -1
    assert L.ga_debug_closed_set(0, 0, None, None, None, 4, 4, 8, 0, None, None, None, None, None) == -1
    empty = np.full((4, 2), 0xFFFFFFFF, dtype=np.uint32)
    assert cs(ck.SET_U64, k=empty) == -1  # the empty key
    unwritten = keys.copy()
    unwritten[2, 3] = 0x80000000
    assert cs(ck.SET_BOARD, k=unwritten) == -1  # a board's hi never has its top bit set
    ops = np.array([1 << 32, ck.HOP_POP], dtype=np.uint64)
    buf = np.zeros(64, dtype=np.uint64)
    b32 = np.zeros(64, dtype=np.uint32)
    hp = lambda arity=4, nh=1, o=ops, opl=1, cap=8, track=0: L.ga_debug_heap(
        0, arity, nh, p(o), len(o), opl, cap, track, p(buf), p(b32), p(buf), p(b32), p(b32))
    assert hp(arity=32) == -1 and hp(arity=2) == -1
    assert hp(nh=0) == -1 and hp(nh=17) == -1 and hp(arity=8, nh=9) == -1 and hp(arity=16, nh=5) == -1
    assert hp(cap=0) == -1 and hp(cap=(1 << 20) + 1) == -1 and hp(opl=0) == -1 and hp(track=2) == -1
    assert hp(arity=16, track=1) == -1  # no dirty mask above 32 top entries
    assert hp(arity=64) == -1  # the wave-wide heap is switched off
    assert hp(o=np.array([0xFFFFFFFF << 32], dtype=np.uint64)) == -1  # f = the infinite cost
    with pytest.raises(ga.GAStarError):
        ck.heap(32, ops, 1, 8)
    out = ctypes.c_double()
    hr = lambda arity=4, waves=8, fill=16, iters=1, o=ctypes.byref(out): L.ga_heap_rate(0, arity, waves, fill, iters, o)
    assert hr(arity=2) == -1 and hr(arity=32) == -1 and hr(arity=64) == -1 and hr(waves=0) == -1 and hr(waves=6) == -1
    assert hr(waves=(1 << 16) + 4) == -1 and hr(fill=1) == -1 and hr(fill=(1 << 16) + 1) == -1
    assert hr(iters=0) == -1 and hr(iters=(1 << 16) + 1) == -1 and hr(o=None) == -1
    assert hr(waves=1 << 16, fill=1 << 16) == -1  # above 16 GiB of heaps
    with pytest.raises(ga.GAStarError):
        ck.heap_rate(32)


def test_status_strings():
    for code, name in [(0, "GA_OK"), (1, "GA_NO_PATH"), (-1, "GA_E_INVALID"), (-2, "GA_E_NOMEM"),
                       (-3, "GA_E_BUDGET"), (-4, "GA_E_HIP"), (-5, "GA_E_TIMEOUT"), (-6, "GA_E_OVERFLOW")]:
        assert ga.status_str(code) == name


LAYOUT_PROGRAM = r"""
#include <stdio.h>
#include <stddef.h>
#include "gastar.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
#define S(T) printf(#T " %zu\n", sizeof(T));
int main(void) {
  S(ga_request) F(ga_request, start) F(ga_request, cells) F(ga_request, cells_device)
  F(ga_request, num_queues) F(ga_request, table_slots) F(ga_request, user_tag)
  S(ga_result) F(ga_result, expanded) F(ga_result, path) F(ga_result, latency_ns) F(ga_result, user_tag)
  S(ga_config) F(ga_config, max_instances) F(ga_config, profile) F(ga_config, rounds) F(ga_config, grow_thresh) F(ga_config, rounds_grid) F(ga_config, lanes) F(ga_config, deep_cells) F(ga_config, slack_grid) F(ga_config, pdb24) F(ga_config, grow_in_place) F(ga_config, block)
  S(ga_engine_stats) F(ga_engine_stats, step_ms_total) F(ga_engine_stats, peak_bytes) F(ga_engine_stats, timed_launches) F(ga_engine_stats, heap_groups) F(ga_engine_stats, work_groups) F(ga_engine_stats, push_in) F(ga_engine_stats, counted_launches) F(ga_engine_stats, lane_launches) F(ga_engine_stats, lane_counted_launches) F(ga_engine_stats, lane_step_ms) F(ga_engine_stats, lane_algo_bytes) F(ga_engine_stats, lane_atomics) F(ga_engine_stats, lane_busy_ms) F(ga_engine_stats, host_admit_ms) F(ga_engine_stats, host_path_ms)
  S(ga_serve_stats) F(ga_serve_stats, solves_per_sec) F(ga_serve_stats, max_inflight)
  return 0;
}
"""


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    src.write_text(LAYOUT_PROGRAM)
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    mirror = {"ga_request": ga.Request, "ga_result": ga.Result, "ga_config": ga.Config,
              "ga_engine_stats": ga.EngineStats, "ga_serve_stats": ga.ServeStats}
    for line in filter(None, out):
        name, val = line.split()
        if "." in name:
            t, m = name.split(".")
            assert getattr(mirror[t], m).offset == int(val), line
        else:
            assert ctypes.sizeof(mirror[name]) == int(val), line


def test_splitmix64_known_values():
    # splitmix64 from state 0: the published first outputs of the generator
    v1, s = ga.splitmix64(0)
    v2, s = ga.splitmix64(s)
    assert v1 == 0xE220A8397B1DCDAF
    assert v2 == 0x6E789E6AA1B965F4


def test_puzzle_generator_is_deterministic_and_solvable():
    for side in (3, 4):
        goal = ga.puzzle_goal(side)
        for seed in range(20):
            a = ga.gen_puzzle(side, 40, seed)
            assert a == ga.gen_puzzle(side, 40, seed)
            assert ga.puzzle_solvable(side, a, goal)
        assert ga.gen_puzzle(side, 0, 7) == goal


def test_puzzle_solvable_parity():
    g = ga.puzzle_goal(4)
    t = ga.state_to_board(g, 4)
    t[0], t[1] = t[1], t[0]
    assert not ga.puzzle_solvable(4, ga.board_to_state(t), g)
    t[2], t[3] = t[3], t[2]
    assert ga.puzzle_solvable(4, ga.board_to_state(t), g)


def test_grid_generator_properties():
    cells, draws = ga.gen_grid(200, 150, 0.2, 42)
    again, _ = ga.gen_grid(200, 150, 0.2, 42)
    assert np.array_equal(cells, again) and draws >= 1
    assert cells[0] != 0 and cells[-1] != 0
    open_w = cells[cells != 0]
    assert open_w.min() >= 1 and open_w.max() <= 9
    assert abs((cells == 0).mean() - 0.2) < 0.02
    with pytest.raises(ValueError):
        ga.gen_grid(0, 10, 0.2, 1)


def test_engine_refuses_without_gpu_or_builds():
    """On a host without a GPU, engine creation reports GA_E_HIP instead of crashing."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(ga.GAStarError) as ei:
        ga.Engine(0)
    assert ei.value.code == ga.E_HIP


def test_workload_requests_build_on_cpu():
    from cuda_a_star_amd import tuning
    from cuda_a_star_amd import workloads as wl

    cfg = tuning.load()
    insts = wl.config5(5, 24, 1500.0, threads=4)
    for x in insts:
        r = x.request(num_queues=cfg.heaps(x.kind, x.side), cells_device=None)
        assert r.num_queues == cfg.heaps(x.kind, x.side)
        assert r.kind in (ga.PUZZLE, ga.GRID)
        r2 = x.request(heap_capacity=16)
        assert r2.heap_capacity == 16
    assert [x.arrival_ns for x in insts] == sorted(x.arrival_ns for x in insts)
    p = wl.puzzle15(3, num_queues=512, table_slots=1 << 20)
    assert p.request().table_slots == 1 << 20 and p.request(table_slots=1 << 12).table_slots == 1 << 12


def test_gen_board_equals_gen_puzzle_on_sides_3_and_4():
    """ga_gen_board follows ga_gen_puzzle draw for draw."""
    for side in (3, 4):
        for seed in range(6):
            for depth in (0, 9, 31, 60):
                for nb in (True, False):
                    b = ga.gen_board(side, depth, seed, no_backtrack=nb)
                    assert list(b) == ga.state_to_board(ga.gen_puzzle(side, depth, seed, no_backtrack=nb), side)


def test_board_solvable_matches_puzzle_solvable_and_parity():
    for side in (3, 4):
        g = ga.puzzle_goal(side)
        for seed in range(8):
            s = ga.gen_puzzle(side, 30, seed)
            t = ga.state_to_board(s, side)
            a, b = [i for i, x in enumerate(t) if x][:2]
            t[a], t[b] = t[b], t[a]  # one transposition of tiles flips reachability
            for board in (ga.state_to_board(s, side), t):
                assert ga.board_solvable(side, board, ga.state_to_board(g, side)) == \
                    ga.puzzle_solvable(side, ga.board_to_state(board), g)
    for seed in range(8):
        b = ga.gen_board(5, 25 + seed, seed)
        assert ga.board_solvable(5, b, ga.board_goal(5))
        t = list(b)
        i, j = t.index(3), t.index(4)
        t[i], t[j] = t[j], t[i]
        assert not ga.board_solvable(5, t, ga.board_goal(5))


def test_puzzle5_request_carries_both_boards():
    b = ga.gen_board(5, 33, 4)
    r = ga.puzzle5_request(b, num_queues=2048)
    raw = ctypes.string_at(r.cells, 50)
    assert r.kind == ga.PUZZLE and r.side == 5 and r.num_queues == 2048
    assert tuple(raw[:25]) == b and tuple(raw[25:]) == ga.board_goal(5)
    arr = (ga.Request * 1)(r)  # the batch array copies the pointer, the buffer stays on r
    assert ctypes.string_at(arr[0].cells, 50) == raw


def _fail_run(reqs, finish_first, code):
    n = len(reqs)
    arr = (ga.Request * n)(*reqs)
    outs = (ga.Result * n)()
    rc = ga.lib().ga_debug_fail_run(arr, n, outs, finish_first, code)
    return rc, [outs[i].status for i in range(n)]


def _mixed_requests():
    good = [ga.puzzle_request(4, ga.gen_puzzle(4, 10, s)) for s in range(4)]
    bad = ga.puzzle_request(4, 0x1111111111111111)        # not a permutation: GA_E_INVALID
    g = ga.puzzle_goal(4)
    t = ga.state_to_board(g, 4)
    t[0], t[1] = t[1], t[0]
    unsolvable = ga.puzzle_request(4, ga.board_to_state(t))  # odd permutation: GA_NO_PATH
    return [good[0], bad, good[1], unsolvable, good[2], good[3]]


def test_failed_run_leaves_no_result_reading_as_success():
    # A run that fails part-way (here: an injected HIP error after two requests
    # finished) returns the error, and every request that had not finished
    # carries it: none is left GA_OK with a cost of 0, none stays GA_PENDING.
    rc, st = _fail_run(_mixed_requests(), 2, ga.E_HIP)
    assert rc == ga.E_HIP
    assert st == [ga.OK, ga.E_INVALID, ga.OK, ga.NO_PATH, ga.E_HIP, ga.E_HIP]
    rc, st = _fail_run(_mixed_requests(), 0, ga.E_NOMEM)
    assert rc == ga.E_NOMEM and st == [ga.E_NOMEM, ga.E_INVALID, ga.E_NOMEM, ga.NO_PATH, ga.E_NOMEM, ga.E_NOMEM]
    assert ga.PENDING not in st
    assert ga.status_str(ga.PENDING) == "GA_PENDING"
    assert ga.lib().ga_debug_fail_run(None, 0, None, 0, 0) == ga.E_INVALID  # the injected code must be an error


def test_python_wrappers_raise_on_a_failed_run(monkeypatch):
    # Engine.solve / solve_batch raise GAStarError when the C run fails, with
    # the per-request results attached (the C entry point is replaced by the
    # fault-injecting one; no device is touched).
    L = ga.lib()

    def failing_batch(h, arr, n, outs):
        return L.ga_debug_fail_run(arr, n, outs, 1, ga.E_HIP)

    monkeypatch.setattr(L, "ga_solve_batch", failing_batch)
    e = object.__new__(ga.Engine)
    e._h = None
    reqs = _mixed_requests()
    with pytest.raises(ga.GAStarError) as ei:
        e.solve_batch(reqs)
    assert ei.value.code == ga.E_HIP
    assert [s.status for s in ei.value.results] == [ga.OK, ga.E_INVALID, ga.E_HIP, ga.NO_PATH, ga.E_HIP, ga.E_HIP]
    with pytest.raises(ga.GAStarError) as ei:
        e.solve(reqs[0])
    assert ei.value.code == ga.E_HIP and ei.value.results[0].status == ga.OK
    e._h = None  # nothing to destroy


def test_engine_create_rejects_invalid_config_before_any_device_call():
    cfg = ga.default_config()
    h = ctypes.c_void_p()
    for field, value in (("heap_arity", 16), ("heap_arity", 5), ("lanes", 5), ("block", 100), ("block", 1024)):
        c = ga.default_config()
        setattr(c, field, value)
        assert ga.lib().ga_engine_create(0, ctypes.byref(c), ctypes.byref(h)) == ga.E_INVALID, field
        assert not h.value
    assert cfg.heap_arity == 4
    assert ga.lib().ga_engine_last_error(None) == b""


def test_solve_noalloc_rejects_bad_arguments_before_any_device_call():
    res = ga.Result()
    req = ga.puzzle_request(3, ga.gen_puzzle(3, 10, 1))
    buf = (ctypes.c_uint64 * 4)()
    assert ga.lib().ga_solve_noalloc(None, ctypes.byref(req), ctypes.byref(res), buf, 4) == ga.E_INVALID


def test_server_entry_points_reject_bad_arguments_before_any_device_call():
    L = ga.lib()
    h = ctypes.c_void_p()
    assert L.ga_serve_start(None, ctypes.byref(h)) == ga.E_INVALID and not h.value
    t = ctypes.c_uint64()
    req = ga.puzzle_request(3, ga.gen_puzzle(3, 10, 1))
    res = ga.Result()
    ss = ga.ServeStats()
    assert L.ga_serve_submit(None, ctypes.byref(req), ctypes.byref(t)) == ga.E_INVALID
    assert L.ga_serve_poll(None, 1, ctypes.byref(res)) == ga.E_INVALID
    assert L.ga_serve_drain(None) == ga.E_INVALID
    assert L.ga_serve_status(None, ctypes.byref(ss)) == ga.E_INVALID
    assert L.ga_serve_stop(None) == ga.E_INVALID
