"""gastar_cli: the reference's flags and text formats (main.cu:31-78,
a_star_sliding_puzzle.h:21-74, a_star_pathfinding.h:38-128) over the header
drop-in a_star<>() (include/gastar/a_star.h).

CPU tests cover argument handling (no HIP call is made before the flags are
parsed); GPU tests solve through the drop-in and compare with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import cuda_a_star_amd as ga
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cuda_a_star_amd", "gastar_cli")


def run(args, stdin=""):
    return subprocess.run([CLI, *args], input=stdin, capture_output=True, text=True, timeout=120)


def test_unknown_option():
    r = run(["--bogus", "1"])
    assert r.returncode == 255 and "Unknown option --bogus" in r.stdout


def test_parameter_without_argument():
    r = run(["--version"])
    assert r.returncode == 255 and "Parameter without argument." in r.stdout


def test_time_unit_must_be_ms_or_us():
    r = run(["--time-unit", "s"])
    assert r.returncode == 255 and "Unknown option --time-unit" in r.stdout


def tiles_text(state, side):
    return ",".join(str(t) if t else "_" for t in ga.state_to_board(state, side))


def parse_tiles(line, side):
    return ga.board_to_state([0 if t == "_" else int(t) for t in line.split(",")])


@pytest.mark.gpu
@pytest.mark.parametrize("side,depth,seed", [(3, 30, 0), (3, 30, 5), (4, 40, 1), (4, 50, 2)])
def test_sliding_matches_oracle(side, depth, seed):
    goal = ga.puzzle_goal(side)
    start = ga.gen_puzzle(side, depth, seed)
    r = run(["--version", "sliding", "--size", str(side), "--num-queues", "1024"],
            f"{tiles_text(start, side)} {tiles_text(goal, side)}\n")
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    int(lines[0])  # elapsed milliseconds
    path = [parse_tiles(l, side) for l in lines[1:]]
    rc, ref, _ = oracle.astar_puzzle(side, start, goal, want_path=False)
    assert len(path) - 1 == ref.cost
    assert oracle.replay_puzzle(side, path, start, goal) == ref.cost


@pytest.mark.gpu
def test_sliding_each_pair_in_microseconds():
    """--each-pair 1 answers every start/target pair of the input in turn, each
    as a single pair is answered; --time-unit us gives the elapsed time in
    microseconds (the bench times the drop-in with these)."""
    goal = ga.puzzle_goal(4)
    starts = [ga.gen_puzzle(4, d, s) for d, s in ((30, 11), (44, 12), (0, 13), (52, 14))]
    text = "".join(f"{tiles_text(s, 4)} {tiles_text(goal, 4)}\n" for s in starts)
    r = run(["--version", "sliding", "--size", "4", "--num-queues", "2048", "--each-pair", "1", "--time-unit", "us"],
            text)
    assert r.returncode == 0, r.stdout + r.stderr
    results, cur = [], None
    for line in r.stdout.split():
        if line.isdigit():
            cur = []
            results.append(cur)
        else:
            cur.append(parse_tiles(line, 4))
    assert len(results) == len(starts)
    for start, path in zip(starts, results):
        rc, ref, _ = oracle.astar_puzzle(4, start, goal, want_path=False)
        assert len(path) - 1 == ref.cost
        assert oracle.replay_puzzle(4, path, start, goal) == ref.cost
    # without the flag only the first pair is answered, as by the reference
    r1 = run(["--version", "sliding", "--size", "4", "--num-queues", "2048"], text)
    assert r1.returncode == 0 and sum(1 for l in r1.stdout.split() if l.isdigit()) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("depth,seed", [(20, 3), (36, 4)])
def test_sliding_size5_matches_oracle(depth, seed):
    """The reference's default size (main.cu:72): 5x5 Position structs through
    the drop-in's wide-key closed set."""
    start = ga.gen_board(5, depth, seed)
    goal = ga.board_goal(5)
    text = lambda b: ",".join(str(t) if t else "_" for t in b)
    r = run(["--version", "sliding", "--size", "5", "--num-queues", "1024"], f"{text(start)} {text(goal)}\n")
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    int(lines[0])  # elapsed milliseconds
    path = [tuple(0 if t == "_" else int(t) for t in l.split(",")) for l in lines[1:]]
    rc, ref, _ = oracle.astar_board(5, start, goal, want_path=False)
    assert rc == oracle.OK and len(path) - 1 == ref.cost
    assert oracle.replay_board(5, path, start, goal) == ref.cost


def sparse_grid(w, h, seed, wall_p=0.2):
    """A grid in the reference's sparse format and as a dense array (0 = wall)."""
    cells, _ = ga.gen_grid(w, h, wall_p, seed, connectivity=8)
    lines = [f"{w},{h}", "0,0", f"{w - 1},{h - 1}"]
    walls = [(i % w, i // w) for i in np.flatnonzero(cells == 0)]
    lines.append(str(len(walls)))
    lines += [f"{x},{y}" for x, y in walls]
    costed = [(i % w, i // w, int(cells[i])) for i in np.flatnonzero(cells > 1)]
    lines.append(str(len(costed)))
    lines += [f"{x},{y},{c}" for x, y, c in costed]
    return "\n".join(lines) + "\n", cells


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,seed", [(20, 15, 1), (64, 64, 2), (300, 280, 3)])
def test_pathfinding_matches_oracle(w, h, seed):
    text, cells = sparse_grid(w, h, seed)
    r = run(["--version", "pathfinding", "--num-queues", "512"], text)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    path = [int(l.split(",")[1]) * w + int(l.split(",")[0]) for l in lines[1:]]
    ref = oracle.dial_grid(cells, w, h, 0, w * h - 1, 8)
    assert oracle.replay_grid(cells, w, h, 8, path, 0, w * h - 1) == ref


@pytest.mark.gpu
def test_pathfinding_wide_board_no_coordinate_wrap():
    # boards wider than 255: the reference's 8-bit coordinates would wrap here
    text, cells = sparse_grid(600, 20, 7, wall_p=0.1)
    r = run(["--version", "pathfinding", "--num-queues", "256"], text)
    assert r.returncode == 0
    lines = r.stdout.strip().split("\n")
    path = [int(l.split(",")[1]) * 600 + int(l.split(",")[0]) for l in lines[1:]]
    assert oracle.replay_grid(cells, 600, 20, 8, path, 0, 600 * 20 - 1) == oracle.dial_grid(cells, 600, 20, 0,
                                                                                              600 * 20 - 1, 8)


@pytest.mark.gpu
def test_pathfinding_no_path_prints_empty_path():
    text = "4,4\n0,0\n3,3\n3\n2,3\n3,2\n2,2\n0\n"
    r = run(["--version", "pathfinding", "--num-queues", "64"], text)
    assert r.returncode == 0
    assert r.stdout.strip().split("\n")[1:] == []
