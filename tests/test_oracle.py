"""Pins the CPU oracle (oracle/oracle.c) against facts it did not produce.

- 8-puzzle: full BFS histogram = OEIS A089473 (181,440 states, max depth 31);
  A* cost = BFS depth for every state of a seeded sample.
- grids: A* cost = Dial's Dijkstra (independent code, no heuristic) on seeded
  grids, 4- and 8-connected; closed forms on wall-free grids.
- 15-puzzle: h <= cost <= walk length, cost = walk length = h (mod 2).
- replay: hand-built legal and illegal paths.
"""
import itertools

import numpy as np
import pytest

import cuda_a_star_amd as ga
import oracle

# OEIS A089473: 8-puzzle configurations at distance n from the goal with the blank in a corner.
A089473 = [1, 2, 4, 8, 16, 20, 39, 62, 116, 152, 286, 396, 748, 1024, 1893, 2512, 4485, 5638, 9529,
           10878, 16993, 17110, 23952, 20224, 24047, 15578, 14560, 6274, 3910, 760, 221, 2]

GOAL8 = 0x087654321
GOAL15 = 0x0FEDCBA987654321


def test_goal_encodings():
    assert ga.puzzle_goal(3) == GOAL8
    assert ga.puzzle_goal(4) == GOAL15


def test_bfs_histogram_is_a089473():
    n, depth, hist = oracle.bfs_puzzle8(GOAL8)
    assert n == 181440 == sum(A089473)
    assert [int(x) for x in hist[:32]] == A089473
    assert int(hist[32:].sum()) == 0


def test_perm_rank_is_a_bijection():
    ranks = {oracle.perm_rank9(ga.board_to_state(p)) for p in itertools.permutations(range(9))}
    assert ranks == set(range(362880))


def test_astar_equals_bfs_on_sample():
    _, depth, _ = oracle.bfs_puzzle8(GOAL8)
    rng = np.random.default_rng(1234)
    checked = 0
    for perm in rng.permuted(np.tile(np.arange(9), (1500, 1)), axis=1):
        s = ga.board_to_state(perm)
        d = depth[oracle.perm_rank9(s)]
        rc, res, path = oracle.astar_puzzle(3, s, GOAL8)
        if d == 0xFF:
            assert rc == oracle.NO_PATH
            continue
        assert rc == oracle.OK and res.cost == d
        assert oracle.replay_puzzle(3, path, s, GOAL8) == d
        checked += 1
    assert checked > 600


def test_astar_deepest_states():
    _, depth, _ = oracle.bfs_puzzle8(GOAL8)
    deep = [ga.board_to_state(p) for p in itertools.permutations(range(9))
            if depth[oracle.perm_rank9(ga.board_to_state(p))] == 31]
    assert len(deep) == 2
    for s in deep:
        rc, res, _ = oracle.astar_puzzle(3, s, GOAL8)
        assert rc == oracle.OK and res.cost == 31


def test_manhattan_excludes_blank():
    # one blank move from the goal: tile 8 is one cell off, the blank is not counted
    s = ga.gen_puzzle(3, 1, 0)
    assert oracle.manhattan(3, s, GOAL8) == 1
    rc, res, _ = oracle.astar_puzzle(3, s, GOAL8)
    assert res.cost == 1


@pytest.mark.parametrize("seed", range(12))
def test_puzzle15_invariants(seed):
    depth = 40 + seed
    s = ga.gen_puzzle(4, depth, seed)
    rc, res, path = oracle.astar_puzzle(4, s, GOAL15)
    assert rc == oracle.OK
    h = oracle.manhattan(4, s, GOAL15)
    assert h <= res.cost <= depth
    assert res.cost % 2 == depth % 2 == h % 2
    assert oracle.replay_puzzle(4, path, s, GOAL15) == res.cost


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("seed", range(6))
def test_grid_astar_equals_dial(conn, seed):
    w, h = 57 + seed * 13, 41 + seed * 7
    cells, _ = ga.gen_grid(w, h, 0.1 + 0.03 * seed, seed, connectivity=conn)
    rc, res, path = oracle.astar_grid(cells, w, h, 0, w * h - 1, conn)
    d = oracle.dial_grid(cells, w, h, 0, w * h - 1, conn)
    assert rc == oracle.OK and res.cost == d
    assert oracle.replay_grid(cells, w, h, conn, path, 0, w * h - 1) == d


@pytest.mark.parametrize("w,h,wt", [(1, 1, 5), (10, 1, 2), (33, 17, 7), (64, 64, 1)])
def test_grid_closed_forms(w, h, wt):
    cells = np.full(w * h, wt, dtype=np.uint8)
    for conn, expect in ((4, wt * (w - 1 + h - 1)), (8, wt * max(w - 1, h - 1))):
        rc, res, _ = oracle.astar_grid(cells, w, h, 0, w * h - 1, conn)
        assert res.cost == expect
        assert oracle.dial_grid(cells, w, h, 0, w * h - 1, conn) == expect


def test_grid_maze_detour():
    # a wall with one gap forces a detour whose cost is known by hand
    w, h = 5, 5
    cells = np.ones((h, w), dtype=np.uint8)
    cells[2, :4] = 0           # row 2 blocked except x = 4
    cells = cells.reshape(-1)
    # (0,0) -> (4,0) 4 moves, -> (4,4) 4 moves, -> (0,4) 4 moves = 12
    assert oracle.dial_grid(cells, w, h, 0, 4 * w + 0, 4) == 12
    rc, res, _ = oracle.astar_grid(cells, w, h, 0, 4 * w + 0, 4)
    assert res.cost == 12


def test_grid_unreachable():
    cells = np.ones(25, dtype=np.uint8)
    cells[[1, 5, 6]] = 0
    assert oracle.dial_grid(cells, 5, 5, 0, 24, 4) == -1
    rc, _, _ = oracle.astar_grid(cells, 5, 5, 0, 24, 4)
    assert rc == oracle.NO_PATH
    assert oracle.dial_grid(cells, 5, 5, 0, 24, 8) == -1


def test_replay_rejects_bad_paths():
    s1 = ga.gen_puzzle(3, 1, 0)
    assert oracle.replay_puzzle(3, [s1, GOAL8], s1, GOAL8) == 1
    assert oracle.replay_puzzle(3, [s1, s1, GOAL8], s1, GOAL8) == -1      # not a move
    assert oracle.replay_puzzle(3, [GOAL8], s1, GOAL8) == -1              # wrong start
    cells = np.array([1, 2, 3, 4], dtype=np.uint8)                        # 2x2
    assert oracle.replay_grid(cells, 2, 2, 4, [0, 1, 3], 0, 3) == 2 + 4
    assert oracle.replay_grid(cells, 2, 2, 4, [0, 3], 0, 3) == -1          # diagonal in 4-conn
    assert oracle.replay_grid(cells, 2, 2, 8, [0, 3], 0, 3) == 4
    cells[1] = 0
    assert oracle.replay_grid(cells, 2, 2, 4, [0, 1, 3], 0, 3) == -1       # through a wall


def test_planted_mistakes_are_caught():
    """A plausible oracle bug (blank counted in h, as the reference does) must
    fail the checks above: counting the blank overestimates and breaks the BFS
    equality on some state."""
    _, depth, _ = oracle.bfs_puzzle8(GOAL8)
    s = ga.gen_puzzle(3, 1, 0)
    blank_counted = oracle.manhattan(3, s, GOAL8) + 1  # the blank is one cell off its goal
    assert blank_counted > depth[oracle.perm_rank9(s)]  # inadmissible: h > h*


# ---------------------------------------------------------------------------
# boards (up to the 24-puzzle): a second encoding of the same search
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("side,depth,seed", [(3, 25, 1), (3, 40, 2), (4, 30, 3), (4, 45, 4), (4, 60, 5)])
def test_board_astar_equals_packed_astar(side, depth, seed):
    """oracle_astar_board (5 bits per cell, 128-bit keys) and oracle_astar_puzzle
    (4 bits per cell, 64-bit keys) are independent codings: same cost, and the
    board path replays under the packed replay too."""
    s = ga.gen_puzzle(side, depth, seed)
    g = ga.puzzle_goal(side)
    sb, gb = ga.state_to_board(s, side), ga.state_to_board(g, side)
    rc1, r1, _ = oracle.astar_puzzle(side, s, g, want_path=False)
    rc2, r2, path = oracle.astar_board(side, sb, gb)
    assert rc1 == rc2 == oracle.OK and r1.cost == r2.cost
    assert oracle.board_manhattan(side, sb, gb) == oracle.manhattan(side, s, g)
    assert oracle.replay_puzzle(side, [ga.board_to_state(b) for b in path], s, g) == r2.cost
    assert oracle.replay_board(side, path, sb, gb) == r2.cost


def test_board_bfs_equals_8puzzle_table():
    """The board BFS reproduces the full 8-puzzle BFS (itself pinned to OEIS A089473)."""
    _, depth, _ = oracle.bfs_puzzle8(GOAL8)
    gb = ga.state_to_board(GOAL8, 3)
    rng = np.random.default_rng(77)
    seen = 0
    for perm in rng.permuted(np.tile(np.arange(9), (60, 1)), axis=1):
        d = depth[oracle.perm_rank9(ga.board_to_state(perm))]
        if d == 0xFF:
            assert oracle.bfs_board(3, perm, gb, 31) == -1
            continue
        assert oracle.bfs_board(3, perm, gb, 31) == d
        seen += 1
    assert seen >= 20


@pytest.mark.parametrize("depth,seed", [(8, 1), (11, 2), (13, 3), (14, 4), (15, 5), (16, 6)])
def test_board_astar_equals_bfs_5x5(depth, seed):
    start, goal = oracle.walk_board(5, depth, seed)
    d = oracle.bfs_board(5, start, goal, depth, max_states=20_000_000)
    assert 0 <= d <= depth
    rc, res, path = oracle.astar_board(5, start, goal)
    assert rc == oracle.OK and res.cost == d
    assert oracle.replay_board(5, path, start, goal) == d


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_board_closed_form_straight_slide(k):
    """The blank slid k cells left along the bottom row (and k cells up the last
    column): Manhattan = k and a k-move path exists, so the optimum is k."""
    goal = list(range(1, 25)) + [0]
    row = goal[:]
    for _ in range(k):
        p = row.index(0)
        row[p], row[p - 1] = row[p - 1], 0
    col = goal[:]
    for _ in range(k):
        p = col.index(0)
        col[p], col[p - 5] = col[p - 5], 0
    for start in (row, col):
        assert oracle.board_manhattan(5, start, goal) == k
        rc, res, path = oracle.astar_board(5, start, goal)
        assert rc == oracle.OK and res.cost == k and len(path) == k + 1


@pytest.mark.parametrize("depth,seed", [(30, 11), (40, 12), (50, 13)])
def test_board_5x5_invariants(depth, seed):
    start, goal = oracle.walk_board(5, depth, seed)
    h = oracle.board_manhattan(5, start, goal)
    rc, res, path = oracle.astar_board(5, start, goal, max_states=30_000_000)
    assert rc == oracle.OK
    assert h <= res.cost <= depth
    assert res.cost % 2 == depth % 2 == h % 2
    assert oracle.replay_board(5, path, start, goal) == res.cost


def test_board_rejects_bad_inputs_and_paths():
    start, goal = oracle.walk_board(5, 6, 9)
    bad = list(goal)
    bad[0] = bad[1]  # not a permutation
    assert oracle.astar_board(5, bad, goal)[0] == oracle.E_INVALID
    rc, res, path = oracle.astar_board(5, start, goal)
    assert oracle.replay_board(5, path[:-1], start, goal) == -1  # does not reach the goal
    jump = [path[0]] + path[2:]
    if len(path) > 2:
        assert oracle.replay_board(5, jump, start, goal) == -1  # two moves at once


# ---- 15-puzzle pattern databases (oracle_pdb15_*) ----
GOAL4 = tuple(range(1, 16)) + (0,)


def _neighbours4(b):
    b = list(b)
    z = b.index(0)
    r, c = divmod(z, 4)
    for nr, nc in ((r + 1, c), (r - 1, c), (r, c + 1), (r, c - 1)):
        if 0 <= nr < 4 and 0 <= nc < 4:
            t = b[:]
            t[z], t[nr * 4 + nc] = t[nr * 4 + nc], 0
            yield tuple(t)


def test_pdb15_tables_structure():
    t = oracle.pdb15()
    for p in range(3):
        seg = t[p * 16 ** 5:(p + 1) * 16 ** 5]
        reach = seg[seg != 255]
        # every placement of 5 distinct cells of 16 is reached, and nothing else
        assert len(reach) == 16 * 15 * 14 * 13 * 12
        assert reach.min() == 0 and (reach == 0).sum() == 1 and reach.max() <= 60
    assert oracle.pdb15_h(GOAL4) == 0


def test_pdb15_dominates_manhattan_and_is_consistent():
    boards = [ga.state_to_board(ga.gen_puzzle(4, d, 700 + d), 4) for d in range(4, 120, 3)]
    rng = np.random.default_rng(8)
    for _ in range(60):  # random solvable permutations
        b = list(rng.permutation(16))
        if not ga.board_solvable(4, b, GOAL4):
            i, j = [k for k in range(16) if b[k]][:2]
            b[i], b[j] = b[j], b[i]
        boards.append(tuple(int(x) for x in b))
    for b in boards:
        h = oracle.pdb15_h(b)
        assert h >= oracle.board_manhattan(4, b, GOAL4)
        for t in _neighbours4(b):  # one move changes h by exactly one tile step, up or down
            assert abs(oracle.pdb15_h(t) - h) == 1


@pytest.mark.parametrize("depth,seed", [(10, 41), (14, 42), (18, 43), (20, 44)])
def test_pdb15_below_exact_distance(depth, seed):
    start = ga.state_to_board(ga.gen_puzzle(4, depth, seed), 4)
    d = oracle.bfs_board(4, start, GOAL4, depth, max_states=20_000_000)
    assert 0 <= d <= depth and oracle.pdb15_h(start) <= d
    rc, res, _ = oracle.astar_board(4, start, GOAL4, want_path=False)
    assert rc == oracle.OK and res.cost == d


def test_pdb15_closed_forms():
    # the blank slid k cells along the bottom row: every displaced tile is one free step from home
    for k in (1, 2, 3):
        row = list(GOAL4)
        for _ in range(k):
            p = row.index(0)
            row[p], row[p - 1] = row[p - 1], 0
        assert oracle.pdb15_h(row) == k
    # tiles 1 and 2 (one pattern) swapped in the top row: they cannot pass each other
    # in the relaxation, so the table adds at least a detour of two moves to Manhattan's 2
    b = list(GOAL4)
    b[0], b[1] = 2, 1
    assert oracle.pdb15_h(b) >= oracle.board_manhattan(4, b, GOAL4) + 2
    # tiles 3 and 4 belong to different patterns and step home through each other's cell,
    # which neither pattern holds: exactly their Manhattan distance, 2
    b = list(GOAL4)
    b[2], b[3] = 4, 3
    assert oracle.pdb15_h(b) == oracle.board_manhattan(4, b, GOAL4) == 2


@pytest.mark.parametrize("depth,seed", [(30, 51), (45, 52), (60, 53)])
def test_pdb15_astar_equals_manhattan_astar(depth, seed):
    start, goal = ga.gen_puzzle(4, depth, seed), ga.puzzle_goal(4)
    rc_m, m, _ = oracle.astar_puzzle(4, start, goal, want_path=False)
    rc_p, p, path = oracle.astar_puzzle_pdb(start)
    assert rc_m == rc_p == oracle.OK and p.cost == m.cost
    assert p.expanded <= m.expanded
    assert oracle.replay_puzzle(4, path, start, goal) == p.cost


# ---- 24-puzzle pattern databases (oracle_pdb24_*) ----
GOAL5 = tuple(range(1, 25)) + (0,)


def _neighbours5(b):
    b = list(b)
    z = b.index(0)
    r, c = divmod(z, 5)
    for nr, nc in ((r + 1, c), (r - 1, c), (r, c + 1), (r, c - 1)):
        if 0 <= nr < 5 and 0 <= nc < 5:
            t = b[:]
            t[z], t[nr * 5 + nc] = t[nr * 5 + nc], 0
            yield tuple(t)


def test_pdb24_tables_structure():
    t = oracle.pdb24()
    for p, nt in enumerate((5, 5, 5, 5, 4)):
        seg = t[p * 25 ** 5:(p + 1) * 25 ** 5]
        reach = seg[seg != 255]
        # every placement of nt distinct cells is reached (the blank-free moves connect them all)
        assert len(reach) == int(np.prod(range(25 - nt + 1, 26)))
        assert reach.min() == 0 and (reach == 0).sum() == 1 and reach.max() <= 8 * nt + 16
    assert oracle.pdb24_h(GOAL5) == 0


def test_pdb24_dominates_manhattan_and_is_consistent():
    boards = [oracle.walk_board(5, d, 500 + d)[0] for d in range(4, 120, 3)]
    rng = np.random.default_rng(7)
    for _ in range(40):  # random solvable permutations
        b = list(rng.permutation(25))
        if not ga.board_solvable(5, b, GOAL5):
            i, j = [k for k in range(25) if b[k]][:2]
            b[i], b[j] = b[j], b[i]
        boards.append(tuple(int(x) for x in b))
    for b in boards:
        h = oracle.pdb24_h(b)
        assert h >= oracle.board_manhattan(5, b, GOAL5)
        for t in _neighbours5(b):  # one move changes h by exactly one tile step, up or down
            assert abs(oracle.pdb24_h(t) - h) == 1


@pytest.mark.parametrize("depth,seed", [(8, 21), (12, 22), (14, 23), (16, 24)])
def test_pdb24_below_bfs_depth_and_astar_exact(depth, seed):
    start, goal = oracle.walk_board(5, depth, seed)
    d = oracle.bfs_board(5, start, goal, depth, max_states=20_000_000)
    assert 0 <= oracle.pdb24_h(start) <= d
    rc, res, path = oracle.astar_board_pdb(start)
    assert rc == oracle.OK and res.cost == d
    assert oracle.replay_board(5, path, start, goal) == d


def test_pdb24_closed_forms():
    # the blank slid k cells: every displaced tile is one free step from home
    for k in (1, 3, 4):
        row = list(GOAL5)
        for _ in range(k):
            p = row.index(0)
            row[p], row[p - 1] = row[p - 1], 0
        assert oracle.pdb24_h(row) == k
    # tiles 1 and 2 (one pattern) swapped in the top row: they cannot pass each other
    # in the relaxation, so the table adds at least a detour of two moves to Manhattan's 2
    b = list(GOAL5)
    b[0], b[1] = 2, 1
    assert oracle.pdb24_h(b) >= oracle.board_manhattan(5, b, GOAL5) + 2
    # tiles 19 and 20 belong to different patterns and step home through each other's cell,
    # which neither pattern holds: exactly their Manhattan distance, 2
    b = list(GOAL5)
    b[18], b[19] = 20, 19
    assert oracle.pdb24_h(b) == oracle.board_manhattan(5, b, GOAL5) == 2


@pytest.mark.parametrize("depth,seed", [(30, 31), (40, 32), (46, 33)])
def test_pdb24_astar_equals_manhattan_astar(depth, seed):
    start, goal = oracle.walk_board(5, depth, seed)
    rc_m, m, _ = oracle.astar_board(5, start, goal, max_states=20_000_000, want_path=False)
    rc_p, p, path = oracle.astar_board_pdb(start)
    assert rc_m == rc_p == oracle.OK and p.cost == m.cost
    assert p.expanded <= m.expanded
    assert oracle.replay_board(5, path, start, goal) == p.cost


def test_pdb24_planted_mistake_is_caught():
    # counting one pattern twice (an overestimate) breaks h <= the true distance
    # on the boards of the BFS check above: that check would catch it
    t = oracle.pdb24()

    def pattern_a(b):
        cell = {tile: c for c, tile in enumerate(b)}
        return int(t[sum(cell[tile] * 25 ** i for i, tile in enumerate((1, 2, 3, 6, 7)))])

    over = 0
    for depth, seed in [(8, 21), (12, 22), (14, 23), (16, 24)]:
        start, goal = oracle.walk_board(5, depth, seed)
        d = oracle.bfs_board(5, start, goal, depth, max_states=20_000_000)
        over += oracle.pdb24_h(start) + pattern_a(start) > d
    assert over > 0
