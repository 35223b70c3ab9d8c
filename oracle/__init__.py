"""CPU oracle for the GA* search path -- TEST INFRASTRUCTURE ONLY.

Only the test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker or the timed CPU
reference.  The product (``cuda_a_star_amd``) never imports it.

``liboracle.so`` is built from ``oracle.c`` by ``oracle/Makefile`` (see
``build()``).  The algorithms and the reference lines they restate are listed in
``oracle.h``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OK, NO_PATH, E_INVALID, E_LIMIT, E_NOMEM = 0, 1, -1, -2, -3


class OracleResult(ctypes.Structure):
    _fields_ = [
        ("cost", ctypes.c_int64),
        ("expanded", ctypes.c_uint64),
        ("generated", ctypes.c_uint64),
        ("stored", ctypes.c_uint64),
        ("path_len", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native(out_dir: str = "/tmp/gastar_oracle") -> str:
    """For the timed CPU baseline only: compile the oracle with -march=native
    for the host this runs on (the ``native`` target of the Makefile, output
    outside the repository) and route every call of this module through it.
    Returns "native", or "portable" when that build fails (liboracle.so is
    then used as is)."""
    global _lib
    out = os.path.join(out_dir, "liboracle_native.so")
    try:
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", f"NATIVE_OUT={out}"], check=True,
                       capture_output=True, timeout=120)
    except (subprocess.SubprocessError, OSError):
        return "portable"
    _lib = None
    lib(out)
    return "native"


def lib(path: str = _LIB_PATH) -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_astar_puzzle.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.POINTER(OracleResult), u64p, ctypes.c_uint32]
        L.oracle_astar_puzzle.restype = ctypes.c_int
        L.oracle_puzzle_manhattan.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_puzzle_manhattan.restype = ctypes.c_uint32
        L.oracle_astar_grid.argtypes = [u8p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(OracleResult), u32p,
                                        ctypes.c_uint32]
        L.oracle_astar_grid.restype = ctypes.c_int
        L.oracle_dial_grid.argtypes = [u8p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_int]
        L.oracle_dial_grid.restype = ctypes.c_int64
        L.oracle_bfs_puzzle8.argtypes = [ctypes.c_uint64, u8p, u32p]
        L.oracle_bfs_puzzle8.restype = ctypes.c_uint32
        L.oracle_perm_rank9.argtypes = [ctypes.c_uint64]
        L.oracle_perm_rank9.restype = ctypes.c_uint32
        L.oracle_replay_puzzle.argtypes = [ctypes.c_int, u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_replay_puzzle.restype = ctypes.c_int64
        L.oracle_replay_grid.argtypes = [u8p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, u32p,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_replay_grid.restype = ctypes.c_int64
        L.oracle_astar_board.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_uint64, ctypes.POINTER(OracleResult),
                                         u8p, ctypes.c_uint32]
        L.oracle_astar_board.restype = ctypes.c_int
        L.oracle_board_manhattan.argtypes = [ctypes.c_int, u8p, u8p]
        L.oracle_board_manhattan.restype = ctypes.c_uint32
        L.oracle_bfs_board.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_uint32, ctypes.c_uint64]
        L.oracle_bfs_board.restype = ctypes.c_int64
        L.oracle_pdb24_build.argtypes = [u8p]
        L.oracle_pdb24_build.restype = ctypes.c_int
        L.oracle_pdb24_h.argtypes = [u8p, u8p]
        L.oracle_pdb24_h.restype = ctypes.c_uint32
        L.oracle_pdb15_build.argtypes = [u8p]
        L.oracle_pdb15_build.restype = ctypes.c_int
        L.oracle_pdb15_h.argtypes = [u8p, u8p]
        L.oracle_pdb15_h.restype = ctypes.c_uint32
        L.oracle_astar_puzzle_pdb.argtypes = [ctypes.c_uint64, u8p, ctypes.c_uint64, ctypes.POINTER(OracleResult),
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32]
        L.oracle_astar_puzzle_pdb.restype = ctypes.c_int
        L.oracle_astar_board_pdb.argtypes = [u8p, u8p, ctypes.c_uint64, ctypes.POINTER(OracleResult), u8p,
                                             ctypes.c_uint32]
        L.oracle_astar_board_pdb.restype = ctypes.c_int
        L.oracle_replay_board.argtypes = [ctypes.c_int, u8p, ctypes.c_uint32, u8p, u8p]
        L.oracle_replay_board.restype = ctypes.c_int64
        _lib = L
    return _lib


def _u8(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def astar_puzzle(side: int, start: int, goal: int, max_states: int = 0, want_path: bool = True):
    """Optimal cost by sequential A*.  Returns (rc, OracleResult, path list)."""
    res = OracleResult()
    cap = 4096 if want_path else 0
    buf = (ctypes.c_uint64 * max(cap, 1))()
    rc = lib().oracle_astar_puzzle(side, start, goal, max_states, ctypes.byref(res),
                                   buf if want_path else None, cap)
    path = [buf[i] for i in range(res.path_len)] if (want_path and rc == OK and res.path_len <= cap) else []
    return rc, res, path


def astar_puzzle_pdb(start: int, max_states: int = 0, want_path: bool = True):
    """astar_puzzle on the 15-puzzle towards the ordered goal with the pattern databases (pdb15)."""
    res = OracleResult()
    cap = 4096 if want_path else 0
    buf = (ctypes.c_uint64 * max(cap, 1))()
    rc = lib().oracle_astar_puzzle_pdb(start, _u8(pdb15()), max_states, ctypes.byref(res),
                                       buf if want_path else None, cap)
    path = [buf[i] for i in range(res.path_len)] if (want_path and rc == OK and res.path_len <= cap) else []
    return rc, res, path


def manhattan(side: int, state: int, goal: int) -> int:
    return lib().oracle_puzzle_manhattan(side, state, goal)


def astar_grid(cells: np.ndarray, w: int, h: int, start: int, goal: int, connectivity: int = 4,
               w_min: int = 0, want_path: bool = True):
    cells = np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1)
    res = OracleResult()
    cap = (w * h) if want_path else 0
    buf = np.zeros(max(cap, 1), dtype=np.uint32)
    rc = lib().oracle_astar_grid(_u8(cells), w, h, start, goal, connectivity, w_min, ctypes.byref(res),
                                 buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if want_path else None, cap)
    path = buf[: res.path_len].tolist() if (want_path and rc == OK) else []
    return rc, res, path


def dial_grid(cells: np.ndarray, w: int, h: int, start: int, goal: int, connectivity: int = 4) -> int:
    cells = np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1)
    return lib().oracle_dial_grid(_u8(cells), w, h, start, goal, connectivity)


def bfs_puzzle8(goal: int):
    depth = np.zeros(362880, dtype=np.uint8)
    hist = np.zeros(64, dtype=np.uint32)
    n = lib().oracle_bfs_puzzle8(goal, _u8(depth), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return n, depth, hist


def perm_rank9(state: int) -> int:
    return lib().oracle_perm_rank9(state)


def replay_puzzle(side: int, path, start: int, goal: int) -> int:
    arr = (ctypes.c_uint64 * max(len(path), 1))(*path)
    return lib().oracle_replay_puzzle(side, arr, len(path), start, goal)


def replay_grid(cells: np.ndarray, w: int, h: int, connectivity: int, path, start: int, goal: int) -> int:
    cells = np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1)
    arr = np.ascontiguousarray(np.asarray(path, dtype=np.uint32))
    if arr.size == 0:
        arr = np.zeros(1, dtype=np.uint32)
    return lib().oracle_replay_grid(_u8(cells), w, h, connectivity,
                                    arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(path), start, goal)


# ---------------------------------------------------------------------------
# boards of side 2..5 (the 24-puzzle): tuples of side*side tile numbers, 0 = blank
# ---------------------------------------------------------------------------
def _board(b, side: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(b, dtype=np.uint8).reshape(-1))
    assert a.size == side * side
    return a


def astar_board(side: int, start, goal, max_states: int = 0, want_path: bool = True):
    """Optimal cost by sequential A* on boards.  Returns (rc, OracleResult, path of boards)."""
    s, g = _board(start, side), _board(goal, side)
    res = OracleResult()
    cap = 4096 if want_path else 0
    buf = np.zeros(max(cap, 1) * side * side, dtype=np.uint8)
    rc = lib().oracle_astar_board(side, _u8(s), _u8(g), max_states, ctypes.byref(res),
                                  _u8(buf) if want_path else None, cap)
    path = []
    if want_path and rc == OK and res.path_len <= cap:
        path = [tuple(int(x) for x in buf[i * side * side:(i + 1) * side * side]) for i in range(res.path_len)]
    return rc, res, path


def board_manhattan(side: int, state, goal) -> int:
    s, g = _board(state, side), _board(goal, side)
    return lib().oracle_board_manhattan(side, _u8(s), _u8(g))


def bfs_board(side: int, start, goal, max_depth: int, max_states: int = 0) -> int:
    s, g = _board(start, side), _board(goal, side)
    return lib().oracle_bfs_board(side, _u8(s), _u8(g), max_depth, max_states)


def replay_board(side: int, path, start, goal) -> int:
    if not path:
        return -1
    arr = np.ascontiguousarray(np.concatenate([_board(b, side) for b in path]))
    s, g = _board(start, side), _board(goal, side)
    return lib().oracle_replay_board(side, _u8(arr), len(path), _u8(s), _u8(g))


def walk_board(side: int, depth: int, seed: int, no_backtrack: bool = True):
    """A random walk of the blank from the ordered goal board (test inputs only;
    moves in the order of a_star_sliding_puzzle.h:88-105, xorshift64* choices)."""
    goal = list(range(1, side * side)) + [0]
    b = goal[:]
    x = (seed * 0x9E3779B97F4A7C15 + 1) & 0xFFFFFFFFFFFFFFFF
    prev = -1
    for _ in range(depth):
        p = b.index(0)
        r, c = divmod(p, side)
        moves = []
        for i, (dr, dc) in enumerate(((1, 0), (-1, 0), (0, 1), (0, -1))):
            nr, nc = r + dr, c + dc
            if 0 <= nr < side and 0 <= nc < side and not (no_backtrack and nr * side + nc == prev):
                moves.append(nr * side + nc)
        x ^= x >> 12
        x ^= (x << 25) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 27
        k = ((x * 0x2545F4914F6CDD1D) & 0xFFFFFFFFFFFFFFFF) >> 33
        nb = moves[k % len(moves)]
        b[p], b[nb] = b[nb], 0
        prev = p
    return tuple(b), tuple(goal)


PDB24_BYTES = 5 * 25 ** 5
_pdb24 = None


def pdb24() -> np.ndarray:
    """The 24-puzzle pattern databases (built once per process by BFS, ~1 s per table)."""
    global _pdb24
    if _pdb24 is None:
        t = np.empty(PDB24_BYTES, dtype=np.uint8)
        rc = lib().oracle_pdb24_build(_u8(t))
        if rc != OK:
            raise RuntimeError(f"oracle_pdb24_build failed ({rc})")
        _pdb24 = t
    return _pdb24


def pdb24_h(board) -> int:
    b = np.ascontiguousarray(np.asarray(board, dtype=np.uint8))
    return lib().oracle_pdb24_h(_u8(pdb24()), _u8(b))


PDB15_BYTES = 3 * 16 ** 5
_pdb15 = None


def pdb15() -> np.ndarray:
    """The 15-puzzle pattern databases (three 5-tile tables, built once per process by BFS)."""
    global _pdb15
    if _pdb15 is None:
        t = np.empty(PDB15_BYTES, dtype=np.uint8)
        rc = lib().oracle_pdb15_build(_u8(t))
        if rc != OK:
            raise RuntimeError(f"oracle_pdb15_build failed ({rc})")
        _pdb15 = t
    return _pdb15


def pdb15_h(board) -> int:
    """h of a 4x4 board (16 tile numbers row-major, 0 = blank) under the 15-puzzle tables."""
    b = np.ascontiguousarray(np.asarray(board, dtype=np.uint8))
    return lib().oracle_pdb15_h(_u8(pdb15()), _u8(b))


def astar_board_pdb(start, max_states: int = 0, want_path: bool = True):
    """oracle_astar_board towards the ordered 5x5 goal with the pattern databases."""
    b = np.ascontiguousarray(np.asarray(start, dtype=np.uint8))
    res = OracleResult()
    cap = 4096 if want_path else 0
    buf = np.zeros(max(cap, 1) * 25, dtype=np.uint8)
    rc = lib().oracle_astar_board_pdb(_u8(b), _u8(pdb24()), max_states, ctypes.byref(res),
                                      _u8(buf) if want_path else None, cap)
    path = ([tuple(buf[25 * k:25 * k + 25]) for k in range(res.path_len)]
            if (want_path and rc == OK and res.path_len <= cap) else [])
    return rc, res, path
