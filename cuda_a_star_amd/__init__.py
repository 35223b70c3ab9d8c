"""MI355X-native GA* (massively parallel A*) engine.

Python host mirror of the C ABI in ``include/gastar.h`` (``libgastar.so``,
HIP for gfx950).  The reference (michalborkowski96/cuda_a_star) exposes
``a_star<T, K, gen, h>(start, target, num_queues) -> {path, elapsed}``
(a_star.h:223-224), called by ``SlidingPuzzle::solve`` and
``Pathfinding::solve`` (a_star_sliding_puzzle.h:165-171,
a_star_pathfinding.h:223-238).  Here the same searches are
``Engine.solve_puzzle`` / ``Engine.solve_grid`` (plain GA*, one instance per
call), plus ``Engine.solve_batch`` (many instances per launch) and
``Engine.serve`` (continuous batching over an arrival stream).

The engine runs only on the GPU through ``libgastar.so``; there is no CPU
fallback.  Building: ``python -m cuda_a_star_amd.build``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgastar.so")

# status codes (include/gastar.h)
OK, NO_PATH, PENDING = 0, 1, 2
E_INVALID, E_NOMEM, E_BUDGET, E_HIP, E_TIMEOUT, E_OVERFLOW = -1, -2, -3, -4, -5, -6
PUZZLE, GRID = 1, 2
MAX_LANES = 4  # GA_MAX_LANES

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class Request(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("side", ctypes.c_uint32),
        ("start", ctypes.c_uint64),
        ("goal", ctypes.c_uint64),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("connectivity", ctypes.c_uint32),
        ("w_min", ctypes.c_uint32),
        ("cells", ctypes.c_void_p),
        ("cells_device", ctypes.c_void_p),
        ("num_queues", ctypes.c_uint32),
        ("heap_capacity", ctypes.c_uint32),
        ("table_slots", ctypes.c_uint64),
        ("user_tag", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("cost", ctypes.c_uint32),
        ("expanded", ctypes.c_uint64),
        ("stored", ctypes.c_uint64),
        ("iterations", ctypes.c_uint32),
        ("retries", ctypes.c_uint32),
        ("path_len", ctypes.c_uint32),
        ("grows", ctypes.c_uint32),
        ("path", _u64p),
        ("search_ns", ctypes.c_uint64),
        ("latency_ns", ctypes.c_uint64),
        ("user_tag", ctypes.c_uint64),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("budget_bytes", ctypes.c_uint64),
        ("chunk_bytes", ctypes.c_uint64),
        ("max_instances", ctypes.c_uint32),
        ("chunk_iters", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("max_retries", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("heap_arity", ctypes.c_uint32),
        ("max_load_pct", ctypes.c_uint32),
        ("pops_per_iter", ctypes.c_uint32),
        ("profile", ctypes.c_uint32),
        ("rounds", ctypes.c_uint32),
        ("adaptive_start", ctypes.c_uint32),
        ("grow_thresh", ctypes.c_uint32),
        ("rounds_grid", ctypes.c_uint32),
        ("lanes", ctypes.c_uint32),
        ("deep_cells", ctypes.c_uint64),
        ("slack_puzzle", ctypes.c_uint32),
        ("slack_grid", ctypes.c_uint32),
        ("pdb24", ctypes.c_uint32),
        ("grow_in_place", ctypes.c_uint32),
        ("block", ctypes.c_uint32),
        ("step_flags", ctypes.c_uint32),
        ("persist", ctypes.c_uint32),
        ("persist_block", ctypes.c_uint32),
        ("persist_cells", ctypes.c_uint64),
        ("pdb15", ctypes.c_uint32),
    ]


class EngineStats(ctypes.Structure):
    _fields_ = [
        ("step_launches", ctypes.c_uint64),
        ("step_ms_total", ctypes.c_double),
        ("expanded", ctypes.c_uint64),
        ("generated", ctypes.c_uint64),
        ("pushed", ctypes.c_uint64),
        ("probes", ctypes.c_uint64),
        ("sift_down", ctypes.c_uint64),
        ("sift_up", ctypes.c_uint64),
        ("pops", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("peak_bytes", ctypes.c_uint64),
        ("groups", ctypes.c_uint64),
        ("inserted", ctypes.c_uint64),
        ("lookups", ctypes.c_uint64),
        ("algo_bytes", ctypes.c_double),
        ("timed_launches", ctypes.c_uint64),
        ("heap_groups", ctypes.c_uint64),
        ("work_groups", ctypes.c_uint64),
        ("sift_down_hbm", ctypes.c_uint64),
        ("sift_up_hbm", ctypes.c_uint64),
        ("push_in", ctypes.c_uint64),
        ("counted_launches", ctypes.c_uint64),
        ("lane_launches", ctypes.c_uint64 * (MAX_LANES + 1)),
        ("lane_timed_launches", ctypes.c_uint64 * (MAX_LANES + 1)),
        ("lane_counted_launches", ctypes.c_uint64 * (MAX_LANES + 1)),
        ("lane_step_ms", ctypes.c_double * (MAX_LANES + 1)),
        ("lane_algo_bytes", ctypes.c_double * (MAX_LANES + 1)),
        ("lane_atomics", ctypes.c_uint64 * (MAX_LANES + 1)),
        ("lane_busy_ms", ctypes.c_double * (MAX_LANES + 1)),
        ("host_admit_ms", ctypes.c_double),
        ("host_snapshot_ms", ctypes.c_double),
        ("host_launch_ms", ctypes.c_double),
        ("host_path_ms", ctypes.c_double),
        ("persist_launches", ctypes.c_uint64),
        ("persist_instances", ctypes.c_uint64),
        ("persist_early_restarts", ctypes.c_uint64),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_uint32),
        ("failed", ctypes.c_uint32),
        ("wall_ns", ctypes.c_uint64),
        ("solves_per_sec", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("max_inflight", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


# Exported symbols declared in include/gastar.h (checked by the CPU tests).
EXPORTS = (
    "ga_config_default", "ga_engine_create", "ga_engine_destroy", "ga_engine_get_stats",
    "ga_engine_reset_stats", "ga_solve", "ga_solve_puzzle", "ga_solve_grid", "ga_solve_batch",
    "ga_serve_run", "ga_serve_start", "ga_serve_submit", "ga_serve_poll", "ga_serve_drain", "ga_serve_status",
    "ga_serve_stop", "ga_solve_noalloc", "ga_result_free", "ga_status_str", "ga_engine_last_error",
    "ga_pdb24_tables",
    "ga_pdb15_tables",
    "ga_debug_fail_run", "ga_debug_closed_set", "ga_debug_heap", "ga_heap_rate",
    "ga_device_upload", "ga_device_free", "ga_atomic_rate",
    "ga_splitmix64", "ga_puzzle_goal", "ga_gen_puzzle", "ga_gen_grid", "ga_puzzle_solvable",
    "ga_gen_board", "ga_board_solvable",
)

_lib = None


def lib() -> ctypes.CDLL:
    """Load libgastar.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cuda_a_star_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    L.ga_config_default.argtypes = [P(Config)]
    L.ga_engine_create.argtypes = [ctypes.c_int, P(Config), P(ctypes.c_void_p)]
    L.ga_engine_destroy.argtypes = [ctypes.c_void_p]
    L.ga_engine_destroy.restype = None
    L.ga_engine_get_stats.argtypes = [ctypes.c_void_p, P(EngineStats)]
    L.ga_engine_reset_stats.argtypes = [ctypes.c_void_p]
    L.ga_engine_reset_stats.restype = None
    L.ga_solve.argtypes = [ctypes.c_void_p, P(Request), P(Result)]
    L.ga_solve_puzzle.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_uint32, P(Result)]
    L.ga_solve_grid.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, P(Result)]
    L.ga_solve_batch.argtypes = [ctypes.c_void_p, P(Request), ctypes.c_uint32, P(Result)]
    L.ga_serve_run.argtypes = [ctypes.c_void_p, P(Request), _u64p, ctypes.c_uint32, P(Result), P(ServeStats)]
    L.ga_serve_start.argtypes = [ctypes.c_void_p, P(ctypes.c_void_p)]
    L.ga_serve_submit.argtypes = [ctypes.c_void_p, P(Request), P(ctypes.c_uint64)]
    L.ga_serve_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, P(Result)]
    L.ga_serve_drain.argtypes = [ctypes.c_void_p]
    L.ga_serve_status.argtypes = [ctypes.c_void_p, P(ServeStats)]
    L.ga_serve_stop.argtypes = [ctypes.c_void_p]
    L.ga_pdb24_tables.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_size_t]
    L.ga_pdb15_tables.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_size_t]
    L.ga_solve_noalloc.argtypes = [ctypes.c_void_p, P(Request), P(Result), _u64p, ctypes.c_uint32]
    L.ga_result_free.argtypes = [P(Result)]
    L.ga_result_free.restype = None
    L.ga_status_str.argtypes = [ctypes.c_int]
    L.ga_status_str.restype = ctypes.c_char_p
    L.ga_engine_last_error.argtypes = [ctypes.c_void_p]
    L.ga_engine_last_error.restype = ctypes.c_char_p
    L.ga_debug_fail_run.argtypes = [P(Request), ctypes.c_uint32, P(Result), ctypes.c_uint32, ctypes.c_int]
    _vp = ctypes.c_void_p
    L.ga_debug_closed_set.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
    L.ga_debug_heap.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
    L.ga_heap_rate.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               P(ctypes.c_double)]
    L.ga_device_upload.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, P(ctypes.c_void_p)]
    L.ga_device_free.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.ga_atomic_rate.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                 P(ctypes.c_double)]
    L.ga_splitmix64.argtypes = [_u64p]
    L.ga_splitmix64.restype = ctypes.c_uint64
    L.ga_puzzle_goal.argtypes = [ctypes.c_uint32]
    L.ga_puzzle_goal.restype = ctypes.c_uint64
    L.ga_gen_puzzle.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64]
    L.ga_gen_puzzle.restype = ctypes.c_uint64
    L.ga_gen_grid.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                              ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                              _u8p]
    L.ga_gen_grid.restype = ctypes.c_uint32
    L.ga_puzzle_solvable.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]
    L.ga_gen_board.argtypes = [ctypes.c_uint32, _u8p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, _u8p]
    L.ga_board_solvable.argtypes = [ctypes.c_uint32, _u8p, _u8p]
    _lib = L
    return L


def status_str(code: int) -> str:
    return lib().ga_status_str(code).decode()


class GAStarError(RuntimeError):
    """A failed call or run.  ``results`` holds the per-request results of a
    failed run: the requests that had finished keep their status, the others
    carry the run's error code."""

    def __init__(self, code: int, what: str = "", results=None):
        super().__init__(f"{status_str(code)} ({code}) {what}".strip())
        self.code = code
        self.results = results


# ---------------------------------------------------------------------------
# instance generators (host code in libgastar.so; see csrc/instances.cpp)
# ---------------------------------------------------------------------------
def splitmix64(state: int) -> tuple[int, int]:
    """One splitmix64 draw: returns (value, next_state)."""
    s = ctypes.c_uint64(state)
    v = lib().ga_splitmix64(ctypes.byref(s))
    return v, s.value


def puzzle_goal(side: int) -> int:
    return lib().ga_puzzle_goal(side)


def gen_puzzle(side: int, depth: int, seed: int, no_backtrack: bool = True, goal: Optional[int] = None) -> int:
    g = puzzle_goal(side) if goal is None else goal
    return lib().ga_gen_puzzle(side, g, depth, 1 if no_backtrack else 0, seed)


def puzzle_solvable(side: int, start: int, goal: int) -> bool:
    return bool(lib().ga_puzzle_solvable(side, start, goal))


def gen_grid(width: int, height: int, wall_p: float, seed: int, w_lo: int = 1, w_hi: int = 9,
             start: Optional[int] = None, goal: Optional[int] = None, connectivity: int = 4,
             max_attempts: int = 1000) -> tuple[np.ndarray, int]:
    """Returns (cells[height*width] uint8, draws used); start/goal default to the corners."""
    s = 0 if start is None else start
    g = width * height - 1 if goal is None else goal
    out = np.empty(width * height, dtype=np.uint8)
    ppm = int(round(wall_p * 1_000_000))
    n = lib().ga_gen_grid(width, height, ppm, w_lo, w_hi, s, g, connectivity, seed, max_attempts,
                          out.ctypes.data_as(_u8p))
    if n == 0:
        raise ValueError("grid generation failed (invalid arguments or no reachable draw)")
    return out, n


def board_to_state(tiles: Sequence[int]) -> int:
    """Row-major tiles (0 = blank) -> nibble-packed state."""
    s = 0
    for c, t in enumerate(tiles):
        s |= int(t) << (4 * c)
    return s


def state_to_board(state: int, side: int) -> List[int]:
    return [(state >> (4 * c)) & 0xF for c in range(side * side)]


# Boards of side 2..5 (side*side tile numbers row-major, 0 = blank): the form
# in which the engine takes 24-puzzles (side 5).
def board_goal(side: int) -> tuple:
    return tuple(range(1, side * side)) + (0,)


def _u8buf(values) -> ctypes.Array:
    return (ctypes.c_uint8 * len(values))(*values)


def gen_board(side: int, depth: int, seed: int, no_backtrack: bool = True, goal=None) -> tuple:
    """The walk of gen_puzzle on a board (same draws and moves for sides 3 and 4)."""
    g = _u8buf(board_goal(side) if goal is None else goal)
    out = (ctypes.c_uint8 * (side * side))()
    if not lib().ga_gen_board(side, g, depth, 1 if no_backtrack else 0, seed, out):
        raise ValueError("invalid board generator arguments")
    return tuple(out)


def board_solvable(side: int, start, goal) -> bool:
    return bool(lib().ga_board_solvable(side, _u8buf(start), _u8buf(goal)))


# ---------------------------------------------------------------------------
# engine
# ---------------------------------------------------------------------------
@dataclass
class Solution:
    """One request's result.  The path's states start..goal are copied out of
    the library's result into `path_states` (a uint64 array; 24-puzzles: a
    list of board tuples); `path` gives them as a list, made on first use (a
    Python int per state costs far more than the copy, and a stream of large
    grids returns millions of states)."""
    status: int
    cost: int
    path_states: object = None
    expanded: int = 0
    stored: int = 0
    iterations: int = 0
    retries: int = 0
    search_ns: int = 0
    latency_ns: int = 0
    user_tag: int = 0
    grows: int = 0

    @property
    def path(self) -> list:
        p = self.path_states
        if p is None:
            return []
        if isinstance(p, np.ndarray):
            p = self.path_states = p.tolist()
        return p

    @property
    def ok(self) -> bool:
        return self.status == OK


def _take(res: Result, board_side: int = 0) -> Solution:
    """board_side 5: the path holds 32-byte boards (4 words per state)."""
    if board_side and res.path_len:
        raw = ctypes.string_at(res.path, 32 * res.path_len)
        n = board_side * board_side
        path = [tuple(raw[32 * k:32 * k + n]) for k in range(res.path_len)]
    else:
        path = np.ctypeslib.as_array(res.path, shape=(res.path_len,)).copy() if res.path_len else None
    sol = Solution(res.status, res.cost, path, res.expanded, res.stored, res.iterations, res.retries,
                   res.search_ns, res.latency_ns, res.user_tag, res.grows)
    lib().ga_result_free(ctypes.byref(res))
    return sol


def default_config() -> Config:
    c = Config()
    lib().ga_config_default(ctypes.byref(c))
    return c


def puzzle_request(side: int, start: int, goal: Optional[int] = None, num_queues: int = 0, **kw) -> Request:
    r = Request()
    r.kind = PUZZLE
    r.side = side
    r.start = start
    r.goal = puzzle_goal(side) if goal is None else goal
    r.num_queues = num_queues
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def puzzle5_request(start, goal=None, num_queues: int = 0, **kw) -> Request:
    """A 24-puzzle request: start and goal boards of 25 tiles (0 = blank) go in
    ``cells`` (50 bytes); the buffer is kept on the request object."""
    g = board_goal(5) if goal is None else tuple(goal)
    r = Request()
    r.kind = PUZZLE
    r.side = 5
    r._boards = _u8buf(tuple(start) + g)
    r.cells = ctypes.addressof(r._boards)
    r.num_queues = num_queues
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def _board_side(req: Request) -> int:
    return 5 if (req.kind == PUZZLE and req.side == 5) else 0


def grid_request(cells: np.ndarray, width: int, height: int, start: int = 0, goal: Optional[int] = None,
                 connectivity: int = 4, num_queues: int = 0, cells_device: Optional[int] = None, **kw) -> Request:
    """``cells`` must stay alive until the request has been solved."""
    r = Request()
    r.kind = GRID
    r.width = width
    r.height = height
    r.start = start
    r.goal = width * height - 1 if goal is None else goal
    r.connectivity = connectivity
    r.num_queues = num_queues
    if cells_device is not None:
        r.cells_device = cells_device
    else:
        assert cells.dtype == np.uint8 and cells.flags.c_contiguous and cells.size == width * height
        r.cells = cells.ctypes.data
    for k, v in kw.items():
        setattr(r, k, v)
    return r


class Engine:
    """One device, one HIP stream.  Not thread-safe."""

    def __init__(self, device: int = 0, **config):
        cfg = default_config()
        for k, v in config.items():
            setattr(cfg, k, v)
        self.config = cfg
        h = ctypes.c_void_p()
        rc = lib().ga_engine_create(device, ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise GAStarError(rc, "ga_engine_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ga_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check_run(self, rc: int, what: str, sols):
        if rc != OK:
            msg = lib().ga_engine_last_error(self._h).decode(errors="replace")
            raise GAStarError(rc, f"{what}: {msg}", sols)

    # -- plain GA*: one instance per call (a_star<>(), a_star.h:223-282) --
    def solve(self, req: Request) -> Solution:
        """One request.  The request's own outcome is in the Solution's status;
        a run that fails (HIP error, device memory) raises GAStarError.  (Runs
        through ga_solve_batch with one request, whose return code tells the
        two apart; ga_solve returns the request's status for both.)"""
        res = Result()
        rc = lib().ga_solve_batch(self._h, ctypes.byref(req), 1, ctypes.byref(res))
        sol = _take(res, _board_side(req))
        self._check_run(rc, "ga_solve", [sol])
        return sol

    def solve_puzzle(self, side: int, start: int, goal: Optional[int] = None, num_queues: int = 0,
                     **kw) -> Solution:
        return self.solve(puzzle_request(side, start, goal, num_queues, **kw))

    def solve_grid(self, cells: np.ndarray, width: int, height: int, start: int = 0, goal: Optional[int] = None,
                   connectivity: int = 4, num_queues: int = 0, **kw) -> Solution:
        cells = np.ascontiguousarray(cells, dtype=np.uint8).reshape(-1)
        return self.solve(grid_request(cells, width, height, start, goal, connectivity, num_queues, **kw))

    def solve_noalloc(self, req: Request, path_cap: int = 1 << 16) -> Solution:
        """Plain GA* of one request with every device buffer in the library's
        static scratch (ga_solve_noalloc): no device allocation."""
        words = 4 if _board_side(req) else 1
        buf = (ctypes.c_uint64 * (path_cap * words))()
        res = Result()
        lib().ga_solve_noalloc(self._h, ctypes.byref(req), ctypes.byref(res), buf, path_cap)
        n = res.path_len if res.status == OK else 0
        if _board_side(req):
            raw = ctypes.string_at(buf, 32 * n)
            path = [tuple(raw[32 * k:32 * k + 25]) for k in range(n)]
        else:
            path = list(buf[:n])
        return Solution(res.status, res.cost, path, res.expanded, res.stored, res.iterations, res.retries,
                        res.search_ns, res.latency_ns, res.user_tag, res.grows)

    # -- batched and served forms --
    def solve_batch(self, reqs: Sequence[Request]) -> List[Solution]:
        n = len(reqs)
        arr = (Request * n)(*reqs)
        outs = (Result * n)()
        rc = lib().ga_solve_batch(self._h, arr, n, outs)
        sols = [_take(outs[i], _board_side(reqs[i])) for i in range(n)]
        self._check_run(rc, "ga_solve_batch", sols)
        return sols

    def serve(self, reqs: Sequence[Request], arrival_ns: Sequence[int]):
        n = len(reqs)
        arr = (Request * n)(*reqs)
        arrivals = (ctypes.c_uint64 * n)(*[int(a) for a in arrival_ns])
        outs = (Result * n)()
        ss = ServeStats()
        rc = lib().ga_serve_run(self._h, arr, arrivals, n, outs, ctypes.byref(ss))
        sols = [_take(outs[i], _board_side(reqs[i])) for i in range(n)]
        self._check_run(rc, "ga_serve_run", sols)
        return sols, ss

    def server(self) -> "Server":
        """An online server on this engine (ga_serve_start): submit, poll, drain."""
        return Server(self)

    def pdb24_tables(self) -> np.ndarray:
        """The device's 24-puzzle pattern databases (ga_pdb24_tables)."""
        out = np.empty(5 * 25 ** 5, dtype=np.uint8)
        rc = lib().ga_pdb24_tables(self._h, out.ctypes.data_as(_u8p), out.nbytes)
        if rc != OK:
            raise GAStarError(rc, "ga_pdb24_tables")
        return out

    def pdb15_tables(self) -> np.ndarray:
        """The device's 15-puzzle pattern databases (ga_pdb15_tables)."""
        out = np.empty(3 * 16 ** 5, dtype=np.uint8)
        rc = lib().ga_pdb15_tables(self._h, out.ctypes.data_as(_u8p), out.nbytes)
        if rc != OK:
            raise GAStarError(rc, "ga_pdb15_tables")
        return out

    def stats(self) -> EngineStats:
        s = EngineStats()
        lib().ga_engine_get_stats(self._h, ctypes.byref(s))
        return s

    def reset_stats(self):
        lib().ga_engine_reset_stats(self._h)

    def atomic_rate(self, table_bytes: int = 1 << 30, ops: int = 1 << 26, kind: int = 0) -> float:
        """Random 64-bit global atomics per second on this device (kind 0:
        fetch-min, 1: CAS from the empty key), or random 64-bit plain loads
        (kind 2), over a table of table_bytes."""
        out = ctypes.c_double()
        rc = lib().ga_atomic_rate(self._h, table_bytes, ops, kind, ctypes.byref(out))
        if rc != OK:
            raise GAStarError(rc, "ga_atomic_rate")
        return out.value

    def upload(self, arr: np.ndarray) -> int:
        """Copies a host array into device memory owned by the caller; returns the device address."""
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        rc = lib().ga_device_upload(self._h, arr.ctypes.data, arr.nbytes, ctypes.byref(p))
        if rc != OK:
            raise GAStarError(rc, "ga_device_upload")
        return p.value

    def free_device(self, addr: int):
        lib().ga_device_free(self._h, addr)


class Server:
    """Online serving on one engine (ga_serve_start .. ga_serve_stop): requests
    are submitted while others run and are admitted as slots and the budget
    free up.  Usable as a context manager (stop on exit)."""

    def __init__(self, engine: Engine):
        h = ctypes.c_void_p()
        rc = lib().ga_serve_start(engine._h, ctypes.byref(h))
        if rc != OK:
            raise GAStarError(rc, "ga_serve_start")
        self._h = h
        self._sides = {}

    def submit(self, req: Request) -> int:
        t = ctypes.c_uint64()
        rc = lib().ga_serve_submit(self._h, ctypes.byref(req), ctypes.byref(t))
        if rc != OK:
            raise GAStarError(rc, "ga_serve_submit")
        self._sides[t.value] = _board_side(req)
        return t.value

    def poll(self, ticket: int) -> Optional[Solution]:
        """The Solution once the request has finished, else None."""
        res = Result()
        rc = lib().ga_serve_poll(self._h, ticket, ctypes.byref(res))
        if rc == PENDING:
            return None
        if rc != OK:
            raise GAStarError(rc, "ga_serve_poll")
        return _take(res, self._sides.pop(ticket, 0))

    def drain(self):
        rc = lib().ga_serve_drain(self._h)
        if rc != OK:
            raise GAStarError(rc, "ga_serve_drain")

    def status(self) -> ServeStats:
        ss = ServeStats()
        lib().ga_serve_status(self._h, ctypes.byref(ss))
        return ss

    def stop(self):
        if getattr(self, "_h", None):
            rc = lib().ga_serve_stop(self._h)
            self._h = None
            if rc != OK:
                raise GAStarError(rc, "ga_serve_stop")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()
