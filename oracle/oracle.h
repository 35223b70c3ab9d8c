/*
 * oracle.h -- sequential CPU oracle for the GA* search path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP engine, the C ABI,
 * the Python package) links, loads or calls this code.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load liboracle.so,
 * and only as the checker / the timed CPU reference, never as the path that is
 * measured or shipped.
 *
 * What it restates (reference = michalborkowski96/cuda_a_star):
 *   - A* with re-opening and the priority order of QueueValue
 *     (f ascending, ties to the larger g): a_star.h:55-66.
 *   - goal test at pop time, the popped goal is not expanded: a_star.h:80-83.
 *   - g' = g + cost of the move: a_star.h:90-95.
 *   - closed-set rule "accept only if g_new < g_stored": fancy_hashmap.h:56-61
 *     with FancyHashmapValue::operator<= comparing g (a_star.h:31-33).
 *   - path = states from start to target through prev links: a_star.h:266-281.
 *   - sliding-puzzle moves in the reference order (blank row+1, row-1, col+1,
 *     col-1), unit cost: a_star_sliding_puzzle.h:83-123.
 *   - grid moves in the reference order (x+1, x-1, y+1, y-1, then the four
 *     diagonals), cost = weight of the entered cell: a_star_pathfinding.h:147-209.
 * Heuristics are the admissible variants the workloads specify (Manhattan
 * without the blank; Manhattan or Chebyshev times the minimum cell weight),
 * not the reference's inadmissible ones (a_star_sliding_puzzle.h:152-163
 * counts the blank; a_star_pathfinding.h:219-221 uses Manhattan with unit
 * diagonals).
 *
 * How each function is pinned (checked by tests/test_oracle.py):
 *   oracle_astar_puzzle  - against the full 8-puzzle BFS (every one of the
 *                          181,440 reachable states, or a seeded sample), whose
 *                          depth histogram must equal OEIS A089473; parity and
 *                          bound invariants on 15-puzzle walks.
 *   oracle_bfs_puzzle8   - histogram = OEIS A089473, total 181,440, max 31.
 *   oracle_astar_grid    - against oracle_dial_grid (an independent Dijkstra
 *                          with a bucket queue) and the closed form
 *                          w*(|dx|+|dy|) / w*max(|dx|,|dy|) on wall-free grids.
 *   oracle_dial_grid     - closed forms above and hand-built mazes.
 *   oracle_replay_*      - hand-built legal and illegal paths.
 *   oracle_astar_board   - equal to oracle_astar_puzzle on 3x3 and 4x4 boards
 *                          (a second encoding of the same search); equal to
 *                          oracle_bfs_board on 5x5 walks; closed form k for a
 *                          blank slid k cells along a row or column; parity
 *                          and bound invariants on deeper 5x5 walks.
 *   oracle_bfs_board     - equal to the 8-puzzle BFS table on 3x3 boards.
 *   oracle_pdb24_build   - h = 0 at the goal; h >= Manhattan and h <= the
 *   oracle_pdb24_h         BFS depth on sampled boards; |h(s) - h(t)| <= 1
 *                          across every move; a one-tile and a two-tile
 *                          closed form; equal to the engine's device-built
 *                          tables byte for byte (a GPU test).
 *   oracle_astar_puzzle_pdb - equal to oracle_astar_puzzle (Manhattan) on
 *                          15-puzzle walks, expanding no more.
 *   oracle_pdb15_build   - the same checks on 4x4 boards (the 15-puzzle):
 *   oracle_pdb15_h         h = 0 at the goal, h >= Manhattan, h <= the BFS
 *                          depth, |h(s) - h(t)| = 1 across every move, closed
 *                          forms; equal to the device's tables byte for byte.
 *   oracle_astar_board_pdb - equal to oracle_astar_board (Manhattan) wherever
 *                          both finish; equal to oracle_bfs_board on walks.
 *
 * State encodings (shared with the engine's C ABI):
 *   puzzle: uint64, nibble i (bits 4i..4i+3) = tile at cell i in row-major
 *           order, 0 = blank; side 3 or 4.
 *   board:  uint8 board[side*side], tile at each cell row-major, 0 = blank;
 *           side 2..5 (the engine takes 5x5 boards this way).
 *   grid:   uint8 cells[w*h] row-major, 0 = wall, 1..255 = cost of entering;
 *           a state is the cell index y*w + x.
 */
#ifndef GASTAR_ORACLE_H
#define GASTAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_OK 0
#define ORACLE_NO_PATH 1
#define ORACLE_E_INVALID (-1)
#define ORACLE_E_LIMIT (-2)
#define ORACLE_E_NOMEM (-3)

typedef struct {
    int64_t cost;        /* optimal path cost, -1 if none */
    uint64_t expanded;   /* nodes popped and expanded */
    uint64_t generated;  /* successors generated */
    uint64_t stored;     /* distinct states stored in the closed set */
    uint32_t path_len;   /* number of states written to path (start..goal) */
    uint32_t pad;
} oracle_result;

/* Sequential A* on a side x side sliding puzzle.  path may be NULL.
 * max_states bounds the closed set (0 = no bound); ORACLE_E_LIMIT if hit. */
int oracle_astar_puzzle(int side, uint64_t start, uint64_t goal,
                        uint64_t max_states, oracle_result* res,
                        uint64_t* path, uint32_t path_cap);

/* Admissible heuristic used by the oracle (Manhattan, blank excluded). */
uint32_t oracle_puzzle_manhattan(int side, uint64_t state, uint64_t goal);

/* Sequential A* on a dense grid; connectivity 4 or 8.  w_min = 0 means
 * "compute the minimum open-cell weight".  path holds cell indices. */
int oracle_astar_grid(const uint8_t* cells, uint32_t w, uint32_t h,
                      uint32_t start, uint32_t goal, int connectivity,
                      uint32_t w_min, oracle_result* res,
                      uint32_t* path, uint32_t path_cap);

/* Dijkstra with a circular bucket queue (Dial), independent of the A* code. */
int64_t oracle_dial_grid(const uint8_t* cells, uint32_t w, uint32_t h,
                         uint32_t start, uint32_t goal, int connectivity);

/* Full BFS of the 8-puzzle from goal.  depth_by_rank[362880] receives the
 * distance of every permutation (0xFF = unreachable), indexed by
 * oracle_perm_rank9; hist[64] receives the depth histogram.  Returns the
 * number of reachable states. */
uint32_t oracle_bfs_puzzle8(uint64_t goal, uint8_t* depth_by_rank, uint32_t* hist);
uint32_t oracle_perm_rank9(uint64_t state);

/* Move-by-move replay.  Returns the summed move cost, or -1 if the path does
 * not start at start, end at goal, or contains an illegal move. */
int64_t oracle_replay_puzzle(int side, const uint64_t* path, uint32_t n,
                             uint64_t start, uint64_t goal);
int64_t oracle_replay_grid(const uint8_t* cells, uint32_t w, uint32_t h,
                           int connectivity, const uint32_t* path, uint32_t n,
                           uint32_t start, uint32_t goal);

/* Boards of side 2..5 (up to the 24-puzzle), given as side*side bytes (tile
 * numbers row-major, 0 = blank), packed internally 5 bits per cell into 128
 * bits.  Same search rules as oracle_astar_puzzle.  path receives path_len
 * boards of side*side bytes each (start first). */
int oracle_astar_board(int side, const uint8_t* start, const uint8_t* goal,
                       uint64_t max_states, oracle_result* res,
                       uint8_t* path, uint32_t path_cap);
uint32_t oracle_board_manhattan(int side, const uint8_t* state, const uint8_t* goal);
/* Breadth-first distance from start to goal, or -1 if it exceeds max_depth,
 * -2 for an invalid board, -3 if max_states (0 = none) is exceeded. */
int64_t oracle_bfs_board(int side, const uint8_t* start, const uint8_t* goal,
                         uint32_t max_depth, uint64_t max_states);
/* Additive pattern databases of 5x5 boards with the ordered goal (tile t on
 * cell t-1, blank last): ORACLE_PDB_BYTES bytes, five tables of 25^5 bytes
 * (the fifth uses its first 25^4), built by breadth-first search. */
#define ORACLE_PDB_PATTERNS 5
#define ORACLE_PDB_BYTES (5ull * 9765625ull)
int oracle_pdb24_build(uint8_t* out);
/* Sum of the five tables for a 5x5 board (an admissible, consistent h). */
uint32_t oracle_pdb24_h(const uint8_t* pdb, const uint8_t* board);
/* The same for 4x4 boards (the 15-puzzle): ORACLE_PDB15_BYTES bytes, three
 * tables of 16^5 bytes for the tile groups {1,2,3,5,6}, {4,7,8,11,12},
 * {9,10,13,14,15}; placement index: sum of cell(tile i) * 16^i. */
#define ORACLE_PDB15_PATTERNS 3
#define ORACLE_PDB15_BYTES (3ull << 20)
int oracle_pdb15_build(uint8_t* out);
/* Sum of the three tables for a 4x4 board (16 tile numbers, 0 = blank). */
uint32_t oracle_pdb15_h(const uint8_t* pdb, const uint8_t* board);
/* oracle_astar_puzzle on the 15-puzzle towards the ordered goal with these
 * tables instead of Manhattan. */
#define ORACLE_PUZZLE4_GOAL 0x0FEDCBA987654321ull
int oracle_astar_puzzle_pdb(uint64_t start, const uint8_t* pdb, uint64_t max_states, oracle_result* res,
                            uint64_t* path, uint32_t path_cap);
/* oracle_astar_board on 5x5 boards towards the ordered goal with the
 * pattern-database heuristic instead of Manhattan. */
int oracle_astar_board_pdb(const uint8_t* start, const uint8_t* pdb, uint64_t max_states, oracle_result* res,
                           uint8_t* path, uint32_t path_cap);
int64_t oracle_replay_board(int side, const uint8_t* path, uint32_t n,
                            const uint8_t* start, const uint8_t* goal);

#ifdef __cplusplus
}
#endif

#endif
