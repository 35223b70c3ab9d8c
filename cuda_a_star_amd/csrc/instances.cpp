// instances.cpp -- seeded instance generators and host-side problem checks.
//
// All randomness comes from splitmix64 (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014), one 64-bit draw at a time,
// so a (seed, parameters) pair names one instance on every host.
//
// Sliding-puzzle moves follow the reference order (a_star_sliding_puzzle.h:88-105):
// blank row+1, row-1, col+1, col-1.  Grid moves follow a_star_pathfinding.h:150-182.
#include <cstdint>
#include <cstring>
#include <vector>

#include "gastar.h"

namespace {

int blank_cell(uint32_t side, uint64_t s) {
    for (uint32_t c = 0; c < side * side; ++c)
        if (((s >> (4 * c)) & 0xF) == 0) return (int)c;
    return -1;
}

bool valid_board(uint32_t side, uint64_t s) {
    if (side != 3 && side != 4) return false;
    uint32_t n = side * side, seen = 0;
    for (uint32_t c = 0; c < n; ++c) {
        uint32_t t = (s >> (4 * c)) & 0xF;
        if (t >= n || ((seen >> t) & 1)) return false;
        seen |= 1u << t;
    }
    return n == 16 || (s >> (4 * n)) == 0;
}

// Applies reference move i (0: row+1, 1: row-1, 2: col+1, 3: col-1) to the blank.
bool apply_move(uint32_t side, uint64_t s, int blank, int i, uint64_t* out) {
    int r = blank / (int)side, c = blank % (int)side;
    switch (i) {
        case 0: ++r; break;
        case 1: --r; break;
        case 2: ++c; break;
        default: --c; break;
    }
    if (r < 0 || c < 0 || r >= (int)side || c >= (int)side) return false;
    int nb = r * (int)side + c;
    uint64_t tile = (s >> (4 * nb)) & 0xF;
    *out = (s & ~(0xFull << (4 * nb))) | (tile << (4 * blank));
    return true;
}

const int kDX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
const int kDY[8] = {0, 0, 1, -1, 1, -1, 1, -1};

bool grid_reachable(const uint8_t* cells, uint32_t w, uint32_t h, uint64_t start, uint64_t goal,
                    uint32_t conn) {
    size_t n = (size_t)w * h;
    std::vector<uint8_t> seen(n, 0);
    std::vector<uint32_t> stack;
    stack.reserve(1024);
    stack.push_back((uint32_t)start);
    seen[start] = 1;
    while (!stack.empty()) {
        uint32_t c = stack.back();
        stack.pop_back();
        if (c == goal) return true;
        uint32_t x = c % w, y = c / w;
        for (uint32_t i = 0; i < conn; ++i) {
            int64_t nx = (int64_t)x + kDX[i], ny = (int64_t)y + kDY[i];
            if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
            uint32_t t = (uint32_t)ny * w + (uint32_t)nx;
            if (!cells[t] || seen[t]) continue;
            seen[t] = 1;
            stack.push_back(t);
        }
    }
    return false;
}

}  // namespace

extern "C" {

uint64_t ga_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t ga_puzzle_goal(uint32_t side) {
    if (side != 3 && side != 4) return 0;
    uint32_t n = side * side;
    uint64_t s = 0;
    for (uint32_t c = 0; c + 1 < n; ++c) s |= (uint64_t)(c + 1) << (4 * c);
    return s;  // last cell holds the blank (0)
}

uint64_t ga_gen_puzzle(uint32_t side, uint64_t goal, uint32_t depth, int no_backtrack, uint64_t seed) {
    if (!valid_board(side, goal)) return 0;
    uint64_t rng = seed;
    uint64_t s = goal;
    int last = -1;
    static const int kInverse[4] = {1, 0, 3, 2};
    for (uint32_t step = 0; step < depth; ++step) {
        int blank = blank_cell(side, s);
        int moves[4];
        uint64_t next[4];
        int nm = 0;
        for (int i = 0; i < 4; ++i) {
            if (no_backtrack && last >= 0 && i == kInverse[last]) continue;
            uint64_t t;
            if (apply_move(side, s, blank, i, &t)) {
                moves[nm] = i;
                next[nm] = t;
                ++nm;
            }
        }
        uint64_t r = ga_splitmix64(&rng);
        int k = (int)(r % (uint64_t)nm);
        s = next[k];
        last = moves[k];
    }
    return s;
}

int ga_puzzle_solvable(uint32_t side, uint64_t start, uint64_t goal) {
    if (!valid_board(side, start) || !valid_board(side, goal)) return 0;
    // Every move is a transposition of the blank with a neighbour and moves the
    // blank one cell, so reachable <=> parity(permutation start->goal) equals
    // parity(Manhattan displacement of the blank).
    uint32_t n = side * side;
    int pos_goal[16];
    for (uint32_t c = 0; c < n; ++c) pos_goal[(goal >> (4 * c)) & 0xF] = (int)c;
    int perm[16];
    for (uint32_t c = 0; c < n; ++c) perm[c] = pos_goal[(start >> (4 * c)) & 0xF];
    int parity = 0;
    bool seen[16] = {false};
    for (uint32_t c = 0; c < n; ++c) {
        if (seen[c]) continue;
        int len = 0;
        for (int j = (int)c; !seen[j]; j = perm[j]) {
            seen[j] = true;
            ++len;
        }
        parity ^= (len - 1) & 1;
    }
    int bs = blank_cell(side, start), bg = blank_cell(side, goal);
    int d = abs(bs / (int)side - bg / (int)side) + abs(bs % (int)side - bg % (int)side);
    return parity == (d & 1);
}

// Boards (side 2..5): the same walk as ga_gen_puzzle -- the same draws and
// the same move choices -- on a board of side*side bytes, so that sides 3 and
// 4 give exactly the boards of ga_gen_puzzle.
int ga_gen_board(uint32_t side, const uint8_t* goal, uint32_t depth, int no_backtrack, uint64_t seed,
                 uint8_t* out) {
    if (side < 2 || side > 5 || !goal || !out) return 0;
    const uint32_t n = side * side;
    uint32_t seen = 0;
    int blank = -1;
    for (uint32_t c = 0; c < n; ++c) {
        if (goal[c] >= n || ((seen >> goal[c]) & 1u)) return 0;
        seen |= 1u << goal[c];
        if (goal[c] == 0) blank = (int)c;
        out[c] = goal[c];
    }
    uint64_t rng = seed;
    int last = -1;
    static const int kInverse[4] = {1, 0, 3, 2};
    for (uint32_t step = 0; step < depth; ++step) {
        const int r = blank / (int)side, c = blank % (int)side;
        int moves[4], cells[4], nm = 0;
        for (int i = 0; i < 4; ++i) {
            if (no_backtrack && last >= 0 && i == kInverse[last]) continue;
            const int nr = r + (i == 0) - (i == 1), nc = c + (i == 2) - (i == 3);  // a_star_sliding_puzzle.h:88-105
            if (nr < 0 || nc < 0 || nr >= (int)side || nc >= (int)side) continue;
            moves[nm] = i;
            cells[nm] = nr * (int)side + nc;
            ++nm;
        }
        const int k = (int)(ga_splitmix64(&rng) % (uint64_t)nm);
        out[blank] = out[cells[k]];
        out[cells[k]] = 0;
        blank = cells[k];
        last = moves[k];
    }
    return 1;
}

int ga_board_solvable(uint32_t side, const uint8_t* start, const uint8_t* goal) {
    if (side < 2 || side > 5 || !start || !goal) return 0;
    const uint32_t n = side * side;
    int pos_goal[25], perm[25];
    uint32_t seen_s = 0, seen_g = 0;
    for (uint32_t c = 0; c < n; ++c) {
        if (start[c] >= n || goal[c] >= n || ((seen_s >> start[c]) & 1u) || ((seen_g >> goal[c]) & 1u)) return 0;
        seen_s |= 1u << start[c];
        seen_g |= 1u << goal[c];
        pos_goal[goal[c]] = (int)c;
    }
    for (uint32_t c = 0; c < n; ++c) perm[c] = pos_goal[start[c]];
    // same rule as ga_puzzle_solvable: permutation parity = parity of the blank's displacement
    int parity = 0, bs = 0, bg = 0;
    bool done[25] = {false};
    for (uint32_t c = 0; c < n; ++c) {
        if (start[c] == 0) bs = (int)c;
        if (goal[c] == 0) bg = (int)c;
        if (done[c]) continue;
        int len = 0;
        for (int j = (int)c; !done[j]; j = perm[j]) {
            done[j] = true;
            ++len;
        }
        parity ^= (len - 1) & 1;
    }
    const int d = abs(bs / (int)side - bg / (int)side) + abs(bs % (int)side - bg % (int)side);
    return parity == (d & 1);
}

uint32_t ga_gen_grid(uint32_t width, uint32_t height, uint32_t wall_ppm, uint32_t w_lo, uint32_t w_hi,
                     uint64_t start, uint64_t goal, uint32_t connectivity, uint64_t seed,
                     uint32_t max_attempts, uint8_t* cells_out) {
    size_t n = (size_t)width * height;
    if (!width || !height || !cells_out || start >= n || goal >= n || w_lo < 1 || w_hi > 255 || w_lo > w_hi ||
        wall_ppm > 1000000 || (connectivity != 4 && connectivity != 8) || max_attempts == 0)
        return 0;
    // wall iff the high 32 bits of the cell's draw fall below wall_ppm/1e6 * 2^32;
    // weight = w_lo + (low 32 bits mod span).
    uint64_t threshold = ((uint64_t)wall_ppm << 32) / 1000000ull;
    uint32_t span = w_hi - w_lo + 1;
    uint64_t rng = seed;
    for (uint32_t attempt = 1; attempt <= max_attempts; ++attempt) {
        for (size_t c = 0; c < n; ++c) {
            uint64_t r = ga_splitmix64(&rng);
            uint8_t wgt = (uint8_t)(w_lo + (uint32_t)(r & 0xFFFFFFFFu) % span);
            bool wall = (r >> 32) < threshold && c != start && c != goal;
            cells_out[c] = wall ? 0 : wgt;
        }
        if (grid_reachable(cells_out, width, height, start, goal, connectivity)) return attempt;
    }
    return 0;
}

}  // extern "C"
