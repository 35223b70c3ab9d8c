// cli.hip -- command-line front end with the reference's flags and text formats.
//
//   gastar_cli --version sliding|pathfinding [--input-data F] [--output-data F]
//              [--num-queues N] [--size S] [--device D] [--each-pair 0|1]
//              [--time-unit ms|us]
//
// Flags, input and output formats follow main.cu:31-78 (output: elapsed
// milliseconds, then one state per line, start..target; an empty path means
// no solution), a_star_sliding_puzzle.h:21-74 (comma-separated tiles, '_' =
// blank; two tokens, start then target) and a_star_pathfinding.h:38-128
// ("W,H", "sx,sy", "tx,ty", obstacle count and "x,y" lines, cost count and
// "x,y,c" lines; unlisted cells cost 1).  Additions: --num-queues (the
// reference hard-codes 4096, main.cu:18), --size (the reference hard-codes 5,
// main.cu:72), --device (the reference hard-codes 1, main.cu:17), and for the
// sliding puzzle --each-pair 1 (solve every start/target pair of the input in
// turn, each answered as one pair is; the reference reads the first) and
// --time-unit us (the elapsed time in microseconds instead of milliseconds).
//
// The problems are plugins over the functor interface (gastar/functional_types.h)
// solved by the header drop-in a_star<>() (gastar/a_star.h).  Differences from
// the reference plugins, all for admissibility or range:
//   sliding:     Manhattan distance without the blank (the reference counts the
//                blank, a_star_sliding_puzzle.h:152-163, which overestimates).
//   pathfinding: Chebyshev distance times the smallest move cost (the reference
//                uses Manhattan with unit-cost diagonals, a_star_pathfinding.h:
//                219-221); 16-bit coordinates (the reference truncates them to
//                8 bits, :185-186); costs must be 1..255 (the reference stores
//                them modulo 256, :110-123).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gastar/a_star.h"

namespace SlidingPuzzle {

constexpr int kMaxCells = 25;
__constant__ signed char goal_row[kMaxCells];
__constant__ signed char goal_col[kMaxCells];

template <unsigned char size>
struct Position {
    unsigned char pos_zero_x;  // blank row
    unsigned char pos_zero_y;  // blank column
    unsigned char board[size][size];
    unsigned char padding[(4 - ((size * size + 2) % 4)) % 4];

    Position() = delete;
    explicit Position(const std::vector<int>& tiles) {
        memset(this, 0, sizeof(*this));
        for (int i = 0; i < size * size; ++i) {
            board[i / size][i % size] = (unsigned char)tiles[i];
            if (tiles[i] == 0) {
                pos_zero_x = (unsigned char)(i / size);
                pos_zero_y = (unsigned char)(i % size);
            }
        }
        fill_padding(*this);
    }
    __host__ __device__ static void fill_padding(Position& p) {
        for (unsigned i = 0; i < sizeof(p.padding); ++i) p.padding[i] = 69;  // fixed padding byte, as the reference
    }
};

// a_star_sliding_puzzle.h:21-48 format: size*size comma-separated integers, '_' = blank.
std::vector<int> parse_tiles(const std::string& tok, int size) {
    std::vector<int> out;
    std::stringstream ss(tok);
    std::string item;
    while (std::getline(ss, item, ',')) out.push_back(item == "_" ? 0 : std::stoi(item));
    if ((int)out.size() != size * size) throw std::runtime_error("Sliding puzzle incorrect input.");
    std::vector<int> seen(size * size, 0);
    for (int t : out) {
        if (t < 0 || t >= size * size || seen[t]++) throw std::runtime_error("Sliding puzzle incorrect input.");
    }
    return out;
}

template <unsigned char size>
std::string format(const Position<size>& p) {
    std::string s;
    for (int i = 0; i < size * size; ++i) {
        int t = p.board[i / size][i % size];
        s += t ? std::to_string(t) : std::string("_");
        if (i + 1 < size * size) s += ',';
    }
    return s;
}

// Moves in the reference order (a_star_sliding_puzzle.h:88-105): blank row+1,
// row-1, column+1, column-1; unit cost.
template <unsigned char size>
__device__ void neighbour_generator(const Position<size>& source, Position<size>& target, bool& target_exists,
                                    uint32_t& cost, unsigned char i) {
    int r = source.pos_zero_x, c = source.pos_zero_y;
    if (i == 0) ++r;
    else if (i == 1) --r;
    else if (i == 2) ++c;
    else --c;
    cost = 1;
    if (r < 0 || c < 0 || r >= size || c >= size) {
        target_exists = false;
        return;
    }
    target = source;
    target.board[source.pos_zero_x][source.pos_zero_y] = source.board[r][c];
    target.board[r][c] = 0;
    target.pos_zero_x = (unsigned char)r;
    target.pos_zero_y = (unsigned char)c;
    Position<size>::fill_padding(target);
    target_exists = true;
}

template <unsigned char size>
__device__ uint32_t manhattan_heuristic(const Position<size>& p) {
    uint32_t h = 0;
    for (int r = 0; r < size; ++r)
        for (int c = 0; c < size; ++c) {
            int t = p.board[r][c];
            if (!t) continue;  // blank excluded
            h += abs(r - goal_row[t]) + abs(c - goal_col[t]);
        }
    return h;
}

template <unsigned char size>
void set_target(const Position<size>& target) {
    signed char rows[kMaxCells] = {0}, cols[kMaxCells] = {0};
    for (int r = 0; r < size; ++r)
        for (int c = 0; c < size; ++c) {
            rows[target.board[r][c]] = (signed char)r;
            cols[target.board[r][c]] = (signed char)c;
        }
    gastar::dropin::check(hipMemcpyToSymbol(HIP_SYMBOL(goal_row), rows, sizeof(rows)), "hipMemcpyToSymbol");
    gastar::dropin::check(hipMemcpyToSymbol(HIP_SYMBOL(goal_col), cols, sizeof(cols)), "hipMemcpyToSymbol");
}

template <unsigned char size>
void solve(std::istream& in, std::ostream& out, uint32_t num_queues, bool each_pair, bool micros) {
    std::string a, b;
    while (in >> a >> b) {
        Position<size> start(parse_tiles(a, size)), target(parse_tiles(b, size));
        set_target(target);
        auto res = a_star<Position<size>, 4, neighbour_generator<size>, manhattan_heuristic<size>>(start, target,
                                                                                                   num_queues);
        if (micros) out << std::chrono::duration_cast<std::chrono::microseconds>(res.second).count() << '\n';
        else out << std::chrono::duration_cast<std::chrono::milliseconds>(res.second).count() << '\n';
        for (const auto& p : res.first) out << format(p) << '\n';
        if (!each_pair) break;
    }
}

}  // namespace SlidingPuzzle

namespace Pathfinding {

struct Position {
    uint16_t x;
    uint16_t y;
    __host__ __device__ uint32_t word() const { return (uint32_t)x | ((uint32_t)y << 16); }
};

__constant__ Position board_size;
__constant__ Position goal;
__constant__ uint32_t w_min;
__constant__ const uint32_t* obstacles;  // sorted words
__constant__ uint32_t obstacles_size;
__constant__ const uint32_t* cost_keys;  // sorted words
__constant__ const uint8_t* cost_values;
__constant__ uint32_t cost_size;

__device__ inline uint32_t find_word(const uint32_t* data, uint32_t n, uint32_t w) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (data[mid] < w) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && data[lo] == w) ? lo : 0xFFFFFFFFu;
}

// Moves in the reference order (a_star_pathfinding.h:150-182): x+1, x-1, y+1,
// y-1, (+1,+1), (+1,-1), (-1,+1), (-1,-1); cost = cost of the entered cell.
__device__ void neighbour_generator(const Position& source, Position& target, bool& target_exists, uint32_t& cost,
                                    unsigned char i) {
    const int DX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
    const int DY[8] = {0, 0, 1, -1, 1, -1, 1, -1};
    const int nx = (int)source.x + DX[i], ny = (int)source.y + DY[i];
    cost = 1;
    if (nx < 0 || ny < 0 || nx >= board_size.x || ny >= board_size.y) {
        target_exists = false;
        return;
    }
    target.x = (uint16_t)nx;
    target.y = (uint16_t)ny;
    if (find_word(obstacles, obstacles_size, target.word()) != 0xFFFFFFFFu) {
        target_exists = false;
        return;
    }
    const uint32_t k = find_word(cost_keys, cost_size, target.word());
    cost = k != 0xFFFFFFFFu ? cost_values[k] : 1u;
    target_exists = true;
}

__device__ uint32_t chebyshev_heuristic(const Position& p) {
    const uint32_t dx = p.x > goal.x ? p.x - goal.x : goal.x - p.x;
    const uint32_t dy = p.y > goal.y ? p.y - goal.y : goal.y - p.y;
    return (dx > dy ? dx : dy) * w_min;
}

Position parse_pair(const std::string& tok) {
    auto d = tok.find(',');
    if (d == std::string::npos) throw std::runtime_error("Pathfinding incorrect input.");
    long x = std::stol(tok.substr(0, d)), y = std::stol(tok.substr(d + 1));
    if (x < 0 || y < 0 || x > 65535 || y > 65535) throw std::runtime_error("Pathfinding incorrect input.");
    return Position{(uint16_t)x, (uint16_t)y};
}

void solve(std::istream& in, std::ostream& out, uint32_t num_queues) {
    using gastar::dropin::check;
    std::string tok;
    in >> tok;
    Position size = parse_pair(tok);
    in >> tok;
    Position start = parse_pair(tok);
    in >> tok;
    Position target = parse_pair(tok);
    if (start.x >= size.x || start.y >= size.y || target.x >= size.x || target.y >= size.y)
        throw std::runtime_error("Pathfinding incorrect input.");
    long n_obst;
    in >> n_obst;
    std::vector<uint32_t> obst;
    for (long i = 0; i < n_obst; ++i) {
        in >> tok;
        obst.push_back(parse_pair(tok).word());
    }
    long n_cost;
    in >> n_cost;
    std::vector<std::pair<uint32_t, uint8_t>> costs;
    uint32_t min_cost = 255;
    for (long i = 0; i < n_cost; ++i) {
        in >> tok;
        auto d2 = tok.rfind(',');
        if (d2 == std::string::npos) throw std::runtime_error("Pathfinding incorrect input.");
        Position p = parse_pair(tok.substr(0, d2));
        long c = std::stol(tok.substr(d2 + 1));
        if (c < 1 || c > 255) throw std::runtime_error("Pathfinding costs must be 1..255.");
        costs.emplace_back(p.word(), (uint8_t)c);
        min_cost = std::min<uint32_t>(min_cost, (uint32_t)c);
    }
    std::sort(obst.begin(), obst.end());
    std::sort(costs.begin(), costs.end());
    std::vector<uint32_t> keys;
    std::vector<uint8_t> vals;
    for (auto& kv : costs) {
        if (!keys.empty() && keys.back() == kv.first) {
            vals.back() = kv.second;  // the last listing of a cell wins
            continue;
        }
        keys.push_back(kv.first);
        vals.push_back(kv.second);
    }
    const uint64_t cells = (uint64_t)size.x * size.y;
    uint32_t wmin = keys.size() + obst.size() >= cells ? min_cost : std::min<uint32_t>(1u, min_cost);
    if (keys.empty()) wmin = 1;
    gastar::dropin::Buffers b;
    uint32_t* d_obst = (uint32_t*)b.get(obst.size() * 4);
    uint32_t* d_keys = (uint32_t*)b.get(keys.size() * 4);
    uint8_t* d_vals = (uint8_t*)b.get(vals.size());
    check(hipMemcpy(d_obst, obst.data(), obst.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(d_keys, keys.data(), keys.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemcpy(d_vals, vals.data(), vals.size(), hipMemcpyHostToDevice), "hipMemcpy");
    uint32_t no = (uint32_t)obst.size(), nc = (uint32_t)keys.size();
    const uint32_t* po = d_obst;
    const uint32_t* pk = d_keys;
    const uint8_t* pv = d_vals;
    check(hipMemcpyToSymbol(HIP_SYMBOL(board_size), &size, sizeof(size)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(goal), &target, sizeof(target)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(w_min), &wmin, sizeof(wmin)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(obstacles), &po, sizeof(po)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(obstacles_size), &no, sizeof(no)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(cost_keys), &pk, sizeof(pk)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(cost_values), &pv, sizeof(pv)), "symbol");
    check(hipMemcpyToSymbol(HIP_SYMBOL(cost_size), &nc, sizeof(nc)), "symbol");
    std::pair<std::vector<Position>, std::chrono::nanoseconds> res;
    if (std::binary_search(obst.begin(), obst.end(), target.word()) ||
        std::binary_search(obst.begin(), obst.end(), start.word())) {
        res.second = std::chrono::nanoseconds(0);  // an obstacle at either end: no path
    } else {
        res = a_star<Position, 8, neighbour_generator, chebyshev_heuristic>(start, target, num_queues);
    }
    out << std::chrono::duration_cast<std::chrono::milliseconds>(res.second).count() << '\n';
    for (const auto& p : res.first) out << p.x << ',' << p.y << '\n';
}

}  // namespace Pathfinding

int main(int argc, char** argv) {
    std::string version, input, output;
    uint32_t num_queues = 4096;  // main.cu:18
    int size = 5;                // main.cu:72
    int device = 0;
    bool each_pair = false, micros = false;
    for (int i = 1; i < argc; i += 2) {
        if (i == argc - 1) {
            std::cout << "Parameter without argument." << std::endl;
            return 255;
        }
        std::string k = argv[i], v = argv[i + 1];
        if (k == "--version") version = v;
        else if (k == "--input-data") input = v;
        else if (k == "--output-data") output = v;
        else if (k == "--num-queues") num_queues = (uint32_t)std::stoul(v);
        else if (k == "--size") size = std::stoi(v);
        else if (k == "--device") device = std::stoi(v);
        else if (k == "--each-pair") each_pair = v != "0";
        else if (k == "--time-unit" && (v == "ms" || v == "us")) micros = v == "us";
        else {
            std::cout << "Unknown option " << k << std::endl;
            return 255;
        }
    }
    try {
        std::ifstream fin;
        std::ofstream fout;
        if (!input.empty()) {
            fin.open(input);
            if (!fin) throw std::runtime_error("cannot open " + input);
        }
        if (!output.empty()) fout.open(output);
        std::istream& in = input.empty() ? std::cin : fin;
        std::ostream& out = output.empty() ? std::cout : fout;
        gastar::dropin::check(hipSetDevice(device), "hipSetDevice");
        if (version == "sliding") {
            switch (size) {
                case 3: SlidingPuzzle::solve<3>(in, out, num_queues, each_pair, micros); break;
                case 4: SlidingPuzzle::solve<4>(in, out, num_queues, each_pair, micros); break;
                case 5: SlidingPuzzle::solve<5>(in, out, num_queues, each_pair, micros); break;
                default: throw std::runtime_error("--size must be 3, 4 or 5");
            }
        } else if (version == "pathfinding") {
            Pathfinding::solve(in, out, num_queues);
        } else {
            std::cout << "Unknown version " << version << std::endl;
            return 255;
        }
    } catch (const std::exception& ex) {
        std::cout << ex.what() << std::endl;
        return 255;
    }
    return 0;
}
