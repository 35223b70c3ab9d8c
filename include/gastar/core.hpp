// core.hpp -- device side of the GA* engine for gfx950 (CDNA4).
//
// One GA* iteration of every resident instance is ONE launch of step_kernel.
// A workgroup serves heaps of one instance; W consecutive lanes ("a heap
// group") own one W-ary heap.  Per iteration a heap group
//   1. pushes the successor records addressed to it by the previous
//      iteration (round-robin, as insert_nodes_into_queues, a_star.h:101-115),
//   2. contributes its top f to the instance's min-f reduction
//      (the termination test of a_star.h:260-262, evaluated one launch later),
//   3. pops its minimum, skipping stale entries (lazy deletion replaces the
//      eager erase of remove_updated_elements, a_star.h:192-203),
//   4. tests the goal at pop (a_star.h:80-83) and expands the node through the
//      problem's successor function, one successor per lane (a_star.h:84-97),
//   5. inserts each successor into the closed set with a 64-bit CAS on the key
//      and an atomicMin on (g << 32 | parent slot): a successor survives only if
//      it lowered g (the INSERTED/UPDATED/REJECTED contract of
//      fancy_hashmap.h:32-81,155-167), so duplicates inside one iteration
//      resolve in the table and no sort is needed (bitonic_sort.h and
//      mark_duplicates, a_star.h:137-146, have no counterpart here),
//   6. writes K successor records to its slice of the candidate buffer.
// Candidate buffers are double-buffered by iteration parity, min-f
// reductions are triple-buffered, so one launch per iteration needs no grid
// barrier and the host never reads back inside the search loop (the reference
// blocks six times per iteration, a_star.h:247-264).
//
// This header is shared by the C ABI library (fixed problem instantiations)
// and by the header drop-in gastar/a_star.h (user problem functors).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gastar {

constexpr uint32_t kNone = 0xFFFFFFFFu;   // no parent / no instance
constexpr uint32_t kInf = 0xFFFFFFFFu;    // infinite cost
constexpr uint64_t kEmptyKey = ~0ull;     // free closed-set slot
constexpr uint64_t kInfVal = ~0ull;       // g = inf, parent = none
constexpr uint64_t kNoRecord = ~0ull;     // empty candidate record

// Instance memory as the kernels see it.  Pointers read out of InstDesc are
// generic to the compiler, which would access them with flat instructions;
// a flat access also counts against the LDS counter, so every LDS wait (the
// wave shuffles included) would wait for all outstanding memory accesses.
// The device code therefore goes through pointers typed as global.  (The
// host pass only parses device code; there the type is a plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
using lptr = __attribute__((address_space(3))) T*;  // LDS
#else
template <class T>
using gptr = T*;
template <class T>
using lptr = T*;
#endif
template <class T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
    return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ lptr<T> as_local(T* p) {
    return (lptr<T>)p;
}

// Instance status.
enum : uint32_t {
    ST_FREE = 0,
    ST_RUNNING = 1,
    ST_DONE = 2,
    ST_OVERFLOW_TABLE = 3,
    ST_OVERFLOW_HEAP = 4,
    ST_ABORTED = 5,
    ST_PAUSED = 6,  // stopped at an iteration boundary so that the host grows its heaps or closed set
};

// Grow requests (InstState::grow): raised during an iteration, acted on at the
// next one, whose heaps all see them (they were written in the previous launch).
enum : uint32_t {
    GROW_HEAP = 1,   // a heap ended the iteration within one iteration's pushes of its capacity
    GROW_TABLE = 2,  // the closed set passed its soft load limit
};

// Problem kinds compiled into the C ABI library.
enum : uint32_t {
    KIND_PUZZLE3 = 1,
    KIND_PUZZLE4 = 2,
    KIND_GRID4 = 3,
    KIND_GRID8 = 4,
    KIND_PUZZLE5 = 5,  // 24-puzzle: 128-bit keys, its own kernel instantiation
    KIND_USER = 16,  // header drop-in: one user problem per kernel instantiation
};

// Heap entry and candidate record, 16 B (one dwordx4).
// prio = f << 32 | ~g: one unsigned compare orders by f ascending and, on equal
// f, by g descending -- QueueValue::operator< of a_star.h:58-63.
struct alignas(16) HeapEntry {
    uint64_t prio;
    uint32_t slot;  // closed-set slot of the state
    uint32_t aux;   // problem scratch (grid: y << 16 | x)
};

// Closed-set slot, 16 B: the key and its value in one line.
// val = g << 32 | parent slot; atomicMin keeps the smallest g (and, for equal
// g, the smallest parent slot, so the stored parent is deterministic per g).
struct alignas(16) Slot {
    uint64_t key;
    uint64_t val;
};

__host__ __device__ inline uint64_t make_prio(uint32_t f, uint32_t g) {
    return ((uint64_t)f << 32) | (uint64_t)(uint32_t)~g;
}
__host__ __device__ inline uint32_t prio_f(uint64_t p) { return (uint32_t)(p >> 32); }
__host__ __device__ inline uint32_t prio_g(uint64_t p) { return ~(uint32_t)p; }
__host__ __device__ inline uint32_t val_g(uint64_t v) { return (uint32_t)(v >> 32); }
// g of a closed-set value in memory: its high word, read on its own
__device__ inline uint32_t g_of(gptr<uint64_t> v) { return ((gptr<uint32_t>)v)[1]; }
__host__ __device__ inline uint32_t val_parent(uint64_t v) { return (uint32_t)v; }

// 64-bit finaliser of MurmurHash3 (public domain, Austin Appleby).  The
// reference hashes with MurmurHash3_x86_32 into (h*B)>>32 buckets
// (helpers.h:10-35, fancy_hashmap.h:151-153); any well-mixed hash gives the
// same search result, only the slot placement differs.
__host__ __device__ inline uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Tiled layout of a grid's closed set and cell weights: 4 x 4 tiles of cells,
// the tiles in row-major order (tile_cols per row), the cells of a tile
// row-major inside it.  A tile of closed-set values (8 B each) fills one
// 128-byte L2 line and a line of weights (1 B each) holds a 32 x 4 strip, so
// the four neighbours of a node lie on the node's own line or one more (with
// rows laid out whole, the cells above and below are on lines of their own).
// Grid states, parent links and heap slots are tiled indices; the host API
// speaks row-major cell indices, converted at admission and in the path walk.
__host__ __device__ inline uint32_t grid_tile_cols(uint32_t width) { return (width + 3u) >> 2; }
__host__ __device__ inline uint64_t grid_tiled_cells(uint32_t width, uint32_t height) {
    return (uint64_t)grid_tile_cols(width) * 16u * ((height + 3u) >> 2);
}
__host__ __device__ inline uint32_t grid_index(uint32_t x, uint32_t y, uint32_t tile_cols) {
    return (((y >> 2) * tile_cols + (x >> 2)) << 4) | ((y & 3u) << 2) | (x & 3u);
}

// Read-only description of one resident instance (written once at admission).
struct InstDesc {
    uint32_t kind;
    uint32_t width;    // puzzle side, or grid width
    uint32_t height;   // grid height
    uint32_t w_min;    // grid heuristic scale
    uint64_t start;
    uint64_t goal;
    const uint8_t* cells;  // grid cells
    HeapEntry* heaps;      // n_heaps * heap_stride entries
    uint32_t* heap_size;   // n_heaps
    HeapEntry* cand;       // 2 * n_heaps * K records
    Slot* table;           // hashed closed set (puzzles, user problems)
    uint64_t* dense;       // direct-indexed closed set (grids)
    uint64_t table_mask;   // hashed: slots - 1
    uint64_t load_limit;   // hashed: max occupied slots before overflow
    uint32_t n_heaps;      // power of two
    uint32_t heap_cap;     // max entries per heap
    uint32_t heap_stride;  // physical entries per heap
    uint32_t start_iter;   // value of the global iteration counter at the first step
    uint32_t a0;           // initial frontier width (heaps in use), a power of two <= n_heaps
    uint32_t rounds;       // expansion rounds per launch for this instance (0 = the launch's default)
    uint32_t f_slack;      // expand a heap's top only if f <= previous min f + f_slack (kInf = always)
    uint32_t tile_cols;    // grid: 4x4 tiles per row of the tiled layout (grid_index)
    uint8_t goal_rc[32];   // puzzle: row << 4 | col of each tile in the goal
    const void* user;      // user problems: device copy of {start, goal}
    void* keys;            // user problems with states wider than 8 bytes: key array
    const uint8_t* pdb;    // 24- and 15-puzzles: the pattern databases (kPdbBytes, kPdb15Bytes), or null for Manhattan
    uint32_t grow_heap_at; // request GROW_HEAP when a heap ends an iteration above this many entries (0 = never)
    uint32_t pad3;
    uint64_t grow_table_at;  // hashed: request GROW_TABLE when more slots than this are occupied (0 = never)
    uint32_t start_xy;     // grid: start cell as y << 16 | x (start and goal hold tiled indices)
    uint32_t goal_xy;      // grid: goal cell as y << 16 | x
};
static_assert(offsetof(InstDesc, goal_rc) % 4 == 0, "goal_rc is read as 32-bit words");

// Mutable per-instance state.
struct InstState {
    uint32_t status;
    uint32_t best;        // smallest g of a popped goal
    uint32_t min_top[3];  // min f over heap tops, by iteration mod 3
    uint32_t iterations;  // iterations started
    uint32_t done_iter;
    uint32_t pad;
    uint32_t active[3];   // heaps in use (frontier width), by iteration mod 3
    uint32_t open_sum[3]; // heap entries after the pushes, by iteration mod 3
    uint32_t active_now;  // latest frontier width (for the host)
    uint32_t pad2;
    uint32_t grow[3];     // grow requests raised during an iteration, by iteration mod 3
    uint32_t paused_iter; // the iteration at which the instance paused (ST_PAUSED)
    unsigned long long occupied;  // closed-set slots in use
    unsigned long long expanded;
    unsigned long long generated;
    unsigned long long pushed;
    unsigned long long pops;
    unsigned long long probes;
    unsigned long long sift_down;
    unsigned long long sift_up;
    unsigned long long groups;    // heap-group iterations that ran (live)
    unsigned long long inserted;  // successors whose atomicMin lowered g
    unsigned long long lookups;   // closed-set find-or-insert calls (one atomicMin each)
    unsigned long long work_groups;  // heap-group iterations with work (pushes or a pop within the gate)
    unsigned long long sift_down_hbm;  // sift-down levels whose children were read from global memory
    unsigned long long sift_up_hbm;    // sift-up levels whose parent was read from global memory
    unsigned long long push_in;        // records pushed from the candidate buffer (written by another heap)
};

// Per-workgroup tallies of one step, in LDS.  Lanes add to them with LDS
// atomics (at most one per wave and event where the event is per lane), and
// each workgroup flushes them to its InstState with one global atomic per
// tally at the end of the step -- rather than every wave of an instance
// hitting the same few InstState words.  Indices below kCounters follow the
// counter fields of InstState from `occupied` on.
enum : uint32_t {
    TL_OCCUPIED, TL_EXPANDED, TL_GENERATED, TL_PUSHED, TL_POPS, TL_PROBES, TL_SIFT_DOWN, TL_SIFT_UP,
    TL_GROUPS, TL_INSERTED, TL_LOOKUPS, TL_WORK, TL_SIFT_DOWN_HBM, TL_SIFT_UP_HBM, TL_PUSH_IN,
    kCounters,
    TL_MIN_TOP = kCounters,  // min f over the heap tops
    TL_OPEN,                 // heap entries after the pushes
    TL_WAVES_DONE,           // waves of the workgroup that have finished the step
    kTallies
};
static_assert(offsetof(InstState, push_in) == offsetof(InstState, occupied) + 8 * (kCounters - 1),
              "InstState counters must be contiguous and in tally order");
static_assert(kTallies <= 64, "tally_end flushes one tally per lane of a wave");

// Workgroup -> (instance, first heap) map of one launch.  `cover` is the
// number of heaps of the instance that the map covers (a power of two): the
// frontier width may grow on the device up to it.
struct WgEntry {
    uint32_t inst;
    uint32_t first_heap;
    uint32_t cover;
    uint32_t pad;
};

// StepParams::flags
enum : uint32_t {
    STEP_DIRTY_TOPS = 1,     // write back only the LDS heap-top entries a launch changed
    STEP_SKIP_SETTLED = 2,   // grids: no atomicMin for a successor whose cell's stored g is already no larger
    STEP_PAIRED_POPS = 4,    // 4-connected grids: two pops per round, both nodes' atomics in flight together
                             // (selects a kernel instantiation on the host; the kernels do not read it)
    STEP_PAIRED_ALL = 8,     // with STEP_PAIRED_POPS: so do the 4-connected grids of a lane that holds several
                             // kinds (the all-kinds instantiation with two pops; host only as well)
};

struct StepParams {
    uint32_t max_pops;     // pops tried per expansion round while entries are stale
    uint32_t max_probes;   // closed-set probe limit (overflow guard)
    uint32_t profile;      // count probes and sift levels
    uint32_t rounds;       // expansion rounds per launch (>= 1)
    uint32_t grow_thresh;  // double the frontier width when open entries exceed this many per heap
    uint32_t flags;        // STEP_* switches
    uint32_t pad[2];
};

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// exchanges inside a heap group of W lanes
// ---------------------------------------------------------------------------
// A group of 4 lanes is a DPP quad: quad permutes run on the VALU without the
// LDS round trip (and the LDS-counter wait) of ds_bpermute, which __shfl uses.
// Wider groups use __shfl.  Every lane of the group must be active.
template <int CTRL>
__device__ __forceinline__ uint32_t quad_perm(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ uint64_t quad_perm64(uint64_t v) {
    return ((uint64_t)quad_perm<CTRL>((uint32_t)(v >> 32)) << 32) | quad_perm<CTRL>((uint32_t)v);
}

// A DPP move over the rows in ROWS (bit r: lanes 16r..16r+15); the other
// rows take `identity`.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_rows(uint32_t v, uint32_t identity) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, CTRL, ROWS, 0xF, false);
}
// Reduction over the wave's 64 lanes on the VALU, returned to every lane.
// Two quad permutes give each quad its value; the half-row mirror (0x141) and
// the row mirror (0x140) each pair it with a disjoint part of its row, so that
// every row of 16 holds its total; row_bcast:15 (0x142) brings row 0's into
// row 1 and row 2's into row 3, row_bcast:31 (0x143) brings lane 31's into
// rows 2 and 3, and lane 63 holds the wave's total.  Every lane must be
// active.  (A __shfl_xor ladder takes six LDS permutes, each waited for.)
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t v, uint32_t identity, Op op) {
    v = op(v, dpp_rows<0xB1, 0xF>(v, identity));  // quad_perm [1,0,3,2]
    v = op(v, dpp_rows<0x4E, 0xF>(v, identity));  // quad_perm [2,3,0,1]
    v = op(v, dpp_rows<0x141, 0xF>(v, identity));
    v = op(v, dpp_rows<0x140, 0xF>(v, identity));
    v = op(v, dpp_rows<0x142, 0xA>(v, identity));
    v = op(v, dpp_rows<0x143, 0xC>(v, identity));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ inline uint32_t wave_min_u32(uint32_t v) {
    return wave_reduce(v, ~0u, [](uint32_t a, uint32_t b) { return a < b ? a : b; });
}
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
    return wave_reduce(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

// value of lane (lane ^ O) of the group, O a power of two below W
template <int W, int O>
__device__ __forceinline__ uint32_t grp_xor(uint32_t v) {
    if constexpr (W == 4) return quad_perm<O == 1 ? 0xB1 : 0x4E>(v);  // [1,0,3,2] / [2,3,0,1]
    else return (uint32_t)__shfl_xor((int)v, O, W);
}
template <int W, int O>
__device__ __forceinline__ uint64_t grp_xor64(uint64_t v) {
    if constexpr (W == 4) return quad_perm64<O == 1 ? 0xB1 : 0x4E>(v);
    else return __shfl_xor(v, O, W);
}

// value of lane l of the group
template <int W>
__device__ __forceinline__ uint32_t grp_bcast(uint32_t v, int l) {
    if constexpr (W == 4) {
        switch (l) {
            case 0: return quad_perm<0x00>(v);
            case 1: return quad_perm<0x55>(v);
            case 2: return quad_perm<0xAA>(v);
            default: return quad_perm<0xFF>(v);
        }
    } else {
        return (uint32_t)__shfl((int)v, l, W);
    }
}
template <int W>
__device__ __forceinline__ uint64_t grp_bcast64(uint64_t v, int l) {
    return ((uint64_t)grp_bcast<W>((uint32_t)(v >> 32), l) << 32) | grp_bcast<W>((uint32_t)v, l);
}

// minimum of (prio, index) over the group, ties to the smaller index
template <int W, int O = W / 2>
__device__ __forceinline__ void grp_min_idx(uint64_t& mp, uint32_t& mi) {
    if constexpr (O > 0) {
        const uint64_t op = grp_xor64<W, O>(mp);
        const uint32_t oi = grp_xor<W, O>(mi);
        if (op < mp || (op == mp && oi < mi)) {
            mp = op;
            mi = oi;
        }
        grp_min_idx<W, O / 2>(mp, mi);
    }
}
template <int W, int O = W / 2>
__device__ __forceinline__ uint32_t grp_min_u32(uint32_t v) {
    if constexpr (O > 0) {
        const uint32_t w = grp_xor<W, O>(v);
        return grp_min_u32<W, O / 2>(w < v ? w : v);
    }
    return v;
}
template <int W, int O = W / 2>
__device__ __forceinline__ uint32_t grp_sum_u32(uint32_t v) {
    if constexpr (O > 0) return grp_sum_u32<W, O / 2>(v + grp_xor<W, O>(v));
    return v;
}

__device__ inline uint32_t load_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long load_relaxed(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_relaxed(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders lane 0's heap writes before the other lanes of its group read them
// (all lanes of a group are in one wave on one CU).
__device__ inline void group_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline void set_status(InstState* st, uint32_t s) {
    atomicCAS(&st->status, (uint32_t)ST_RUNNING, s);
}

// The workgroup's tallies (kTallies words of LDS).  `prof` (uniform) turns
// on the counters that only profiling needs.
struct Tally {
    uint32_t* c;
    bool prof;

    // v of every active lane (always counted)
    __device__ __forceinline__ void add(uint32_t i, uint32_t v) const {
        if (v) atomicAdd(&c[i], v);
    }
    // profile only: v of every active lane
    __device__ __forceinline__ void prof_add(uint32_t i, uint32_t v) const {
        if (prof && v) atomicAdd(&c[i], v);
    }
    // one per active lane where `cond` holds, one LDS atomic per wave (always counted)
    __device__ __forceinline__ void count(uint32_t i, bool cond) const {
        const uint64_t b = __ballot(cond);
        const uint64_t act = __ballot(1);
        if (b && __lane_id() == (uint32_t)__builtin_ctzll(act)) atomicAdd(&c[i], (uint32_t)__popcll(b));
    }
    // profile only: as count()
    __device__ __forceinline__ void prof_count(uint32_t i, bool cond) const {
        if (prof) count(i, cond);
    }
};

// Heaps of one workgroup with work this launch: filled in phase 1 of
// run_group, read in phase 2 (at most G heap groups per workgroup).
template <int G>
struct WorkListT {
    uint32_t n;
    uint32_t q[G];  // heap
    uint32_t S[G];  // its size before the pushes
};

// Threads per step-kernel workgroup: the default, and the sizes the step
// kernels are instantiated for (ga_config.block).  The work list holds one
// entry per heap group, so W >= 4 gives at most B / 4 entries.
constexpr int kStepBlock = 256;

// LDS of one step-kernel workgroup of B threads: the heap tops of its
// (B / W) groups, HeapRef<W>::kTop entries each (most for W = 4: 21 KB at
// B = 256, so seven workgroups share a CU), the tallies and the work list.
template <int B>
struct StepSharedT {
    static constexpr uint32_t kHeapTopEntries = (B / 4) * (1 + 4 + 16);
    HeapEntry heap_top[kHeapTopEntries];
    uint32_t dirty[B / 4];  // per heap group: the heap-top entries its launch changed
    uint32_t tally[kTallies];
    WorkListT<B / 4> work;
};
using StepShared = StepSharedT<kStepBlock>;
using WorkList = WorkListT<kStepBlock / 4>;

// Clears the tallies and the work list; every thread of the workgroup calls
// it.  The kernels read the instance's status just before (each wave its own
// load, issued as soon as the instance is known, so that it travels with the
// instance descriptor's loads) and pass it to run_group, which every wave
// calls whatever the status (it holds a barrier).
template <int B>
__device__ inline void step_begin(StepSharedT<B>& sh) {
    if (threadIdx.x < kTallies) sh.tally[threadIdx.x] = threadIdx.x == TL_MIN_TOP ? kInf : 0u;
    if (threadIdx.x == 0) sh.work.n = 0;
    __syncthreads();
}

// Adds the tallies to the instance's state for iteration `it`; every thread
// of the workgroup calls it.  The last wave to arrive flushes, so waves that
// finish early leave at once instead of waiting at a barrier.
__device__ inline void tally_end(uint32_t* c, InstState* st, uint32_t it, uint64_t grow_table_at = 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // this wave's tally atomics first
    uint32_t arrived = 0;
    if (__lane_id() == 0) arrived = atomicAdd(&c[TL_WAVES_DONE], 1u);
    arrived = __builtin_amdgcn_readfirstlane(arrived);
    if (arrived + 1 != (blockDim.x + 63) / 64) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const uint32_t i = __lane_id();
    if (i >= TL_WAVES_DONE) return;
    const uint32_t v = __hip_atomic_load(&c[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (i < kCounters) {
        if (v) {
            auto* ctr = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(st) + offsetof(InstState, occupied));
            const unsigned long long before = atomicAdd(ctr + i, (unsigned long long)v);
            // the closed set passed its soft limit during this iteration: grow it before the next
            if (i == TL_OCCUPIED && grow_table_at && before + v > grow_table_at) atomicOr(&st->grow[it % 3], GROW_TABLE);
        }
    } else if (i == TL_MIN_TOP) {
        if (v != kInf) atomicMin(&st->min_top[it % 3], v);
    } else if (v) {
        atomicAdd(&st->open_sum[it % 3], v);
    }
}

// ---------------------------------------------------------------------------
// W-ary heap, logical index j at physical j + W - 1 so that the W children
// W*j+1 .. W*j+W of node j start on a W-entry boundary (W*16 B aligned).
// ---------------------------------------------------------------------------
// One 16-byte load of a heap entry.  Read field by field, the compiler may
// load the priority alone, compare, and fetch slot and aux only on the branch
// that moves the entry: a second round trip per level.
__device__ __forceinline__ HeapEntry load_entry(gptr<HeapEntry> p) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = *reinterpret_cast<gptr<const u32x4>>(p);
    HeapEntry e;
    e.prio = ((uint64_t)v.y << 32) | v.x;
    e.slot = v.z;
    e.aux = v.w;
    return e;
}

// A heap group's heap during one launch.  Logical entries [0, kTop) -- at
// W = 4 the root and two full levels (1 + W + W^2 = 21 entries), at W = 8 the
// root, its children and the children of nodes 1 and 2 (1 + 3W = 25, within
// the 32-bit dirty mask) -- are held in the group's LDS slice, copied in from
// global memory before the launch works on the heap and written back after;
// deeper entries stay in global memory.  The heaps of a search are shallow
// (tens of entries), so most sift steps and parent checks are LDS accesses
// instead of memory round trips.  A node's W children lie on one side of the
// boundary, so the lanes that read one level take the same path.  (Measured
// on the serving stream: with one pop per launch 1 + 3W entries beat 1 + 2W
// and 1 + 4W by 1-2 %, 1 + W and 1 + 5W by 3-4 %; with two pops per round
// and three rounds per grid launch, 1 + 5W = 21 beat 1 + 3W by 2.6 %: the
// copy in and out is paid once per launch, the accesses it saves per pop.)
template <int W>
struct HeapRef {
    static constexpr uint32_t kTop = W == 4 ? 1 + W + W * W : 1 + 3 * W;
    static constexpr int kPer = (int)((kTop + W - 1) / W);  // LDS entries per lane in fill/write_back
    gptr<HeapEntry> g;  // logical entry 0 in global memory
    lptr<HeapEntry> t;  // logical entry 0 in LDS
    lptr<uint32_t> dirty;  // bit j: LDS entry j changed since the fill
    bool track;            // keep `dirty` (else every entry is written back)

    __device__ __forceinline__ HeapEntry get(uint32_t j) const {
        if (j < kTop) return t[j];
        return load_entry(g + j);
    }
    __device__ __forceinline__ uint64_t prio(uint32_t j) const { return j < kTop ? t[j].prio : g[j].prio; }
    __device__ __forceinline__ void put(uint32_t j, const HeapEntry& e) const {
        if (j < kTop) {
            t[j] = e;
            if constexpr (kTop <= 32)  // wider tops keep no dirty mask
                if (track) __hip_atomic_fetch_or(dirty, 1u << j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            g[j] = e;
        }
    }
    // Copies entries [0, min(S, kTop)) into LDS (lane l: l, l+W, ...).  The
    // loads carry no branch (positions at or past S read entry 0, which every
    // heap's storage holds, and are dropped), so they share one round trip.
    __device__ __forceinline__ void fill(uint32_t S, int lane) const {
        HeapEntry v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const uint32_t j = (uint32_t)lane + (uint32_t)i * W;
            v[i] = load_entry(g + (j < S && j < kTop ? j : 0u));
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const uint32_t j = (uint32_t)lane + (uint32_t)i * W;
            if (j < S && j < kTop) t[j] = v[i];
        }
    }
    // Writes entries [0, min(S, kTop)) back to global memory: all of them, or
    // (with a dirty mask, tops of at most 32 entries) those the launch changed.
    __device__ __forceinline__ void write_back(uint32_t S, int lane) const {
        if constexpr (kTop <= 32) {
            const uint32_t mask = track ? *dirty : ~0u;
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const uint32_t j = (uint32_t)lane + (uint32_t)i * W;
                if (j < S && j < kTop && ((mask >> j) & 1u)) g[j] = t[j];
            }
        } else {
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const uint32_t j = (uint32_t)lane + (uint32_t)i * W;
                if (j < S && j < kTop) g[j] = t[j];
            }
        }
    }
};
static_assert(HeapRef<4>::kTop <= 32 && HeapRef<8>::kTop <= 32, "dirty masks are 32-bit");
static_assert(StepSharedT<256>::kHeapTopEntries >= 64 * HeapRef<4>::kTop &&
                  StepSharedT<256>::kHeapTopEntries >= 32 * HeapRef<8>::kTop,
              "the LDS slice of every heap group of a workgroup");

// Removes the minimum.  The root, the last entry and the root's W children
// are read together; the popped state's key and stored g (via `load`) are
// then read while the sift-down runs.  If the root's f exceeds gate_f the
// heap is left as it is and false is returned (`top` still holds the root).
template <int W, class Load, class Pre, class Key>
__device__ inline bool heap_pop(const HeapRef<W>& H, uint32_t& S, int lane, uint32_t& levels, uint32_t& hbm_levels,
                                Load load, Pre pre, Key& key, uint32_t& g, uint32_t gate_f, HeapEntry& top) {
    top = H.get(0);
    const uint32_t n = S - 1;
    HeapEntry last;
    last.prio = ~0ull;
    last.slot = 0;
    last.aux = 0;
    if (n > 0) last = H.get(n);
    uint32_t idx = 1 + (uint32_t)lane;
    HeapEntry e;
    e.prio = ~0ull;
    e.slot = 0;
    e.aux = 0;
    if (idx < n) e = H.get(idx);
    if (prio_f(top.prio) > gate_f) return false;
    S = n;
    load(top, key, g);
    // The first level's choice uses entries read with the root; it is made
    // before pre() issues its memory operations, so that no wait in the
    // sift-down below covers them (waits on loads issued later still do).
    uint64_t mp = e.prio;
    uint32_t mi = idx;
    grp_min_idx<W>(mp, mi);
    pre(top);  // the problem's early work for expanding top
    if (n == 0) return true;
    uint32_t j = 0;
    for (;;) {
        // e holds child W*j + 1 + lane of node j (or +inf past the end); mp/mi its group minimum
        if (W * j + 1 >= n) break;
        if (mp >= last.prio) break;
        if (idx == mi) H.put(j, e);
        j = mi;
        ++levels;
        if (W * j + 1 >= HeapRef<W>::kTop) ++hbm_levels;  // this level's children are not in the LDS copy
        idx = W * j + 1 + (uint32_t)lane;
        e.prio = ~0ull;
        if (idx < n) e = H.get(idx);
        mp = e.prio;
        mi = idx;
        grp_min_idx<W>(mp, mi);
    }
    if (lane == 0) H.put(j, last);
    return true;
}

template <int W>
__device__ inline void heap_push(const HeapRef<W>& H, uint32_t S, const HeapEntry& x, uint32_t& levels,
                                 uint32_t& hbm_levels) {
    uint32_t j = S;
    while (j > 0) {
        const uint32_t p = (j - 1) / W;
        const HeapEntry pe = H.get(p);
        if (x.prio < pe.prio) {
            H.put(j, pe);
            j = p;
            ++levels;
            if (p >= HeapRef<W>::kTop) ++hbm_levels;  // the parent was read from global memory
        } else {
            break;
        }
    }
    H.put(j, x);
}

// ---------------------------------------------------------------------------
// closed sets
// ---------------------------------------------------------------------------
// Open addressing with linear probing.  A key is claimed with a 64-bit CAS;
// keys are never removed, so a non-empty key read with a plain load is exact
// and a stale "empty" read is corrected by the CAS.
__device__ inline bool hash_find_or_insert(gptr<Slot> table, uint64_t mask, uint64_t key, uint32_t max_probes,
                                           uint32_t& slot, bool& isnew, uint32_t& probes) {
    // Every probe is one CAS(empty -> key): it claims a free slot, or returns the
    // key already there (found, or occupied by another key: probe on).  One
    // round trip per probe instead of a load followed by a CAS.
    uint64_t i = mix64(key) & mask;
    for (uint32_t p = 0; p < max_probes; ++p) {
        ++probes;
        uint64_t prev = kEmptyKey;  // receives the key found in the slot
        __hip_atomic_compare_exchange_strong(&table[i].key, &prev, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        if (prev == kEmptyKey || prev == key) {
            slot = (uint32_t)i;
            isnew = prev == kEmptyKey;
            return true;
        }
        i = (i + 1) & mask;
    }
    return false;
}

__device__ inline bool hash_find(gptr<const Slot> table, uint64_t mask, uint64_t key, uint64_t max_probes,
                                 uint32_t& slot) {
    uint64_t i = mix64(key) & mask;
    for (uint64_t p = 0; p < max_probes; ++p) {
        const uint64_t k = table[i].key;
        if (k == key) {
            slot = (uint32_t)i;
            return true;
        }
        if (k == kEmptyKey) return false;
        i = (i + 1) & mask;
    }
    return false;
}

// ---------------------------------------------------------------------------
// closed-set layouts (a problem policy derives from one of them)
// ---------------------------------------------------------------------------
// Every policy provides: Key, kDense, load (key and g|parent of a slot),
// val_ptr, find_or_insert, find, is_goal, start_key, goal_key.

// Hashed set of 64-bit keys: Slot{key, val}, open addressing.
struct HashedU64Set {
    using Key = uint64_t;
    static constexpr bool kDense = false;
    // successor i's prefetched value, or 0 (no successor) when it is the
    // popped node's stored parent; the sets without parents keep every move
    __device__ static inline uint32_t drop_parent(const InstDesc&, uint32_t, int, uint32_t w, uint32_t) { return w; }
    // The key and the stored g (the high word of val; the parent half is not
    // needed at a pop, and a load of it would only hold a register).
    // (par: the stored parent, which only the dense sets read: kNone here)
    __device__ static inline void load(const InstDesc& d, uint32_t slot, Key& key, uint32_t& g, uint32_t& par) {
        par = kNone;
        // The key may have been claimed by a CAS earlier in this launch, so it
        // is read at agent scope (a plain load could see a stale empty L1
        // line, and a re-read on that case would be a branch that waits for
        // the load).  A stale g only makes the staleness test conservative.
        key = __hip_atomic_load(&as_global(d.table)[slot].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        g = g_of(&as_global(d.table)[slot].val);
    }
    __device__ static inline gptr<uint64_t> val_ptr(const InstDesc& d, uint32_t slot) {
        return &as_global(d.table)[slot].val;
    }
    __device__ static inline bool find_or_insert(const InstDesc& d, Key key, uint32_t max_probes, uint32_t& slot,
                                                 bool& isnew, uint32_t& probes) {
        return hash_find_or_insert(as_global(d.table), d.table_mask, key, max_probes, slot, isnew, probes);
    }
    __device__ static inline bool find(const InstDesc& d, Key key, uint32_t& slot) {
        return hash_find(as_global<const Slot>(d.table), d.table_mask, key, d.table_mask + 1, slot);
    }
    __device__ static inline bool is_goal(const InstDesc& d, Key key) { return key == d.goal; }
    __device__ static inline Key start_key(const InstDesc& d) { return d.start; }
    __device__ static inline Key goal_key(const InstDesc& d) { return d.goal; }
};

// Goal position (row << 4 | col) of puzzle tile t, t < TILES.  The table is
// read as whole words at uniform addresses (scalar loads, once per wave) and
// the byte picked in registers: a byte load indexed by the lane's tile would
// put a memory round trip inside every expansion.
template <int TILES>
__device__ inline uint32_t goal_of(const InstDesc& d, uint32_t t) {
    static_assert(TILES <= (int)sizeof(d.goal_rc), "goal table too small");
    // (readfirstlane makes each word an opaque uniform value, so that the
    // selection below is not turned back into a lane-indexed load)
    const uint32_t* w = reinterpret_cast<const uint32_t*>(d.goal_rc);
    uint32_t word = __builtin_amdgcn_readfirstlane(w[0]);
#pragma unroll
    for (int k = 1; k < (TILES + 3) / 4; ++k) {
        const uint32_t wk = __builtin_amdgcn_readfirstlane(w[k]);
        word = (t >> 2) == (uint32_t)k ? wk : word;
    }
    return (word >> (8 * (t & 3))) & 255u;
}

// Direct-indexed set over grid cells: slot = the cell's tiled index
// (grid_index), val in d.dense.
struct DenseCellSet {
    using Key = uint64_t;
    static constexpr bool kDense = true;
    // successor i's prefetched value, or 0 (no successor) when it is the
    // popped node's stored parent; the sets without parents keep every move
    __device__ static inline uint32_t drop_parent(const InstDesc&, uint32_t, int, uint32_t w, uint32_t) { return w; }
    // the stored g and parent cell in one 8-byte load (the parent lets the
    // expansion skip the move back to it: see GridOps::drop_parent)
    __device__ static inline void load(const InstDesc& d, uint32_t slot, Key& key, uint32_t& g, uint32_t& par) {
        key = slot;
        const uint64_t v = as_global(d.dense)[slot];
        g = val_g(v);
        par = val_parent(v);
    }
    __device__ static inline gptr<uint64_t> val_ptr(const InstDesc& d, uint32_t slot) {
        return as_global(d.dense) + slot;
    }
    __device__ static inline bool find_or_insert(const InstDesc&, Key key, uint32_t, uint32_t& slot, bool& isnew,
                                                 uint32_t& probes) {
        slot = (uint32_t)key;
        isnew = false;  // dense sets do not count occupancy
        ++probes;
        return true;
    }
    __device__ static inline bool find(const InstDesc&, Key key, uint32_t& slot) {
        slot = (uint32_t)key;
        return true;
    }
    __device__ static inline bool is_goal(const InstDesc& d, Key key) { return key == d.goal; }
    __device__ static inline Key start_key(const InstDesc& d) { return d.start; }
    __device__ static inline Key goal_key(const InstDesc& d) { return d.goal; }
};

// ---------------------------------------------------------------------------
// keys wider than 64 bits (24-puzzle, user states of more than 8 bytes)
// ---------------------------------------------------------------------------
constexpr uint64_t kBusyKey = ~0ull - 1;  // wide-key slot being written

// The bytes of a state as 32-bit words.  Keys are handled in this form so
// that T needs no default constructor (the reference's states delete it).
template <typename T>
struct RawKey {
    static_assert(sizeof(T) % 4 == 0, "state size must be a multiple of 4 bytes");
    uint32_t w[sizeof(T) / 4];
};

template <typename T>
__host__ __device__ inline RawKey<T> to_raw(const T& t) {
    RawKey<T> k;
    __builtin_memcpy(&k, &t, sizeof(T));
    return k;
}

template <typename T>
__host__ __device__ inline T from_raw(const RawKey<T>& k) {
    return *reinterpret_cast<const T*>(&k);
}

template <typename T>
__host__ __device__ inline uint64_t hash_words(const RawKey<T>& k) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ sizeof(T);
    for (size_t i = 0; i < sizeof(T) / 4; ++i) h = mix64(h ^ ((uint64_t)k.w[i] << 7) ^ i);
    return h;
}

// States wider than 8 bytes: Slot.key holds a tag (empty / busy / a 62-bit
// fingerprint) and d.keys[slot] the state, each 32-bit word in a 64-bit cell
// that holds (1 << 32) | word once written (the side array is filled with
// ones).  A lane that claims an empty slot stores the words, then publishes the
// fingerprint, in the same loop iteration; a lane that meets a busy slot, or
// its fingerprint with a word still unwritten, looks again on its next
// iteration, so no lane waits inside an inner loop on another lane of its own
// wave.  Every word is read and written whole with an agent-scope atomic, so
// no fence orders the words and the tag (an agent-scope release and acquire
// are an L2 write-back and invalidate on this part, whose XCDs each have their
// own L2).  The loop is bounded; running out counts as a closed-set overflow.
template <typename T>
struct WideKeySet {
    using Key = RawKey<T>;
    static constexpr bool kDense = false;
    static constexpr size_t kWords = sizeof(T) / 4;
    struct Cells {
        uint64_t c[kWords];
    };
    static constexpr uint64_t kUnwritten = ~0ull;  // a cell before its word is stored
    // successor i's prefetched value, or 0 (no successor) when it is the
    // popped node's stored parent; the sets without parents keep every move
    __device__ static inline uint32_t drop_parent(const InstDesc&, uint32_t, int, uint32_t w, uint32_t) { return w; }
    static constexpr uint32_t kMaxIters = 1u << 20;

    __device__ static inline gptr<uint64_t> cells(const InstDesc& d, uint64_t slot) {
        return as_global(reinterpret_cast<uint64_t*>(d.keys)) + slot * kWords;
    }
    // the key in a slot whose fingerprint was seen: 1 (it is `key`), 0 (another
    // key), -1 (a word not yet visible: look again)
    __device__ static inline int match(const InstDesc& d, uint64_t slot, const Key& key) {
        const gptr<uint64_t> c = cells(d, slot);
        uint64_t v[kWords];
        for (size_t i = 0; i < kWords; ++i) v[i] = __hip_atomic_load(c + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool unwritten = false, same_words = true;
        for (size_t i = 0; i < kWords; ++i) {
            unwritten |= v[i] == kUnwritten;
            same_words &= (uint32_t)v[i] == key.w[i];
        }
        return unwritten ? -1 : same_words ? 1 : 0;
    }
    // the key of a slot that entered a heap: its words were complete when the
    // lane that pushed it wrote or matched them (a later launch, or that lane)
    __device__ static inline Key load_key(const InstDesc& d, uint32_t slot) {
        const gptr<uint64_t> c = cells(d, slot);
        Key k;
        for (size_t i = 0; i < kWords; ++i)
            k.w[i] = (uint32_t)__hip_atomic_load(c + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return k;
    }
    __device__ static inline bool same(const Key& a, const Key& b) {
        for (size_t i = 0; i < kWords; ++i)
            if (a.w[i] != b.w[i]) return false;
        return true;
    }
    __device__ static inline void load(const InstDesc& d, uint32_t slot, Key& key, uint32_t& g, uint32_t& par) {
        par = kNone;
        key = load_key(d, slot);
        g = g_of(&as_global(d.table)[slot].val);
    }
    __device__ static inline gptr<uint64_t> val_ptr(const InstDesc& d, uint32_t slot) {
        return &as_global(d.table)[slot].val;
    }
    // Every probe is one CAS(empty -> busy): it claims a free slot, or returns
    // the tag already there (one round trip, where a load of the tag followed
    // by the claim took two for every new key).  All side effects (claim, key
    // words, publish) happen inside the loop body of the iteration that claims
    // the slot; the loop exit carries none, so a lane of the same wave that
    // saw the slot busy finds it published on its next iteration (divergent
    // exits are only taken once the wave leaves).
    __device__ static inline bool find_or_insert(const InstDesc& d, const Key& key, uint32_t max_probes,
                                                 uint32_t& slot, bool& isnew, uint32_t& probes) {
        const uint64_t h = hash_words(key);
        const uint64_t fp = h >> 2;  // never equal to the empty or busy tag
        uint64_t i = h & d.table_mask;
        uint32_t probed = 0;
        int outcome = 0;  // 0 searching, 1 found, 2 inserted, 3 table exhausted
        for (uint32_t iter = 0; iter < kMaxIters; ++iter) {
            ++probes;
            const gptr<uint64_t> tagp = &as_global(d.table)[i].key;
            uint64_t tag = kEmptyKey;  // receives the tag found in the slot
            __hip_atomic_compare_exchange_strong(tagp, &tag, kBusyKey, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if (tag == kEmptyKey) {  // claimed: store the words, then publish the tag
                const gptr<uint64_t> c = cells(d, i);
                for (size_t w = 0; w < kWords; ++w)
                    __hip_atomic_store(c + w, (1ull << 32) | key.w[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(tagp, fp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                outcome = 2;
            }
            bool again = tag == kBusyKey;  // a busy slot is looked at again next iteration
            if (outcome == 0 && !again) {
                if (tag == fp) {
                    const int m = match(d, i, key);
                    if (m > 0) outcome = 1;
                    again = m < 0;  // published, a word not yet visible
                }
                if (outcome == 0 && !again) {
                    if (++probed >= max_probes || probed > d.table_mask) outcome = 3;
                    else i = (i + 1) & d.table_mask;
                }
            }
            if (outcome != 0) break;
        }
        slot = (uint32_t)i;
        isnew = outcome == 2;
        return outcome == 1 || outcome == 2;
    }
    __device__ static inline bool find(const InstDesc& d, const Key& key, uint32_t& slot) {
        const uint64_t h = hash_words(key);
        const uint64_t fp = h >> 2;
        uint64_t i = h & d.table_mask;
        for (uint64_t p = 0; p <= d.table_mask; ++p) {
            const uint64_t tag = as_global(d.table)[i].key;
            if (tag == kEmptyKey) return false;
            if (tag == fp && match(d, i, key) > 0) {
                slot = (uint32_t)i;
                return true;
            }
            i = (i + 1) & d.table_mask;
        }
        return false;
    }
    __device__ static inline bool is_goal(const InstDesc& d, const Key& key) { return same(key, goal_key(d)); }
    __device__ static inline Key start_key(const InstDesc& d) { return as_global(reinterpret_cast<const Key*>(d.user))[0]; }
    __device__ static inline Key goal_key(const InstDesc& d) { return as_global(reinterpret_cast<const Key*>(d.user))[1]; }
    __host__ __device__ static inline T unpack(const Key& k) { return from_raw<T>(k); }
    __host__ __device__ static inline Key pack(const T& t) { return to_raw<T>(t); }
};

// Additive pattern databases of the 15-puzzle towards the ordered goal (tile
// t on cell t - 1, blank on cell 15): the relaxation of the 24-puzzle's tables
// below (a pattern's tiles step into any neighbouring cell not held by a tile
// of the same pattern), with the 15 tiles in three disjoint patterns of five.
// Pattern p's table (kPdb15Table bytes at p * kPdb15Table) is indexed by the
// sum over its tiles i of cell * 16^i; the sum over the patterns is
// admissible and consistent.
constexpr int kPdb15Patterns = 3;
constexpr uint32_t kPdb15Tiles[kPdb15Patterns][5] = {{1, 2, 3, 5, 6}, {4, 7, 8, 11, 12}, {9, 10, 13, 14, 15}};
constexpr uint64_t kPdb15Table = 1ull << 20;  // 16^5
constexpr uint64_t kPdb15Bytes = kPdb15Patterns * kPdb15Table;
constexpr uint64_t kPuzzle4OrderedGoal = 0x0FEDCBA987654321ull;
// pattern p: nibble t holds i + 1 for tile t at position i of the pattern, else 0
constexpr uint64_t pdb15_pattern_word(int p) {
    uint64_t w = 0;
    for (int i = 0; i < 5; ++i) w |= (uint64_t)(i + 1) << (4 * kPdb15Tiles[p][i]);
    return w;
}
// pattern of tile t (0..2), 3 for the blank: two bits per tile
constexpr uint32_t pdb15_pattern_of_word() {
    uint32_t w = 3u;  // tile 0, the blank
    for (int p = 0; p < kPdb15Patterns; ++p)
        for (int i = 0; i < 5; ++i) w |= (uint32_t)p << (2 * kPdb15Tiles[p][i]);
    return w;
}

// ---------------------------------------------------------------------------
// built-in problems
// ---------------------------------------------------------------------------
// Sliding puzzle of side S, state = 16 nibbles (tile of cell i in nibble i).
// Moves in the reference order (a_star_sliding_puzzle.h:88-105); unit cost
// (:84); h = Manhattan distance of the tiles, blank excluded (the reference's
// manhattan_heuristic, :152-163, counts the blank and is not admissible).
template <int S>
struct PuzzleOps : HashedU64Set {
    static constexpr int K = 4;
    static constexpr int kCells = S * S;
    static constexpr uint64_t kLowBits = (kCells == 16) ? 0x1111111111111111ull
                                                        : ((1ull << (4 * kCells)) - 1) & 0x1111111111111111ull;

    __device__ static inline int blank_of(uint64_t s) {
        uint64_t y = s | (s >> 1);
        y |= y >> 2;  // bit 4c = OR of the bits of nibble c
        const uint64_t z = ~y & kLowBits;
        return (int)(__builtin_ctzll(z) >> 2);
    }

    __device__ static inline uint32_t md(const InstDesc& d, uint32_t tile, int cell) {
        const uint32_t g = goal_of<kCells>(d, tile);
        const int gr = (int)(g >> 4), gc = (int)(g & 15u);
        const int r = cell / S, c = cell % S;
        return (uint32_t)(abs(r - gr) + abs(c - gc));
    }

    // 15-puzzles towards the ordered goal with pattern databases (d.pdb): the
    // placement index of pattern p's tiles in state s, and a table entry
    __device__ static inline uint32_t pdb_index(uint64_t s, uint32_t p) {
        constexpr uint64_t w0 = pdb15_pattern_word(0), w1 = pdb15_pattern_word(1), w2 = pdb15_pattern_word(2);
        const uint64_t w = p == 0 ? w0 : p == 1 ? w1 : w2;
        uint32_t idx = 0;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint32_t at = (uint32_t)(w >> (4 * ((s >> (4 * c)) & 15u))) & 15u;  // position + 1, or 0
            idx += at ? (uint32_t)c << (4 * (at - 1)) : 0u;
        }
        return idx;
    }
    __device__ static inline uint32_t pdb_value(const InstDesc& d, uint32_t p, uint32_t idx) {
        return as_global(d.pdb)[(uint64_t)p * kPdb15Table + idx];
    }

    __device__ static inline uint32_t heuristic(const InstDesc& d, uint64_t s, uint32_t) {
        uint32_t h = 0;
        if constexpr (S == 4) {
            if (d.pdb) {
                for (uint32_t p = 0; p < kPdb15Patterns; ++p) h += pdb_value(d, p, pdb_index(s, p));
                return h;
            }
        }
#pragma unroll
        for (int c = 0; c < kCells; ++c) {
            const uint32_t t = (uint32_t)(s >> (4 * c)) & 15u;
            if (t) h += md(d, t, c);
        }
        return h;
    }

    __device__ static inline uint32_t aux_of(const InstDesc&, uint64_t) { return 0; }

    // successors are computed from the state alone: nothing to read early
    __device__ static inline uint32_t prefetch(const InstDesc&, const HeapEntry&, int) { return 0; }

    // h of the successor from h of the parent: only the moved tile changes.
    // aux = the cell the blank left to reach this state, + 1 (0 at the start):
    // moving it back would recreate the parent, whose stored g is lower, so
    // that move is not generated (its closed-set atomic could not succeed).
    __device__ static inline bool succ(const InstDesc& d, uint64_t s, uint32_t from, uint32_t h, int i, uint32_t,
                                       uint64_t& t, uint32_t& taux, uint32_t& cost, uint32_t& th) {
        const int b = blank_of(s);
        int r = b / S, c = b % S;
        if (i == 0) ++r;
        else if (i == 1) --r;
        else if (i == 2) ++c;
        else --c;
        if (r < 0 || r >= S || c < 0 || c >= S) return false;
        const int nb = r * S + c;
        if ((uint32_t)nb + 1 == from) return false;
        const uint32_t tile = (uint32_t)(s >> (4 * nb)) & 15u;
        t = (s & ~(0xFull << (4 * nb))) | ((uint64_t)tile << (4 * b));
        taux = (uint32_t)b + 1;
        cost = 1;
        if constexpr (S == 4) {
            if (d.pdb) {  // only the moved tile's pattern changes: its placement index moves by (b - nb) * 16^i
                constexpr uint32_t pw = pdb15_pattern_of_word();
                const uint32_t p = (pw >> (2 * tile)) & 3u;
                const uint32_t before = pdb_index(s, p);
                const uint32_t i = ((uint32_t)((p == 0 ? pdb15_pattern_word(0) : p == 1 ? pdb15_pattern_word(1)
                                                                                         : pdb15_pattern_word(2)) >>
                                               (4 * tile)) & 15u) - 1u;
                const uint32_t after = (uint32_t)((int)before + (b - nb) * (1 << (4 * i)));
                th = h - pdb_value(d, p, before) + pdb_value(d, p, after);
                return true;
            }
        }
        th = h - md(d, tile, nb) + md(d, tile, b);
        return true;
    }
};

// 4- or 8-connected grid, state = the cell's tiled index (grid_index), aux =
// y << 16 | x.  Moves in the reference order (a_star_pathfinding.h:150-182);
// cost = weight of the entered cell (:203-208), 0 = wall; h = Manhattan (4) or
// Chebyshev (8) distance times the minimum cell weight, both admissible (the
// reference's plain Manhattan with unit diagonals, :219-221, is not).  The
// closed set and the weights are both in the tiled layout.
template <int CONN>
struct GridOps : DenseCellSet {
    static constexpr int K = CONN;

    __device__ static inline uint32_t hxy(const InstDesc& d, uint32_t x, uint32_t y) {
        const uint32_t gx = d.goal_xy & 0xFFFFu, gy = d.goal_xy >> 16;
        const uint32_t dx = x > gx ? x - gx : gx - x;
        const uint32_t dy = y > gy ? y - gy : gy - y;
        if (CONN == 8) return (dx > dy ? dx : dy) * d.w_min;
        return (dx + dy) * d.w_min;
    }

    // aux of the start state (the only state whose aux is not made by succ)
    __device__ static inline uint32_t aux_of(const InstDesc& d, uint64_t) { return d.start_xy; }

    __device__ static inline uint32_t heuristic(const InstDesc& d, uint64_t s, uint32_t aux) {
        return hxy(d, aux & 0xFFFFu, aux >> 16);
    }

    // Weight of the cell that move i enters (0: wall or off the grid), read as
    // soon as the node is popped, while its heap is still being sifted.
    __device__ static inline uint32_t prefetch(const InstDesc& d, const HeapEntry& e, int i) {
        uint32_t nx, ny;
        if (!neighbour(d, e.aux, i, nx, ny)) return 0;
        return as_global(d.cells)[grid_index(nx, ny, d.tile_cols)];
    }

    // Stored g of the cell that move i enters (kInf off the grid), read with its weight.
    __device__ static inline uint32_t prefetch_g(const InstDesc& d, const HeapEntry& e, int i) {
        uint32_t nx, ny;
        if (!neighbour(d, e.aux, i, nx, ny)) return kInf;
        return g_of(as_global(d.dense) + grid_index(nx, ny, d.tile_cols));
    }

    __device__ static inline bool neighbour(const InstDesc& d, uint32_t aux, int i, uint32_t& nx, uint32_t& ny) {
        // move i: dx = {1,-1,0,0,1,1,-1,-1}[i], dy = {0,0,1,-1,1,-1,1,-1}[i],
        // each stored +1 in nibble i of an immediate (no table in memory)
        const uint32_t dx = (0x00221102u >> (4 * i)) & 15u, dy = (0x02020211u >> (4 * i)) & 15u;
        nx = (aux & 0xFFFFu) + dx - 1u;
        ny = (aux >> 16) + dy - 1u;
        return nx < d.width && ny < d.height;  // also catches -1 wrap-around
    }

    // The stored parent cell of a node popped with its stored g: the move back
    // to it would give that cell g + its weight, above its stored g (every
    // weight >= 1), so it is not generated.
    __device__ static inline uint32_t drop_parent(const InstDesc& d, uint32_t aux, int i, uint32_t w, uint32_t par) {
        uint32_t nx, ny;
        if (!neighbour(d, aux, i, nx, ny)) return w;
        return grid_index(nx, ny, d.tile_cols) == par ? 0u : w;
    }

    // `w` = prefetch(d, aux, i)
    __device__ static inline bool succ(const InstDesc& d, uint64_t, uint32_t aux, uint32_t, int i, uint32_t w,
                                       uint64_t& t, uint32_t& taux, uint32_t& cost, uint32_t& th) {
        uint32_t nx, ny;
        if (w == 0 || !neighbour(d, aux, i, nx, ny)) return false;
        t = grid_index(nx, ny, d.tile_cols);
        taux = (ny << 16) | nx;
        cost = w;
        th = hxy(d, nx, ny);
        return true;
    }
};

// 24-puzzle (5x5 board).  Board5 packs cells 0..11 into lo and 12..23 into hi,
// 5 bits each (bits 5c..5c+4), and cell 24 into the spare bits: its low 4
// bits in lo[60..63], its high bit in hi[60].  aux = the blank's cell (and,
// from bit 8, the cell it left to reach the state, + 1).  Moves,
// cost and heuristic as PuzzleOps (a_star_sliding_puzzle.h:83-123, Manhattan
// without the blank).
struct Board5 {
    uint64_t lo, hi;
};

// Closed set of 24-puzzle boards: a 32-byte slot {lo, hi, val, unused} holds
// the board and its g | parent in one line.  lo is claimed by CAS(empty ->
// lo); it never has all bits set (its 5-bit fields hold tiles 0..24), so it
// cannot be taken for an empty slot.  hi is stored right after the claim; a
// board's hi has its top bit clear and the table is filled with all bits
// set, so a reader that finds its own lo with hi's top bit set knows the
// board is still being written and looks again on its next iteration.  Each
// word is read and written whole, so no fence orders the two: a claim is one
// atomic, a lookup one atomic and one load issued together (WideKeySet, for
// states of any width, keeps a tag and a side array of marked words instead).
struct Board5KeySet {
    using Key = RawKey<Board5>;
    static constexpr bool kDense = false;
    static constexpr uint32_t kMaxIters = 1u << 20;
    struct alignas(32) PairSlot {
        uint64_t lo, hi, val, unused;
    };
    static constexpr uint64_t kUnwritten = 1ull << 63;  // hi's top bit: the board is not yet stored

    __device__ static inline gptr<PairSlot> slots(const InstDesc& d) {
        return as_global(reinterpret_cast<PairSlot*>(d.table));
    }
    __host__ __device__ static inline uint64_t lo_of(const Key& k) { return (uint64_t)k.w[0] | (uint64_t)k.w[1] << 32; }
    __host__ __device__ static inline uint64_t hi_of(const Key& k) { return (uint64_t)k.w[2] | (uint64_t)k.w[3] << 32; }
    __device__ static inline uint32_t drop_parent(const InstDesc&, uint32_t, int, uint32_t w, uint32_t) { return w; }
    __device__ static inline void load(const InstDesc& d, uint32_t slot, Key& key, uint32_t& g, uint32_t& par) {
        // agent-scope loads: with several rounds per launch a board stored by
        // another wave earlier in this launch can be popped
        const gptr<PairSlot> s = slots(d) + slot;
        const uint64_t lo = __hip_atomic_load(&s->lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t hi = __hip_atomic_load(&s->hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        g = g_of(&s->val);
        par = kNone;
        key.w[0] = (uint32_t)lo;
        key.w[1] = (uint32_t)(lo >> 32);
        key.w[2] = (uint32_t)hi;
        key.w[3] = (uint32_t)(hi >> 32);
    }
    __device__ static inline gptr<uint64_t> val_ptr(const InstDesc& d, uint32_t slot) {
        return &(slots(d) + slot)->val;
    }
    // As in WideKeySet, all side effects happen in the loop body of the
    // iteration that claims the slot, so a lane that met the slot half
    // written finds it complete on its next iteration.
    __device__ static inline bool find_or_insert(const InstDesc& d, const Key& key, uint32_t max_probes,
                                                 uint32_t& slot, bool& isnew, uint32_t& probes) {
        const uint64_t lo = lo_of(key), hi = hi_of(key);
        uint64_t i = hash_words(key) & d.table_mask;
        uint32_t probed = 0;
        int outcome = 0;  // 0 searching, 1 found, 2 inserted, 3 table exhausted
        for (uint32_t iter = 0; iter < kMaxIters; ++iter) {
            ++probes;
            const gptr<PairSlot> s = slots(d) + i;
            // hi is read with the claim: a board stored earlier compares at once
            const uint64_t seen_hi = __hip_atomic_load(&s->hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t prev = kEmptyKey;  // receives the lo found in the slot
            __hip_atomic_compare_exchange_strong(&s->lo, &prev, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            bool again = false;  // the same slot next iteration: its board is being written
            if (prev == kEmptyKey) {
                __hip_atomic_store(&s->hi, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                outcome = 2;
            } else if (prev == lo) {
                if (seen_hi & kUnwritten) again = true;
                else if (seen_hi == hi) outcome = 1;
            }
            if (outcome == 0 && !again) {
                if (++probed >= max_probes || probed > d.table_mask) outcome = 3;
                else i = (i + 1) & d.table_mask;
            }
            if (outcome != 0) break;
        }
        slot = (uint32_t)i;
        isnew = outcome == 2;
        return outcome == 1 || outcome == 2;
    }
    // After the search (no concurrent writers).
    __device__ static inline bool find(const InstDesc& d, const Key& key, uint32_t& slot) {
        const uint64_t lo = lo_of(key), hi = hi_of(key);
        uint64_t i = hash_words(key) & d.table_mask;
        for (uint64_t p = 0; p <= d.table_mask; ++p) {
            const PairSlot& s = reinterpret_cast<const PairSlot*>(d.table)[i];
            if (s.lo == kEmptyKey) return false;
            if (s.lo == lo && s.hi == hi) {
                slot = (uint32_t)i;
                return true;
            }
            i = (i + 1) & d.table_mask;
        }
        return false;
    }
    __device__ static inline bool is_goal(const InstDesc& d, const Key& key) {
        const Key g = goal_key(d);
        return lo_of(key) == lo_of(g) && hi_of(key) == hi_of(g);
    }
    __device__ static inline Key start_key(const InstDesc& d) { return as_global(reinterpret_cast<const Key*>(d.user))[0]; }
    __device__ static inline Key goal_key(const InstDesc& d) { return as_global(reinterpret_cast<const Key*>(d.user))[1]; }
};

// Additive pattern databases of the 24-puzzle towards the ordered goal (tile
// t on cell t - 1, blank on cell 24).  The 24 tiles form five disjoint
// patterns; pattern p's table (kPdbTable bytes at p * kPdbTable) holds, for
// each placement of its tiles (index: sum over its tiles i of cell * 25^i),
// the fewest moves of those tiles that bring them to their goal cells when a
// tile may step into any neighbouring cell not held by a tile of the same
// pattern.  That relaxes the puzzle, and each real move moves one tile of one
// pattern, so the sum over the patterns is admissible and changes by at most
// 1 per move (consistent).
constexpr int kPdbPatterns = 5;
constexpr uint32_t kPdbTiles[kPdbPatterns][5] = {
    {1, 2, 3, 6, 7}, {4, 5, 8, 9, 10}, {11, 12, 16, 17, 21}, {13, 14, 18, 19, 22}, {15, 20, 23, 24, 0}};
constexpr int kPdbCount[kPdbPatterns] = {5, 5, 5, 5, 4};
constexpr uint64_t kPdbTable = 9765625;  // 25^5
constexpr uint64_t kPdbBytes = kPdbPatterns * kPdbTable;

// Pattern (high 3 bits) and position in the pattern (low 3 bits) of tile t;
// 63 for the blank.
constexpr uint32_t pdb_code_of(uint32_t t) {
    for (int p = 0; p < kPdbPatterns; ++p)
        for (int i = 0; i < kPdbCount[p]; ++i)
            if (kPdbTiles[p][i] == t) return (uint32_t)(p << 3 | i);
    return 63u;
}
// the codes of tiles 10k .. 10k+9, 6 bits each
constexpr uint64_t pdb_code_word(int k) {
    uint64_t w = 0;
    for (int j = 0; j < 10; ++j) w |= (uint64_t)pdb_code_of((uint32_t)(10 * k + j)) << (6 * j);
    return w;
}
__host__ __device__ inline uint32_t pdb_code(uint32_t t) {
    constexpr uint64_t w0 = pdb_code_word(0), w1 = pdb_code_word(1), w2 = pdb_code_word(2);
    const uint64_t w = t < 10 ? w0 : t < 20 ? w1 : w2;
    return (uint32_t)(w >> (6 * (t % 10))) & 63u;
}
__host__ __device__ inline uint32_t pow25(uint32_t i) {
    return i == 0 ? 1u : i == 1 ? 25u : i == 2 ? 625u : i == 3 ? 15625u : 390625u;
}

struct Puzzle5Ops : Board5KeySet {
    static constexpr int K = 4;
    using Key = RawKey<Board5>;

    __host__ __device__ static inline uint32_t get(const Board5& b, int c) {
        if (c < 12) return (uint32_t)(b.lo >> (5 * c)) & 31u;
        if (c < 24) return (uint32_t)(b.hi >> (5 * (c - 12))) & 31u;
        return (uint32_t)(b.lo >> 60) | ((uint32_t)(b.hi >> 60) & 1u) << 4;
    }
    __host__ __device__ static inline void set(Board5& b, int c, uint32_t t) {
        if (c < 12) {
            b.lo = (b.lo & ~(31ull << (5 * c))) | ((uint64_t)t << (5 * c));
        } else if (c < 24) {
            b.hi = (b.hi & ~(31ull << (5 * (c - 12)))) | ((uint64_t)t << (5 * (c - 12)));
        } else {
            b.lo = (b.lo & ~(15ull << 60)) | ((uint64_t)(t & 15u) << 60);
            b.hi = (b.hi & ~(1ull << 60)) | ((uint64_t)(t >> 4) << 60);
        }
    }
    __device__ static inline Board5 board(const Key& k) { return from_raw<Board5>(k); }

    __device__ static inline uint32_t md(const InstDesc& d, uint32_t tile, int cell) {
        const uint32_t g = goal_of<25>(d, tile);
        const int gr = (int)(g >> 4), gc = (int)(g & 15u);
        const int r = cell / 5, c = cell % 5;
        return (uint32_t)((r > gr ? r - gr : gr - r) + (c > gc ? c - gc : gc - c));
    }
    // placement index of pattern p's tiles on board b
    __device__ static inline uint32_t pdb_index(const Board5& b, uint32_t p) {
        uint32_t idx = 0;
        for (int c = 0; c < 25; ++c) {
            const uint32_t code = pdb_code(get(b, c));
            if ((code >> 3) == p) idx += (uint32_t)c * pow25(code & 7u);
        }
        return idx;
    }
    __device__ static inline uint32_t pdb_value(const InstDesc& d, uint32_t p, uint32_t idx) {
        return as_global(d.pdb)[(uint64_t)p * kPdbTable + idx];
    }
    __device__ static inline uint32_t heuristic(const InstDesc& d, const Key& k, uint32_t) {
        const Board5 b = board(k);
        uint32_t h = 0;
        if (d.pdb) {
            for (uint32_t p = 0; p < kPdbPatterns; ++p) h += pdb_value(d, p, pdb_index(b, p));
            return h;
        }
        for (int c = 0; c < 25; ++c) {
            const uint32_t t = get(b, c);
            if (t) h += md(d, t, c);
        }
        return h;
    }
    __device__ static inline uint32_t aux_of(const InstDesc&, const Key& k) {
        const Board5 b = board(k);
        for (int c = 0; c < 25; ++c)
            if (get(b, c) == 0) return (uint32_t)c;
        return 0;
    }
    __device__ static inline uint32_t prefetch(const InstDesc&, const HeapEntry&, int) { return 0; }
    // h of the successor from h of the parent: only the moved tile changes.
    // aux = the blank's cell | (the cell it left to reach this state + 1) << 8
    // (0 at the start): the move back to the parent is not generated.
    __device__ static inline bool succ(const InstDesc& d, const Key& k, uint32_t aux, uint32_t h, int i, uint32_t,
                                       Key& t, uint32_t& taux, uint32_t& cost, uint32_t& th) {
        const int b = (int)(aux & 0xFFu);
        int r = b / 5, c = b % 5;
        if (i == 0) ++r;
        else if (i == 1) --r;
        else if (i == 2) ++c;
        else --c;
        if (r < 0 || r >= 5 || c < 0 || c >= 5) return false;
        const int nb = r * 5 + c;
        if ((uint32_t)nb + 1 == (aux >> 8)) return false;
        Board5 s = board(k);
        const uint32_t tile = get(s, nb);
        if (d.pdb) {  // only the moved tile's pattern changes: its placement index moves by (b - nb) * 25^i
            const uint32_t code = pdb_code(tile), p = code >> 3;
            const uint32_t before = pdb_index(s, p);
            const uint32_t after = (uint32_t)((int)before + (b - nb) * (int)pow25(code & 7u));
            th = h - pdb_value(d, p, before) + pdb_value(d, p, after);
        } else {
            th = h - md(d, tile, nb) + md(d, tile, b);
        }
        set(s, b, tile);
        set(s, nb, 0u);
        t = to_raw(s);
        taux = (uint32_t)nb | ((uint32_t)b + 1) << 8;
        cost = 1;
        return true;
    }
};

// ---------------------------------------------------------------------------
// closed-set update of one successor
// ---------------------------------------------------------------------------
// The contract of the reference's mass insert (fancy_hashmap.h:4-8,56-61):
// a successor is INSERTED when its state was not in the set, UPDATED when it
// was and the successor lowers its g, and REJECTED otherwise; the first two
// survive and are pushed.  Here the state is found or claimed (find_or_insert)
// and then atomicMin(val, g << 32 | parent) decides: the successor survives
// iff the minimum it replaced had a larger g.  Successors of one state that
// meet in the same launch are ordered by the atomic, so exactly one claim per
// state reports it new, the stored value ends as the smallest (g, parent) of
// all of them, and at least the one that stored it survives.  A claim whose
// atomicMin found a smaller or equal g already there (another successor of
// the same state came first) is REJECTED.  CS_FULL: the probe limit was
// reached without finding or claiming a slot (a closed-set overflow).
enum : int { CS_REJECTED = 0, CS_INSERTED = 1, CS_UPDATED = 2, CS_FULL = 3 };

template <class Ops>
__device__ __forceinline__ int closed_set_update(const InstDesc& d, const typename Ops::Key& key, uint32_t g,
                                                 uint32_t parent, uint32_t max_probes, uint32_t& slot, bool& isnew,
                                                 uint32_t& probes) {
    if (!Ops::find_or_insert(d, key, max_probes, slot, isnew, probes)) return CS_FULL;
    const uint64_t mine = ((uint64_t)g << 32) | parent;
    const uint64_t old =
        __hip_atomic_fetch_min(Ops::val_ptr(d, slot), mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g >= val_g(old)) return CS_REJECTED;
    return isnew ? CS_INSERTED : CS_UPDATED;
}

// ---------------------------------------------------------------------------
// one GA* iteration of one heap group
// ---------------------------------------------------------------------------
// Pushes the K records held by the group's lanes (record j in lane j % W,
// register j / W) into the group's own heap; lane 0 performs the pushes.
template <int W, int KW>
__device__ inline void push_records(const HeapRef<W>& H, uint32_t& S, const HeapEntry (&rec)[KW], int K, int lane,
                                    uint32_t best, uint32_t cap, bool& overflow, const Tally& tl, bool from_cand) {
    // Record j = r*W + l sits in lane l, register r.  Valid records take the
    // positions S, S+1, ... in j order.  When every new position's parent is an
    // old entry (true unless the heap is nearly empty), the lanes check their
    // parents in parallel and place the records that already satisfy the heap
    // order in one round trip; lane 0 then sifts up only the others.  Sifting
    // one record up only lowers values on its own path to the root, so a record
    // that satisfied its parent before still does.
    const uint32_t gbase = __lane_id() & ~(uint32_t)(W - 1);
    const uint64_t gmask = (W == 64) ? ~0ull : ((1ull << W) - 1);
    bool v[KW];
    uint32_t rank[KW];
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < KW; ++r) {
        v[r] = (r * W + lane < K) && rec[r].prio != kNoRecord && prio_f(rec[r].prio) < best;
        const uint64_t bits = (__ballot(v[r]) >> gbase) & gmask;
        rank[r] = c + (uint32_t)__popcll(bits & ((1ull << lane) - 1));
        c += (uint32_t)__popcll(bits);
    }
    if (c == 0) return;
    if (S + c > cap) {
        overflow = true;
        return;
    }
    if (S == 0 || (int)c - 2 >= (W - 1) * (int)S) {
        // parents may be among the new records: plain sequential pushes
#pragma unroll
        for (int r = 0; r < KW; ++r) {
#pragma unroll
            for (int l = 0; l < W; ++l) {
                if (r * W + l >= K) break;
                const int vl = (int)grp_bcast<W>(v[r] ? 1u : 0u, l);
                if (!vl) continue;
                HeapEntry x;
                x.prio = grp_bcast64<W>(rec[r].prio, l);
                x.slot = grp_bcast<W>(rec[r].slot, l);
                x.aux = grp_bcast<W>(rec[r].aux, l);
                if (lane == 0) {
                    uint32_t lv = 0, hl = 0;
                    heap_push<W>(H, S, x, lv, hl);
                    tl.prof_add(TL_SIFT_UP, lv);
                    tl.prof_add(TL_SIFT_UP_HBM, hl);
                    ++S;
                }
            }
        }
        S = grp_bcast<W>(S, 0);
        if (lane == 0) {
            tl.prof_add(TL_PUSHED, c);
            if (from_cand) tl.prof_add(TL_PUSH_IN, c);
        }
        return;
    }
    bool bad[KW];
#pragma unroll
    for (int r = 0; r < KW; ++r) {
        bad[r] = false;
        if (v[r]) {
            const uint32_t pos = S + rank[r];
            const uint64_t pp = H.prio((pos - 1) / W);
            if (rec[r].prio >= pp) H.put(pos, rec[r]);
            else bad[r] = true;
        }
    }
#pragma unroll
    for (int r = 0; r < KW; ++r) {
#pragma unroll
        for (int l = 0; l < W; ++l) {
            if (r * W + l >= K) break;
            const int bl = (int)grp_bcast<W>(bad[r] ? 1u : 0u, l);
            if (!bl) continue;
            HeapEntry x;
            x.prio = grp_bcast64<W>(rec[r].prio, l);
            x.slot = grp_bcast<W>(rec[r].slot, l);
            x.aux = grp_bcast<W>(rec[r].aux, l);
            const uint32_t pos = S + grp_bcast<W>(rank[r], l);
            if (lane == 0) {  // sift up from its leaf position
                uint32_t lv = 0, hl = 0;
                heap_push<W>(H, pos, x, lv, hl);
                tl.prof_add(TL_SIFT_UP, lv);
                tl.prof_add(TL_SIFT_UP_HBM, hl);
            }
        }
    }
    if (lane == 0) {
        tl.prof_add(TL_PUSHED, c);
        if (from_cand) tl.prof_add(TL_PUSH_IN, c);
    }
    S += c;
}

// One launch = one GA* iteration for the heaps of one workgroup, in two phases.
//
// Phase 1 (every heap group) reads, in one round trip, the instance's best
// and previous minimum f, its heap's size and root and the records addressed
// to it (the instance's status was loaded by the kernel beforehand; a wave
// that finds the instance finished takes no part).  It contributes the heap's top after
// the pushes -- min(root f, f of the incoming records), the pushes have not
// happened yet -- to the min-f reduction (HeapsHost::minimum, heaps.h:291-294)
// and its size to the open count, and it decides whether the heap has work:
// records to push, or a root within the f gate.  Idle heaps clear their
// candidate slice.
// Phase 2: the heaps with work are packed into the first groups of the
// workgroup through an LDS list, and those groups copy their heap's top levels
// into LDS (HeapRef), push, pop and expand, and write the top levels back;
// waves left without a heap are done.  A wave lives as long as its slowest
// group, so packing spares the other waves from waiting on a few busy heaps.
//
// Expansion: `rounds` rounds per launch.  Successors of rounds 0..rounds-2 go
// into the group's own heap and can be expanded in the same launch; successors
// of the last round go to the candidate buffer and are spread over all heaps
// at the next launch.  The termination test stays exact: each heap's top is
// its top after the distributed pushes and before its first pop, and no other
// group touches the heap, so the tops of all heaps form one snapshot of the
// open set.  A root whose f is more than f_slack above the previous
// iteration's minimum is left in place (the heap sits the launch out) and
// still enters the reduction.
//
// Res (resident heaps, the one-launch form, where a workgroup owns its
// instance's heaps for the whole launch and heap q is the workgroup's heap
// group q): each group keeps its own heap's top levels in its own LDS slice
// and the heap's size in *res_size across iterations, so phase 1 reads size
// and root without a memory access, phase 2 needs neither the copy into LDS
// nor the write-back (the entries at or past the LDS levels live in global
// memory as always), heaps are not packed, and the records phase 1 read stay
// in the group's registers.  At it = 0 the slice is filled from the seeded
// heap.
//
// Every thread of the workgroup calls run_group (it holds a barrier).
template <int W, class Ops, class WL, bool Res = false, int P = 1>
__device__ __forceinline__ void run_group(const InstDesc& d, InstState* st, uint32_t status, uint32_t it, uint32_t q,
                                          bool active, uint32_t cover, const StepParams& prm, const Tally& tl,
                                          WL& wl, HeapEntry* heap_top, uint32_t* dirty_masks,
                                          uint32_t* res_size = nullptr) {
    constexpr int K = Ops::K;
    constexpr int KW = (K + W - 1) / W;
    const int lane = (int)(threadIdx.x & (W - 1));
    const uint32_t N = d.n_heaps;  // heaps allocated; candidate buffers are sized for all of them

    // Frontier width: heaps q < A take part in this iteration; the records read
    // now were written by the A_w heaps of the previous iteration.  A only
    // grows, so a heap that joins starts empty and loses nothing.
    // With a fixed width (the default) every heap is in use from the start and
    // the incoming records need not wait for these loads.
    const bool fixed = d.a0 == 0 || d.a0 >= N;
    const uint32_t A = fixed ? N : __builtin_amdgcn_readfirstlane(load_relaxed(&st->active[it % 3]));
    const uint32_t Aw =
        (fixed || it == 0) ? A : __builtin_amdgcn_readfirstlane(load_relaxed(&st->active[(it - 1) % 3]));

    // ---------------- phase 1 ----------------
    // The loads of phase 1 travel together: best and the previous min f, then
    // (for a running instance: the status was loaded before the descriptor
    // arrived) the heap's size and root and the records addressed to it.
    // Waves may see different statuses (another workgroup may end the
    // instance meanwhile); each only decides whether its own heaps take part.
    const uint32_t best_ld = load_relaxed(&st->best);
    const uint32_t mt_ld = load_relaxed(&st->min_top[(it + 2) % 3]);  // (it - 1) mod 3; unused at it = 0
    uint64_t occupied = 0;
    if constexpr (!Ops::kDense) occupied = load_relaxed(&st->occupied);
    bool running = __builtin_amdgcn_readfirstlane(status) == ST_RUNNING;
    // grow requests of the previous iteration (every workgroup of this launch reads the same
    // value: it was written in the previous launch): the instance pauses here, before any of
    // its heaps does anything in this iteration, and the host resumes it at this iteration
    if (running && it > 0) {
        const uint32_t req = __builtin_amdgcn_readfirstlane(load_relaxed(&st->grow[(it + 2) % 3]));
        if (req) {
            if (threadIdx.x == 0) {
                store_relaxed(&st->paused_iter, it);
                set_status(st, ST_PAUSED);
            }
            running = false;
        }
    }
    active = active && running && q < A;
    const gptr<const HeapEntry> cin = as_global<const HeapEntry>(d.cand) + (size_t)((it - 1) & 1) * N * K;
    uint32_t S = 0;
    uint64_t root_prio = ~0ull;
    HeapEntry rec[KW];
#pragma unroll
    for (int r = 0; r < KW; ++r) {
        rec[r].prio = kNoRecord;
        rec[r].slot = 0;
        rec[r].aux = 0;
    }
    if (active) {
        if constexpr (Res) {
            S = *res_size;
            if (lane == 0 && S > 0) {
                root_prio = it == 0 ? (as_global(d.heaps) + (size_t)q * d.heap_stride + (W - 1))[0].prio
                                    : heap_top[(size_t)(threadIdx.x / W) * HeapRef<W>::kTop].prio;
            }
        } else {
            S = as_global(d.heap_size)[q];
            // the root slot always lies inside the heap's storage; it is used only if S > 0
            if (lane == 0) root_prio = (as_global(d.heaps) + (size_t)q * d.heap_stride + (W - 1))[0].prio;
        }
        if (it > 0 && q < Aw) {
            // insert_nodes_into_queues, a_star.h:101-115: heap q takes slots (q + epoch) mod N + j*N
            const uint32_t base = (q + it) & (Aw - 1);
#pragma unroll
            for (int r = 0; r < KW; ++r) {
                const int j = r * W + lane;
                if (j < K) rec[r] = cin[base + (size_t)j * Aw];
            }
        }
    }

    uint32_t best = __builtin_amdgcn_readfirstlane(best_ld);
    const uint32_t mt = it > 0 ? __builtin_amdgcn_readfirstlane(mt_ld) : 0u;

    // -- termination test on the previous iteration's reduction (a_star.h:260)
    bool live = active;
    if (it > 0 && (mt == kInf || (best != kInf && best <= mt))) {
        if (live && lane == 0) {
            set_status(st, ST_DONE);
            store_relaxed(&st->done_iter, it);
        }
        live = false;
    }
    if (live && q == 0 && lane == 0) {
        store_relaxed(&st->min_top[(it + 1) % 3], kInf);
        store_relaxed(&st->iterations, it + 1);
        // next iteration's width: double while the open list holds more than
        // grow_thresh entries per heap in use (counted one launch ago)
        uint32_t an = A;
        if (it > 0 && A < cover) {
            const uint32_t open = load_relaxed(&st->open_sum[(it - 1) % 3]);
            if (open > prm.grow_thresh * A) an = 2 * A < cover ? 2 * A : cover;
        }
        store_relaxed(&st->active[(it + 1) % 3], an);
        store_relaxed(&st->active_now, an);
        store_relaxed(&st->open_sum[(it + 1) % 3], 0u);
        store_relaxed(&st->grow[(it + 1) % 3], 0u);  // raised during the next iteration, read at the one after
    }
    if constexpr (!Ops::kDense) {
        if (live && occupied > d.load_limit) {
            if (lane == 0) set_status(st, ST_OVERFLOW_TABLE);
            live = false;
        }
    }
    const uint32_t gate = (it > 0 && mt != kInf && d.f_slack != kInf)
                              ? (uint32_t)min((uint64_t)mt + d.f_slack, (uint64_t)kInf - 1)
                              : kInf;

    // the heap's top after its pushes, and its size after them
    uint32_t rec_f = kInf, n_rec = 0;
#pragma unroll
    for (int r = 0; r < KW; ++r) {
        const bool v = live && rec[r].prio != kNoRecord && prio_f(rec[r].prio) < best;  // the pushes' filter
        if (v) {
            rec_f = min(rec_f, prio_f(rec[r].prio));
            ++n_rec;
        }
    }
    rec_f = grp_min_u32<W>(rec_f);
    n_rec = grp_sum_u32<W>(n_rec);
    const uint32_t root_f0 = grp_bcast<W>(prio_f(root_prio), 0);
    const uint32_t root_f = (live && S > 0) ? root_f0 : kInf;
    const bool work = live && (n_rec > 0 || root_f <= gate);  // root_f = kInf when empty
    {
        const uint32_t top_f = wave_min_u32(lane == 0 ? min(root_f, rec_f) : kInf);
        const uint32_t open = wave_sum_u32((live && lane == 0) ? S + n_rec : 0u);
        if ((threadIdx.x & 63) == 0) {
            if (top_f != kInf) atomicMin(&tl.c[TL_MIN_TOP], top_f);
            tl.add(TL_OPEN, open);
        }
    }
    tl.prof_count(TL_GROUPS, live && lane == 0);
    if (live && !work) {  // an idle heap sends nothing: clear its candidate slice
        const gptr<HeapEntry> cout = as_global(d.cand) + (size_t)(it & 1) * N * K + (size_t)q * K;
        HeapEntry none;
        none.prio = kNoRecord;
        none.slot = 0;
        none.aux = 0;
#pragma unroll
        for (int r = 0; r < KW; ++r) {
            const int i = r * W + lane;
            if (i < K) cout[i] = none;
        }
    }
    // append the heaps with work to the list: one LDS atomic per wave
    if (!Res) {
        const bool lead = work && lane == 0;
        const uint64_t m = __ballot(lead);
        if (m) {
            const int first = __builtin_ctzll(m);
            uint32_t base = 0;
            if ((int)__lane_id() == first) base = atomicAdd(&wl.n, (uint32_t)__popcll(m));
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, first);  // first: uniform (from a ballot)
            if (lead) {
                const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1));
                wl.q[k] = q;
                wl.S[k] = S;
            }
        }
    }
    __syncthreads();

    // ---------------- phase 2 ----------------
    uint32_t g;
    bool mine;
    if constexpr (Res) {  // every group works its own heap in its own slice
        if (!__ballot(work)) return;  // no heap of this wave has work
        g = threadIdx.x / W;
        mine = work;
    } else {
        const uint32_t n_work = wl.n;
        if ((threadIdx.x & ~63u) / W >= n_work) return;  // no heap for this wave
        g = threadIdx.x / W;
        mine = g < n_work;
        q = mine ? wl.q[g] : 0u;
        S = mine ? wl.S[g] : 0u;
    }
    const bool track = !Res && (prm.flags & STEP_DIRTY_TOPS) != 0;
    const HeapRef<W> H{as_global(d.heaps) + (size_t)q * d.heap_stride + (W - 1),
                       as_local(heap_top) + (size_t)g * HeapRef<W>::kTop, as_local(dirty_masks) + g, track};
    if constexpr (Res) {
        if (!mine) {  // (the records phase 1 read stay in the registers of a group with work)
#pragma unroll
            for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
        }
        if (mine && it == 0) H.fill(S, lane);  // the seeded heap's entries into the slice, once
    } else {
#pragma unroll
        for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
        if (mine && it > 0 && q < Aw) {  // the records phase 1 read for this heap (cached)
            const uint32_t base = (q + it) & (Aw - 1);
#pragma unroll
            for (int r = 0; r < KW; ++r) {
                const int j = r * W + lane;
                if (j < K) rec[r] = cin[base + (size_t)j * Aw];
            }
        }
        if (mine) {
            if (track && lane == 0) *H.dirty = 0u;
            H.fill(S, lane);  // the heap's top levels into LDS, with the records' loads
        }
    }
    group_fence();
    bool overflow_heap = false, overflow_table = false;
    // always-on counts, per lane: one wave reduction and one LDS atomic each at
    // the end (a ballot and an LDS atomic per event cost ~15 instructions each)
    uint32_t n_expanded = 0, n_generated = 0, n_occupied = 0;

    // -- 1. push the records addressed to this heap
    if (mine && it > 0 && q < Aw) {
        push_records<W, KW>(H, S, rec, K, lane, best, d.heap_cap, overflow_heap, tl, true);
#pragma unroll
        for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
    }
    group_fence();

    // -- 2. first pop of this launch, unless the root lies above the f gate
    uint32_t top_par = kNone;  // the popped node's stored parent (dense sets)
    auto load = [&](const HeapEntry& t, typename Ops::Key& k, uint32_t& gs) { Ops::load(d, t.slot, k, gs, top_par); };
    uint32_t pre[KW];  // per successor: what the problem reads early (grid: the entered cell's weight)
    uint32_t pre_g[KW];  // dense sets with STEP_SKIP_SETTLED: the entered cell's stored g, read as early
    const bool skip_settled = Ops::kDense && (prm.flags & STEP_SKIP_SETTLED) != 0;
    auto prefetch = [&](const HeapEntry& t) {
#pragma unroll
        for (int r = 0; r < KW; ++r) {
            const int i = r * W + lane;
            pre[r] = i < K ? Ops::prefetch(d, t, i) : 0u;
            pre_g[r] = kInf;
            if constexpr (Ops::kDense)
                if (skip_settled && i < K) pre_g[r] = Ops::prefetch_g(d, t, i);
        }
    };
    HeapEntry top;
    top.prio = 0;
    top.slot = 0;
    top.aux = 0;
    typename Ops::Key top_key{};
    uint32_t top_g = 0;  // g stored in the closed set for the popped node
    bool popped = false, gated = false;
    if (mine && S > 0) {
        uint32_t lv = 0, hl = 0;
        popped = heap_pop<W>(H, S, lane, lv, hl, load, prefetch, top_key, top_g, gate, top);
        gated = !popped;
        group_fence();
        if (lane == 0) {
            tl.prof_add(TL_SIFT_DOWN, lv);
            tl.prof_add(TL_SIFT_DOWN_HBM, hl);
            tl.prof_add(TL_POPS, popped ? 1u : 0u);
        }
    }

    // -- 3..6. expansion rounds
    if (mine) {
        const uint32_t rounds = d.rounds ? d.rounds : prm.rounds;
        if constexpr (P >= 2) {
            // Several pops per round (grids, STEP_PAIRED_POPS): each round pops
            // the heap's next P nodes (a stale one is dropped, not replaced)
            // and issues the closed-set atomics of all of their successors
            // before it reads any of their results, so every lane has P
            // independent atomics in flight where a round of single pops has
            // one.  The later nodes' successors go into the heap; in the last
            // round the first node's go to the candidate buffer.  Exactness is
            // unaffected: any order of expansions is, with re-opening and the
            // termination test.
            static_assert(Ops::kDense, "several pops per round are for the dense grid sets");
            constexpr int Q = P - 1;  // the nodes after the first
            uint32_t parq[Q], preq[Q][KW], preq_g[Q][KW];
            HeapEntry recq[Q][KW];
            for (uint32_t round = 0; round < rounds && !gated; ++round) {
                bool have1 = popped;
                popped = false;
                if (!have1 && S > 0) {
                    uint32_t lv = 0, hl = 0;
                    have1 = heap_pop<W>(H, S, lane, lv, hl, load, prefetch, top_key, top_g, gate, top);
                    gated = !have1;
                    group_fence();
                    if (lane == 0) {
                        tl.prof_add(TL_SIFT_DOWN, lv);
                        tl.prof_add(TL_SIFT_DOWN_HBM, hl);
                        tl.prof_add(TL_POPS, have1 ? 1u : 0u);
                    }
                }
                HeapEntry topq[Q];
                typename Ops::Key keyq[Q];
                uint32_t gq[Q];
                bool haveq[Q], any = have1;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    topq[q].prio = 0;
                    topq[q].slot = 0;
                    topq[q].aux = 0;
                    keyq[q] = typename Ops::Key{};
                    gq[q] = 0;
                    haveq[q] = false;
                    if (!gated && S > 0) {
                        auto loadq = [&](const HeapEntry& t, typename Ops::Key& k, uint32_t& gs) {
                            Ops::load(d, t.slot, k, gs, parq[q]);
                        };
                        auto prefetchq = [&](const HeapEntry& t) {
#pragma unroll
                            for (int r = 0; r < KW; ++r) {
                                const int i = r * W + lane;
                                preq[q][r] = i < K ? Ops::prefetch(d, t, i) : 0u;
                                preq_g[q][r] = (skip_settled && i < K) ? Ops::prefetch_g(d, t, i) : kInf;
                            }
                        };
                        uint32_t lv = 0, hl = 0;
                        haveq[q] = heap_pop<W>(H, S, lane, lv, hl, loadq, prefetchq, keyq[q], gq[q], gate, topq[q]);
                        gated = !haveq[q];
                        group_fence();
                        if (lane == 0) {
                            tl.prof_add(TL_SIFT_DOWN, lv);
                            tl.prof_add(TL_SIFT_DOWN_HBM, hl);
                            tl.prof_add(TL_POPS, haveq[q] ? 1u : 0u);
                        }
                    }
                    any = any || haveq[q];
                }
                if (!any) break;
                // stale entries (g above the stored one) are dropped; the goal
                // is tested at pop (a_star.h:80-83); a node with f >= best
                // cannot improve it
                const uint32_t g0a = prio_g(top.prio), fa = prio_f(top.prio);
                uint32_t g0q[Q], fq[Q];
                have1 = have1 && g0a <= top_g;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    g0q[q] = prio_g(topq[q].prio);
                    fq[q] = prio_f(topq[q].prio);
                    haveq[q] = haveq[q] && g0q[q] <= gq[q];
                }
                if (have1 && Ops::is_goal(d, top_key)) {
                    if (lane == 0) atomicMin(&st->best, g0a);
                    best = g0a < best ? g0a : best;
                    have1 = false;
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    if (haveq[q] && Ops::is_goal(d, keyq[q])) {
                        if (lane == 0) atomicMin(&st->best, g0q[q]);
                        best = g0q[q] < best ? g0q[q] : best;
                        haveq[q] = false;
                    }
                }
                have1 = have1 && (best == kInf || fa < best);
                uint32_t nq = 0;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    haveq[q] = haveq[q] && (best == kInf || fq[q] < best);
                    nq += haveq[q] ? 1u : 0u;
                }
                n_expanded += lane == 0 ? (have1 ? 1u : 0u) + nq : 0u;
                // every node's successors; their atomics issued together, their results read after
#pragma unroll
                for (int r = 0; r < KW; ++r) {
                    const int i = r * W + lane;
                    rec[r].prio = kNoRecord;
                    typename Ops::Key sa, sq[Q];
                    uint32_t aaux = 0, acost = 0, ath = 0, qaux[Q], qcost[Q], qth[Q];
                    const uint32_t wa = i < K ? Ops::drop_parent(d, top.aux, i, pre[r], top_par) : 0u;
                    const bool gena = have1 && i < K && Ops::succ(d, top_key, top.aux, fa - g0a, i, wa, sa, aaux, acost, ath);
                    const uint32_t ga = g0a + acost;
                    const bool ua = gena && ga + ath < best && pre_g[r] > ga;
                    n_generated += gena ? 1u : 0u;
                    uint32_t gqs[Q];
                    bool uq[Q];
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        recq[q][r].prio = kNoRecord;
                        qaux[q] = qcost[q] = qth[q] = 0;
                        const uint32_t wq = i < K ? Ops::drop_parent(d, topq[q].aux, i, preq[q][r], parq[q]) : 0u;
                        const bool genq = haveq[q] && i < K &&
                                          Ops::succ(d, keyq[q], topq[q].aux, fq[q] - g0q[q], i, wq, sq[q], qaux[q], qcost[q], qth[q]);
                        gqs[q] = g0q[q] + qcost[q];
                        uq[q] = genq && gqs[q] + qth[q] < best && preq_g[q][r] > gqs[q];
                        n_generated += genq ? 1u : 0u;
                    }
                    uint64_t olda = ~0ull, oldq[Q];
                    if (ua)
                        olda = __hip_atomic_fetch_min(Ops::val_ptr(d, (uint32_t)sa), ((uint64_t)ga << 32) | top.slot,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        oldq[q] = ~0ull;
                        if (uq[q])
                            oldq[q] = __hip_atomic_fetch_min(Ops::val_ptr(d, (uint32_t)sq[q]),
                                                             ((uint64_t)gqs[q] << 32) | topq[q].slot, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
                    }
                    const bool ia = ua && ga < val_g(olda);
                    if (ia) {
                        rec[r].prio = make_prio(ga + ath, ga);
                        rec[r].slot = (uint32_t)sa;
                        rec[r].aux = aaux;
                    }
                    tl.prof_count(TL_PROBES, ua);
                    tl.prof_count(TL_LOOKUPS, ua);
                    tl.prof_count(TL_INSERTED, ia);
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        const bool iq = uq[q] && gqs[q] < val_g(oldq[q]);
                        if (iq) {
                            recq[q][r].prio = make_prio(gqs[q] + qth[q], gqs[q]);
                            recq[q][r].slot = (uint32_t)sq[q];
                            recq[q][r].aux = qaux[q];
                        }
                        tl.prof_count(TL_PROBES, uq[q]);
                        tl.prof_count(TL_LOOKUPS, uq[q]);
                        tl.prof_count(TL_INSERTED, iq);
                    }
                }
                // every node's records into the heap in one push (one round
                // trip for the parents' checks), except that in the last round
                // the first node's go to the candidate buffer; records past K
                // in a register block are "none" (set above)
                const bool last = round + 1 == rounds;
                HeapEntry all[(Q + 1) * KW];
#pragma unroll
                for (int q = 0; q < Q; ++q)
#pragma unroll
                    for (int r = 0; r < KW; ++r) all[q * KW + r] = recq[q][r];
#pragma unroll
                for (int r = 0; r < KW; ++r) {
                    all[Q * KW + r] = rec[r];
                    if (last) all[Q * KW + r].prio = kNoRecord;
                }
                push_records<W, (Q + 1) * KW>(H, S, all, (Q + 1) * KW * W, lane, best, d.heap_cap, overflow_heap, tl,
                                              false);
                group_fence();
                if (last) break;
#pragma unroll
                for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
            }
        } else
        for (uint32_t round = 0; round < rounds && !gated; ++round) {
            // pop, skipping entries whose g is no longer the stored one
            bool have = false;
            for (uint32_t t = 0; t < prm.max_pops; ++t) {
                if (!popped) {
                    if (S == 0) break;
                    uint32_t lv = 0, hl = 0;
                    const bool ok = heap_pop<W>(H, S, lane, lv, hl, load, prefetch, top_key, top_g, gate, top);
                    if (lane == 0) {
                        tl.prof_add(TL_SIFT_DOWN, lv);
                        tl.prof_add(TL_SIFT_DOWN_HBM, hl);
                    }
                    if (!ok) {
                        gated = true;
                        break;
                    }
                    group_fence();
                    if (lane == 0) tl.prof_add(TL_POPS, 1u);
                }
                popped = false;
                if (prio_g(top.prio) > top_g) continue;  // stale
                have = true;
                break;
            }
            if (!have) break;
            // goal test at pop (a_star.h:80-83); a node with f >= best cannot improve it
            const uint32_t g0 = prio_g(top.prio), f = prio_f(top.prio);
            if (Ops::is_goal(d, top_key)) {
                if (lane == 0) atomicMin(&st->best, g0);
                best = g0 < best ? g0 : best;
                continue;
            }
            if (best != kInf && f >= best) continue;
            n_expanded += lane == 0 ? 1u : 0u;
#pragma unroll
            for (int r = 0; r < KW; ++r) {
                const int i = r * W + lane;
                if (i < K) pre[r] = Ops::drop_parent(d, top.aux, i, pre[r], top_par);
            }
            // expansion: successors, closed-set insertion (a successor survives
            // only if its atomicMin of (g << 32 | parent) lowered g)
            const uint32_t h = f - g0;
#pragma unroll
            for (int r = 0; r < KW; ++r) {
                const int i = r * W + lane;
                rec[r].prio = kNoRecord;
                typename Ops::Key skey;
                uint32_t taux, cost, th;
                const bool gen = i < K && Ops::succ(d, top_key, top.aux, h, i, pre[r], skey, taux, cost, th);
                n_generated += gen ? 1u : 0u;
                if (gen) {
                    const uint32_t g2 = g0 + cost, f2 = g2 + th;
                    // a cell's stored g only decreases, so a read of it that is already no larger
                    // than g2 means the atomicMin below could not succeed
                    if (f2 < best && pre_g[r] > g2) {
                        uint32_t slot;
                        bool isnew;
                        uint32_t np = 0;
                        const int cs = closed_set_update<Ops>(d, skey, g2, top.slot, prm.max_probes, slot, isnew, np);
                        const bool found = cs != CS_FULL;
                        const bool ins = cs == CS_INSERTED || cs == CS_UPDATED;
                        if (ins) {
                            rec[r].prio = make_prio(f2, g2);
                            rec[r].slot = slot;
                            rec[r].aux = taux;
                        }
                        overflow_table |= !found;
                        if constexpr (!Ops::kDense) n_occupied += (found && isnew) ? 1u : 0u;
                        tl.prof_count(TL_PROBES, true);  // first probe of every lookup
                        if (np > 1) tl.prof_add(TL_PROBES, np - 1);
                        tl.prof_count(TL_LOOKUPS, found);
                        tl.prof_count(TL_INSERTED, ins);
                    }
                }
            }
            if (round + 1 == rounds) break;  // last round: records go to the candidate buffer
            push_records<W, KW>(H, S, rec, K, lane, best, d.heap_cap, overflow_heap, tl, false);
#pragma unroll
            for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
            group_fence();
        }
        group_fence();
        if constexpr (Res) {
            *res_size = S;
        } else {
            H.write_back(S, lane);
            if (lane == 0) {
                as_global(d.heap_size)[q] = S;
                // one more iteration could bring this heap past its capacity: grow before it
                if (d.grow_heap_at && S > d.grow_heap_at) atomicOr(&st->grow[it % 3], GROW_HEAP);
            }
        }
        // candidate slice of this heap: K records
        const gptr<HeapEntry> cout = as_global(d.cand) + (size_t)(it & 1) * N * K + (size_t)q * K;
#pragma unroll
        for (int r = 0; r < KW; ++r) {
            const int i = r * W + lane;
            if (i < K) cout[i] = rec[r];
        }
    }
    if (overflow_heap) set_status(st, ST_OVERFLOW_HEAP);
    if (overflow_table) set_status(st, ST_OVERFLOW_TABLE);
    {
        const uint32_t e_sum = wave_sum_u32(n_expanded), g_sum = wave_sum_u32(n_generated);
        const uint32_t o_sum = Ops::kDense ? 0u : wave_sum_u32(n_occupied);
        if ((threadIdx.x & 63) == 0) {
            tl.add(TL_EXPANDED, e_sum);
            tl.add(TL_GENERATED, g_sum);
            tl.add(TL_OCCUPIED, o_sum);
        }
    }

    tl.prof_count(TL_WORK, mine && lane == 0);
}

// Per-workgroup prologue shared by every step kernel: returns false (for the
// whole workgroup) when the launch map has no instance for it.
__device__ inline bool step_prologue(const WgEntry* wgmap, const InstDesc* descs, const uint32_t* iter_base,
                                     uint32_t r, uint32_t& inst, uint32_t& it, uint32_t& first_heap,
                                     uint32_t& cover) {
    const WgEntry we = wgmap[blockIdx.x];
    if (we.inst == kNone) return false;
    inst = we.inst;
    first_heap = we.first_heap;
    cover = we.cover;
    it = *iter_base + r - descs[we.inst].start_iter;
    // The descriptor's first two 64-byte lines hold every field that phase 1
    // of the step addresses its loads with; start_iter (above) and kind are
    // read here, with the instance's status, so that both lines are in the
    // scalar cache before the workgroup barrier that the reads after it wait
    // behind.  (The empty statement keeps the kind load from being dropped.)
    static_assert(offsetof(InstDesc, start_iter) / 64 == 1 && offsetof(InstDesc, a0) / 64 == 1 &&
                      offsetof(InstDesc, heap_size) / 64 == 0,
                  "descriptor layout: phase-1 fields in lines 0 and 1");
    const uint32_t kind = descs[we.inst].kind;
    asm volatile("" ::"s"(kind));
    return true;
}

// Advances the global iteration counter after a chunk of steps (a vector
// store; the next launch reads it after the kernel boundary).
template <int Unused = 0>
__global__ void advance_kernel(uint32_t* iter_base, uint32_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *iter_base += by;
}

// Grid-stride fill (a kernel rather than hipMemsetAsync so that it may sit in
// captured graphs).
template <typename T>
__global__ void fill_kernel(T* p, uint64_t n, T v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Resets the instance counters and seeds the search: the start state enters
// the closed set with g = 0 (the reference never stores it, a_star.h:226 vs
// :253, so it can be re-expanded) and heap 0 (insert_first, a_star.h:68-71).
// The closed set must already be cleared and every heap size zeroed.
template <class Ops>
__device__ inline bool seed_instance(const InstDesc& d, InstState* st, int W) {
    InstState z;
    z.status = ST_RUNNING;
    z.best = kInf;
    z.min_top[0] = z.min_top[1] = z.min_top[2] = kInf;
    z.iterations = 0;
    z.done_iter = 0;
    z.pad = 0;
    const uint32_t a0 = (d.a0 && d.a0 <= d.n_heaps) ? d.a0 : d.n_heaps;
    z.active[0] = z.active[1] = z.active[2] = a0;
    z.open_sum[0] = z.open_sum[1] = z.open_sum[2] = 0;
    z.active_now = a0;
    z.pad2 = 0;
    z.grow[0] = z.grow[1] = z.grow[2] = 0;
    z.paused_iter = 0;
    z.occupied = Ops::kDense ? 0 : 1;
    z.expanded = z.generated = z.pushed = z.pops = z.probes = z.sift_down = z.sift_up = 0;
    z.groups = z.inserted = z.lookups = z.work_groups = 0;
    z.sift_down_hbm = z.sift_up_hbm = z.push_in = 0;
    uint32_t slot, probes = 0;
    bool isnew;
    const typename Ops::Key start = Ops::start_key(d);
    if (!Ops::find_or_insert(d, start, 1u << 20, slot, isnew, probes)) {
        z.status = ST_OVERFLOW_TABLE;
        *st = z;
        return false;
    }
    *Ops::val_ptr(d, slot) = kNone;  // g = 0, no parent
    const uint32_t aux = Ops::aux_of(d, start);
    HeapEntry e;
    e.prio = make_prio(Ops::heuristic(d, start, aux), 0);
    e.slot = slot;
    e.aux = aux;
    d.heaps[W - 1] = e;
    d.heap_size[0] = 1;
    *st = z;
    return true;
}

// Walks the parent links from the goal; writes states goal..start and returns
// their number, or kNone if cap is exceeded / the goal is not stored.
template <class Ops>
__device__ inline uint32_t walk_path(const InstDesc& d, typename Ops::Key* out, uint32_t cap) {
    uint32_t slot;
    if (!Ops::find(d, Ops::goal_key(d), slot)) return kNone;
    uint32_t n = 0;
    while (slot != kNone) {
        if (n >= cap) return kNone;
        typename Ops::Key key;
        uint32_t g, par;
        Ops::load(d, slot, key, g, par);
        out[n++] = key;
        slot = val_parent(*Ops::val_ptr(d, slot));
    }
    return n;
}

// walk_path for dense grids with every thread of a kPathThreads-thread
// workgroup (all of them call it, with `tab` pointing at kPathTabWords 16-bit
// words of LDS, 4-byte aligned).  A node's parent is one of its neighbours, so the threads
// load the parent links of a kPathWin x kPathWin window of cells in one round
// trip (windows aligned to the 4 x 4 tiles of the layout: 16 threads read one
// tile's 128-byte line; padding cells lie inside the allocation; a window
// reaches 7/8 of its width from the current cell towards the start, where the
// walk heads) and turn them, in parallel, into link tables: each cell's parent
// as an index inside the window (or a code: no parent, a parent outside the
// window and the direction to it, or a parent that is not a neighbour), then
// the cell 2, 4, ... steps up where every step stays inside.  The walk follows
// the longest links, kPathHop steps per dependent LDS read, recording where
// each hop starts, and the last few single steps lead to the window's edge
// (or the start); then every thread writes the cells of its hops, and the
// next window is loaded around the parent.  Same result as walk_path, with
// the cells goal..start written as row-major indices y * width + x, the form
// the host API uses; their number, or kNone past cap.
constexpr uint32_t kPathWin = 128;
constexpr uint32_t kPathThreads = 1024;
constexpr uint32_t kPathCells = kPathWin * kPathWin;  // in-window indices: below 2^14
constexpr int kPathLevels = 2;                        // link tables of 1, 2, ... steps
constexpr uint32_t kPathHop = 1u << (kPathLevels - 1);
constexpr uint32_t kPathTabWords = kPathLevels * kPathCells + kPathCells / kPathHop + 8;  // tables, hop starts, chain
constexpr uint32_t kLinkRoot = 0xFFFFu;  // no parent (the start); in the longer tables: no link
constexpr uint32_t kLinkFar = 0xFFFEu;   // a parent that is not a neighbour: read again from memory
constexpr uint32_t kLinkOut = 0x8000u;   // | direction (dx + 1) + 3 (dy + 1): a neighbour outside the window

// Origin of a window along one axis: c the current cell, t the start's, w the
// window's width, extent the tiled layout's cells along the axis (w and extent
// multiples of 4, w <= extent).  The result is a multiple of 4 with
// o <= c < o + w for every c < extent.
__device__ inline uint32_t path_window_origin(uint32_t c, uint32_t t, uint32_t w, uint32_t extent) {
    const uint32_t back = t < c ? w - w / 8 : w / 8;  // window cells before c
    const uint32_t o = (c > back ? c - back : 0u) & ~3u;
    return min(o, extent - w);
}

__device__ inline uint32_t walk_grid_path(const InstDesc& d, uint64_t* out, uint32_t cap, uint16_t* tab) {
    constexpr int kPer = (int)(kPathCells / kPathThreads);  // window cells per thread
    constexpr uint32_t kMaxHops = kPathCells / kPathHop;
    static_assert(kPathWin % 4 == 0 && kPathCells % kPathThreads == 0 && kPathCells <= kLinkOut, "path window");
    static_assert((kPathLevels * kPathCells + kMaxHops) % 2 == 0, "the chain's 32-bit words: 4-byte aligned in tab");
    uint16_t* const link = tab;                                   // one step
    uint16_t* const top = tab + (kPathLevels - 1) * kPathCells;  // kPathHop steps
    uint16_t* const hops = tab + kPathLevels * kPathCells;       // where each hop starts
    uint32_t* const chain = reinterpret_cast<uint32_t*>(hops + kMaxHops);  // the chain's outcome, for every wave
    const uint32_t tid = threadIdx.x;
    const uint32_t W = d.width, H = d.height, tc = d.tile_cols;
    const uint32_t ex = tc * 4u, ey = ((H + 3u) / 4u) * 4u;  // the tiled layout's extent
    const uint32_t ww = ex < kPathWin ? ex : kPathWin, wh = ey < kPathWin ? ey : kPathWin;
    const uint32_t sx = d.start_xy & 0xFFFFu, sy = d.start_xy >> 16;
    // the parent half of each closed-set value (g << 32 | parent, little-endian)
    const gptr<const uint32_t> par = as_global<const uint32_t>(reinterpret_cast<const uint32_t*>(d.dense));
    uint32_t x = d.goal_xy & 0xFFFFu, y = d.goal_xy >> 16;
    uint32_t n = 0;
    for (;;) {
        const uint32_t ox = path_window_origin(x, sx, ww, ex), oy = path_window_origin(y, sy, wh, ey);
        auto coord = [&](uint32_t c) { return (uint64_t)(oy + c / kPathWin) * W + (ox + c % kPathWin); };
        uint32_t v[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r) {  // all loads in flight together; positions outside read the origin
            const uint32_t k = tid + kPathThreads * (uint32_t)r, t = k >> 4, o = k & 15u;
            const uint32_t wx = (t % (kPathWin / 4)) * 4u + (o & 3u), wy = (t / (kPathWin / 4)) * 4u + (o >> 2);
            const bool in = wx < ww && wy < wh;
            v[r] = par[2 * (size_t)grid_index(ox + (in ? wx : 0u), oy + (in ? wy : 0u), tc)];
        }
        __syncthreads();  // every thread is done with the previous window's tables
        // branch-free: a thread's cells share their offsets inside the tile
        // (ix, iy) and their column; row r lies 8 r cells below the first
        const uint32_t ix = tid & 3u, iy = (tid >> 2) & 3u;
        const uint32_t wx = ((tid >> 4) % (kPathWin / 4)) * 4u + ix;
        const uint32_t wy0 = (tid >> 4) / (kPathWin / 4) * 4u + iy;
        static_assert(kPathThreads % (4 * kPathWin) == 0, "every row of a thread's cells: the same column");
        constexpr uint32_t kRowStep = kPathThreads / kPathWin;  // cells (a multiple of 4) between a thread's rows
        const uint32_t tile0 = ((oy + wy0) >> 2) * tc + ((ox + wx) >> 2);  // the tile of the thread's first cell
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint32_t wy = wy0 + kRowStep * (uint32_t)r, cx = ox + wx, cy = oy + wy, p = v[r];
            // the parent's offset per axis, from the offsets inside the tile:
            // -1, 0 or +1 for a neighbour, -2 for none
            const int dx = __builtin_amdgcn_sbfe((int)(p - ix), 0, 2), dy = __builtin_amdgcn_sbfe((int)(p - (iy << 2)), 2, 2);
            const uint32_t px = cx + (uint32_t)dx, py = cy + (uint32_t)dy;
            // the parent's tile: the cell's, or the one next to it where the step crosses a tile's edge
            const uint32_t tile = tile0 + (kRowStep / 4 * (uint32_t)r) * tc;
            const uint32_t ptile = tile + (uint32_t)(((int)ix + dx) >> 2) + (uint32_t)__mul24(((int)iy + dy) >> 2, (int)tc);
            const bool nb = (dx != -2) & (dy != -2) & (px < W) & (py < H) & ((p >> 4) == ptile);
            const uint32_t rx = wx + (uint32_t)dx, ry = wy + (uint32_t)dy;
            uint32_t code = ((rx < ww) & (ry < wh)) ? ry * kPathWin + rx : kLinkOut | (uint32_t)(dx + 3 * dy + 4);
            code = nb ? code : kLinkFar;
            code = p == kNone ? kLinkRoot : code;
            code = ((wx < ww) & (wy < wh)) ? code : kLinkFar;  // positions outside the window: never reached
            link[wy * kPathWin + wx] = (uint16_t)code;
        }
#pragma unroll
        for (int l = 1; l < kPathLevels; ++l) {  // two links of the level below, where both stay inside
            const uint16_t* const lo = tab + (l - 1) * kPathCells;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const uint32_t k = tid + kPathThreads * (uint32_t)r;
                const uint32_t a = lo[k];
                const uint32_t b = a < kPathCells ? (uint32_t)lo[a] : kLinkRoot;
                tab[l * kPathCells + k] = (uint16_t)(b < kPathCells ? b : kLinkRoot);
            }
        }
        __syncthreads();
        // wave 0 follows the chain (each read is one LDS broadcast, its value
        // uniform): it records where each hop starts, then takes the last
        // single steps (fewer than kPathHop links, since no hop of kPathHop
        // steps starts where they begin) to the chain's end inside the window
        if (tid < 64) {
            uint32_t c = (y - oy) * kPathWin + (x - ox), m = 0, nt = 0, code = kLinkFar;
            for (uint32_t h = (uint32_t)__builtin_amdgcn_readfirstlane((int)top[c]); h < kPathCells;
                 h = (uint32_t)__builtin_amdgcn_readfirstlane((int)top[c])) {
                if (m == kMaxHops) break;  // links in a cycle
                if (tid == 0) hops[m] = (uint16_t)c;
                ++m;
                c = h;
            }
            if (m < kMaxHops) {
                for (; nt < kPathHop; ++nt) {
                    const uint32_t i = n + kPathHop * m + nt;
                    if (tid == 0 && i < cap) out[i] = coord(c);
                    code = (uint32_t)__builtin_amdgcn_readfirstlane((int)link[c]);
                    if (code >= kPathCells) break;
                    c = code;
                }
            }
            if (tid == 0) {
                chain[0] = m < kMaxHops && nt < kPathHop ? m : kNone;
                chain[1] = nt + 1;
                chain[2] = c;
                chain[3] = code;
            }
        }
        __syncthreads();  // the hop starts and the chain's outcome are visible
        const uint32_t m = chain[0], nt = chain[1], c = chain[2], code = chain[3];
        if (m == kNone || n + kPathHop * m + nt > cap) return kNone;
        for (uint32_t j = tid; j < m; j += kPathThreads) {
            uint32_t a = hops[j];
#pragma unroll
            for (uint32_t i = 0; i < kPathHop; ++i) {
                out[n + kPathHop * j + i] = coord(a);
                if (i + 1 < kPathHop) a = link[a];
            }
        }
        n += kPathHop * m + nt;
        if (code == kLinkRoot) return n;  // the start
        // the walk leaves the window at c: on to its parent
        const uint32_t cx = ox + c % kPathWin, cy = oy + c / kPathWin;
        if (code == kLinkFar) {  // decoded from the index; one outside the grid ends the walk with an error
            const uint32_t p = par[2 * (size_t)grid_index(cx, cy, tc)], t = p >> 4;
            x = (t % tc) * 4u + (p & 3u);
            y = (t / tc) * 4u + ((p >> 2) & 3u);
            if (x >= W || y >= H) return kNone;
        } else {
            const uint32_t dir = code & 15u;
            x = cx + dir % 3u - 1u;
            y = cy + dir / 3u - 1u;
        }
    }
}

}  // namespace gastar
