// engine.hip -- host runtime and C ABI of the GA* engine (libgastar.so).
//
// One engine owns a HIP stream for the search and one for path walks, a
// device arena for heaps and closed sets, and a slot table of resident
// instances.  Plain GA* (ga_solve), batched GA* (ga_solve_batch) and serving
// (ga_serve_run) all run the same loop:
//   admit -> launch a chunk of R iterations -> read back the previous chunk's
//   per-instance state -> retire finished instances -> admit more.
// Inside a chunk the host never synchronises; one launch is one iteration of
// every resident instance (see gastar/core.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gastar.h"
#include "gastar/core.hpp"

using namespace gastar;

namespace {

// ---------------------------------------------------------------------------
// kernels instantiated for the built-in problems
// ---------------------------------------------------------------------------
constexpr int kBlock = kStepBlock;

constexpr uint32_t kind_bit(uint32_t k) { return 1u << k; }
constexpr uint32_t kAllKinds =
    kind_bit(KIND_PUZZLE3) | kind_bit(KIND_PUZZLE4) | kind_bit(KIND_GRID4) | kind_bit(KIND_GRID8);

// The step kernel for the kinds in KINDS.  The register count of a kernel is
// the maximum over the kinds it can run, so a lane that holds one kind runs an
// instantiation for that kind alone: the 15-puzzle and 4-connected grid paths
// fit 64 VGPRs (8 waves per SIMD) where the kernel for all kinds needs 77 (6).
template <int W, uint32_t KINDS, int B, int P = 1>
__global__ __launch_bounds__(B) void step_kernel(const WgEntry* __restrict__ wgmap,
                                                 const InstDesc* __restrict__ descs, InstState* states,
                                                 const uint32_t* __restrict__ iter_base, uint32_t r,
                                                 StepParams prm) {
    static_assert((B / W) * HeapRef<W>::kTop <= StepSharedT<B>::kHeapTopEntries && W >= 4, "LDS heap tops");
    __shared__ StepSharedT<B> sh;
    uint32_t inst, it, first, cover;
    if (!step_prologue(wgmap, descs, iter_base, r, inst, it, first, cover)) return;
    const uint32_t status = load_relaxed(&states[inst].status);
    step_begin(sh);
    const InstDesc& d = descs[inst];
    const uint32_t q = first + threadIdx.x / W;
    const bool active = q < d.n_heaps;
    const Tally tl{sh.tally, prm.profile != 0};
    {
        bool ran = false;
        if constexpr ((KINDS & kind_bit(KIND_PUZZLE3)) != 0)
            if (d.kind == KIND_PUZZLE3) {
                run_group<W, PuzzleOps<3>>(d, states + inst, status, it, q, active, cover, prm, tl, sh.work,
                                           sh.heap_top, sh.dirty);
                ran = true;
            }
        if constexpr ((KINDS & kind_bit(KIND_PUZZLE4)) != 0)
            if (d.kind == KIND_PUZZLE4) {
                run_group<W, PuzzleOps<4>>(d, states + inst, status, it, q, active, cover, prm, tl, sh.work,
                                           sh.heap_top, sh.dirty);
                ran = true;
            }
        if constexpr ((KINDS & kind_bit(KIND_GRID4)) != 0)
            if (d.kind == KIND_GRID4) {
                run_group<W, GridOps<4>, decltype(sh.work), false, P>(d, states + inst, status, it, q, active, cover,
                                                                      prm, tl, sh.work, sh.heap_top, sh.dirty);
                ran = true;
            }
        if constexpr ((KINDS & kind_bit(KIND_GRID8)) != 0)
            if (d.kind == KIND_GRID8) {
                run_group<W, GridOps<8>>(d, states + inst, status, it, q, active, cover, prm, tl, sh.work,
                                         sh.heap_top, sh.dirty);
                ran = true;
            }
        // a kind this instantiation cannot run: stop the instance rather than let
        // it stall (set_status changes only a running instance)
        if (!ran && threadIdx.x == 0) set_status(states + inst, ST_ABORTED);
    }
    tally_end(sh.tally, states + inst, it, d.grow_table_at);
}

// Step kernel of the wide-key kinds (24-puzzle): its own instantiation so that
// the 128-bit keys do not raise the register count of step_kernel.  Its
// instances run on a launch lane of their own.
template <int W, int B>
__global__ __launch_bounds__(B) void step_kernel_wide(const WgEntry* __restrict__ wgmap,
                                                      const InstDesc* __restrict__ descs, InstState* states,
                                                      const uint32_t* __restrict__ iter_base, uint32_t r,
                                                      StepParams prm) {
    static_assert((B / W) * HeapRef<W>::kTop <= StepSharedT<B>::kHeapTopEntries && W >= 4, "LDS heap tops");
    __shared__ StepSharedT<B> sh;
    uint32_t inst, it, first, cover;
    if (!step_prologue(wgmap, descs, iter_base, r, inst, it, first, cover)) return;
    const uint32_t status = load_relaxed(&states[inst].status);
    step_begin(sh);
    const InstDesc& d = descs[inst];
    const uint32_t q = first + threadIdx.x / W;
    const Tally tl{sh.tally, prm.profile != 0};
    if (d.kind == KIND_PUZZLE5)
        run_group<W, Puzzle5Ops>(d, states + inst, status, it, q, q < d.n_heaps, cover, prm, tl, sh.work,
                                 sh.heap_top, sh.dirty);
    tally_end(sh.tally, states + inst, it, d.grow_table_at);
}

__device__ inline void seed_one(const InstDesc& d, InstState* st, int W) {
    switch (d.kind) {
        case KIND_PUZZLE3: seed_instance<PuzzleOps<3>>(d, st, W); break;
        case KIND_PUZZLE4: seed_instance<PuzzleOps<4>>(d, st, W); break;
        case KIND_GRID4: seed_instance<GridOps<4>>(d, st, W); break;
        case KIND_GRID8: seed_instance<GridOps<8>>(d, st, W); break;
        case KIND_PUZZLE5: seed_instance<Puzzle5Ops>(d, st, W); break;
        default: break;
    }
}

__global__ void seed_kernel(InstDesc d, InstDesc* descs, InstState* states, uint32_t slot, int W) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    descs[slot] = d;
    seed_one(d, states + slot, W);
}

// Seeds instances 0..n-1 of a one-launch batch (their descriptors already in
// `descs`), one thread each.
__global__ void seed_batch_kernel(const InstDesc* descs, InstState* states, uint32_t n, int W) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seed_one(descs[i], states + i, W);
}

// Batched GA* with every iteration of an instance inside one launch
// (ga_config.persist).  Workgroup b runs instance order[b] from its seed to
// its end; the instance's heaps are the workgroup's B / W heap groups, so the
// exchange of successor records between its heaps (the candidate buffer) and
// the min-f termination test cross only waves of this workgroup, and an
// iteration needs a workgroup barrier where the launch lanes need a kernel
// boundary and a host step per chunk.  Nothing crosses workgroups, so no
// workgroup waits for another to be resident: those beyond what the device
// holds at once start as earlier ones finish (the host orders the instances
// largest first).  Global stores and atomics a wave issues before the
// barrier are seen by the other waves' loads after it: the waves of a
// workgroup share one CU, whose vector memory pipeline keeps them in order.
// Grow requests are off (the host cannot step in): an instance that fills
// its heaps or closed set ends in an overflow status and is restarted on the
// launch lanes.  `max_iters` bounds every loop (ST_ABORTED).  When an
// instance ends, its final status | kDoneFlag goes to done[inst] in host
// memory (coherent pinned memory, a system-scope store), where the host reads
// it while the launch still runs: an instance that outgrew its structures is
// restarted on the lanes at once rather than after the whole wave.
constexpr uint32_t kDoneFlag = 0x80000000u;

template <int W, int B>
__global__ __launch_bounds__(B) void persist_kernel(const InstDesc* __restrict__ descs, InstState* states,
                                                    const uint32_t* __restrict__ order, uint32_t max_iters,
                                                    StepParams prm, uint32_t* done) {
    static_assert((B / W) * HeapRef<W>::kTop <= StepSharedT<B>::kHeapTopEntries && W >= 4, "LDS heap tops");
    __shared__ StepSharedT<B> sh;
    const uint32_t inst = order[blockIdx.x];
    const InstDesc& d = descs[inst];
    InstState* st = states + inst;
    const uint32_t q = threadIdx.x / W;
    const bool active = q < d.n_heaps;
    const Tally tl{sh.tally, false};
    using WL = decltype(sh.work);
    // the heap group's heap size, kept in registers across iterations (its
    // top levels stay in the group's LDS slice: run_group's resident form)
    uint32_t size = active ? as_global(d.heap_size)[q] : 0u;
    for (uint32_t it = 0;; ++it) {
        // Every thread reads the same status: the barrier that ended the
        // previous iteration orders it after every write to it by this
        // workgroup, and no other workgroup writes it.
        const uint32_t status = load_relaxed(&st->status);
        if (status != ST_RUNNING) break;
        if (it >= max_iters) {
            if (threadIdx.x == 0) set_status(st, ST_ABORTED);
            break;
        }
        step_begin(sh);
        switch (d.kind) {
            case KIND_PUZZLE3:
                run_group<W, PuzzleOps<3>, WL, true>(d, st, status, it, q, active, d.n_heaps, prm, tl, sh.work,
                                                     sh.heap_top, sh.dirty, &size);
                break;
            case KIND_PUZZLE4:
                run_group<W, PuzzleOps<4>, WL, true>(d, st, status, it, q, active, d.n_heaps, prm, tl, sh.work,
                                                     sh.heap_top, sh.dirty, &size);
                break;
            case KIND_GRID4:  // paired pops: two nodes' atomics in flight per lane (the kernel's register
                              // count, set by its other kinds, leaves room for them)
                run_group<W, GridOps<4>, WL, true, 2>(d, st, status, it, q, active, d.n_heaps, prm, tl, sh.work,
                                                      sh.heap_top, sh.dirty, &size);
                break;
            case KIND_GRID8:
                run_group<W, GridOps<8>, WL, true>(d, st, status, it, q, active, d.n_heaps, prm, tl, sh.work,
                                                   sh.heap_top, sh.dirty, &size);
                break;
            default:
                if (threadIdx.x == 0) set_status(st, ST_ABORTED);
                break;
        }
        tally_end(sh.tally, st, it, 0);
        __syncthreads();  // the last wave's flush of the tallies, and every heap and record write
    }
    if (threadIdx.x == 0 && done)
        __hip_atomic_store(done + inst, kDoneFlag | load_relaxed(&st->status), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

struct PathJob {
    uint32_t inst;
    uint32_t cap;     // states
    uint64_t offset;  // words of the output buffer
};

// One kPathThreads-thread workgroup per finished instance: grids walk with
// all of its threads through an LDS window of the closed set, the hashed sets
// with thread 0 (their parent slots have no locality to exploit).
__global__ __launch_bounds__(kPathThreads) void path_kernel(const InstDesc* descs, const PathJob* jobs, uint32_t n,
                                                            uint64_t* out, uint32_t* lens) {
    __shared__ __align__(16) uint16_t tab[kPathTabWords];  // 32-bit words sit at its end (walk_grid_path)
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const PathJob j = jobs[i];
    const InstDesc& d = descs[j.inst];
    uint32_t len = kNone;
    if (d.kind == KIND_GRID4 || d.kind == KIND_GRID8) {
        len = walk_grid_path(d, out + j.offset, j.cap, tab);
    } else if (threadIdx.x == 0) {
        switch (d.kind) {
            case KIND_PUZZLE3: len = walk_path<PuzzleOps<3>>(d, out + j.offset, j.cap); break;
            case KIND_PUZZLE4: len = walk_path<PuzzleOps<4>>(d, out + j.offset, j.cap); break;
            case KIND_PUZZLE5:
                len = walk_path<Puzzle5Ops>(d, reinterpret_cast<Puzzle5Ops::Key*>(out + j.offset), j.cap);
                break;
            default: break;
        }
    }
    if (threadIdx.x == 0) lens[i] = len;
}

// One level of the breadth-first search that builds a pattern database of
// NT tiles on an S x S board (core.hpp: kPdbTiles for S = 5, kPdb15Tiles for
// S = 4; placement index: sum of cell * (S*S)^i): every placement at distance
// d marks the unvisited placements one tile move away with d + 1.  Several
// threads may mark the same byte with the same value, which is harmless; a
// level whose predecessor marked nothing returns at once.
template <int NT, int S>
__global__ void pdb_level_kernel(uint8_t* dist, uint32_t d, uint32_t* changed) {
    constexpr uint32_t C = S * S;
    if (d > 0 && __hip_atomic_load(&changed[d - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    uint32_t n = 1, pw[NT];
    for (int i = 0; i < NT; ++i) {
        pw[i] = n;
        n *= C;
    }
    bool any = false;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        if (dist[idx] != d) continue;
        int pos[NT];
        uint32_t rest = idx;
        for (int i = 0; i < NT; ++i) {
            pos[i] = (int)(rest % C);
            rest /= C;
        }
        for (int i = 0; i < NT; ++i) {
            const int r = pos[i] / S, c = pos[i] % S;
            for (int m = 0; m < 4; ++m) {  // row +1, row -1, column +1, column -1
                const int nr = r + (m == 0) - (m == 1), nc = c + (m == 2) - (m == 3);
                if (nr < 0 || nr >= S || nc < 0 || nc >= S) continue;
                const int np = nr * S + nc;
                bool taken = false;
                for (int k = 0; k < NT; ++k) taken |= (k != i && pos[k] == np);
                if (taken) continue;
                const uint32_t nidx = (uint32_t)((int)idx + (np - pos[i]) * (int)pw[i]);
                if (dist[nidx] == 0xFF) {
                    dist[nidx] = (uint8_t)(d + 1);
                    any = true;
                }
            }
        }
    }
    if (any) __hip_atomic_store(&changed[d], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Copies a grid's row-major cell weights into the tiled layout of grid_index,
// one 4 x 4 tile (16 bytes) per thread; padding cells beyond the grid's width
// or height become walls.
__global__ void tile_cells_kernel(const uint8_t* src, uint8_t* dst, uint32_t w, uint32_t h, uint32_t tile_cols,
                                  uint64_t tiles) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tiles; t += stride) {
        const uint32_t x0 = (uint32_t)(t % tile_cols) * 4u, y0 = (uint32_t)(t / tile_cols) * 4u;
        uint32_t row[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            uint32_t v = 0;
            const uint32_t y = y0 + r;
#pragma unroll
            for (uint32_t c = 0; c < 4; ++c)
                if (x0 + c < w && y < h) v |= (uint32_t)src[(uint64_t)y * w + x0 + c] << (8 * c);
            row[r] = v;
        }
        reinterpret_cast<uint4*>(dst)[t] = make_uint4(row[0], row[1], row[2], row[3]);
    }
}

__global__ void put_byte_kernel(uint8_t* p, uint8_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---- growth of a paused instance's structures (ga_engine::grow_in_place) ----
// Inserts every key of the old closed set `od` into the cleared larger one
// `nd`, with its value, and records where it went: map[old slot] = new slot
// (kNone for empty slots).  Keys are unique in the old set, so every insert
// claims a fresh slot.
template <class Set>
__global__ void rehash_kernel(InstDesc od, InstDesc nd, uint32_t* map, uint32_t* failed) {
    const uint64_t n = od.table_mask + 1, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        typename Set::Key key;
        uint32_t g, par;
        Set::load(od, (uint32_t)i, key, g, par);
        const uint64_t first = *reinterpret_cast<const uint64_t*>(&key);
        if (first == kEmptyKey) {  // the first key word of an empty slot (hashed and board sets alike)
            map[i] = kNone;
            continue;
        }
        uint32_t slot, probes = 0;
        bool isnew;
        if (!Set::find_or_insert(nd, key, 1u << 20, slot, isnew, probes)) {
            map[i] = kNone;
            atomicAdd(failed, 1u);
            continue;
        }
        *Set::val_ptr(nd, slot) = *Set::val_ptr(od, (uint32_t)i);
        map[i] = slot;
    }
}

// Parent links of the rehashed set: old slots to new slots.
template <class Set>
__global__ void remap_parents_kernel(InstDesc nd, const uint32_t* map) {
    const uint64_t n = nd.table_mask + 1, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const gptr<uint64_t> vp = Set::val_ptr(nd, (uint32_t)j);
        const uint64_t v = *vp;
        if (v == kInfVal) continue;  // empty slot
        const uint32_t par = val_parent(v);
        if (par != kNone) *vp = (v & ~0xFFFFFFFFull) | map[par];
    }
}

// Copies heap q's entries 0 .. size[q]-1 from the old layout to the new one
// (map, if given, renumbers their closed-set slots); one thread per entry
// position of the old stride.
__global__ void copy_heaps_kernel(const HeapEntry* from, HeapEntry* to, const uint32_t* size, uint32_t n_heaps,
                                  uint32_t old_stride, uint32_t new_stride, uint32_t W, const uint32_t* map,
                                  uint64_t n_map) {
    const uint64_t total = (uint64_t)n_heaps * old_stride, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
        const uint32_t q = (uint32_t)(k / old_stride), pos = (uint32_t)(k % old_stride);
        if (pos < W - 1 || pos - (W - 1) >= size[q]) continue;  // logical entry j sits at physical j + W - 1
        HeapEntry e = from[k];
        if (map && e.slot < n_map) e.slot = map[e.slot];
        to[(uint64_t)q * new_stride + pos] = e;
    }
}

// Renumbers the closed-set slots of the candidate records the next
// iteration reads (the first n of the buffer).
__global__ void remap_records_kernel(HeapEntry* rec, uint64_t n, const uint32_t* map, uint64_t n_map) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        if (rec[k].prio != kNoRecord && rec[k].slot < n_map) rec[k].slot = map[rec[k].slot];
}

// Resumes a paused instance with its grown descriptor: clears its grow
// requests and sets it running again (its heaps, records and closed set are
// those of the iteration it paused at).
__global__ void resume_kernel(InstDesc d, InstDesc* descs, InstState* states, uint32_t slot) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    descs[slot] = d;
    InstState* st = states + slot;
    st->grow[0] = st->grow[1] = st->grow[2] = 0;
    st->status = ST_RUNNING;
}

// Launch map of a single resident instance (slot 0) with n_heaps heaps.
__global__ void map_kernel(WgEntry* map, uint32_t n_wg, uint32_t heaps_per_wg, uint32_t n_heaps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_wg) map[i] = WgEntry{0u, i * heaps_per_wg, n_heaps, 0u};
}

// The walked path (states goal..start in path_kernel's output, `wide`: two
// words of Board5 per state) in start..goal order; 24-puzzle boards are
// unpacked into 32 bytes each (25 tile bytes, 7 zero bytes), the layout of
// ga_result paths.
__global__ void path_order_kernel(const uint64_t* src, const uint32_t* lens, uint64_t* dst, uint32_t wide) {
    const uint32_t n = lens[0];
    if (n == kNone) return;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t from = n - 1 - k;
        if (!wide) {
            dst[k] = src[from];
            continue;
        }
        const Board5 x{src[2 * (size_t)from], src[2 * (size_t)from + 1]};
        uint64_t w[4] = {0, 0, 0, 0};
        for (int c = 0; c < 25; ++c) w[c >> 3] |= (uint64_t)Puzzle5Ops::get(x, c) << (8 * (c & 7));
        for (int q = 0; q < 4; ++q) dst[4 * (size_t)k + q] = w[q];
    }
}

__global__ void abort_kernel(InstState* states, uint32_t slot) {
    if (threadIdx.x == 0 && blockIdx.x == 0) atomicCAS(&states[slot].status, (uint32_t)ST_RUNNING, (uint32_t)ST_ABORTED);
}

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct GaError : std::runtime_error {
    int code;
    GaError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw GaError(GA_E_HIP, std::string(what) + ": " + hipGetErrorName(e));
}

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

uint64_t pow2_floor(uint64_t v) {
    uint64_t p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

uint64_t pow2_ceil(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p *= 2;
    return p;
}

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// device arena: chunks from hipMalloc, first-fit with coalescing
// ---------------------------------------------------------------------------
class Arena {
   public:
    explicit Arena(uint64_t chunk_bytes) : chunk_bytes_(chunk_bytes) {}
    ~Arena() {
        for (auto& c : chunks_) (void)hipFree(c.base);
    }

    void* alloc(uint64_t bytes) {
        bytes = round_up(bytes ? bytes : 1, kAlign);
        for (auto& c : chunks_) {
            void* p = take(c, bytes);
            if (p) return p;
        }
        uint64_t sz = std::max<uint64_t>(chunk_bytes_, round_up(bytes, 1ull << 21));
        char* base = nullptr;
        if (hipMalloc(&base, sz) != hipSuccess) {
            (void)hipGetLastError();
            // Chunks are kept once made, so the device can be held by chunks with
            // nothing in them, each too small for this request (a grown closed set
            // is larger than one chunk).  Give those back and try once more: the
            // only hipFree calls outside engine destruction, on this path alone.
            if (!trim() || hipMalloc(&base, sz) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        chunks_.push_back(Chunk{base, sz, {}});
        chunks_.back().free[0] = sz;
        return take(chunks_.back(), bytes);
    }

    void release(void* p) {
        auto it = sizes_.find(p);
        if (it == sizes_.end()) return;
        uint64_t bytes = it->second;
        sizes_.erase(it);
        in_use_ -= bytes;
        for (auto& c : chunks_) {
            char* cp = (char*)p;
            if (cp < c.base || cp >= c.base + c.size) continue;
            uint64_t off = (uint64_t)(cp - c.base);
            auto nx = c.free.lower_bound(off);
            if (nx != c.free.end() && off + bytes == nx->first) {
                bytes += nx->second;
                nx = c.free.erase(nx);
            }
            if (nx != c.free.begin()) {
                auto pv = std::prev(nx);
                if (pv->first + pv->second == off) {
                    pv->second += bytes;
                    return;
                }
            }
            c.free[off] = bytes;
            return;
        }
    }

    uint64_t in_use() const { return in_use_; }
    uint64_t peak() const { return peak_; }

    // Frees the chunks that hold no allocation; returns whether any was freed.
    bool trim() {
        bool freed = false;
        for (size_t i = chunks_.size(); i-- > 0;) {
            const Chunk& c = chunks_[i];
            if (c.free.size() == 1 && c.free.begin()->first == 0 && c.free.begin()->second == c.size) {
                (void)hipFree(c.base);
                chunks_.erase(chunks_.begin() + (ptrdiff_t)i);
                freed = true;
            }
        }
        return freed;
    }

   private:
    static constexpr uint64_t kAlign = 256;
    struct Chunk {
        char* base;
        uint64_t size;
        std::map<uint64_t, uint64_t> free;  // offset -> bytes
    };
    void* take(Chunk& c, uint64_t bytes) {
        for (auto it = c.free.begin(); it != c.free.end(); ++it) {
            if (it->second < bytes) continue;
            uint64_t off = it->first, rest = it->second - bytes;
            c.free.erase(it);
            if (rest) c.free[off + bytes] = rest;
            void* p = c.base + off;
            sizes_[p] = bytes;
            in_use_ += bytes;
            peak_ = std::max(peak_, in_use_);
            return p;
        }
        return nullptr;
    }
    uint64_t chunk_bytes_;
    std::vector<Chunk> chunks_;
    std::map<void*, uint64_t> sizes_;
    uint64_t in_use_ = 0, peak_ = 0;
};

// Pinned host buffer that grows on demand.
struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) (void)hipHostFree(p);
        size_t sz = std::max<size_t>({n, bytes + bytes / 2, 4096});  // geometric: few reallocations
        p = nullptr;
        bytes = 0;
        hip_check(hipHostMalloc(&p, sz, hipHostMallocDefault), "hipHostMalloc");
        bytes = sz;
    }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

// Device bytes of a grid's cell weights: the tiled copy the kernels read and,
// for cells in host memory, the row-major upload it is made from.
uint64_t grid_cell_bytes(const ga_request& r) {
    return round_up(grid_tiled_cells(r.width, r.height), 256) + (r.cells_device ? 0 : (uint64_t)r.width * r.height);
}

int kind_of(const ga_request& r) {
    if (r.kind == GA_PUZZLE) return r.side == 3 ? KIND_PUZZLE3 : r.side == 4 ? KIND_PUZZLE4 : KIND_PUZZLE5;
    return r.connectivity == 8 ? KIND_GRID8 : KIND_GRID4;
}

int branching_of(int kind) { return kind == KIND_GRID8 ? 8 : 4; }

bool is_wide(int kind) { return kind == KIND_PUZZLE5; }

// 64-bit words per path state in the path buffers
uint32_t words_of(int kind) { return is_wide(kind) ? 2 : 1; }

// a 24-puzzle board (25 bytes) in the device packing
Puzzle5Ops::Key pack_board5(const uint8_t* b) {
    Board5 x{0, 0};
    for (int c = 0; c < 25; ++c) Puzzle5Ops::set(x, c, b[c]);
    return to_raw(x);
}

// ---------------------------------------------------------------------------
// one request moving through the engine
// ---------------------------------------------------------------------------
struct Job {
    const ga_request* req = nullptr;
    ga_result* out = nullptr;
    uint64_t arrival_ns = 0;  // relative to the start of the run
    uint32_t heap_shift = 0, table_shift = 0, retries = 0;
    int kind = 0;
    uint32_t w_min = 1;
    // residency
    int slot = -1;
    uint64_t first_chunk = 0;  // first chunk whose read-back reflects this job
    InstDesc desc{};
    uint64_t bytes = 0;
    std::vector<void*> allocs;
    const uint8_t* cells_dev = nullptr;
    uint64_t t_admit = 0;
    uint32_t last_iterations = 0;
    uint32_t cover = 0;  // heaps covered by the launch map (power of two)
    uint32_t lane = 0;   // launch lane
    bool finished = false;
    Puzzle5Ops::Key ends[2] = {};  // 24-puzzle: start and goal in the device packing (uploaded at admission)
    void* owner = nullptr;         // the online server's entry holding this job
    uint32_t grows = 0;            // in-place growths of its heaps or closed set
    uint64_t expanded_before = 0;  // expansions of the attempts given up by restarts
};

struct PathBatch {
    std::vector<Job*> jobs;
    std::vector<uint64_t> offsets;
    std::vector<InstState> snaps;
    void* dev = nullptr;   // PathJob[n] | lens[n] | out[total]
    Pinned* host = nullptr;  // lens[n] | out[total]; from the engine's pool
    hipEvent_t ev = nullptr;
    uint64_t total = 0;
};

}  // namespace

namespace {

// Step-kernel instantiation a launch runs: a kind mask (kind_bit of one kind
// that has a single-kind instantiation, or kAllKinds), or kWideVariant for
// the 24-puzzle kernel.
constexpr uint32_t kWideVariant = kind_bit(KIND_PUZZLE5);
// the 4-connected grid instantiation with paired pops (ga_config.step_flags & STEP_PAIRED_POPS), and the
// all-kinds one whose 4-connected grids pop in pairs (arity 4 only)
constexpr uint32_t kGrid4Paired = kind_bit(KIND_GRID4) | (1u << 31);
constexpr uint32_t kAllPaired = kAllKinds | (1u << 30);

template <int W, int B>
void launch_step_w(uint32_t variant, hipStream_t s, uint32_t grid, const WgEntry* map, const InstDesc* descs,
                   InstState* states, const uint32_t* iter, uint32_t r, const StepParams& prm) {
    const dim3 g(grid), b(B);
    switch (variant) {
        case kind_bit(KIND_PUZZLE3):
            step_kernel<W, kind_bit(KIND_PUZZLE3), B><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            break;
        case kind_bit(KIND_PUZZLE4):
            step_kernel<W, kind_bit(KIND_PUZZLE4), B><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            break;
        case kind_bit(KIND_GRID4):
            step_kernel<W, kind_bit(KIND_GRID4), B><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            break;
        case kGrid4Paired:  // W = 4 only (upload_map)
            if constexpr (W == 4)
                step_kernel<W, kind_bit(KIND_GRID4), B, 2><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            else
                throw GaError(GA_E_INVALID, "two pops per round need heap arity 4");
            break;
        case kAllPaired:  // W = 4 only (upload_map)
            if constexpr (W == 4)
                step_kernel<W, kAllKinds, B, 2><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            else
                throw GaError(GA_E_INVALID, "two pops per round need heap arity 4");
            break;
        case kWideVariant:
            step_kernel_wide<W, B><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            break;
        default:
            step_kernel<W, kAllKinds, B><<<g, b, 0, s>>>(map, descs, states, iter, r, prm);
            break;
    }
}

template <int B>
void launch_step_b(uint32_t arity, uint32_t variant, hipStream_t s, uint32_t grid, const WgEntry* map,
                   const InstDesc* descs, InstState* states, const uint32_t* iter, uint32_t r, const StepParams& prm) {
    if (arity == 8)
        launch_step_w<8, B>(variant, s, grid, map, descs, states, iter, r, prm);
    else
        launch_step_w<4, B>(variant, s, grid, map, descs, states, iter, r, prm);
}

// `block`: threads per step workgroup (ga_config.block: 128, 256 or 512)
void launch_step(uint32_t arity, uint32_t block, uint32_t variant, hipStream_t s, uint32_t grid, const WgEntry* map,
                 const InstDesc* descs, InstState* states, const uint32_t* iter, uint32_t r, const StepParams& prm) {
    if (block == 128)
        launch_step_b<128>(arity, variant, s, grid, map, descs, states, iter, r, prm);
    else if (block == 512)
        launch_step_b<512>(arity, variant, s, grid, map, descs, states, iter, r, prm);
    else
        launch_step_b<256>(arity, variant, s, grid, map, descs, states, iter, r, prm);
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// scratch of the no-allocation entry point (ga_solve_noalloc)
// ---------------------------------------------------------------------------
// All device memory that ga_solve_noalloc's kernels use beyond the caller's
// inputs and outputs -- instance descriptor and state, launch map, iteration
// counter, heaps, candidate records, closed set, a grid's cells, the path
// walk's buffers -- is carved out of this array.  Each GPU the library runs on
// holds its own copy until the process ends.  Its host-side name is not a
// device address: host code uses the address hipGetSymbolAddress returns for
// the current device (scratch_address), never the name.
constexpr uint64_t kScratchBytes = 256ull << 20;
static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `device`'s scratch; `device` must be current.
char* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw GaError(GA_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw GaError(GA_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[device] = (char*)p;
    }
    return g_scratch_ptr[device];
}

}  // namespace

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
// A launch lane: one stream with its own iteration clock, launch map and up
// to two chunks in flight.  An instance advances one iteration per launch of
// its lane.  Lanes run concurrently, so an instance that needs many
// iterations (a large grid) is not paced by the launch that carries all the
// others.  Slots are partitioned by lane (slot % lanes) so that a slot is only
// ever reused in stream order on the lane that last ran it.
struct Lane {
    hipStream_t stream = nullptr;
    WgEntry* d_map = nullptr;
    uint32_t map_cap = 0, n_wg = 0;
    uint32_t grid = 0;  // workgroups launched: n_wg, or in graph mode n_wg rounded up to a sixteenth of map_cap
    uint64_t n_heaps = 0;     // heaps the map covers
    uint32_t variant = 0;     // kind mask of the step-kernel instantiation the lane launches
    bool map_dirty = true;
    Pinned h_map[2];
    hipEvent_t ev_map[2] = {nullptr, nullptr};
    int map_buf = 0;
    uint32_t* d_iter = nullptr;
    uint64_t host_base = 0;   // iteration counter at the start of the next chunk
    uint64_t launched = 0;    // chunks launched
    uint64_t processed = 0;   // chunks whose read-back has been handled
    Pinned h_states[2];
    hipEvent_t ev_chunk[2] = {nullptr, nullptr};
    uint32_t chunk_reads[2] = {0, 0};
    std::vector<hipEvent_t> prof_ev[2];  // profile: 2 events per step, per chunk parity
    hipEvent_t busy_ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // profile: around every chunk
    bool timed[2] = {false, false};      // profile: whether that chunk's steps were timed
    // graph mode: the captured chunks, by step-kernel instantiation (all kinds,
    // 8-puzzle, 15-puzzle, 4-connected grid), without / with device counters,
    // and launch grid; valid while the map buffer stays the same
    std::map<uint64_t, hipGraphExec_t> graphs;
    uint32_t running = 0;     // resident jobs not yet finished
    std::vector<int> free_slots;
};

struct ga_engine {
    int device = 0;
    ga_config cfg{};
    hipStream_t path_stream = nullptr;
    // The one-launch batch form's stream, created at its first use: an engine
    // that never uses the form keeps one stream per launch lane and the path
    // stream only (HIP gives a process a few hardware queues; streams beyond
    // them share queues, and lanes that share one run one after the other).
    hipStream_t persist_stream = nullptr;
    hipStream_t ensure_persist_stream() {
        if (!persist_stream) hip_check(hipStreamCreateWithFlags(&persist_stream, hipStreamNonBlocking), "hipStreamCreate");
        return persist_stream;
    }
    Arena* arena = nullptr;
    InstDesc* d_descs = nullptr;
    InstState* d_states = nullptr;
    std::vector<Lane> lanes;  // cfg.lanes lanes, then the wide-key lane
    uint32_t n_lanes = 1;     // cfg.lanes + 1; slots are partitioned by slot % n_lanes
    std::vector<Job*> resident;
    ga_engine_stats stats{};
    // algorithmic bytes per lane: profiled-launch terms, all-launch terms
    double algo_sampled[GA_MAX_LANES + 1] = {}, algo_always[GA_MAX_LANES + 1] = {};
    std::string last_error;
    bool broken = false;  // a HIP error ended a run: the device state is unknown, later runs are refused
    std::atomic<bool> serving{false};  // an online server owns the engine (ga_serve_start .. ga_serve_stop)
    char* scratch_dev = nullptr;  // this device's scratch (ga_solve_noalloc)
    uint8_t* d_pdb = nullptr;     // 24-puzzle pattern databases (kPdbBytes + level flags), built on first use
    uint8_t* d_pdb15 = nullptr;   // 15-puzzle pattern databases (kPdb15Bytes + level flags), built on first use
    // the one-launch form's completion flags: coherent pinned host memory the
    // device writes (one word per instance of a wave), grown on demand
    uint32_t* persist_done = nullptr;
    uint32_t* persist_done_dev = nullptr;
    size_t persist_done_cap = 0;
    uint32_t* ensure_persist_done(size_t n) {
        if (n > persist_done_cap) {
            if (persist_done) (void)hipHostFree(persist_done);
            persist_done = persist_done_dev = nullptr;
            persist_done_cap = 0;
            const size_t cap = std::max<size_t>(n, 1024);
            void* p = nullptr;
            hip_check(hipHostMalloc(&p, cap * sizeof(uint32_t), hipHostMallocCoherent | hipHostMallocMapped),
                      "hipHostMalloc (completion flags)");
            persist_done = (uint32_t*)p;
            void* dp = nullptr;
            hip_check(hipHostGetDevicePointer(&dp, p, 0), "hipHostGetDevicePointer");
            persist_done_dev = (uint32_t*)dp;
            persist_done_cap = cap;
        }
        return persist_done;
    }

    void init(int dev, const ga_config* c) {
        device = dev;
        ga_config_default(&cfg);
        if (c) cfg = *c;
        if (cfg.heap_arity == 0) cfg.heap_arity = 4;
        // Heap arity 16 was taken out: it served slower than 4 and 8, and a
        // served run with it (4,096 heaps per 15-puzzle) once stalled with no
        // cause shown.
        if (cfg.heap_arity != 4 && cfg.heap_arity != 8) throw GaError(GA_E_INVALID, "heap_arity must be 4 or 8");
        if (cfg.block == 0) cfg.block = kStepBlock;
        if (cfg.block != 128 && cfg.block != 256 && cfg.block != 512)
            throw GaError(GA_E_INVALID, "block must be 128, 256 or 512");
        if (cfg.persist_block == 0) cfg.persist_block = 1024;
        if (cfg.persist_block != 256 && cfg.persist_block != 512 && cfg.persist_block != 1024)
            throw GaError(GA_E_INVALID, "persist_block must be 256, 512 or 1024");
        if (cfg.lanes > GA_MAX_LANES) throw GaError(GA_E_INVALID, "lanes must be at most 4");
        if (cfg.max_instances == 0) cfg.max_instances = 1;
        if (cfg.chunk_iters == 0) cfg.chunk_iters = 1;
        if (cfg.max_load_pct == 0 || cfg.max_load_pct > 95) cfg.max_load_pct = 50;
        if (cfg.pops_per_iter == 0) cfg.pops_per_iter = 1;
        if (cfg.rounds == 0) cfg.rounds = 1;
        if (cfg.grow_thresh == 0) cfg.grow_thresh = 32;
        if (cfg.lanes == 0) cfg.lanes = 1;
        // Under rocprofv3 (which exports ROCPROF_OUTPUT_PATH to the program it
        // runs) chunks are launched one kernel at a time: a kernel trace of
        // graph replay ended once in an HSA queue abort (malformed AQL
        // packet) and once in a host segmentation fault inside the runtime.
        if (cfg.use_graph && getenv("ROCPROF_OUTPUT_PATH")) cfg.use_graph = 0;
        n_lanes = cfg.lanes + 1;
        if (cfg.max_instances < n_lanes) cfg.max_instances = n_lanes;
        hip_check(hipSetDevice(device), "hipSetDevice");
        scratch_dev = scratch_address(device);
        hip_check(hipStreamCreateWithFlags(&path_stream, hipStreamNonBlocking), "hipStreamCreate");
        arena = new Arena(cfg.chunk_bytes);
        hip_check(hipMalloc(&d_descs, sizeof(InstDesc) * cfg.max_instances), "hipMalloc descs");
        hip_check(hipMalloc(&d_states, sizeof(InstState) * cfg.max_instances), "hipMalloc states");
        hip_check(hipMemset(d_states, 0, sizeof(InstState) * cfg.max_instances), "hipMemset states");
        lanes.resize(n_lanes);
        for (uint32_t l = 0; l < n_lanes; ++l) {
            Lane& L = lanes[l];
            hip_check(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipMalloc(&L.d_iter, sizeof(uint32_t)), "hipMalloc iter");
            hip_check(hipMemset(L.d_iter, 0, sizeof(uint32_t)), "hipMemset iter");
            for (int i = 0; i < 2; ++i) {
                hip_check(hipEventCreateWithFlags(&L.ev_map[i], hipEventDisableTiming), "hipEventCreate");
                hip_check(hipEventCreateWithFlags(&L.ev_chunk[i], hipEventDisableTiming), "hipEventCreate");
                L.h_states[i].ensure(sizeof(InstState) * cfg.max_instances);
            }
            for (int s = (int)cfg.max_instances - 1; s >= 0; --s)
                if ((uint32_t)s % n_lanes == l) L.free_slots.push_back(s);
        }
        resident.assign(cfg.max_instances, nullptr);
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }

    ~ga_engine() {
        if (path_stream) (void)hipStreamSynchronize(path_stream);
        for (Lane& L : lanes)
            if (L.stream) (void)hipStreamSynchronize(L.stream);
        for (Deferred& df : deferred) {  // their memory goes with the arena
            delete df.fail;
            (void)hipEventDestroy(df.ev);
        }
        for (Pinned* p : pinned_pool) delete p;
        for (hipEvent_t ev : event_pool) (void)hipEventDestroy(ev);
        for (Lane& L : lanes) {
            if (L.stream) (void)hipStreamSynchronize(L.stream);
            for (auto& kv : L.graphs)
                if (kv.second) (void)hipGraphExecDestroy(kv.second);
            for (auto& v : L.prof_ev)
                for (auto ev : v) (void)hipEventDestroy(ev);
            for (auto& pair : L.busy_ev)
                for (hipEvent_t ev : pair)
                    if (ev) (void)hipEventDestroy(ev);
            for (int i = 0; i < 2; ++i) {
                if (L.ev_map[i]) (void)hipEventDestroy(L.ev_map[i]);
                if (L.ev_chunk[i]) (void)hipEventDestroy(L.ev_chunk[i]);
            }
            if (L.d_map) (void)hipFree(L.d_map);
            if (L.d_iter) (void)hipFree(L.d_iter);
            if (L.stream) (void)hipStreamDestroy(L.stream);
        }
        if (path_stream) (void)hipStreamSynchronize(path_stream);
        if (d_descs) (void)hipFree(d_descs);
        if (d_states) (void)hipFree(d_states);
        if (d_pdb) (void)hipFree(d_pdb);
        if (d_pdb15) (void)hipFree(d_pdb15);
        if (persist_done) (void)hipHostFree(persist_done);
        delete arena;
        if (path_stream) (void)hipStreamDestroy(path_stream);
        if (persist_stream) {
            (void)hipStreamSynchronize(persist_stream);
            (void)hipStreamDestroy(persist_stream);
        }
    }

    // Lane of a request: 24-puzzles take the wide-key lane; puzzles and small
    // grids share lane 0; the largest grids (>= deep_cells cells) take the last
    // configured lane and, with three or more lanes, grids of >= deep_cells/4
    // cells the one before it.
    uint32_t lane_of(const Job& j) const {
        if (is_wide(j.kind)) return n_lanes - 1;
        const uint32_t n = cfg.lanes;
        if (n == 1 || j.req->kind != GA_GRID || !cfg.deep_cells) return 0;
        const uint64_t cells = (uint64_t)j.req->width * j.req->height;
        uint64_t t = cfg.deep_cells;
        uint32_t lane = n - 1;
        while (lane > 0 && cells < t) {
            --lane;
            t /= 4;
        }
        return lane;
    }

    // ---- planning -----------------------------------------------------------
    struct Plan {
        uint32_t n_heaps, heap_cap, heap_stride;
        uint64_t table_slots;  // hashed sets; dense sets use width*height
        uint64_t bytes;
    };

    // `heaps` (nonzero): that many heaps, whatever the request asks for, each
    // holding at least `cap_min` entries (the one-launch batch form).
    Plan plan(const Job& j, uint32_t heaps = 0, uint32_t cap_min = 0) const {
        const ga_request& r = *j.req;
        const uint32_t W = cfg.heap_arity;
        Plan p{};
        uint64_t n_req = heaps ? heaps : r.num_queues;
        uint64_t cells = (uint64_t)r.width * r.height;
        if (j.kind == KIND_PUZZLE3) {
            p.n_heaps = (uint32_t)pow2_floor(n_req ? n_req : 256);
            p.heap_cap = r.heap_capacity ? r.heap_capacity : 1024;
            p.table_slots = r.table_slots ? pow2_ceil(r.table_slots) : (1ull << 19);
        } else if (j.kind == KIND_PUZZLE4 || j.kind == KIND_PUZZLE5) {
            p.n_heaps = (uint32_t)pow2_floor(n_req ? n_req : 16384);
            p.heap_cap = r.heap_capacity ? r.heap_capacity : 512;
            p.table_slots = r.table_slots ? pow2_ceil(r.table_slots) : (1ull << 22);
        } else {
            uint64_t dflt = std::min<uint64_t>(16384, std::max<uint64_t>(64, pow2_floor(std::max<uint64_t>(cells / 64, 1))));
            p.n_heaps = (uint32_t)pow2_floor(n_req ? n_req : dflt);
            p.heap_cap = r.heap_capacity ? r.heap_capacity : 256;
            p.table_slots = 0;
        }
        p.n_heaps = std::min<uint32_t>(std::max<uint32_t>(p.n_heaps, 1), 1u << 22);
        p.heap_cap = std::max(p.heap_cap, cap_min);
        uint64_t cap = (uint64_t)p.heap_cap << j.heap_shift;
        p.heap_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(cap, 16), 1u << 28);
        p.heap_stride = (uint32_t)round_up(p.heap_cap + W - 1, W);
        if (p.table_slots) p.table_slots = std::min<uint64_t>(p.table_slots << j.table_shift, 1ull << 32);
        const int K = branching_of(j.kind);
        uint64_t b = (uint64_t)p.n_heaps * p.heap_stride * sizeof(HeapEntry);  // heaps
        b += (uint64_t)p.n_heaps * sizeof(uint32_t);                          // heap sizes
        b += 2ull * p.n_heaps * K * sizeof(HeapEntry);                         // candidates
        if (is_wide(j.kind))  // 32-byte board slots, start and goal
            b += p.table_slots * sizeof(Board5KeySet::PairSlot) + 2 * sizeof(Puzzle5Ops::Key);
        else
            b += p.table_slots ? p.table_slots * sizeof(Slot) : grid_tiled_cells(r.width, r.height) * sizeof(uint64_t);
        p.bytes = b;
        return p;
    }

    // ---- 24-puzzle pattern databases -------------------------------------------
    static constexpr uint32_t kPdbLevels = 64;  // the deepest table (5 tiles on 5x5) has 30 levels
    static bool ordered_goal5(const ga_request& r) {
        if (r.kind != GA_PUZZLE || r.side != 5 || !r.cells) return false;
        for (int c = 0; c < 25; ++c)
            if (r.cells[25 + c] != (c < 24 ? c + 1 : 0)) return false;
        return true;
    }
    // Builds the tables on `stream` (breadth-first search from each pattern's
    // goal placement, one launch per level) and waits for them; one hipMalloc
    // per engine lifetime.
    const uint8_t* ensure_pdb(hipStream_t stream) {
        if (d_pdb) return d_pdb;
        uint8_t* p = nullptr;
        const uint64_t words = (kPdbBytes + 7) / 8;
        const uint64_t bytes = words * 8 + (uint64_t)kPdbPatterns * kPdbLevels * sizeof(uint32_t);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            throw GaError(GA_E_NOMEM, "pattern databases");
        }
        uint32_t* changed = (uint32_t*)(p + words * 8);
        fill_kernel<uint64_t><<<1024, kBlock, 0, stream>>>((uint64_t*)p, words, ~0ull);
        fill_kernel<uint32_t><<<1, kBlock, 0, stream>>>(changed, (uint64_t)kPdbPatterns * kPdbLevels, 0u);
        for (int q = 0; q < kPdbPatterns; ++q) {
            uint32_t goal = 0;
            for (int i = 0; i < kPdbCount[q]; ++i) goal += (kPdbTiles[q][i] - 1) * pow25((uint32_t)i);
            uint8_t* dist = p + (uint64_t)q * kPdbTable;
            put_byte_kernel<<<1, 64, 0, stream>>>(dist + goal, 0);
            for (uint32_t d = 0; d < kPdbLevels; ++d) {
                uint32_t* ch = changed + (uint64_t)q * kPdbLevels;
                if (kPdbCount[q] == 5)
                    pdb_level_kernel<5, 5><<<4096, kBlock, 0, stream>>>(dist, d, ch);
                else
                    pdb_level_kernel<4, 5><<<512, kBlock, 0, stream>>>(dist, d, ch);
            }
        }
        hip_check(hipGetLastError(), "pattern database build");
        std::vector<uint32_t> flags((size_t)kPdbPatterns * kPdbLevels);
        hip_check(hipMemcpyAsync(flags.data(), changed, flags.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                  "read back level flags");
        hip_check(hipStreamSynchronize(stream), "pattern database build");
        for (int q = 0; q < kPdbPatterns; ++q)
            if (flags[(size_t)q * kPdbLevels + kPdbLevels - 1]) {  // still growing at the last level
                (void)hipFree(p);
                throw GaError(GA_E_HIP, "pattern database deeper than the level bound");
            }
        d_pdb = p;
        return d_pdb;
    }

    // ---- 15-puzzle pattern databases ------------------------------------------
    static bool ordered_goal4(const ga_request& r) {
        return r.kind == GA_PUZZLE && r.side == 4 && r.goal == kPuzzle4OrderedGoal;
    }
    // The same for the 15-puzzle's three 5-tile tables (kPdb15Bytes: 3 MiB).
    const uint8_t* ensure_pdb15(hipStream_t stream) {
        if (d_pdb15) return d_pdb15;
        uint8_t* p = nullptr;
        const uint64_t bytes = kPdb15Bytes + (uint64_t)kPdb15Patterns * kPdbLevels * sizeof(uint32_t);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            throw GaError(GA_E_NOMEM, "pattern databases");
        }
        uint32_t* changed = (uint32_t*)(p + kPdb15Bytes);
        fill_kernel<uint64_t><<<256, kBlock, 0, stream>>>((uint64_t*)p, kPdb15Bytes / 8, ~0ull);
        fill_kernel<uint32_t><<<1, kBlock, 0, stream>>>(changed, (uint64_t)kPdb15Patterns * kPdbLevels, 0u);
        for (int q = 0; q < kPdb15Patterns; ++q) {
            uint32_t goal = 0;
            for (int i = 0; i < 5; ++i) goal += (kPdb15Tiles[q][i] - 1) << (4 * i);
            uint8_t* dist = p + (uint64_t)q * kPdb15Table;
            put_byte_kernel<<<1, 64, 0, stream>>>(dist + goal, 0);
            for (uint32_t d = 0; d < kPdbLevels; ++d)
                pdb_level_kernel<5, 4><<<1024, kBlock, 0, stream>>>(dist, d, changed + (uint64_t)q * kPdbLevels);
        }
        hip_check(hipGetLastError(), "pattern database build");
        std::vector<uint32_t> flags((size_t)kPdb15Patterns * kPdbLevels);
        hip_check(hipMemcpyAsync(flags.data(), changed, flags.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
                  "read back level flags");
        hip_check(hipStreamSynchronize(stream), "pattern database build");
        for (int q = 0; q < kPdb15Patterns; ++q)
            if (flags[(size_t)q * kPdbLevels + kPdbLevels - 1]) {  // still growing at the last level
                (void)hipFree(p);
                throw GaError(GA_E_HIP, "pattern database deeper than the level bound");
            }
        d_pdb15 = p;
        return d_pdb15;
    }

    // ---- admission ------------------------------------------------------------
    // Takes the device buffers of one instance from `take` (the arena, or the
    // no-allocation scratch) and fills its descriptor; returns false when
    // `take` returns nullptr for one of them.
    template <class Take>
    bool carve_instance(const Job& j, const Plan& p, Take take, InstDesc& d) const {
        const ga_request& r = *j.req;
        const int K = branching_of(j.kind);
        d = InstDesc{};
        d.kind = (uint32_t)j.kind;
        d.start = r.start;
        d.goal = r.goal;
        d.heaps = (HeapEntry*)take((uint64_t)p.n_heaps * p.heap_stride * sizeof(HeapEntry));
        d.heap_size = (uint32_t*)take((uint64_t)p.n_heaps * sizeof(uint32_t));
        d.cand = (HeapEntry*)take(2ull * p.n_heaps * K * sizeof(HeapEntry));
        if (!d.heaps || !d.heap_size || !d.cand) return false;
        if (r.kind == GA_PUZZLE) {
            if (is_wide(j.kind)) {
                d.table = (Slot*)take(p.table_slots * sizeof(Board5KeySet::PairSlot));
                d.user = take(2 * sizeof(Puzzle5Ops::Key));
                if (!d.user) return false;
            } else {
                d.table = (Slot*)take(p.table_slots * sizeof(Slot));
            }
            if (!d.table) return false;
        } else {
            d.dense = (uint64_t*)take(grid_tiled_cells(r.width, r.height) * sizeof(uint64_t));
            if (!d.dense) return false;
            // the weights in the tiled layout, then (host cells) the row-major copy they are tiled from
            d.cells = (const uint8_t*)take(grid_cell_bytes(r));
            if (!d.cells) return false;
        }
        d.n_heaps = p.n_heaps;
        d.heap_cap = p.heap_cap;
        d.heap_stride = p.heap_stride;
        d.a0 = cfg.adaptive_start ? (uint32_t)std::min<uint64_t>(p.n_heaps, pow2_floor(cfg.adaptive_start))
                                  : p.n_heaps;
        d.rounds = (r.kind == GA_GRID && cfg.rounds_grid) ? cfg.rounds_grid : cfg.rounds;
        {
            const uint32_t slack = r.kind == GA_GRID ? cfg.slack_grid : cfg.slack_puzzle;
            d.f_slack = slack ? slack - 1 : kInf;  // 0 = no gating; k = slack k-1
        }
        if (r.kind == GA_PUZZLE) {
            d.width = r.side;
            d.height = r.side;
            d.table_mask = p.table_slots - 1;
            d.load_limit = p.table_slots * cfg.max_load_pct / 100;
            for (uint32_t c = 0; c < r.side * r.side; ++c) {
                const uint32_t t = r.side == 5 ? r.cells[25 + c] : (uint32_t)(r.goal >> (4 * c)) & 15u;
                d.goal_rc[t] = (uint8_t)(((c / r.side) << 4) | (c % r.side));
            }
            if (is_wide(j.kind)) d.start = d.goal = 0;  // the keys are in d.user
        } else {
            d.width = r.width;
            d.height = r.height;
            d.w_min = j.w_min;
            d.tile_cols = grid_tile_cols(r.width);
            const uint32_t sx = (uint32_t)(r.start % r.width), sy = (uint32_t)(r.start / r.width);
            const uint32_t gx = (uint32_t)(r.goal % r.width), gy = (uint32_t)(r.goal / r.width);
            d.start = grid_index(sx, sy, d.tile_cols);
            d.goal = grid_index(gx, gy, d.tile_cols);
            d.start_xy = (sy << 16) | sx;
            d.goal_xy = (gy << 16) | gx;
        }
        return true;
    }

    // Clears an instance's closed set and heap sizes and uploads its problem
    // data on `stream` (the descriptor itself is written by seed_kernel).
    // j.ends must stay in place until the stream has run the copy.
    void init_instance(Job& j, const Plan& p, const InstDesc& d, hipStream_t stream) {
        const ga_request& r = *j.req;
        const uint32_t fill_grid = 1024;
        if (r.kind == GA_PUZZLE) {
            if (is_wide(j.kind)) {  // start and goal keys for the seed, the goal test and the path walk
                j.ends[0] = pack_board5(r.cells);
                j.ends[1] = pack_board5(r.cells + 25);
                hip_check(hipMemcpyAsync((void*)d.user, j.ends, sizeof(j.ends), hipMemcpyHostToDevice, stream),
                          "upload ends");
            }
            const uint64_t words = p.table_slots * (is_wide(j.kind) ? sizeof(Board5KeySet::PairSlot) : sizeof(Slot)) / 8;
            fill_kernel<uint64_t><<<fill_grid, kBlock, 0, stream>>>((uint64_t*)d.table, words, ~0ull);
        } else {
            const uint64_t tiled = grid_tiled_cells(r.width, r.height);
            const uint8_t* src = r.cells_device;
            if (!src) {  // host cells: uploaded row-major behind the tiled weights
                uint8_t* stage = (uint8_t*)d.cells + round_up(tiled, 256);
                hip_check(hipMemcpyAsync(stage, r.cells, (uint64_t)r.width * r.height, hipMemcpyHostToDevice, stream),
                          "upload cells");
                src = stage;
            }
            const uint64_t tiles = tiled / 16;
            tile_cells_kernel<<<(uint32_t)std::min<uint64_t>(4096, (tiles + kBlock - 1) / kBlock), kBlock, 0, stream>>>(
                src, (uint8_t*)d.cells, r.width, r.height, d.tile_cols, tiles);
            fill_kernel<uint64_t><<<fill_grid, kBlock, 0, stream>>>(d.dense, tiled, ~0ull);
        }
        fill_kernel<uint32_t><<<std::min<uint32_t>(fill_grid, (p.n_heaps + kBlock - 1) / kBlock), kBlock, 0,
                                stream>>>(d.heap_size, p.n_heaps, 0u);
    }

    // Device bytes an admission of j takes from the budget.
    uint64_t need_of(const Job& j) const {
        return plan(j).bytes + (j.req->kind == GA_GRID ? grid_cell_bytes(*j.req) : 0);
    }

    // Returns 1 if admitted, 0 if it must wait (no slot / budget), or throws.
    // `reserve`: budget bytes held back for requests that wait on other lanes.
    int admit(Job& j, uint64_t reserve = 0) {
        const uint32_t lane = lane_of(j);
        Lane& L = lanes[lane];
        if (L.free_slots.empty()) return 0;
        const ga_request& r = *j.req;
        Plan p = plan(j);
        uint64_t need = p.bytes + (r.kind == GA_GRID ? grid_cell_bytes(r) : 0);
        if (need > cfg.budget_bytes) throw GaError(GA_E_BUDGET, "request exceeds the admission budget");
        if (arena->in_use() + need + reserve > cfg.budget_bytes) return 0;
        const int W = (int)cfg.heap_arity;
        std::vector<void*> allocs;
        bool others_resident = false;
        for (Job* o : resident)
            if (o) {
                others_resident = true;
                break;
            }
        bool short_of_memory = false;
        auto take = [&](uint64_t bytes) -> void* {
            if (short_of_memory) return nullptr;
            void* ptr = arena->alloc(bytes);
            if (!ptr) {
                for (void* a : allocs) arena->release(a);
                allocs.clear();
                if (!others_resident) throw GaError(GA_E_NOMEM, "device arena exhausted");
                short_of_memory = true;  // retry once resident requests have released memory
                return nullptr;
            }
            allocs.push_back(ptr);
            return ptr;
        };
        InstDesc d{};
        if (!carve_instance(j, p, take, d)) return 0;
        hipStream_t stream = L.stream;
        if (cfg.pdb24 && ordered_goal5(r)) d.pdb = ensure_pdb(stream);
        if (cfg.pdb15 && ordered_goal4(r)) d.pdb = ensure_pdb15(stream);
        set_grow_limits(j, p, d);
        init_instance(j, p, d, stream);
        int slot = L.free_slots.back();
        L.free_slots.pop_back();
        d.start_iter = (uint32_t)L.host_base;
        seed_kernel<<<1, 64, 0, stream>>>(d, d_descs, d_states, (uint32_t)slot, W);
        hip_check(hipGetLastError(), "seed launch");
        j.slot = slot;
        j.lane = lane;
        j.desc = d;
        j.allocs = std::move(allocs);
        j.bytes = need;
        j.first_chunk = L.launched;
        j.t_admit = now_ns();
        j.last_iterations = 0;
        // The map covers every allocated heap, so the width can double on the
        // device every iteration; workgroups of heaps not yet in use exit at once.
        j.cover = p.n_heaps;
        resident[slot] = &j;
        ++L.running;
        L.map_dirty = true;
        stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, arena->in_use());
        return 1;
    }

    // The grow thresholds of an instance (InstDesc::grow_heap_at, grow_table_at):
    // a heap may gain K records and K more per expansion round in one
    // iteration, and the closed set K new keys per heap and round, so a
    // request raised at the end of an iteration is acted on before any
    // structure can overflow.  A structure already at its largest size has
    // none (it then overflows, and the request restarts).
    void set_grow_limits(const Job& j, const Plan& p, InstDesc& d) const {
        d.grow_heap_at = 0;
        d.grow_table_at = 0;
        if (!cfg.grow_in_place) return;
        const uint64_t K = (uint64_t)branching_of(j.kind), rounds = d.rounds ? d.rounds : cfg.rounds;
        const uint64_t per_iter = K * (rounds + 1);
        d.grow_heap_at = p.heap_cap < (1u << 26) ? (uint32_t)std::max<int64_t>(1, (int64_t)p.heap_cap - (int64_t)per_iter) : 0u;
        d.grow_table_at = 0;
        if (p.table_slots && p.table_slots < (1ull << 29)) {  // at most 2^31 slots after a growth
            const uint64_t new_keys = (uint64_t)p.n_heaps * rounds * K;
            d.grow_table_at = d.load_limit > new_keys ? d.load_limit - new_keys : 1;
        }
    }

    // Grows the structures a paused instance asked for, in place: the heaps
    // are copied into four times their capacity and/or the closed set
    // rehashed into four times its slots (heap entries, the candidate
    // records the next iteration reads and the parent links follow the new
    // slot numbers), then the instance resumes at the iteration it paused at
    // with all its search state.  Returns false (nothing changed) when the
    // arena cannot hold the larger structures.
    bool grow_in_place(Job& j, const InstState& st, Lane& L) {
        const uint32_t req = st.grow[(st.paused_iter + 2) % 3];
        const Plan op = plan(j);
        const bool dense = j.kind == KIND_GRID4 || j.kind == KIND_GRID8;
        const bool heap = (req & GROW_HEAP) != 0, table = (req & GROW_TABLE) != 0 && !dense;
        Job g = j;
        if (heap) g.heap_shift += 2;
        if (table) g.table_shift += 2;
        const Plan np = plan(g);
        if (np.bytes > op.bytes && arena->in_use() + (np.bytes - op.bytes) > cfg.budget_bytes) return false;
        InstDesc d = j.desc;
        std::vector<void*> fresh;
        auto take = [&](uint64_t bytes) -> void* {
            void* ptr = arena->alloc(bytes);
            if (ptr) fresh.push_back(ptr);
            return ptr;
        };
        HeapEntry* heaps = heap ? (HeapEntry*)take((uint64_t)np.n_heaps * np.heap_stride * sizeof(HeapEntry)) : d.heaps;
        const uint64_t slot_bytes = is_wide(j.kind) ? sizeof(Board5KeySet::PairSlot) : sizeof(Slot);
        Slot* tab = table ? (Slot*)take(np.table_slots * slot_bytes) : d.table;
        const uint64_t n_old = op.table_slots;
        uint32_t* map = table ? (uint32_t*)take(n_old * sizeof(uint32_t) + 256) : nullptr;
        if (!heaps || (table && (!tab || !map))) {  // (grids have no table: d.table is null)
            for (void* f : fresh) arena->release(f);
            return false;
        }
        hipStream_t s = L.stream;
        const uint32_t W = cfg.heap_arity;
        uint32_t* failed = map ? map + n_old : nullptr;  // one word past the map
        if (table) {
            InstDesc nd = d;
            nd.table = tab;
            nd.table_mask = np.table_slots - 1;
            fill_kernel<uint64_t><<<1024, kBlock, 0, s>>>((uint64_t*)tab, np.table_slots * slot_bytes / 8, ~0ull);
            fill_kernel<uint32_t><<<1, 64, 0, s>>>(failed, 1, 0u);
            const uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (n_old + kBlock - 1) / kBlock);
            if (is_wide(j.kind)) {
                rehash_kernel<Board5KeySet><<<grid, kBlock, 0, s>>>(d, nd, map, failed);
                remap_parents_kernel<Board5KeySet><<<4096, kBlock, 0, s>>>(nd, map);
            } else {
                rehash_kernel<HashedU64Set><<<grid, kBlock, 0, s>>>(d, nd, map, failed);
                remap_parents_kernel<HashedU64Set><<<4096, kBlock, 0, s>>>(nd, map);
            }
            // the records iteration paused_iter reads: the first (width of the iteration before) * K
            const uint32_t K = (uint32_t)branching_of(j.kind);
            const uint64_t width = st.active[(st.paused_iter + 2) % 3];
            HeapEntry* cin = d.cand + (uint64_t)((st.paused_iter - 1) & 1) * d.n_heaps * K;
            remap_records_kernel<<<1024, kBlock, 0, s>>>(cin, width * K, map, n_old);
            d.table = tab;
            d.table_mask = np.table_slots - 1;
            d.load_limit = np.table_slots * cfg.max_load_pct / 100;
        }
        if (heap || table) {
            const uint64_t total = (uint64_t)op.n_heaps * op.heap_stride;
            const uint32_t grid = (uint32_t)std::min<uint64_t>(4096, (total + kBlock - 1) / kBlock);
            if (heap)
                copy_heaps_kernel<<<grid, kBlock, 0, s>>>(d.heaps, heaps, d.heap_size, op.n_heaps, op.heap_stride,
                                                          np.heap_stride, W, map, n_old);
            else  // same layout: renumber in place
                copy_heaps_kernel<<<grid, kBlock, 0, s>>>(d.heaps, d.heaps, d.heap_size, op.n_heaps, op.heap_stride,
                                                          op.heap_stride, W, map, n_old);
            d.heaps = heaps;
            d.heap_cap = np.heap_cap;
            d.heap_stride = np.heap_stride;
        }
        set_grow_limits(g, np, d);
        // the next launch of this lane runs iteration paused_iter of the instance
        d.start_iter = (uint32_t)(L.host_base - st.paused_iter);
        resume_kernel<<<1, 64, 0, s>>>(d, d_descs, d_states, (uint32_t)j.slot);
        hip_check(hipGetLastError(), "grow in place");
        // The old structures (and the slot map) are freed once the growth's
        // kernels have run, without the host waiting for them: a wait here held
        // the run loop for the rest of the lane's chunk in flight.
        Deferred df;
        if (failed) {
            df.fail = take_pinned(sizeof(uint32_t));
            hip_check(hipMemcpyAsync(df.fail->p, failed, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "read back");
        }
        auto drop = [&](void* old) {
            df.frees.push_back(old);
            j.allocs.erase(std::remove(j.allocs.begin(), j.allocs.end(), old), j.allocs.end());
        };
        if (heap) drop(j.desc.heaps);
        if (table) drop(j.desc.table);
        if (map) df.frees.push_back(map);
        for (void* f : fresh)
            if (f != map) j.allocs.push_back(f);
        df.ev = take_event();
        hip_check(hipEventRecord(df.ev, s), "record");
        deferred.push_back(df);
        j.heap_shift = g.heap_shift;
        j.table_shift = g.table_shift;
        j.bytes += np.bytes - op.bytes;
        j.desc = d;
        ++j.grows;
        stats.peak_bytes = std::max<uint64_t>(stats.peak_bytes, arena->in_use());
        return true;
    }

    // Marks a running job as no longer launched (its lane's map drops it).
    void retire(Job& j) {
        if (j.finished) return;
        j.finished = true;
        Lane& L = lanes[j.lane];
        --L.running;
        L.map_dirty = true;
    }

    // Pinned host buffers and timing-free events of the path walks, reused from
    // batch to batch: hipHostMalloc / hipHostFree per batch had taken most of
    // the run loop's time (a host buffer's free is slow), starving the lanes.
    std::vector<Pinned*> pinned_pool;
    std::vector<hipEvent_t> event_pool;
    Pinned* take_pinned(size_t bytes) {
        // the smallest pooled buffer that holds `bytes`, else the largest (grown to fit)
        long fit = -1, largest = -1;
        for (size_t i = 0; i < pinned_pool.size(); ++i) {
            const size_t b = pinned_pool[i]->bytes;
            if (b >= bytes && (fit < 0 || b < pinned_pool[(size_t)fit]->bytes)) fit = (long)i;
            if (largest < 0 || b > pinned_pool[(size_t)largest]->bytes) largest = (long)i;
        }
        const long k = fit >= 0 ? fit : largest;
        Pinned* p = nullptr;
        if (k >= 0) {
            p = pinned_pool[(size_t)k];
            pinned_pool.erase(pinned_pool.begin() + k);
        } else {
            p = new Pinned();
        }
        p->ensure(bytes);
        return p;
    }
    void give_pinned(Pinned* p) { pinned_pool.push_back(p); }
    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t ev = event_pool.back();
            event_pool.pop_back();
            return ev;
        }
        hipEvent_t ev = nullptr;
        hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        return ev;
    }
    void give_event(hipEvent_t ev) { event_pool.push_back(ev); }

    // Memory whose last readers are kernels still queued (an in-place
    // growth's old structures): freed once the event after them completes.
    struct Deferred {
        hipEvent_t ev = nullptr;
        std::vector<void*> frees;
        Pinned* fail = nullptr;  // the rehash's lost-key count, read back with the event
    };
    std::vector<Deferred> deferred;
    // Frees what has become free (wait: all of it, after the queued work).
    // Throws when a rehash reports lost keys (never expected: a new table is
    // four times larger than the old).
    void reap_deferred(bool wait) {
        bool lost = false;
        hipError_t err = hipSuccess;
        size_t k = 0;
        for (size_t i = 0; i < deferred.size(); ++i) {
            Deferred& df = deferred[i];
            const hipError_t q = wait ? hipEventSynchronize(df.ev) : hipEventQuery(df.ev);
            if (q == hipErrorNotReady) {
                deferred[k++] = df;
                continue;
            }
            if (q != hipSuccess) err = q;
            else if (df.fail && *(const uint32_t*)df.fail->p) lost = true;
            for (void* f : df.frees) arena->release(f);
            if (df.fail) give_pinned(df.fail);
            give_event(df.ev);
        }
        deferred.resize(k);
        hip_check(err, "growth event");
        if (lost) throw GaError(GA_E_HIP, "closed-set rehash lost keys");
    }

    void release(Job& j) {
        for (void* a : j.allocs) arena->release(a);
        j.allocs.clear();
        if (j.slot >= 0) {
            Lane& L = lanes[j.lane];
            resident[j.slot] = nullptr;
            L.free_slots.push_back(j.slot);
            j.slot = -1;
            L.map_dirty = true;
        }
    }

    // After a run has failed: waits for everything launched, drops the jobs
    // still resident (they belong to the failed run) and resets every lane's
    // iteration clock, so that the engine holds nothing of that run.  A HIP
    // error leaves the device state unknown: the engine is then marked
    // broken and refuses later runs.
    void abandon(int code) {
        bool ok = code != GA_E_HIP;
        for (Lane& L : lanes) ok = hipStreamSynchronize(L.stream) == hipSuccess && ok;
        ok = hipStreamSynchronize(path_stream) == hipSuccess && ok;
        if (persist_stream) ok = hipStreamSynchronize(persist_stream) == hipSuccess && ok;
        for (Deferred& df : deferred) {  // the streams are idle: their memory is free
            for (void* f : df.frees) arena->release(f);
            if (df.fail) give_pinned(df.fail);
            give_event(df.ev);
        }
        deferred.clear();
        for (Job*& j : resident) {
            if (!j) continue;
            for (void* a : j->allocs) arena->release(a);
            j->allocs.clear();
            lanes[j->lane].free_slots.push_back(j->slot);
            j->slot = -1;
            j = nullptr;
        }
        for (Lane& L : lanes) {
            L.processed = L.launched;
            L.running = 0;
            L.n_wg = 0;
            L.map_dirty = true;
            L.host_base = 0;
            if (ok) {
                fill_kernel<uint32_t><<<1, 64, 0, L.stream>>>(L.d_iter, 1, 0u);
                ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(L.stream) == hipSuccess;
            }
        }
        (void)hipGetLastError();
        if (!ok) broken = true;
    }

    // ---- launch ---------------------------------------------------------------
    void upload_map(Lane& L, uint32_t lane) {
        const uint32_t W = cfg.heap_arity;
        const uint32_t heaps_per_wg = cfg.block / W;
        std::vector<WgEntry> map;
        L.n_heaps = 0;
        uint32_t kinds = 0;
        for (uint32_t s = lane; s < cfg.max_instances; s += n_lanes) {
            Job* j = resident[s];
            if (!j || j->finished) continue;
            const uint32_t n = std::min(j->cover, j->desc.n_heaps);
            L.n_heaps += n;
            kinds |= kind_bit((uint32_t)j->kind);
            for (uint32_t h = 0; h < n; h += heaps_per_wg) map.push_back(WgEntry{s, h, n, 0});
        }
        L.n_wg = (uint32_t)map.size();
        L.map_dirty = false;
        if (!L.n_wg) return;
        // a single-kind instantiation when the lane holds one kind that has one
        uint32_t variant =
            (kinds == kind_bit(KIND_PUZZLE3) || kinds == kind_bit(KIND_PUZZLE4) || kinds == kind_bit(KIND_GRID4))
                ? kinds
                : kAllKinds;
        // two pops per round for 4-connected grids, at W = 4 only (at W = 8 half of a group's lanes
        // hold no grid successor)
        if ((cfg.step_flags & STEP_PAIRED_POPS) && W == 4) {
            if (variant == kind_bit(KIND_GRID4)) variant = kGrid4Paired;
            else if (variant == kAllKinds && (cfg.step_flags & STEP_PAIRED_ALL)) variant = kAllPaired;
        }
        L.variant = variant;  // its captured chunk, if any, stays valid (graphs are kept per instantiation)
        if (L.n_wg > L.map_cap) {
            if (L.d_map) {
                hip_check(hipStreamSynchronize(L.stream), "sync");
                hip_check(hipFree(L.d_map), "hipFree map");
            }
            L.map_cap = (uint32_t)pow2_ceil(L.n_wg);
            hip_check(hipMalloc(&L.d_map, sizeof(WgEntry) * L.map_cap), "hipMalloc map");
            drop_graph(L);  // the captured launches refer to the old map and grid size
        }
        // Graph mode launches a captured grid: n_wg rounded up to a sixteenth of
        // map_cap (one graph per such size is kept), padded with entries that
        // exit at once.  Launching map_cap workgroups, a size that only grows,
        // had left more than half of every launch as padding.
        L.grid = cfg.use_graph ? std::min<uint32_t>(L.map_cap, (uint32_t)round_up(L.n_wg, std::max<uint32_t>(64, L.map_cap / 16)))
                               : L.n_wg;
        const uint32_t n_up = L.grid;
        map.resize(n_up, WgEntry{kNone, 0, 0, 0});
        int b = L.map_buf;
        L.map_buf ^= 1;
        hip_check(hipEventSynchronize(L.ev_map[b]), "map event");
        L.h_map[b].ensure(sizeof(WgEntry) * n_up);
        memcpy(L.h_map[b].p, map.data(), sizeof(WgEntry) * n_up);
        hip_check(hipMemcpyAsync(L.d_map, L.h_map[b].p, sizeof(WgEntry) * n_up, hipMemcpyHostToDevice, L.stream),
                  "upload map");
        hip_check(hipEventRecord(L.ev_map[b], L.stream), "record map");
    }

    void drop_graph(Lane& L) {
        if (L.graphs.empty()) return;
        hip_check(hipStreamSynchronize(L.stream), "sync");
        for (auto& kv : L.graphs)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        L.graphs.clear();
    }

    static int variant_slot(uint32_t variant) {
        switch (variant) {
            case kind_bit(KIND_PUZZLE3): return 1;
            case kind_bit(KIND_PUZZLE4): return 2;
            case kind_bit(KIND_GRID4): return 3;
            case kGrid4Paired: return 4;
            case kAllPaired: return 5;
            default: return 0;
        }
    }

    void enqueue_steps(Lane& L, uint32_t grid, const StepParams& prm) {
        const uint32_t R = cfg.chunk_iters;
        hipStream_t stream = L.stream;
        const uint32_t variant = &L == &lanes.back() ? kWideVariant : L.variant;
        for (uint32_t r = 0; r < R; ++r)
            launch_step(cfg.heap_arity, cfg.block, variant, stream, grid, L.d_map, d_descs, d_states, L.d_iter, r, prm);
        advance_kernel<<<1, 64, 0, stream>>>(L.d_iter, R);
    }

    void launch_chunk(Lane& L) {
        const uint32_t R = cfg.chunk_iters;
        // profile n samples two kinds of chunks of the lane: chunk c is timed
        // when c % n == 0 (one pair of HIP events around its R launches, which
        // are issued one by one with nothing between them: no graph node, no
        // event) and counted when c % n == n / 2 (the device counters that only
        // profiling needs are on in its steps; their ballots and LDS atomics
        // slow an expansion, so timed chunks run without them).  n == 1: every
        // chunk is both.
        const uint64_t ph = cfg.profile ? L.launched % cfg.profile : 1;
        const bool timed = cfg.profile && ph == 0;
        const bool counted = cfg.profile && ph == cfg.profile / 2;
        StepParams prm{cfg.pops_per_iter, 1u << 16, counted ? 1u : 0u, cfg.rounds, cfg.grow_thresh, cfg.step_flags, {0, 0}};
        const int b = (int)(L.launched & 1);
        std::vector<hipEvent_t>& pe = L.prof_ev[b];
        if (timed && pe.size() < 2) {
            for (size_t i = pe.size(); i < 2; ++i) {
                hipEvent_t ev;
                hip_check(hipEventCreate(&ev), "hipEventCreate");
                pe.push_back(ev);
            }
        }
        L.timed[b] = timed;
        if (cfg.profile) {
            for (hipEvent_t& ev : L.busy_ev[b])
                if (!ev) hip_check(hipEventCreate(&ev), "hipEventCreate");
            hip_check(hipEventRecord(L.busy_ev[b][0], L.stream), "event");
        }
        if (cfg.use_graph && !timed) {
            const uint64_t key = ((uint64_t)variant_slot(L.variant) << 40) | ((uint64_t)(counted ? 1 : 0) << 32) | L.grid;
            hipGraphExec_t& ge = L.graphs[key];
            if (!ge) {
                hipGraph_t g;
                hip_check(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal), "begin capture");
                enqueue_steps(L, L.grid, prm);
                hip_check(hipStreamEndCapture(L.stream, &g), "end capture");
                hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "instantiate");
                (void)hipGraphDestroy(g);
            }
            hip_check(hipGraphLaunch(ge, L.stream), "graph launch");
        } else {
            if (timed) hip_check(hipEventRecord(pe[0], L.stream), "event");
            enqueue_steps(L, L.n_wg, prm);
            if (timed) hip_check(hipEventRecord(pe[1], L.stream), "event");
        }
        hip_check(hipGetLastError(), "step launch");
        if (cfg.profile) hip_check(hipEventRecord(L.busy_ev[b][1], L.stream), "event");
        L.host_base += R;
        stats.step_launches += R;
        const uint32_t li = (uint32_t)(&L - lanes.data());
        stats.lane_launches[li] += R;
        if (counted) {
            stats.heap_groups += L.n_heaps * R;
            stats.counted_launches += R;
            stats.lane_counted_launches[li] += R;
        }
        uint32_t n_read = 0;
        for (uint32_t s = 0; s < cfg.max_instances; ++s)
            if (resident[s]) n_read = s + 1;
        hip_check(hipMemcpyAsync(L.h_states[b].p, d_states, sizeof(InstState) * n_read, hipMemcpyDeviceToHost,
                                 L.stream),
                  "read back states");
        hip_check(hipEventRecord(L.ev_chunk[b], L.stream), "record chunk");
        L.chunk_reads[b] = n_read;
        ++L.launched;
    }

    // Called once the chunk's read-back event has completed: the time of a
    // timed chunk's R launches (and its one-block counter advance).
    void collect_profile(Lane& L, uint64_t chunk) {
        if (!cfg.profile) return;
        const uint32_t li = (uint32_t)(&L - lanes.data());
        float ms = 0;
        // every chunk: from the GPU reaching it (or the previous chunk's end) to its last launch's end
        hip_check(hipEventElapsedTime(&ms, L.busy_ev[chunk & 1][0], L.busy_ev[chunk & 1][1]), "elapsed");
        stats.lane_busy_ms[li] += ms;
        if (!L.timed[chunk & 1]) return;
        std::vector<hipEvent_t>& pe = L.prof_ev[chunk & 1];
        hip_check(hipEventElapsedTime(&ms, pe[0], pe[1]), "elapsed");
        stats.timed_launches += cfg.chunk_iters;
        stats.lane_timed_launches[li] += cfg.chunk_iters;
        stats.step_ms_total += ms;
        stats.lane_step_ms[li] += ms;
    }

    // ---- results -------------------------------------------------------------
    void fill_counters(Job& j, const InstState& s) {
        ga_result* o = j.out;
        const bool dense = j.kind == KIND_GRID4 || j.kind == KIND_GRID8;
        o->expanded = j.expanded_before + s.expanded;
        o->stored = dense ? 0 : s.occupied;
        o->iterations = s.iterations;
        o->retries = j.retries;
        o->grows = j.grows;
        stats.expanded += s.expanded;
        stats.generated += s.generated;
        stats.pushed += s.pushed;
        stats.probes += s.probes;
        stats.sift_down += s.sift_down;
        stats.sift_up += s.sift_up;
        stats.pops += s.pops;
        stats.groups += s.groups;
        stats.inserted += s.inserted;
        stats.lookups += s.lookups;
        stats.work_groups += s.work_groups;
        stats.sift_down_hbm += s.sift_down_hbm;
        stats.sift_up_hbm += s.sift_up_hbm;
        stats.push_in += s.push_in;
        algo_sampled[j.lane] += algo_bytes(j.kind, s, true);
        algo_always[j.lane] += algo_bytes(j.kind, s, false);
        // one atomicMin per find-or-insert, plus one CAS or atomic load per probe of a hashed set;
        // a dense set's probe is the cell index itself, no memory access
        stats.lane_atomics[j.lane] += s.lookups + (dense ? 0 : s.probes);
    }

    // Algorithmic bytes of the step kernel: the bytes its data structures'
    // operations must read or write in global memory, with E = 16 (heap entry
    // and candidate record) and W = heap arity:
    //   per live heap group and launch  E (its top, read for the min-f test)
    //   per pop                         2E (root + last entry) + the closed-set
    //                                   stale check (key 8 + g 4 hashed, g|parent
    //                                   8 dense, board 16 + g 4 for 24-puzzles)
    //   per sift-down level in global   W*E (children) + E (entry moved)
    //   memory (levels inside the LDS copy of the heap's top are not counted)
    //   per closed-set probe            8 (CAS on the key; 24-puzzle 16: key
    //                                   CAS + board half)
    //   per closed-set lookup           8 (atomicMin on g | parent)
    //   per generated grid move         1 (weight of the entered cell)
    //   per push                        E (entry written)
    //   per record pushed from the      2E (written by its producer, read by
    //   candidate buffer                its heap)
    //   per sift-up level in global     2E (parent read, entry moved)
    //   memory
    // Idle heaps clearing their candidate slice and the copy of heap tops into
    // LDS and back are this design's overhead, not counted here: they show in
    // the measured traffic.  `sampled`: the terms from the counters that only
    // profiled launches count; otherwise the always-counted term (grid moves).
    double algo_bytes(int kind, const InstState& s, bool sampled) const {
        const bool dense = kind == KIND_GRID4 || kind == KIND_GRID8;
        const double E = 16, W = cfg.heap_arity;
        if (!sampled) return dense ? (double)s.generated : 0.0;
        const double stale = dense ? 8 : is_wide(kind) ? 20 : 12;
        double b = (double)s.groups * E;
        b += (double)s.pops * (2 * E + stale);
        b += (double)s.sift_down_hbm * (W * E + E);
        if (!dense) b += (double)s.probes * (is_wide(kind) ? 16 : 8);
        b += (double)s.lookups * 8;
        b += (double)s.pushed * E + (double)s.push_in * 2 * E + (double)s.sift_up_hbm * 2 * E;
        return b;
    }
    // algorithmic bytes of all step launches of lane l: the sampled terms
    // scaled from its counted launches to all of them
    double algo_lane(uint32_t l) const {
        const double scale = stats.lane_counted_launches[l]
                                 ? (double)stats.lane_launches[l] / (double)stats.lane_counted_launches[l]
                                 : 0.0;
        return algo_always[l] + algo_sampled[l] * scale;
    }
    double algo_total() const {
        double t = 0;
        for (uint32_t l = 0; l < n_lanes; ++l) t += algo_lane(l);
        return t;
    }

    void finish(Job& j, int status, uint64_t t0) {
        j.out->status = status;
        j.out->user_tag = j.req->user_tag;
        uint64_t t = now_ns();
        if (j.t_admit) j.out->search_ns = t - j.t_admit;
        j.out->latency_ns = t - t0 - std::min(j.arrival_ns, t - t0);
        j.finished = true;
        ++stats.completed;
    }
};

namespace {

// ---------------------------------------------------------------------------
// the run loop shared by plain, batched and served GA*
// ---------------------------------------------------------------------------
int validate(const ga_request& r) {
    if (r.kind == GA_PUZZLE && r.side == 5) {  // boards in cells: start, then goal
        if (!r.cells) return GA_E_INVALID;
        for (int k = 0; k < 2; ++k) {
            uint32_t seen = 0;
            for (int c = 0; c < 25; ++c) {
                const uint8_t t = r.cells[25 * k + c];
                if (t >= 25 || ((seen >> t) & 1u)) return GA_E_INVALID;
                seen |= 1u << t;
            }
        }
        return GA_OK;
    }
    if (r.kind == GA_PUZZLE) {
        if (r.side != 3 && r.side != 4) return GA_E_INVALID;
        uint32_t n = r.side * r.side;
        for (uint64_t s : {r.start, r.goal}) {
            uint32_t seen = 0;
            for (uint32_t c = 0; c < n; ++c) {
                uint32_t t = (uint32_t)(s >> (4 * c)) & 15u;
                if (t >= n || ((seen >> t) & 1)) return GA_E_INVALID;
                seen |= 1u << t;
            }
            if (n < 16 && (s >> (4 * n)) != 0) return GA_E_INVALID;
        }
        return GA_OK;
    }
    if (r.kind == GA_GRID) {
        if (!r.width || !r.height || r.width > 65535 || r.height > 65535) return GA_E_INVALID;
        if ((uint64_t)r.width * r.height >= 0xFFFFFFFFull) return GA_E_INVALID;
        if (grid_tiled_cells(r.width, r.height) >= 0xFFFFFFFFull) return GA_E_INVALID;  // tiled indices are 32-bit
        if (r.connectivity != 4 && r.connectivity != 8) return GA_E_INVALID;
        if (!r.cells && !r.cells_device) return GA_E_INVALID;
        uint64_t n = (uint64_t)r.width * r.height;
        if (r.start >= n || r.goal >= n) return GA_E_INVALID;
        if (r.cells && (!r.cells[r.start] || !r.cells[r.goal])) return GA_E_INVALID;
        return GA_OK;
    }
    return GA_E_INVALID;
}

uint32_t grid_wmin(const ga_request& r) {
    if (r.w_min) return r.w_min;
    if (!r.cells) return 1;  // weights are >= 1, so 1 is always admissible
    uint64_t n = (uint64_t)r.width * r.height;
    uint32_t m = 256;
    for (uint64_t i = 0; i < n && m > 1; ++i)
        if (r.cells[i] && r.cells[i] < m) m = r.cells[i];
    return m == 256 ? 1 : m;
}

// ---------------------------------------------------------------------------
// the one-launch batch form (ga_config.persist)
// ---------------------------------------------------------------------------
constexpr uint32_t kPersistMaxIters = 1u << 22;

bool persist_eligible(const ga_engine* e, const Job& j) {
    if (j.kind == KIND_PUZZLE3 || j.kind == KIND_PUZZLE4) return true;
    if (j.kind == KIND_GRID4 || j.kind == KIND_GRID8)
        return (uint64_t)j.req->width * j.req->height <= e->cfg.persist_cells;
    return false;
}

// Expected work, for ordering the workgroups of a wave largest first: grids
// by their cells, ahead of puzzles; puzzles by the Manhattan distance of their
// start (a sliding-puzzle search grows with its depth).
double persist_work(const Job& j) {
    const ga_request& r = *j.req;
    if (r.kind == GA_GRID) return 1e6 + (double)r.width * r.height;
    const uint32_t n = r.side * r.side;
    uint32_t pos[16] = {};
    for (uint32_t c = 0; c < n; ++c) pos[(r.goal >> (4 * c)) & 15u] = c;
    double h = 0;
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t t = (uint32_t)(r.start >> (4 * c)) & 15u;
        if (t) h += std::abs((int)(c / r.side) - (int)(pos[t] / r.side)) + std::abs((int)(c % r.side) - (int)(pos[t] % r.side));
    }
    return h;
}

template <int W>
void launch_persist_w(uint32_t block, hipStream_t s, uint32_t n, const InstDesc* descs, InstState* states,
                      const uint32_t* order, uint32_t max_iters, const StepParams& prm, uint32_t* done) {
    switch (block) {
        case 256: persist_kernel<W, 256><<<n, 256, 0, s>>>(descs, states, order, max_iters, prm, done); break;
        case 512: persist_kernel<W, 512><<<n, 512, 0, s>>>(descs, states, order, max_iters, prm, done); break;
        default: persist_kernel<W, 1024><<<n, 1024, 0, s>>>(descs, states, order, max_iters, prm, done); break;
    }
}

// The one-launch form's waves, polled by the run loop (so that the launch
// lanes run beside them, and a lane waiting for memory a wave holds keeps
// polling until the wave has ended and freed it).  A wave takes as many of
// the remaining instances as the budget holds, largest first, carves and
// clears their structures, seeds them with one launch and runs them to their
// ends with one persist_kernel launch on the engine's persist stream; once
// its states are back, the paths of those that reached the goal are walked
// with one path_kernel launch.  Instances that outgrew their heaps or closed
// set go back to the run loop to restart on the lanes with larger structures.
class PersistRun {
   public:
    PersistRun(ga_engine* e, std::vector<Job*> jobs, uint64_t t0) : e(e), jobs(std::move(jobs)), t0(t0) {
        e->ensure_persist_stream();
        std::stable_sort(this->jobs.begin(), this->jobs.end(),
                         [](Job* a, Job* b) { return persist_work(*a) > persist_work(*b); });
    }
    PersistRun(const PersistRun&) = delete;
    PersistRun& operator=(const PersistRun&) = delete;
    ~PersistRun() { drop(); }

    bool idle() const { return stage == IDLE && next == jobs.size(); }

    // One pass: moves a wave whose work has completed on to its next stage, or
    // launches the next wave.  Finished jobs are appended to `finished`, jobs
    // to restart on the lanes to `restart`.  Returns whether anything moved.
    bool poll(std::vector<Job*>& finished, std::vector<Job*>& restart) {
        bool moved = false;
        if (stage == SEARCH) moved = early_restarts(restart);
        if (stage != IDLE) {
            const hipError_t q = hipEventQuery(ev);
            if (q == hipErrorNotReady) return moved;
            hip_check(q, "one-launch batch");
            if (stage == SEARCH) {
                outcomes(finished, restart);
            } else {
                paths(finished);
            }
            return true;
        }
        if (next < jobs.size()) {
            launch(finished);
            return true;
        }
        return false;
    }

    // After a failure: waits for the stream and frees the wave.
    void drop() {
        if (stage != IDLE || !wave.empty()) (void)hipStreamSynchronize(e->persist_stream);
        for (size_t i = 0; i < wave.size(); ++i) {
            // a job restarted early may hold the lanes' allocations by now: its
            // wave memory was moved to `meta` when it was restarted
            if (i < restarted.size() && restarted[i]) continue;
            for (void* a : wave[i]->allocs) e->arena->release(a);
            wave[i]->allocs.clear();
        }
        wave.clear();
        for (void* m : meta) e->arena->release(m);
        meta.clear();
        restarted.clear();
        if (host) e->give_pinned(host);
        if (path_host) e->give_pinned(path_host);
        host = path_host = nullptr;
        if (ev) e->give_event(ev);
        ev = nullptr;
        stage = IDLE;
    }

   private:
    enum Stage { IDLE, SEARCH, PATHS };
    ga_engine* const e;
    std::vector<Job*> jobs;
    const uint64_t t0;
    size_t next = 0;
    Stage stage = IDLE;
    std::vector<Job*> wave;
    std::vector<void*> meta;  // the wave's own device buffers, and those of instances restarted early
    std::vector<char> restarted;  // per wave instance: restarted on the lanes while the wave ran
    Pinned* host = nullptr;   // descriptors and order up, states back
    Pinned* path_host = nullptr;
    hipEvent_t ev = nullptr;
    InstDesc* d_descs = nullptr;
    InstState* snap = nullptr;
    std::vector<uint32_t> done;  // wave indices that reached the goal, and their path offsets
    std::vector<uint64_t> offs;
    uint64_t path_lens_b = 0;

    void launch(std::vector<Job*>& finished) {
        const uint32_t W = e->cfg.heap_arity, B = e->cfg.persist_block, H = B / W;
        hipStream_t s = e->persist_stream;
        // heap entries per heap: a workgroup's few heaps hold the whole open list
        auto cap_min = [](const Job& j) { return (j.kind == KIND_GRID4 || j.kind == KIND_GRID8) ? 1024u : 4096u; };
        std::vector<InstDesc> descs;
        for (; next < jobs.size(); ++next) {
            Job& j = *jobs[next];
            const ga_engine::Plan p = e->plan(j, H, cap_min(j));
            const uint64_t need = p.bytes + (j.req->kind == GA_GRID ? grid_cell_bytes(*j.req) : 0);
            if (need > e->cfg.budget_bytes) {
                e->finish(j, GA_E_BUDGET, t0);
                finished.push_back(&j);
                continue;
            }
            if (e->arena->in_use() + need > e->cfg.budget_bytes) break;
            std::vector<void*> allocs;
            bool short_of_memory = false;
            auto take = [&](uint64_t bytes) -> void* {
                void* ptr = short_of_memory ? nullptr : e->arena->alloc(bytes);
                if (!ptr) short_of_memory = true;
                else allocs.push_back(ptr);
                return ptr;
            };
            InstDesc d{};
            if (!e->carve_instance(j, p, take, d)) {
                for (void* a : allocs) e->arena->release(a);
                // nothing else holds memory that could be freed: waiting would not help
                bool others = !wave.empty();
                for (Job* o : e->resident) others = others || o != nullptr;
                if (!others) throw GaError(GA_E_NOMEM, "device arena exhausted (one-launch batch)");
                break;
            }
            d.grow_heap_at = 0;  // no grow requests: the host cannot step in during the launch
            d.grow_table_at = 0;
            d.start_iter = 0;
            if (e->cfg.pdb15 && ga_engine::ordered_goal4(*j.req)) d.pdb = e->ensure_pdb15(s);
            e->init_instance(j, p, d, s);
            j.allocs = std::move(allocs);
            j.desc = d;
            j.lane = 0;
            j.bytes = need;
            wave.push_back(&j);
            descs.push_back(d);
        }
        if (wave.empty()) return;  // the memory is held by the lanes: the next pass tries again
        const uint32_t n = (uint32_t)wave.size();
        e->stats.peak_bytes = std::max<uint64_t>(e->stats.peak_bytes, e->arena->in_use());
        const uint64_t desc_b = round_up(n * sizeof(InstDesc), 256), state_b = round_up(n * sizeof(InstState), 256);
        char* dm = (char*)e->arena->alloc(desc_b + state_b + round_up(n * sizeof(uint32_t), 256));
        if (!dm) throw GaError(GA_E_NOMEM, "device arena exhausted (one-launch batch)");
        meta.push_back(dm);
        d_descs = (InstDesc*)dm;
        InstState* d_states = (InstState*)(dm + desc_b);
        uint32_t* d_order = (uint32_t*)(dm + desc_b + state_b);
        host = e->take_pinned(desc_b + state_b + n * sizeof(uint32_t));
        char* hm = (char*)host->p;
        std::memcpy(hm, descs.data(), n * sizeof(InstDesc));
        uint32_t* h_order = (uint32_t*)(hm + desc_b + state_b);
        for (uint32_t i = 0; i < n; ++i) h_order[i] = i;  // the wave is already largest first
        hip_check(hipMemcpyAsync(d_descs, hm, n * sizeof(InstDesc), hipMemcpyHostToDevice, s), "upload descs");
        hip_check(hipMemcpyAsync(d_order, h_order, n * sizeof(uint32_t), hipMemcpyHostToDevice, s), "upload order");
        seed_batch_kernel<<<(n + 63) / 64, 64, 0, s>>>(d_descs, d_states, n, (int)W);
        const uint64_t t_start = now_ns();
        for (Job* j : wave) j->t_admit = t_start;
        // the launch's loops end by this many iterations per instance whatever the configuration
        // (an unlimited search would otherwise only be bounded by the 32-bit counter)
        const uint32_t max_iters =
            e->cfg.max_iterations ? std::min(e->cfg.max_iterations, kPersistMaxIters) : kPersistMaxIters;
        const StepParams prm{e->cfg.pops_per_iter, 1u << 16, 0u, e->cfg.rounds, e->cfg.grow_thresh,
                             e->cfg.step_flags, {0, 0}};
        uint32_t* done = e->ensure_persist_done(n);  // the previous wave's launch has ended (drop)
        std::memset(done, 0, n * sizeof(uint32_t));
        restarted.assign(n, 0);
        if (W == 8) launch_persist_w<8>(B, s, n, d_descs, d_states, d_order, max_iters, prm, e->persist_done_dev);
        else launch_persist_w<4>(B, s, n, d_descs, d_states, d_order, max_iters, prm, e->persist_done_dev);
        hip_check(hipGetLastError(), "one-launch batch");
        snap = (InstState*)(hm + desc_b);
        hip_check(hipMemcpyAsync(snap, d_states, n * sizeof(InstState), hipMemcpyDeviceToHost, s), "read back states");
        ev = e->take_event();
        hip_check(hipEventRecord(ev, s), "record");
        stage = SEARCH;
        e->stats.persist_launches += 1;
        e->stats.persist_instances += n;
    }

    // While the wave runs: an instance whose workgroup reported an overflow
    // (its completion flag) is restarted on the lanes now, its structures 4x
    // the one that filled up, instead of after the whole wave.  Its memory in
    // the wave stays until the wave ends.  (The expansions of that attempt are
    // not known yet and are not added to its result's count.)  Returns
    // whether any was restarted.
    bool early_restarts(std::vector<Job*>& restart) {
        const volatile uint32_t* done = e->persist_done;
        bool any = false;
        for (size_t i = 0; i < wave.size(); ++i) {
            if (restarted[i]) continue;
            const uint32_t v = done[i];
            if (!(v & kDoneFlag)) continue;
            const uint32_t st = v & ~kDoneFlag;
            if (st != ST_OVERFLOW_HEAP && st != ST_OVERFLOW_TABLE) continue;
            Job& j = *wave[i];
            const bool dense = j.kind == KIND_GRID4 || j.kind == KIND_GRID8;
            if (j.retries >= e->cfg.max_retries || (dense && st == ST_OVERFLOW_TABLE)) continue;  // outcomes() ends it
            restarted[i] = 1;
            for (void* a : j.allocs) meta.push_back(a);
            j.allocs.clear();
            ++j.retries;
            if (st == ST_OVERFLOW_HEAP) j.heap_shift += 2;
            else j.table_shift += 2;
            restart.push_back(&j);
            ++e->stats.persist_early_restarts;
            any = true;
        }
        return any;
    }

    // The wave's states are back: finishes what did not reach the goal,
    // returns the overflowed to the lanes, and walks the paths of the rest.
    void outcomes(std::vector<Job*>& finished, std::vector<Job*>& restart) {
        const uint32_t n = (uint32_t)wave.size();
        done.clear();
        offs.clear();
        uint64_t total = 0;
        std::vector<Job*> keep;  // jobs whose memory stays until the path walk
        for (uint32_t i = 0; i < n; ++i) {
            if (restarted[i]) continue;  // already back on the lanes
            Job& j = *wave[i];
            const InstState& st = snap[i];
            if (st.status == ST_DONE && st.best != kInf) {
                done.push_back(i);
                offs.push_back(total);
                total += (uint64_t)st.best + 1;  // every move costs >= 1
                continue;
            }
            e->fill_counters(j, st);
            if (st.status == ST_OVERFLOW_HEAP || st.status == ST_OVERFLOW_TABLE) {
                j.expanded_before += st.expanded;
                if (j.retries >= e->cfg.max_retries) {
                    e->finish(j, GA_E_OVERFLOW, t0);
                    finished.push_back(&j);
                    continue;
                }
                ++j.retries;  // the lanes start it again, with their heap count and 4x the structure that filled up
                if (st.status == ST_OVERFLOW_HEAP) j.heap_shift += 2;
                else j.table_shift += 2;
                restart.push_back(&j);
                continue;
            }
            j.out->cost = 0;
            e->finish(j, st.status == ST_DONE ? GA_NO_PATH : GA_E_TIMEOUT, t0);
            finished.push_back(&j);
        }
        e->give_event(ev);
        ev = nullptr;
        if (done.empty()) {
            drop();
            return;
        }
        const uint32_t m = (uint32_t)done.size();
        const uint64_t jobs_b = round_up(m * sizeof(PathJob), 256);
        path_lens_b = round_up(m * sizeof(uint32_t), 16);
        char* pd = (char*)e->arena->alloc(jobs_b + path_lens_b + total * sizeof(uint64_t));
        if (!pd) throw GaError(GA_E_NOMEM, "device arena exhausted (paths)");
        meta.push_back(pd);
        path_host = e->take_pinned(jobs_b + path_lens_b + total * sizeof(uint64_t));
        PathJob* pj = (PathJob*)path_host->p;
        for (uint32_t k = 0; k < m; ++k) pj[k] = PathJob{done[k], snap[done[k]].best + 1, offs[k]};
        hipStream_t s = e->persist_stream;
        hip_check(hipMemcpyAsync(pd, pj, m * sizeof(PathJob), hipMemcpyHostToDevice, s), "upload path jobs");
        uint32_t* d_lens = (uint32_t*)(pd + jobs_b);
        path_kernel<<<m, kPathThreads, 0, s>>>(d_descs, (const PathJob*)pd, m, (uint64_t*)(pd + jobs_b + path_lens_b), d_lens);
        hip_check(hipGetLastError(), "path launch");
        hip_check(hipMemcpyAsync((char*)path_host->p + jobs_b, d_lens, path_lens_b + total * sizeof(uint64_t),
                                 hipMemcpyDeviceToHost, s),
                  "read back paths");
        ev = e->take_event();
        hip_check(hipEventRecord(ev, s), "record");
        stage = PATHS;
    }

    void paths(std::vector<Job*>& finished) {
        const uint32_t m = (uint32_t)done.size();
        const char* base = (const char*)path_host->p + round_up(m * sizeof(PathJob), 256);
        const uint32_t* lens = (const uint32_t*)base;
        const uint64_t* out = (const uint64_t*)(base + path_lens_b);
        for (uint32_t k = 0; k < m; ++k) {
            Job& j = *wave[done[k]];
            const InstState& st = snap[done[k]];
            e->fill_counters(j, st);
            j.out->cost = st.best;
            const uint32_t len = lens[k];
            if (len == kNone || len == 0) {
                e->finish(j, GA_E_OVERFLOW, t0);
            } else {
                j.out->path = (uint64_t*)malloc(sizeof(uint64_t) * len);
                if (!j.out->path) throw std::bad_alloc();
                for (uint32_t i = 0; i < len; ++i) j.out->path[i] = out[offs[k] + (len - 1 - i)];
                j.out->path_len = len;
                e->finish(j, GA_OK, t0);
            }
            finished.push_back(&j);
        }
        drop();
    }
};

// The run loop shared by plain, batched and served GA* and by the online
// server.  Jobs are added in arrival order; each step() admits what has
// arrived and fits (retries first), retires finished path walks, and lets
// every lane handle its chunk read-backs and launch its next chunk.  With
// `timed` arrivals a job is admitted once its arrival_ns (relative to the
// loop's start) has passed; otherwise at once.  Throws GaError when the run
// fails (the caller then calls discard() and ga_engine::abandon()).
class RunLoop {
   public:
    // requests an admission pass looks past while their lanes wait
    static constexpr uint32_t kAdmitScan = 512;

    RunLoop(ga_engine* e, bool timed, uint64_t t0 = now_ns()) : e(e), timed(timed), t0(t0) {}
    RunLoop(const RunLoop&) = delete;
    RunLoop& operator=(const RunLoop&) = delete;

    ga_engine* const e;
    const bool timed;
    const uint64_t t0;                // the loop's start (steady clock, ns)
    uint32_t inflight = 0, max_inflight = 0;
    std::vector<Job*> finished;       // jobs completed since the owner last cleared this list

    PersistRun* persist = nullptr;    // the one-launch batch form's waves, run beside the lanes

    void add(Job* j) { pending.push_back(j); }

    // One pass; returns false once nothing is pending, running or walking.
    bool step() {
        e->reap_deferred(false);  // memory an in-place growth no longer needs
        // -- the one-launch form: the next stage of its wave, or the next wave
        // (ahead of the lanes' admissions, which then see its memory in use)
        bool waved = false;
        if (persist) {
            std::vector<Job*> again;
            waved = persist->poll(finished, again);
            for (Job* j : again) retry.push_back(j);
        }
        // -- admit: retries first, then arrivals in order.  Each lane admits
        // its requests first come, first served; a request that must wait
        // holds up only the requests of its own lane behind it (a big grid
        // waiting for memory on its lane no longer keeps puzzles off a lane
        // that has run dry), and while it waits the other lanes leave the
        // budget bytes it needs free, so that it is not starved.
        const uint64_t t_admit = now_ns();
        uint64_t now = t_admit - t0;
        {
            const uint32_t all = (1u << e->n_lanes) - 1;
            uint32_t blocked = 0;  // lanes whose first waiting request could not be admitted
            uint64_t reserve = 0;  // bytes those requests need
            uint32_t skipped = 0;  // requests passed over in this scan (bounded)
            for (std::deque<Job*>* q : {&retry, &pending}) {
                for (auto it = q->begin(); it != q->end() && blocked != all && skipped < kAdmitScan;) {
                    Job* j = *it;
                    if (timed && j->arrival_ns > now) break;
                    const uint32_t lane = e->lane_of(*j);
                    if ((blocked >> lane) & 1u) {
                        ++it;
                        ++skipped;
                        continue;
                    }
                    int rc;
                    try {
                        rc = e->admit(*j, reserve);
                    } catch (const GaError& err) {
                        if (err.code == GA_E_HIP) throw;
                        it = q->erase(it);
                        complete(*j, err.code);
                        continue;
                    }
                    if (!rc) {
                        blocked |= 1u << lane;
                        if (!e->lanes[lane].free_slots.empty()) reserve += e->need_of(*j);
                        ++it;
                        ++skipped;
                        continue;
                    }
                    it = q->erase(it);
                    ++inflight;
                }
            }
        }
        max_inflight = std::max(max_inflight, inflight);
        uint64_t t_mark = now_ns();
        e->stats.host_admit_ms += (double)(t_mark - t_admit) * 1e-6;
        // host time since t_mark into `acc`, in ms
        auto lap = [&t_mark](double& acc) {
            const uint64_t t = now_ns();
            acc += (double)(t - t_mark) * 1e-6;
            t_mark = t;
        };
        // -- retire path walks that have completed
        bool progressed = waved;
        while (!batches.empty() && hipEventQuery(batches.front()->ev) == hipSuccess) {
            finalize_batch(batches.front());
            batches.erase(batches.begin());
            progressed = true;
        }
        lap(e->stats.host_path_ms);
        // -- every lane keeps up to two chunks in flight and handles read-backs in order
        bool busy = false;
        // A lone busy lane may block on its chunk instead of polling, but only
        // when nothing else waits for the host: a block lasts a whole chunk, and
        // requests arriving meanwhile would not be admitted (another lane that
        // ran dry would stay idle for that long).
        uint32_t lanes_busy = 0;
        for (const Lane& L : e->lanes) lanes_busy += (L.running > 0 || L.launched > L.processed) ? 1u : 0u;
        const bool may_block = lanes_busy == 1 && pending.empty() && retry.empty() && batches.empty();
        for (uint32_t l = 0; l < e->n_lanes; ++l) {
            Lane& L = e->lanes[l];
            if (L.launched > L.processed) {
                const uint64_t c = L.processed;
                const bool must_wait = L.launched - L.processed >= 2 || L.running == 0;
                hipError_t q = hipEventQuery(L.ev_chunk[c & 1]);
                if (q == hipErrorNotReady && must_wait && may_block) {
                    hip_check(hipEventSynchronize(L.ev_chunk[c & 1]), "chunk event");
                    q = hipSuccess;
                }
                if (q == hipSuccess) {
                    lap(e->stats.host_launch_ms);  // (a blocking wait above counts as launch time)
                    handle_snapshot(l, c);
                    ++L.processed;
                    progressed = true;
                    lap(e->stats.host_snapshot_ms);
                } else if (q != hipErrorNotReady) {
                    hip_check(q, "chunk event");
                }
            }
            if (L.running > 0 && L.launched - L.processed < 2) {
                if (L.map_dirty) e->upload_map(L, l);
                if (L.n_wg) {
                    e->launch_chunk(L);
                    progressed = true;
                }
            }
            lap(e->stats.host_launch_ms);
            if (L.running > 0 || L.launched > L.processed) busy = true;
        }
        if (persist && !persist->idle()) busy = true;
        if (busy) {
            if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
            return true;
        }
        if (!batches.empty()) {
            finalize_batch(batches.front());
            batches.erase(batches.begin());
            lap(e->stats.host_path_ms);
            return true;
        }
        if (pending.empty() && retry.empty()) {
            e->reap_deferred(true);
            return false;
        }
        if (retry.empty() && timed) {
            uint64_t wait = pending.front()->arrival_ns;
            now = now_ns() - t0;
            if (wait > now) std::this_thread::sleep_for(std::chrono::nanoseconds(std::min<uint64_t>(wait - now, 200000)));
        }
        return true;
    }

    // After a failure: waits for the path walks in flight and frees them (the
    // lanes and resident jobs are ga_engine::abandon's).
    void discard() {
        if (persist) persist->drop();
        (void)hipStreamSynchronize(e->path_stream);
        for (PathBatch* b : batches) {
            e->arena->release(b->dev);
            if (b->ev) e->give_event(b->ev);
            if (b->host) e->give_pinned(b->host);
            delete b;
        }
        batches.clear();
    }

   private:
    std::deque<Job*> pending, retry;
    std::vector<PathBatch*> batches;

    void complete(Job& j, int status) {
        e->finish(j, status, t0);
        finished.push_back(&j);
    }
    void finalize_batch(PathBatch* b) {
        hip_check(hipEventSynchronize(b->ev), "path event");
        size_t n = b->jobs.size();
        const uint32_t* lens = (const uint32_t*)b->host->p;
        const uint64_t* out = (const uint64_t*)((char*)b->host->p + round_up(n * sizeof(uint32_t), 16));
        for (size_t i = 0; i < n; ++i) {
            Job& j = *b->jobs[i];
            uint32_t len = lens[i];
            e->fill_counters(j, b->snaps[i]);
            j.out->cost = b->snaps[i].best;
            if (len == kNone || len == 0) {
                complete(j, GA_E_OVERFLOW);
            } else if (is_wide(j.kind)) {  // 24-puzzle: 32-byte boards (4 words) per state
                j.out->path = (uint64_t*)calloc((size_t)len * 4, sizeof(uint64_t));
                const uint64_t* src = out + b->offsets[i];
                for (uint32_t k = 0; k < len; ++k) {
                    Board5 x;
                    std::memcpy(&x, src + 2 * (size_t)(len - 1 - k), sizeof(x));
                    uint8_t* dst = (uint8_t*)(j.out->path + 4 * (size_t)k);
                    for (int c = 0; c < 25; ++c) dst[c] = (uint8_t)Puzzle5Ops::get(x, c);
                }
                j.out->path_len = len;
                complete(j, GA_OK);
            } else {
                j.out->path = (uint64_t*)malloc(sizeof(uint64_t) * len);
                for (uint32_t k = 0; k < len; ++k) j.out->path[k] = out[b->offsets[i] + (len - 1 - k)];
                j.out->path_len = len;
                complete(j, GA_OK);
            }
            e->release(j);
            --inflight;
        }
        e->arena->release(b->dev);
        e->give_event(b->ev);
        e->give_pinned(b->host);
        delete b;
    }

    void handle_snapshot(uint32_t lane, uint64_t chunk) {
        Lane& L = e->lanes[lane];
        e->collect_profile(L, chunk);
        const InstState* snap = (const InstState*)L.h_states[chunk & 1].p;
        std::vector<Job*> done;
        for (uint32_t s = lane; s < L.chunk_reads[chunk & 1]; s += e->n_lanes) {
            Job* j = e->resident[s];
            if (!j || j->finished || j->first_chunk > chunk) continue;
            const InstState& st = snap[s];
            if (st.status == ST_RUNNING) {
                j->last_iterations = st.iterations;
                if (e->cfg.max_iterations && st.iterations > e->cfg.max_iterations) {
                    abort_kernel<<<1, 64, 0, L.stream>>>(e->d_states, (uint32_t)s);
                }
                continue;
            }
            if (st.status == ST_DONE && st.best != kInf) {
                e->retire(*j);  // leaves the launch map; memory kept for the path walk
                done.push_back(j);
                continue;
            }
            if (st.status == ST_PAUSED) {  // grow in place and resume; snapshots already in flight still show the pause
                if (e->grow_in_place(*j, st, L)) {
                    j->first_chunk = L.launched;
                    continue;
                }
                // no room for the larger structures: restart as after an overflow
                InstState ov = st;
                ov.status = (st.grow[(st.paused_iter + 2) % 3] & GROW_TABLE) ? ST_OVERFLOW_TABLE : ST_OVERFLOW_HEAP;
                restart_or_fail(j, ov);
                continue;
            }
            if (st.status == ST_OVERFLOW_HEAP || st.status == ST_OVERFLOW_TABLE) {
                restart_or_fail(j, st);
                continue;
            }
            e->retire(*j);
            // DONE without a goal = search space exhausted; ABORTED = iteration limit
            e->fill_counters(*j, st);
            j->out->cost = 0;
            complete(*j, st.status == ST_DONE ? GA_NO_PATH : GA_E_TIMEOUT);
            e->release(*j);
            --inflight;
        }
        launch_paths(lane, snap, done);
    }

    // A request that overflowed a heap or its closed set (st.status) restarts
    // from scratch with larger structures, or fails after max_retries.
    void restart_or_fail(Job* j, const InstState& st) {
        e->retire(*j);
        e->fill_counters(*j, st);
        j->expanded_before += st.expanded;  // the next attempt starts from nothing
        e->release(*j);
        --inflight;
        if (j->retries >= e->cfg.max_retries) {
            complete(*j, GA_E_OVERFLOW);
            return;
        }
        ++j->retries;
        // 4x the structure that overflowed; 2x the other one if it was more than
        // a quarter full (open entries per heap, stored states), so that a search
        // which outgrows both does not use up its retries one structure at a time
        const ga_engine::Plan p = e->plan(*j);
        const uint32_t open = std::max({st.open_sum[0], st.open_sum[1], st.open_sum[2]});
        const bool heaps_full = (uint64_t)open * 4 > (uint64_t)p.heap_cap * p.n_heaps;
        const bool table_full = p.table_slots && st.occupied * 4 > p.table_slots;
        if (st.status == ST_OVERFLOW_HEAP) {
            j->heap_shift += 2;
            if (table_full) j->table_shift += 1;
        } else {
            j->table_shift += 2;
            if (heaps_full) j->heap_shift += 1;
        }
        // Hashed sets: the heaps together hold at least as many entries as the
        // closed set may store states.  The open list stays below the stored
        // count, so a search that has outgrown its table does not fail again,
        // late, on a heap (a restart repeats all the work done so far).
        for (ga_engine::Plan q = e->plan(*j);
             q.table_slots && q.heap_cap < (1u << 28) &&
             (uint64_t)q.n_heaps * q.heap_cap < q.table_slots * e->cfg.max_load_pct / 100;
             q = e->plan(*j))
            ++j->heap_shift;
        j->finished = false;
        retry.push_back(j);
    }

    // Walks the paths of the instances that finished in a chunk (one batch, on
    // the path stream, once this lane's work so far has finished).
    void launch_paths(uint32_t lane, const InstState* snap, const std::vector<Job*>& done) {
        Lane& L = e->lanes[lane];
        if (done.empty()) return;
        // walk the paths on the path stream once this lane's work so far has finished
        PathBatch* b = new PathBatch();
        batches.push_back(b);  // freed by finalize_batch, or by the clean-up of a failed run
        b->jobs = done;
        uint64_t total = 0;
        for (Job* j : done) {
            b->offsets.push_back(total);
            b->snaps.push_back(snap[j->slot]);
            total += ((uint64_t)snap[j->slot].best + 1) * words_of(j->kind);  // every move costs >= 1
        }
        size_t n = done.size();
        size_t lens_bytes = round_up(n * sizeof(uint32_t), 16);
        size_t jobs_bytes = round_up(n * sizeof(PathJob), 16);
        b->dev = e->arena->alloc(jobs_bytes + lens_bytes + total * sizeof(uint64_t));
        if (!b->dev) throw GaError(GA_E_NOMEM, "device arena exhausted (paths)");
        b->host = e->take_pinned(jobs_bytes + lens_bytes + total * sizeof(uint64_t));
        PathJob* pj = (PathJob*)((char*)b->host->p + lens_bytes + total * sizeof(uint64_t));
        for (size_t i = 0; i < n; ++i)
            pj[i] = PathJob{(uint32_t)done[i]->slot, snap[done[i]->slot].best + 1, b->offsets[i]};
        char* dj = (char*)b->dev;
        uint32_t* dl = (uint32_t*)(dj + jobs_bytes);
        uint64_t* dout = (uint64_t*)(dj + jobs_bytes + lens_bytes);
        hipEvent_t ready = e->take_event();
        hip_check(hipEventRecord(ready, L.stream), "record");
        hip_check(hipStreamWaitEvent(e->path_stream, ready, 0), "wait");
        hip_check(hipMemcpyAsync(dj, pj, n * sizeof(PathJob), hipMemcpyHostToDevice, e->path_stream), "upload jobs");
        path_kernel<<<(uint32_t)n, kPathThreads, 0, e->path_stream>>>(e->d_descs, (const PathJob*)dj, (uint32_t)n, dout, dl);
        hip_check(hipGetLastError(), "path launch");
        hip_check(hipMemcpyAsync(b->host->p, dl, lens_bytes + total * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                 e->path_stream),
                  "read back paths");
        b->ev = e->take_event();
        hip_check(hipEventRecord(b->ev, e->path_stream), "record");
        e->give_event(ready);  // the wait above holds the record it was made against
        b->total = total;
    }
};

void run_job_list(ga_engine* e, std::vector<Job*> order, bool timed, ga_serve_stats* ss,
                  PersistRun* persist = nullptr) {
    RunLoop loop(e, timed);
    loop.persist = persist;
    std::stable_sort(order.begin(), order.end(), [](Job* a, Job* b) { return a->arrival_ns < b->arrival_ns; });
    for (Job* j : order) loop.add(j);
    try {
        while (loop.step()) loop.finished.clear();
    } catch (...) {
        loop.discard();
        throw;
    }
    if (ss) {
        std::vector<double> lat;
        uint64_t last = 0, first = ~0ull;
        ss->completed = ss->failed = 0;
        for (Job* j : order) {
            first = std::min(first, j->arrival_ns);
            if (j->out->status == GA_OK || j->out->status == GA_NO_PATH) {
                ++ss->completed;
                lat.push_back(j->out->latency_ns / 1e6);
                last = std::max(last, j->arrival_ns + j->out->latency_ns);
            } else {
                ++ss->failed;
            }
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            if (lat.empty()) return 0.0;
            size_t k = (size_t)std::min<double>(lat.size() - 1, std::ceil(p * lat.size()) - 1);
            return lat[k];
        };
        ss->wall_ns = last > first ? last - first : 0;
        ss->solves_per_sec = ss->wall_ns ? ss->completed / (ss->wall_ns / 1e9) : 0.0;
        ss->p50_ms = pct(0.50);
        ss->p99_ms = pct(0.99);
        ss->max_inflight = loop.max_inflight;
    }
}

void run_jobs(ga_engine* e, std::vector<Job>& jobs, bool timed, ga_serve_stats* ss) {
    std::vector<Job*> order;
    for (auto& j : jobs) order.push_back(&j);
    run_job_list(e, std::move(order), timed, ss);
}

// ga_solve_batch: with cfg.persist on, the eligible requests run in the
// one-launch form beside the launch lanes, which take the rest.
void run_batch(ga_engine* e, std::vector<Job>& jobs) {
    std::vector<Job*> mine, rest;
    for (Job& j : jobs) ((e->cfg.persist && persist_eligible(e, j)) ? mine : rest).push_back(&j);
    if (mine.empty()) {
        run_job_list(e, std::move(rest), false, nullptr);
        return;
    }
    PersistRun pr(e, std::move(mine), now_ns());
    run_job_list(e, std::move(rest), false, nullptr, &pr);
}

// The checks every request passes before it reaches the device: the result
// starts as GA_PENDING; a malformed request gets GA_E_INVALID and an
// unsolvable puzzle GA_NO_PATH at once (the job is then marked finished).
void prepare_job(Job& j, const ga_request& r, ga_result* out, uint64_t arrival_ns) {
    memset(out, 0, sizeof(ga_result));
    out->status = GA_PENDING;
    out->user_tag = r.user_tag;
    j.req = &r;
    j.out = out;
    j.arrival_ns = arrival_ns;
    const int v = validate(r);
    if (v != GA_OK) {
        out->status = v;
        j.finished = true;
        return;
    }
    j.kind = kind_of(r);
    if (r.kind == GA_GRID) j.w_min = grid_wmin(r);
    const bool solvable = r.kind != GA_PUZZLE || (r.side == 5 ? ga_board_solvable(5, r.cells, r.cells + 25)
                                                              : ga_puzzle_solvable(r.side, r.start, r.goal));
    if (!solvable) {
        out->status = GA_NO_PATH;
        j.finished = true;
    }
}

// Runs the live jobs of one call; throws GaError when the run fails.
using Runner = std::function<void(std::vector<Job>&)>;

// The request checks and the result contract shared by every solve and serve
// entry point: each result is GA_PENDING until its job finishes; a run that
// fails leaves its error code in every result still pending.  Returns GA_OK
// once every request has its final status, or the run's error.
int run_checked(ga_engine* e, const ga_request* reqs, uint32_t n, ga_result* outs, const uint64_t* arrivals,
                const Runner& runner) {
    if ((!reqs && n) || (!outs && n)) return GA_E_INVALID;
    std::vector<Job> jobs(n);
    for (uint32_t i = 0; i < n; ++i) prepare_job(jobs[i], reqs[i], &outs[i], arrivals ? arrivals[i] : 0);
    std::vector<Job> live;
    for (uint32_t i = 0; i < n; ++i)
        if (!jobs[i].finished) live.push_back(jobs[i]);
    int rc = GA_OK;
    try {
        runner(live);
    } catch (const GaError& err) {
        if (e) e->last_error = err.what();
        rc = err.code;
    } catch (const std::bad_alloc&) {
        if (e) e->last_error = "host memory exhausted";
        rc = GA_E_NOMEM;
    }
    if (rc != GA_OK && e && !e->broken) e->abandon(rc);
    for (uint32_t i = 0; i < n; ++i)
        if (outs[i].status == GA_PENDING) {
            if (rc == GA_OK) {  // a job the run loop did not finish: never expected, reported as a failure
                rc = GA_E_HIP;
                if (e) e->last_error = "a request was left unfinished by a completed run";
            }
            outs[i].status = rc;
        }
    return rc;
}

int run_engine(ga_engine* e, const ga_request* reqs, const uint64_t* arrivals, uint32_t n, ga_result* outs,
               ga_serve_stats* ss, bool batch = false) {
    if (!e) return GA_E_INVALID;
    if (e->serving) {  // refused before anything runs: every result says so
        e->last_error = "an online server owns the engine";
        if (!reqs || !outs) return GA_E_INVALID;
        for (uint32_t i = 0; i < n; ++i) {
            Job j;
            prepare_job(j, reqs[i], &outs[i], 0);
            if (outs[i].status == GA_PENDING) outs[i].status = GA_E_INVALID;
        }
        return GA_E_INVALID;
    }
    return run_checked(e, reqs, n, outs, arrivals, [&](std::vector<Job>& live) {
        if (e->broken) throw GaError(GA_E_HIP, "the engine was stopped by a HIP error in an earlier run");
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        if (batch) run_batch(e, live);
        else run_jobs(e, live, arrivals != nullptr, ss);
    });
}

// Percentile p of sorted values (the first value at or above p of them).
double percentile(const std::vector<double>& sorted, double p) {
    if (sorted.empty()) return 0.0;
    const size_t k = (size_t)std::min<double>(sorted.size() - 1, std::ceil(p * sorted.size()) - 1);
    return sorted[k];
}

}  // namespace

// ---------------------------------------------------------------------------
// online server: requests submitted while the engine runs
// ---------------------------------------------------------------------------
// One engine-owned thread drives the RunLoop; submissions reach it through
// a queue.  A request is admitted as soon as a slot and the budget allow,
// into the same launches as the requests already resident.
struct ga_server {
    struct Entry {
        ga_request req{};
        std::vector<uint8_t> data;  // copy of the host grid cells, or of the two side-5 boards
        ga_result res{};
        Job job;
        bool done = false;
    };

    ga_engine* e = nullptr;
    uint64_t t0 = 0;  // the loop's clock origin: arrival times are submission times after it
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int error = GA_OK;
    uint64_t next_ticket = 1;
    std::unordered_map<uint64_t, std::unique_ptr<Entry>> entries;  // by ticket, until polled
    std::deque<Entry*> incoming;  // submitted, not yet given to the loop
    uint32_t outstanding = 0;     // submitted and not finished
    uint32_t completed = 0, failed = 0, max_inflight = 0;
    std::vector<double> lat_ms;  // latencies of the completed requests
    uint64_t first_arrival = ~0ull, last_done = 0;

    // with mu held
    void mark_done(Entry& en) {
        en.done = true;
        --outstanding;
        const int st = en.res.status;
        if (st == GA_OK || st == GA_NO_PATH) {
            ++completed;
            lat_ms.push_back(en.res.latency_ns / 1e6);
            last_done = std::max(last_done, en.job.arrival_ns + en.res.latency_ns);
        } else {
            ++failed;
        }
    }

    // with mu held: the run failed with `code`; every request not finished gets it
    void fail_all(int code) {
        error = code;
        for (auto& kv : entries)
            if (!kv.second->done) {
                kv.second->res.status = code;
                mark_done(*kv.second);
            }
        incoming.clear();
    }

    void run() {
        RunLoop loop(e, false, t0);
        try {
            hip_check(hipSetDevice(e->device), "hipSetDevice");
        } catch (const GaError& err) {
            std::lock_guard<std::mutex> lk(mu);
            e->last_error = err.what();
            fail_all(err.code);
            cv_done.notify_all();
            return;
        }
        for (;;) {
            std::deque<Entry*> fresh;
            {
                std::lock_guard<std::mutex> lk(mu);
                fresh.swap(incoming);
            }
            for (Entry* en : fresh) loop.add(&en->job);
            bool more = false;
            int code = GA_OK;
            try {
                more = loop.step();
            } catch (const GaError& err) {
                e->last_error = err.what();
                code = err.code;
            } catch (const std::bad_alloc&) {
                e->last_error = "host memory exhausted";
                code = GA_E_NOMEM;
            }
            std::unique_lock<std::mutex> lk(mu);
            for (Job* j : loop.finished) mark_done(*static_cast<Entry*>(j->owner));
            loop.finished.clear();
            max_inflight = std::max(max_inflight, loop.max_inflight);
            if (code != GA_OK) {
                lk.unlock();
                loop.discard();
                if (!e->broken) e->abandon(code);
                lk.lock();
                fail_all(code);
                cv_done.notify_all();
                return;
            }
            cv_done.notify_all();
            if (!more) {
                if (stop && incoming.empty()) return;
                cv_work.wait_for(lk, std::chrono::milliseconds(5), [&] { return stop || !incoming.empty(); });
            }
        }
    }
};

namespace {

// Random 64-bit global atomics at agent scope, the closed sets' operations:
// kind 0 a fetch-min (the atomicMin on a slot's g | parent), kind 1 a CAS
// from the empty key (a probe).  Each thread issues PER of them to unrelated
// addresses back to back and keeps every returned value, so they are all in
// flight together: the launch measures the device's rate for them.
template <int PER>
__global__ __launch_bounds__(256) void atomic_rate_kernel(uint64_t* table, uint64_t mask, int kind, uint64_t* sink) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const gptr<uint64_t> t = as_global(table);
    uint64_t got[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const uint64_t h = mix64(tid * PER + (uint64_t)i + 1);
        const gptr<uint64_t> p = t + (h & mask);
        if (kind == 0) {
            got[i] = __hip_atomic_fetch_min(p, h >> 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (kind == 2) {
            got[i] = *p;  // a plain load: the random-line rate of the memory system
        } else {
            uint64_t prev = kEmptyKey;
            __hip_atomic_compare_exchange_strong(p, &prev, h >> 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            got[i] = prev;
        }
    }
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) acc ^= got[i];
    if (acc == 0x5A5A5A5A5A5A5A5Aull) *sink = acc;  // keeps the returned values live
}

int atomic_rate(ga_engine* e, uint64_t table_bytes, uint64_t ops, int kind, double* ops_per_sec) {
    constexpr int kPer = 16;
    constexpr uint32_t kThreads = 256;
    const uint64_t words = pow2_floor(table_bytes / sizeof(uint64_t));
    const uint64_t blocks = (ops + (uint64_t)kPer * kThreads - 1) / ((uint64_t)kPer * kThreads);
    if (blocks > (1u << 30)) return GA_E_INVALID;
    uint64_t* table = nullptr;
    uint64_t* sink = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = GA_OK;
    try {
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        if (hipMalloc(&table, words * sizeof(uint64_t)) != hipSuccess || hipMalloc(&sink, sizeof(uint64_t)) != hipSuccess)
            throw GaError(GA_E_NOMEM, "atomic table");
        hip_check(hipEventCreate(&t0), "hipEventCreate");
        hip_check(hipEventCreate(&t1), "hipEventCreate");
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the first pass maps the table's pages; the second is timed
            fill_kernel<uint64_t><<<1024, kThreads>>>(table, words, kEmptyKey);
            hip_check(hipEventRecord(t0, nullptr), "record");
            atomic_rate_kernel<kPer><<<(uint32_t)blocks, kThreads>>>(table, words - 1, kind, sink);
            hip_check(hipEventRecord(t1, nullptr), "record");
            hip_check(hipEventSynchronize(t1), "sync");
            hip_check(hipGetLastError(), "atomic_rate_kernel");
            hip_check(hipEventElapsedTime(&ms, t0, t1), "elapsed");
        }
        *ops_per_sec = ms > 0.f ? (double)(blocks * kThreads * kPer) / (ms * 1e-3) : 0.0;
    } catch (const GaError& err) {
        e->last_error = err.what();
        rc = err.code;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (table) (void)hipFree(table);
    if (sink) (void)hipFree(sink);
    return rc;
}

// ---------------------------------------------------------------------------
// plain GA* of one request with no device allocation (ga_solve_noalloc)
// ---------------------------------------------------------------------------
// Bump allocation inside this device's scratch, 256-byte aligned.
struct ScratchCarve {
    char* base;
    uint64_t off = 0;
    void* operator()(uint64_t bytes) {
        off = round_up(off, 256);
        if (off > kScratchBytes || bytes > kScratchBytes - off) return nullptr;
        void* p = base + off;
        off += bytes;
        return p;
    }
};

// The search of ga_solve on one request, with every device buffer in the
// scratch array (nothing allocated: no hipMalloc/hipFree, no event, stream or
// pinned buffer created) on lane 0's stream, one chunk at a time.  Retries
// after a capacity overflow as ga_solve does, as long as the larger
// structures fit the scratch.  Throws GaError.
void solve_noalloc(ga_engine* e, const ga_request& r, ga_result* out, uint64_t* path_buf, uint32_t path_cap) {
    Job j;
    j.req = &r;
    j.out = out;
    j.kind = kind_of(r);
    if (r.kind == GA_GRID) j.w_min = grid_wmin(r);
    hip_check(hipSetDevice(e->device), "hipSetDevice");
    hipStream_t st = e->lanes[0].stream;
    // the pinned read-back buffer of lane 0, allocated when the engine was created
    InstState* snap = (InstState*)e->lanes[0].h_states[0].p;
    const uint32_t W = e->cfg.heap_arity, R = e->cfg.chunk_iters;
    const uint32_t variant = is_wide(j.kind) ? kWideVariant
                             : j.kind == KIND_GRID8 ? kAllKinds
                                                    : kind_bit((uint32_t)j.kind);
    const uint64_t t0 = now_ns();
    for (;;) {
        const ga_engine::Plan p = e->plan(j);
        ScratchCarve take{e->scratch_dev};
        InstDesc* d_desc = (InstDesc*)take(sizeof(InstDesc));
        InstState* d_state = (InstState*)take(sizeof(InstState));
        uint32_t* d_iter = (uint32_t*)take(sizeof(uint32_t));
        const uint32_t heaps_per_wg = e->cfg.block / W;
        const uint32_t n_wg = (p.n_heaps + heaps_per_wg - 1) / heaps_per_wg;
        WgEntry* d_map = (WgEntry*)take((uint64_t)n_wg * sizeof(WgEntry));
        InstDesc d{};
        if (!d_map || !e->carve_instance(j, p, take, d))
            throw GaError(GA_E_BUDGET, "request needs more device memory than the " +
                                           std::to_string(kScratchBytes >> 20) + " MiB no-allocation scratch");
        // the pattern databases when this engine has built them (building them would allocate)
        if (e->cfg.pdb24 && e->d_pdb && ga_engine::ordered_goal5(r)) d.pdb = e->d_pdb;
        if (e->cfg.pdb15 && e->d_pdb15 && ga_engine::ordered_goal4(r)) d.pdb = e->d_pdb15;
        e->init_instance(j, p, d, st);
        fill_kernel<uint32_t><<<1, 64, 0, st>>>(d_iter, 1, 0u);
        map_kernel<<<(n_wg + 255) / 256, 256, 0, st>>>(d_map, n_wg, heaps_per_wg, p.n_heaps);
        d.start_iter = 0;
        seed_kernel<<<1, 64, 0, st>>>(d, d_desc, d_state, 0u, (int)W);
        hip_check(hipGetLastError(), "seed launch");
        const StepParams prm{e->cfg.pops_per_iter, 1u << 16, 0u, e->cfg.rounds, e->cfg.grow_thresh, e->cfg.step_flags,
                             {0, 0}};
        for (;;) {
            for (uint32_t k = 0; k < R; ++k)
                launch_step(W, e->cfg.block, variant, st, n_wg, d_map, d_desc, d_state, d_iter, k, prm);
            advance_kernel<<<1, 64, 0, st>>>(d_iter, R);
            hip_check(hipGetLastError(), "step launch");
            hip_check(hipMemcpyAsync(snap, d_state, sizeof(InstState), hipMemcpyDeviceToHost, st), "read back state");
            hip_check(hipStreamSynchronize(st), "chunk");
            if (snap->status != ST_RUNNING) break;
            if (e->cfg.max_iterations && snap->iterations > e->cfg.max_iterations) {
                abort_kernel<<<1, 64, 0, st>>>(d_state, 0u);
                hip_check(hipGetLastError(), "abort launch");
            }
        }
        const InstState s = *snap;
        out->expanded = s.expanded;
        out->stored = (j.kind == KIND_GRID4 || j.kind == KIND_GRID8) ? 0 : s.occupied;
        out->iterations = s.iterations;
        out->retries = j.retries;
        if (s.status == ST_OVERFLOW_HEAP || s.status == ST_OVERFLOW_TABLE) {
            if (j.retries >= e->cfg.max_retries) throw GaError(GA_E_OVERFLOW, "capacity exceeded after all retries");
            ++j.retries;
            if (s.status == ST_OVERFLOW_HEAP) j.heap_shift += 2;
            else j.table_shift += 2;
            continue;  // the larger structures must still fit the scratch (else GA_E_BUDGET)
        }
        out->search_ns = now_ns() - t0;
        if (s.status == ST_ABORTED) throw GaError(GA_E_TIMEOUT, "iteration limit reached");
        if (s.status != ST_DONE) throw GaError(GA_E_HIP, "unexpected instance status");
        if (s.best == kInf) {
            out->status = GA_NO_PATH;
            return;
        }
        out->cost = s.best;
        // walk the parent links into the scratch, then order (and unpack) the states
        const uint32_t cap = s.best + 1;  // every move costs >= 1
        const uint32_t words = words_of(j.kind), out_words = is_wide(j.kind) ? 4u : 1u;
        PathJob* d_job = (PathJob*)take(sizeof(PathJob));
        uint32_t* d_len = (uint32_t*)take(sizeof(uint32_t));
        uint64_t* d_walk = (uint64_t*)take((uint64_t)cap * words * sizeof(uint64_t));
        uint64_t* d_path = (uint64_t*)take((uint64_t)cap * out_words * sizeof(uint64_t));
        if (!d_job || !d_len || !d_walk || !d_path)
            throw GaError(GA_E_BUDGET, "the path does not fit the no-allocation scratch");
        const PathJob job{0u, cap, 0u};
        hip_check(hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, st), "upload path job");
        path_kernel<<<1, kPathThreads, 0, st>>>(d_desc, d_job, 1u, d_walk, d_len);
        path_order_kernel<<<std::max<uint32_t>(1, std::min<uint32_t>(1024, (cap + 255) / 256)), 256, 0, st>>>(
            d_walk, d_len, d_path, is_wide(j.kind) ? 1u : 0u);
        hip_check(hipGetLastError(), "path launch");
        uint32_t* h_len = reinterpret_cast<uint32_t*>(snap);  // the pinned buffer again
        hip_check(hipMemcpyAsync(h_len, d_len, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "read back length");
        hip_check(hipStreamSynchronize(st), "path");
        const uint32_t len = *h_len;
        if (len == kNone || len == 0) throw GaError(GA_E_OVERFLOW, "path reconstruction failed");
        out->path_len = len;
        if (len > path_cap) {  // the caller's buffer is too small: path_len says how many states are needed
            out->status = GA_E_INVALID;
            e->last_error = "path buffer too small: " + std::to_string(len) + " states needed";
            return;
        }
        hip_check(hipMemcpyAsync(path_buf, d_path, (uint64_t)len * out_words * sizeof(uint64_t),
                                 hipMemcpyDeviceToHost, st),
                  "read back path");
        hip_check(hipStreamSynchronize(st), "path copy");
        out->status = GA_OK;
        return;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void ga_config_default(ga_config* cfg) {
    memset(cfg, 0, sizeof(*cfg));
    cfg->budget_bytes = 64ull << 30;
    cfg->chunk_bytes = 4ull << 30;
    cfg->max_instances = 4096;
    cfg->chunk_iters = 8;
    cfg->use_graph = 0;
    cfg->max_retries = 6;
    cfg->max_iterations = 0;
    cfg->heap_arity = 4;
    cfg->max_load_pct = 50;
    cfg->pops_per_iter = 2;
    cfg->profile = 0;
    cfg->rounds = 1;
    cfg->adaptive_start = 0;
    cfg->grow_thresh = 32;
    cfg->rounds_grid = 0;
    cfg->lanes = 1;
    cfg->deep_cells = 4ull << 20;
    cfg->slack_puzzle = 0;
    cfg->slack_grid = 0;
    cfg->pdb24 = 1;
    cfg->pdb15 = 1;
    cfg->grow_in_place = 1;
    cfg->block = kStepBlock;
    cfg->persist = 0;
    cfg->persist_block = 1024;
    cfg->persist_cells = 0;
}

int ga_engine_create(int device, const ga_config* cfg, ga_engine** out) {
    if (!out) return GA_E_INVALID;
    *out = nullptr;
    ga_engine* e = new (std::nothrow) ga_engine();
    if (!e) return GA_E_NOMEM;
    try {
        e->init(device, cfg);
    } catch (const GaError& err) {
        delete e;
        return err.code;
    }
    *out = e;
    return GA_OK;
}

void ga_engine_destroy(ga_engine* e) { delete e; }

int ga_engine_get_stats(ga_engine* e, ga_engine_stats* out) {
    if (!e || !out || e->serving) return GA_E_INVALID;
    *out = e->stats;
    out->algo_bytes = e->algo_total();
    for (uint32_t l = 0; l < e->n_lanes; ++l) out->lane_algo_bytes[l] = e->algo_lane(l);
    out->peak_bytes = e->arena->peak();
    return GA_OK;
}

void ga_engine_reset_stats(ga_engine* e) {
    if (!e) return;
    e->stats = ga_engine_stats{};
    for (uint32_t l = 0; l <= GA_MAX_LANES; ++l) e->algo_sampled[l] = e->algo_always[l] = 0;
}

int ga_solve(ga_engine* e, const ga_request* req, ga_result* out) {
    if (!e || !req || !out) return GA_E_INVALID;
    (void)run_engine(e, req, nullptr, 1, out, nullptr);  // sets out->status in every case
    return out->status;
}

int ga_solve_puzzle(ga_engine* e, uint32_t side, uint64_t start, uint64_t goal, uint32_t num_queues,
                    ga_result* out) {
    ga_request r{};
    r.kind = GA_PUZZLE;
    r.side = side;
    r.start = start;
    r.goal = goal;
    r.num_queues = num_queues;
    return ga_solve(e, &r, out);
}

int ga_solve_grid(ga_engine* e, const uint8_t* cells, uint32_t width, uint32_t height, uint64_t start,
                  uint64_t goal, uint32_t connectivity, uint32_t num_queues, ga_result* out) {
    ga_request r{};
    r.kind = GA_GRID;
    r.cells = cells;
    r.width = width;
    r.height = height;
    r.start = start;
    r.goal = goal;
    r.connectivity = connectivity;
    r.num_queues = num_queues;
    return ga_solve(e, &r, out);
}

int ga_solve_batch(ga_engine* e, const ga_request* reqs, uint32_t n, ga_result* outs) {
    return run_engine(e, reqs, nullptr, n, outs, nullptr, true);
}

int ga_serve_run(ga_engine* e, const ga_request* reqs, const uint64_t* arrival_ns, uint32_t n, ga_result* outs,
                 ga_serve_stats* stats) {
    if (!arrival_ns) return GA_E_INVALID;
    ga_serve_stats local{};
    int rc = run_engine(e, reqs, arrival_ns, n, outs, stats ? stats : &local);
    return rc;
}

void ga_result_free(ga_result* r) {
    if (!r) return;
    free(r->path);
    r->path = nullptr;
    r->path_len = 0;
}

const char* ga_status_str(int status) {
    switch (status) {
        case GA_OK: return "GA_OK";
        case GA_NO_PATH: return "GA_NO_PATH";
        case GA_PENDING: return "GA_PENDING";
        case GA_E_INVALID: return "GA_E_INVALID";
        case GA_E_NOMEM: return "GA_E_NOMEM";
        case GA_E_BUDGET: return "GA_E_BUDGET";
        case GA_E_HIP: return "GA_E_HIP";
        case GA_E_TIMEOUT: return "GA_E_TIMEOUT";
        case GA_E_OVERFLOW: return "GA_E_OVERFLOW";
        default: return "GA_E_UNKNOWN";
    }
}

int ga_solve_noalloc(ga_engine* e, const ga_request* req, ga_result* out, uint64_t* path_buf, uint32_t path_cap) {
    if (!e || !req || !out || (!path_buf && path_cap)) return GA_E_INVALID;
    memset(out, 0, sizeof(*out));
    out->status = GA_PENDING;
    out->user_tag = req->user_tag;
    int v = validate(*req);
    if (v != GA_OK) return out->status = v;
    const bool solvable = req->kind != GA_PUZZLE ||
                          (req->side == 5 ? ga_board_solvable(5, req->cells, req->cells + 25)
                                          : ga_puzzle_solvable(req->side, req->start, req->goal));
    if (!solvable) return out->status = GA_NO_PATH;
    if (e->broken) {
        e->last_error = "the engine was stopped by a HIP error in an earlier run";
        return out->status = GA_E_HIP;
    }
    if (e->serving) {
        e->last_error = "an online server owns the engine";
        return out->status = GA_E_INVALID;
    }
    try {
        solve_noalloc(e, *req, out, path_buf, path_cap);
    } catch (const GaError& err) {
        e->last_error = err.what();
        out->status = err.code;
        // nothing of this call may stay in flight on the scratch
        if (hipStreamSynchronize(e->lanes[0].stream) != hipSuccess || err.code == GA_E_HIP) e->broken = true;
        (void)hipGetLastError();
    }
    return out->status;
}

int ga_serve_start(ga_engine* e, ga_server** out) {
    if (!e || !out) return GA_E_INVALID;
    *out = nullptr;
    if (e->broken) return GA_E_HIP;
    bool expected = false;
    if (!e->serving.compare_exchange_strong(expected, true)) {
        e->last_error = "an online server already owns the engine";
        return GA_E_INVALID;
    }
    ga_server* sv = new (std::nothrow) ga_server();
    if (!sv) {
        e->serving = false;
        return GA_E_NOMEM;
    }
    sv->e = e;
    sv->t0 = now_ns();
    try {
        sv->th = std::thread([sv] { sv->run(); });
    } catch (const std::system_error&) {
        delete sv;
        e->serving = false;
        return GA_E_NOMEM;
    }
    *out = sv;
    return GA_OK;
}

int ga_serve_submit(ga_server* sv, const ga_request* req, uint64_t* ticket) {
    if (!sv || !req || !ticket) return GA_E_INVALID;
    std::unique_ptr<ga_server::Entry> en(new (std::nothrow) ga_server::Entry());
    if (!en) return GA_E_NOMEM;
    en->req = *req;
    if (validate(*req) == GA_OK) {  // copy the host data the request points to
        const ga_request& r = *req;
        size_t bytes = 0;
        if (r.kind == GA_GRID && r.cells && !r.cells_device) bytes = (size_t)r.width * r.height;
        if (r.kind == GA_PUZZLE && r.side == 5) bytes = 50;
        if (bytes) {
            try {
                en->data.assign(r.cells, r.cells + bytes);
            } catch (const std::bad_alloc&) {
                return GA_E_NOMEM;
            }
            en->req.cells = en->data.data();
        }
    }
    const uint64_t t = now_ns();
    prepare_job(en->job, en->req, &en->res, t > sv->t0 ? t - sv->t0 : 0);  // host checks only
    en->job.owner = en.get();
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->error != GA_OK) return sv->error;
    if (sv->stop) return GA_E_INVALID;
    sv->first_arrival = std::min(sv->first_arrival, en->job.arrival_ns);
    ++sv->outstanding;
    ga_server::Entry* raw = en.get();
    *ticket = sv->next_ticket++;
    sv->entries.emplace(*ticket, std::move(en));
    if (raw->job.finished) {  // failed its checks: final at once
        sv->mark_done(*raw);
        sv->cv_done.notify_all();
    } else {
        sv->incoming.push_back(raw);
        sv->cv_work.notify_one();
    }
    return GA_OK;
}

int ga_serve_poll(ga_server* sv, uint64_t ticket, ga_result* out) {
    if (!sv || !out) return GA_E_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->entries.find(ticket);
    if (it == sv->entries.end()) return GA_E_INVALID;
    if (!it->second->done) return GA_PENDING;
    *out = it->second->res;  // the path now belongs to the caller (ga_result_free)
    sv->entries.erase(it);
    return GA_OK;
}

int ga_serve_drain(ga_server* sv) {
    if (!sv) return GA_E_INVALID;
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->outstanding == 0; });
    return sv->error;
}

int ga_serve_status(ga_server* sv, ga_serve_stats* out) {
    if (!sv || !out) return GA_E_INVALID;
    std::lock_guard<std::mutex> lk(sv->mu);
    std::vector<double> lat = sv->lat_ms;
    std::sort(lat.begin(), lat.end());
    memset(out, 0, sizeof(*out));
    out->completed = sv->completed;
    out->failed = sv->failed;
    out->wall_ns = sv->last_done > sv->first_arrival && sv->first_arrival != ~0ull ? sv->last_done - sv->first_arrival : 0;
    out->solves_per_sec = out->wall_ns ? sv->completed / (out->wall_ns / 1e9) : 0.0;
    out->p50_ms = percentile(lat, 0.50);
    out->p99_ms = percentile(lat, 0.99);
    out->max_inflight = sv->max_inflight;
    return GA_OK;
}

int ga_serve_stop(ga_server* sv) {
    if (!sv) return GA_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->th.joinable()) sv->th.join();
    const int rc = sv->error;
    for (auto& kv : sv->entries) ga_result_free(&kv.second->res);  // results never polled
    sv->e->serving = false;
    delete sv;
    return rc;
}

int ga_pdb15_tables(ga_engine* e, uint8_t* out, size_t bytes) {
    if (!e || !out || bytes < kPdb15Bytes || e->serving) return GA_E_INVALID;
    if (e->broken) return GA_E_HIP;
    try {
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        const uint8_t* p = e->ensure_pdb15(e->lanes[0].stream);
        hip_check(hipMemcpy(out, p, kPdb15Bytes, hipMemcpyDeviceToHost), "copy pattern databases");
    } catch (const GaError& err) {
        e->last_error = err.what();
        return err.code;
    }
    return GA_OK;
}

int ga_pdb24_tables(ga_engine* e, uint8_t* out, size_t bytes) {
    if (!e || !out || bytes < kPdbBytes || e->serving) return GA_E_INVALID;
    if (e->broken) return GA_E_HIP;
    try {
        hip_check(hipSetDevice(e->device), "hipSetDevice");
        const uint8_t* p = e->ensure_pdb(e->lanes.back().stream);
        hip_check(hipMemcpy(out, p, kPdbBytes, hipMemcpyDeviceToHost), "copy pattern databases");
    } catch (const GaError& err) {
        e->last_error = err.what();
        return err.code;
    }
    return GA_OK;
}

const char* ga_engine_last_error(ga_engine* e) { return e ? e->last_error.c_str() : ""; }

int ga_debug_fail_run(const ga_request* reqs, uint32_t n, ga_result* outs, uint32_t finish_first, int fail_code) {
    if (fail_code >= 0) return GA_E_INVALID;
    return run_checked(nullptr, reqs, n, outs, nullptr, [&](std::vector<Job>& live) {
        for (uint32_t i = 0; i < live.size() && i < finish_first; ++i) {
            live[i].out->status = GA_OK;
            live[i].finished = true;
        }
        throw GaError(fail_code, "injected failure");
    });
}

int ga_device_upload(ga_engine* e, const void* host, size_t bytes, void** device_out) {
    if (!e || !host || !device_out) return GA_E_INVALID;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return GA_E_NOMEM;
    if (hipMemcpy(p, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return GA_E_HIP;
    }
    *device_out = p;
    return GA_OK;
}

int ga_device_free(ga_engine* e, void* device) {
    if (!e) return GA_E_INVALID;
    return hipFree(device) == hipSuccess ? GA_OK : GA_E_HIP;
}

int ga_atomic_rate(ga_engine* e, uint64_t table_bytes, uint64_t ops, int kind, double* ops_per_sec) {
    if (!e || !ops_per_sec || kind < 0 || kind > 2 || table_bytes < sizeof(uint64_t) || ops == 0)
        return GA_E_INVALID;
    return atomic_rate(e, table_bytes, ops, kind, ops_per_sec);
}

}  // extern "C"
