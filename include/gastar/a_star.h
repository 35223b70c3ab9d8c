// a_star.h -- header drop-in for the reference's GA* template.
//
//   template <typename T, unsigned char max_neighbours,
//             NeighbourGenerator<T> neighbour_generator, Heuristic<T> heuristic,
//             uint16_t bucket_depth = 256>
//   std::pair<std::vector<T>, std::chrono::nanoseconds>
//   a_star(const T& start, const T& target, const uint32_t num_queues);
//
// Same signature and return value as the reference (a_star.h:223-224,281):
// the path start..target (empty if the target is unreachable) and the time of
// the search loop (a_star.h:245,266).  The problem functors are __device__
// functions passed as template arguments, exactly as in
// a_star_sliding_puzzle.h:170 and a_star_pathfinding.h:233, and are inlined
// into the engine's step kernel (gastar/core.hpp), which this header
// instantiates in the caller's translation unit.
//
// Requirements on T (as in the reference: helpers.h:10,98,113): trivially
// copyable, sizeof(T) % 4 == 0, deterministic padding bytes; move costs >= 1.
// States of at most 8 bytes are keyed by a 64-bit CAS (an 8-byte state must
// not have all bits set); wider states use a slot tag protocol (see WideKeySet).
// bucket_depth is accepted for source compatibility and unused (the closed set
// is open-addressed, not bucketed).  Errors throw std::runtime_error, as
// cuda_check does (helpers.h:46-50).  num_queues is rounded down to a power of
// two.  Not re-entrant across threads (neither is the reference).
#ifndef GASTAR_A_STAR_H
#define GASTAR_A_STAR_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gastar/core.hpp"
#include "gastar/functional_types.h"

namespace gastar {
namespace dropin {

// States of at most 8 bytes: packed into a 64-bit key (zero-extended).
template <typename T>
struct PackedKeySet : HashedU64Set {
    __host__ __device__ static inline uint64_t pack(const T& t) {
        uint64_t k = 0;
        __builtin_memcpy(&k, &t, sizeof(T));
        return k;
    }
    __host__ __device__ static inline T unpack(uint64_t k) {
        return *reinterpret_cast<const T*>(&k);
    }
};

template <typename T>
using KeySetFor = typename std::conditional<(sizeof(T) <= 8), PackedKeySet<T>, WideKeySet<T>>::type;

// Problem policy over the reference's functor interface.
template <typename T, int K_, NeighbourGenerator<T> gen, Heuristic<T> heur>
struct UserOps : KeySetFor<T> {
    using Set = KeySetFor<T>;
    using Key = typename Set::Key;
    static constexpr int K = K_;
    __device__ static inline uint32_t heuristic(const InstDesc&, const Key& key, uint32_t) {
        return heur(Set::unpack(key));
    }
    __device__ static inline uint32_t aux_of(const InstDesc&, const Key&) { return 0; }
    __device__ static inline uint32_t prefetch(const InstDesc&, const HeapEntry&, int) { return 0; }
    __device__ static inline bool succ(const InstDesc&, const Key& key, uint32_t, uint32_t, int i, uint32_t, Key& t,
                                       uint32_t& taux, uint32_t& cost, uint32_t& th) {
        const T src = Set::unpack(key);
        T dst = src;
        bool exists = false;
        uint32_t c = 0;
        gen(src, dst, exists, c, (unsigned char)i);
        if (!exists) return false;
        t = Set::pack(dst);
        taux = 0;
        cost = c;
        th = heur(dst);
        return true;
    }
};

constexpr int kW = 4;  // lanes per heap group
constexpr int kBlock = kStepBlock;

template <class Ops>
__global__ __launch_bounds__(kBlock) void user_step_kernel(const WgEntry* __restrict__ wgmap,
                                                           const InstDesc* __restrict__ descs, InstState* states,
                                                           const uint32_t* __restrict__ iter_base, uint32_t r,
                                                           StepParams prm) {
    __shared__ StepShared sh;
    uint32_t inst, it, first, cover;
    if (!step_prologue(wgmap, descs, iter_base, r, inst, it, first, cover)) return;
    const uint32_t status = load_relaxed(&states[inst].status);
    step_begin(sh);
    const InstDesc& d = descs[inst];
    const uint32_t q = first + threadIdx.x / kW;
    run_group<kW, Ops>(d, states + inst, status, it, q, q < d.n_heaps, cover, prm,
                       Tally{sh.tally, prm.profile != 0}, sh.work, sh.heap_top, sh.dirty);
    tally_end(sh.tally, states + inst, it);
}

template <class Ops>
__global__ void user_seed_kernel(InstDesc d, InstDesc* descs, InstState* states) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    descs[0] = d;
    seed_instance<Ops>(d, states, kW);
}

template <class Ops>
__global__ void user_path_kernel(const InstDesc* descs, typename Ops::Key* out, uint32_t cap, uint32_t* len) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *len = walk_path<Ops>(descs[0], out, cap);
}

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorName(e));
}

// Device buffers of one search; freed on scope exit.
struct Buffers {
    std::vector<void*> ptrs;
    void* get(size_t bytes) {
        void* p = nullptr;
        check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
        ptrs.push_back(p);
        return p;
    }
    ~Buffers() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

template <class Ops, typename T>
std::pair<std::vector<T>, std::chrono::nanoseconds> solve(const T& start, const T& target, uint32_t num_queues) {
    using Key = typename Ops::Key;
    uint32_t n = 1;
    while (n * 2 <= std::max<uint32_t>(num_queues, 1)) n *= 2;
    uint32_t heap_cap = 1024;
    uint64_t slots = 1ull << 20;
    const uint32_t R = 16;  // launches between checks of the search's status
    for (int attempt = 0; attempt < 8; ++attempt) {
        Buffers b;
        InstDesc d{};
        d.kind = KIND_USER;
        d.n_heaps = n;
        d.a0 = n;  // every queue in use from the start, as in the reference
        d.f_slack = kInf;          // no gating of heap pops
        d.heap_cap = heap_cap;
        d.heap_stride = (heap_cap + kW - 1 + kW - 1) / kW * kW;
        d.heaps = (HeapEntry*)b.get(sizeof(HeapEntry) * (size_t)n * d.heap_stride);
        d.heap_size = (uint32_t*)b.get(sizeof(uint32_t) * n);
        d.cand = (HeapEntry*)b.get(sizeof(HeapEntry) * 2ull * n * Ops::K);
        d.table = (Slot*)b.get(sizeof(Slot) * slots);
        d.table_mask = slots - 1;
        d.load_limit = slots / 2;
        if constexpr (sizeof(T) > 8) {  // the side array of key words, unwritten (all ones)
            const size_t side = sizeof(typename WideKeySet<T>::Cells) * slots;
            d.keys = b.get(side);
            check(hipMemset(d.keys, 0xFF, side), "hipMemset");
        }
        const RawKey<T> ends[2] = {to_raw(start), to_raw(target)};
        void* user = b.get(sizeof(ends));
        check(hipMemcpy(user, ends, sizeof(ends), hipMemcpyHostToDevice), "hipMemcpy");
        d.user = user;
        if constexpr (sizeof(T) <= 8) {
            d.start = Ops::pack(start);
            d.goal = Ops::pack(target);
        }
        InstDesc* descs = (InstDesc*)b.get(sizeof(InstDesc));
        InstState* st = (InstState*)b.get(sizeof(InstState));
        uint32_t* iter = (uint32_t*)b.get(sizeof(uint32_t));
        const uint32_t wgs = (n + kBlock / kW - 1) / (kBlock / kW);
        WgEntry* map = (WgEntry*)b.get(sizeof(WgEntry) * wgs);
        std::vector<WgEntry> hmap(wgs);
        for (uint32_t w = 0; w < wgs; ++w) hmap[w] = WgEntry{0, w * (kBlock / kW), n, 0};
        check(hipMemcpy(map, hmap.data(), sizeof(WgEntry) * wgs, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemset(iter, 0, sizeof(uint32_t)), "hipMemset");
        check(hipMemset(d.heap_size, 0, sizeof(uint32_t) * n), "hipMemset");
        check(hipMemset(d.table, 0xFF, sizeof(Slot) * slots), "hipMemset");
        user_seed_kernel<Ops><<<1, 64>>>(d, descs, st);
        check(hipGetLastError(), "seed");
        check(hipDeviceSynchronize(), "seed");
        StepParams prm{2, 1u << 16, 0, 1, 32, 0, {0, 0}};
        InstState hs{};
        auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            for (uint32_t r = 0; r < R; ++r)
                user_step_kernel<Ops><<<wgs, kBlock>>>(map, descs, st, iter, r, prm);
            advance_kernel<<<1, 64>>>(iter, R);
            check(hipGetLastError(), "step");
            check(hipMemcpy(&hs, st, sizeof(hs), hipMemcpyDeviceToHost), "hipMemcpy");
            if (hs.status != ST_RUNNING) break;
        }
        auto t1 = std::chrono::steady_clock::now();
        if (hs.status == ST_OVERFLOW_HEAP) {
            heap_cap *= 4;
            continue;
        }
        if (hs.status == ST_OVERFLOW_TABLE) {
            slots *= 4;
            heap_cap *= 2;
            continue;
        }
        std::vector<T> path;
        if (hs.status == ST_DONE && hs.best != kInf) {
            const uint32_t cap = hs.best + 1;  // every move costs >= 1
            Key* out = (Key*)b.get(sizeof(Key) * cap);
            uint32_t* len = (uint32_t*)b.get(sizeof(uint32_t));
            user_path_kernel<Ops><<<1, 64>>>(descs, out, cap, len);
            check(hipGetLastError(), "path");
            uint32_t hl = 0;
            check(hipMemcpy(&hl, len, sizeof(hl), hipMemcpyDeviceToHost), "hipMemcpy");
            if (hl == kNone) throw std::runtime_error("GA*: path reconstruction failed");
            std::vector<Key> keys(hl);
            check(hipMemcpy(keys.data(), out, sizeof(Key) * hl, hipMemcpyDeviceToHost), "hipMemcpy");
            for (uint32_t k = hl; k-- > 0;) path.push_back(Ops::unpack(keys[k]));
        }
        return {std::move(path), std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0)};
    }
    throw std::runtime_error("GA*: capacity exceeded");
}

}  // namespace dropin
}  // namespace gastar

template <typename T, unsigned char max_neighbours, NeighbourGenerator<T> neighbour_generator,
          Heuristic<T> heuristic, uint16_t bucket_depth = 256>
std::pair<std::vector<T>, std::chrono::nanoseconds> a_star(const T& start, const T& target,
                                                           const uint32_t num_queues) {
    static_assert(sizeof(T) % 4 == 0, "state size must be a multiple of 4 bytes (as in the reference)");
    (void)bucket_depth;
    using Ops = gastar::dropin::UserOps<T, max_neighbours, neighbour_generator, heuristic>;
    return gastar::dropin::solve<Ops, T>(start, target, num_queues);
}

#endif
