// Kernel-level checks of the engine's device data structures: the closed-set
// update (find or claim, then atomicMin of g << 32 | parent) on each of the
// hashed layouts, and the W-ary heaps with their LDS top levels, run on caller
// data through the same device functions a search step uses
// (closed_set_update, push_records, heap_pop, HeapRef in gastar/core.hpp).
// The reference's counterparts are its mass insert into the bucketed hashmap
// (fancy_hashmap.h:32-81,155-167) and its heap pop/push (heaps.h:132-177); it
// has no such entry points.  Test support only: every call allocates, runs,
// copies back, frees and synchronises.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <vector>

#include "gastar.h"
#include "gastar/core.hpp"

using namespace gastar;

namespace {

struct State12 {
    uint32_t w[3];
};
using Wide12Set = WideKeySet<State12>;

// key i of the caller's word array, in the layout's key type
template <class Set>
struct KeyWords;
template <>
struct KeyWords<HashedU64Set> {
    static constexpr uint32_t kWords = 2;
    __device__ static uint64_t read(const uint32_t* k) { return (uint64_t)k[0] | (uint64_t)k[1] << 32; }
};
template <>
struct KeyWords<Board5KeySet> {
    static constexpr uint32_t kWords = 4;
    __device__ static Board5KeySet::Key read(const uint32_t* k) {
        Board5KeySet::Key key;
        for (int i = 0; i < 4; ++i) key.w[i] = k[i];
        return key;
    }
};
template <>
struct KeyWords<Wide12Set> {
    static constexpr uint32_t kWords = 3;
    __device__ static Wide12Set::Key read(const uint32_t* k) {
        Wide12Set::Key key;
        for (int i = 0; i < 3; ++i) key.w[i] = k[i];
        return key;
    }
};

template <class Set>
__device__ void update_one(const InstDesc& d, const uint32_t* keys, const uint32_t* g, const uint32_t* parent,
                           uint32_t i, uint32_t max_probes, int32_t* outcome, uint32_t* slot_out) {
    const auto key = KeyWords<Set>::read(keys + (size_t)i * KeyWords<Set>::kWords);
    uint32_t slot = kNone, probes = 0;
    bool isnew = false;
    const int cs = closed_set_update<Set>(d, key, g[i], parent[i], max_probes, slot, isnew, probes);
    outcome[i] = cs;
    slot_out[i] = cs == CS_FULL ? kNone : slot;
}

template <class Set>
__global__ void closed_set_kernel(InstDesc d, const uint32_t* keys, const uint32_t* g, const uint32_t* parent,
                                  uint32_t n, uint32_t max_probes, int serial, int32_t* outcome, uint32_t* slot_out) {
    if (serial) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (uint32_t i = 0; i < n; ++i) update_one<Set>(d, keys, g, parent, i, max_probes, outcome, slot_out);
        return;
    }
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) update_one<Set>(d, keys, g, parent, i, max_probes, outcome, slot_out);
}

// Pushes the collected batch (if any) as one set of candidate records, counting
// a refused batch.
template <int W, int KW>
__device__ __forceinline__ void flush_batch(const HeapRef<W>& H, uint32_t& S, HeapEntry (&rec)[KW], uint32_t& nb,
                                            uint32_t& refused, int lane, uint32_t cap, const Tally& tl) {
    if (nb) {
        bool over = false;
        push_records<W, KW>(H, S, rec, KW * W, lane, kInf, cap, over, tl, false);
        refused += over ? 1u : 0u;
#pragma unroll
        for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
        nb = 0;
    }
    group_fence();
}

// One wave; heap h = threadIdx.x / W runs its operations [begin, end).
template <int W>
__global__ void __launch_bounds__(64) heap_kernel(HeapEntry* heaps, uint32_t* sizes, uint32_t stride, uint32_t cap,
                                                  uint32_t n_heaps, const uint64_t* ops, uint32_t n_ops,
                                                  uint32_t begin, uint32_t end, int track, uint64_t* popped,
                                                  uint32_t* popped_tag, uint32_t* overflow) {
    constexpr int KW = 2;
    constexpr int K = KW * W;  // records per batch
    constexpr uint32_t kTop = HeapRef<W>::kTop;
    __shared__ HeapEntry top[(64 / W) * kTop];
    __shared__ uint32_t dirty[64 / W];
    const int lane = (int)(threadIdx.x & (W - 1));
    const uint32_t h = threadIdx.x / W;
    const bool mine = h < n_heaps;
    const Tally tl{nullptr, false};  // counts nothing: push_records only adds profile counts
    const HeapRef<W> H{as_global(heaps) + (size_t)(mine ? h : 0u) * stride + (W - 1), as_local(top) + (size_t)h * kTop,
                       as_local(dirty) + h, track != 0};
    uint32_t S = mine ? sizes[h] : 0u;
    if (mine) {
        if (track && lane == 0) *H.dirty = 0u;
        H.fill(S, lane);
    }
    group_fence();
    HeapEntry rec[KW];
#pragma unroll
    for (int r = 0; r < KW; ++r) rec[r].prio = kNoRecord;
    uint32_t nb = 0, refused = 0;  // record positions in the batch; batches refused
    auto flush = [&]() { flush_batch<W, KW>(H, S, rec, nb, refused, lane, cap, tl); };
    for (uint32_t i = begin; i < end; ++i) {
        const uint64_t op = mine ? ops[(size_t)h * n_ops + i] : GA_HOP_NOP;
        if (op == GA_HOP_NOP) continue;
        if (op == GA_HOP_FLUSH) {
            flush();
        } else if (op == GA_HOP_POP) {
            flush();
            uint64_t got = ~0ull;
            uint32_t tag = kNone;
            if (S > 0) {
                HeapEntry t;
                uint64_t key = 0;
                uint32_t g = 0, lv = 0, hl = 0;
                auto load = [](const HeapEntry&, uint64_t& k, uint32_t& gs) {
                    k = 0;
                    gs = 0;
                };
                auto pre = [](const HeapEntry&) {};
                heap_pop<W>(H, S, lane, lv, hl, load, pre, key, g, kInf, t);
                got = t.prio;
                tag = t.aux == h ? t.slot : 0xFFFFFFFEu;
            }
            group_fence();
            if (lane == 0) {
                popped[(size_t)h * n_ops + i] = got;
                popped_tag[(size_t)h * n_ops + i] = tag;
            }
        } else {
            if (nb == (uint32_t)K) flush();
            HeapEntry x;
            x.prio = op == GA_HOP_GAP ? kNoRecord : op;
            x.slot = i;
            x.aux = h;
#pragma unroll
            for (int r = 0; r < KW; ++r)
                if (nb / W == (uint32_t)r && nb % W == (uint32_t)lane) rec[r] = x;
            ++nb;
        }
    }
    flush();
    if (mine) {
        H.write_back(S, lane);
        if (lane == 0) {
            sizes[h] = S;
            overflow[h] += refused;
        }
    }
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t n) { return hipMalloc(&p, n * sizeof(T) > 0 ? n * sizeof(T) : 1) == hipSuccess; }
};

int hip_rc(hipError_t e) { return e == hipSuccess ? GA_OK : GA_E_HIP; }

template <class Set>
int run_closed_set(const uint32_t* keys, const uint32_t* g, const uint32_t* parent, uint32_t n, uint32_t log2_slots,
                   uint32_t max_probes, int serial, int32_t* outcome, uint32_t* slot, uint64_t* tag_out,
                   uint32_t* key_out, uint64_t* val_out, size_t slot_bytes) {
    constexpr uint32_t kw = KeyWords<Set>::kWords;
    const size_t slots = (size_t)1 << log2_slots;
    DevBuf<uint8_t> table, side;
    DevBuf<uint32_t> dkeys, dg, dpar, dslot;
    DevBuf<int32_t> dout;
    if (!table.alloc(slots * slot_bytes) || !dkeys.alloc((size_t)n * kw) || !dg.alloc(n) || !dpar.alloc(n) ||
        !dslot.alloc(n) || !dout.alloc(n))
        return GA_E_NOMEM;
    const bool wide = std::is_same<Set, Wide12Set>::value;
    constexpr size_t side_bytes = sizeof(Wide12Set::Cells);  // per slot: each key word in a marked 64-bit cell
    if (wide && !side.alloc(slots * side_bytes)) return GA_E_NOMEM;
    int rc = hip_rc(hipMemset(table.p, 0xFF, slots * slot_bytes));
    if (rc == GA_OK && wide) rc = hip_rc(hipMemset(side.p, 0xFF, slots * side_bytes));
    if (rc == GA_OK && n) rc = hip_rc(hipMemcpy(dkeys.p, keys, (size_t)n * kw * 4, hipMemcpyHostToDevice));
    if (rc == GA_OK && n) rc = hip_rc(hipMemcpy(dg.p, g, (size_t)n * 4, hipMemcpyHostToDevice));
    if (rc == GA_OK && n) rc = hip_rc(hipMemcpy(dpar.p, parent, (size_t)n * 4, hipMemcpyHostToDevice));
    if (rc != GA_OK) return rc;
    InstDesc d;
    std::memset(&d, 0, sizeof d);
    d.table = reinterpret_cast<Slot*>(table.p);
    d.table_mask = slots - 1;
    d.keys = side.p;
    if (n) {
        const uint32_t blocks = serial ? 1u : (n + 255) / 256;
        closed_set_kernel<Set><<<blocks, serial ? 1 : 256>>>(d, dkeys.p, dg.p, dpar.p, n, max_probes, serial, dout.p,
                                                            dslot.p);
        rc = hip_rc(hipGetLastError());
        if (rc == GA_OK) rc = hip_rc(hipDeviceSynchronize());
        if (rc != GA_OK) return rc;
        rc = hip_rc(hipMemcpy(outcome, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        if (rc == GA_OK) rc = hip_rc(hipMemcpy(slot, dslot.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        if (rc != GA_OK) return rc;
    }
    if (!tag_out && !key_out && !val_out) return GA_OK;
    std::vector<uint8_t> host(slots * slot_bytes);
    std::vector<Wide12Set::Cells> host_side(wide ? slots : 0);
    rc = hip_rc(hipMemcpy(host.data(), table.p, host.size(), hipMemcpyDeviceToHost));
    if (rc == GA_OK && wide)
        rc = hip_rc(hipMemcpy(host_side.data(), side.p, slots * side_bytes, hipMemcpyDeviceToHost));
    if (rc != GA_OK) return rc;
    for (size_t s = 0; s < slots; ++s) {
        const uint8_t* b = host.data() + s * slot_bytes;
        uint64_t tag, val;
        uint32_t words[4];
        std::memcpy(&tag, b, 8);
        if (std::is_same<Set, Board5KeySet>::value) {
            std::memcpy(words, b, 16);  // lo, hi
            std::memcpy(&val, b + 16, 8);
        } else {
            std::memcpy(&val, b + 8, 8);
            if (wide)
                for (int w = 0; w < 3; ++w) words[w] = (uint32_t)host_side[s].c[w];
            else std::memcpy(words, b, 8);
        }
        if (tag_out) tag_out[s] = tag;
        if (val_out) val_out[s] = val;
        if (key_out) std::memcpy(key_out + s * kw, words, kw * 4);
    }
    return GA_OK;
}

// Heap operation rate over the whole device.  Every W-lane group owns one heap
// of `fill` entries in global memory, with its top entries in LDS for the
// launch as a search step keeps them.  Each of `iters` rounds pops two entries
// and pushes two successors of the first (f raised by 1..8, g by 1) as one
// batch, so every heap keeps its size and each push and pop walks a heap of
// `fill` entries.  Heap h's logical entry j lies at h * stride + W - 1 + j.
template <int W>
__global__ void __launch_bounds__(256) heap_rate_kernel(HeapEntry* heaps, uint32_t stride, uint32_t fill,
                                                        uint32_t iters, uint64_t* sink) {
    constexpr int G = 256 / W;  // heaps per workgroup
    constexpr uint32_t kTop = HeapRef<W>::kTop;
    __shared__ HeapEntry top[G * kTop];
    __shared__ uint32_t dirty[G];
    const int lane = (int)(threadIdx.x & (W - 1));
    const uint32_t hl = threadIdx.x / W;
    const uint32_t h = blockIdx.x * G + hl;
    const Tally tl{nullptr, false};  // counts nothing: push_records only adds profile counts
    const HeapRef<W> H{as_global(heaps) + (size_t)h * stride + (W - 1), as_local(top) + (size_t)hl * kTop,
                       as_local(dirty) + hl, false};
    uint32_t S = fill;
    H.fill(S, lane);
    group_fence();
    auto load = [](const HeapEntry&, uint64_t& k, uint32_t& gs) {
        k = 0;
        gs = 0;
    };
    auto pre = [](const HeapEntry&) {};
    uint64_t acc = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        HeapEntry a, b;
        uint64_t key = 0;
        uint32_t g = 0, lv = 0, hlv = 0;
        heap_pop<W>(H, S, lane, lv, hlv, load, pre, key, g, kInf, a);
        group_fence();
        heap_pop<W>(H, S, lane, lv, hlv, load, pre, key, g, kInf, b);
        group_fence();
        acc += b.prio;
        HeapEntry rec[1];
        rec[0].prio = kNoRecord;
        rec[0].slot = it;
        rec[0].aux = h;
        if (lane < 2) {
            const uint32_t step = 1u + (uint32_t)((mix64(((uint64_t)h << 32) | it) >> (8 * lane)) & 7u);
            rec[0].prio = make_prio(prio_f(a.prio) + step, prio_g(a.prio) + 1u);
        }
        bool over = false;
        push_records<W, 1>(H, S, rec, W, lane, kInf, fill + W, over, tl, false);
        group_fence();
    }
    H.write_back(S, lane);
    if (acc == 0x5A5A5A5A5A5A5A5Aull) *sink = acc;  // keeps the pops live
}

// entries of every heap: logical j < fill holds f = j + 1, g = 0 (ascending,
// so a valid heap of any arity); the rest of the heap's storage is empty
template <int W>
__global__ void heap_rate_init_kernel(HeapEntry* heaps, uint64_t n, uint32_t stride, uint32_t fill) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = (uint32_t)(i % stride);
        HeapEntry x;
        x.prio = (e >= W - 1 && e - (W - 1) < fill) ? make_prio(e - (W - 1) + 1u, 0u) : kNoRecord;
        x.slot = 0;
        x.aux = 0;
        heaps[i] = x;
    }
}

template <int W>
int run_heap_rate(uint32_t waves, uint32_t fill, uint32_t iters, double* pops_per_sec) {
    const uint32_t blocks = waves / 4;  // 256-thread workgroups
    const uint32_t n_heaps = blocks * (256 / W);
    const uint32_t stride = fill + 2 * W;  // capacity fill + W, and the W - 1 entries before logical entry 0
    const uint64_t n = (uint64_t)n_heaps * stride;
    DevBuf<HeapEntry> heaps;
    DevBuf<uint64_t> sink;
    if (!heaps.alloc(n) || !sink.alloc(1)) return GA_E_NOMEM;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = hip_rc(hipEventCreate(&t0));
    if (rc == GA_OK) rc = hip_rc(hipEventCreate(&t1));
    float ms = 0.f;
    for (int rep = 0; rep < 2 && rc == GA_OK; ++rep) {  // the first pass maps the pages; the second is timed
        heap_rate_init_kernel<W><<<1024, 256>>>(heaps.p, n, stride, fill);
        rc = hip_rc(hipEventRecord(t0, nullptr));
        if (rc == GA_OK) {
            heap_rate_kernel<W><<<blocks, 256>>>(heaps.p, stride, fill, iters, sink.p);
            rc = hip_rc(hipGetLastError());
        }
        if (rc == GA_OK) rc = hip_rc(hipEventRecord(t1, nullptr));
        if (rc == GA_OK) rc = hip_rc(hipEventSynchronize(t1));
        if (rc == GA_OK) rc = hip_rc(hipEventElapsedTime(&ms, t0, t1));
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (rc == GA_OK) *pops_per_sec = ms > 0.f ? 2.0 * n_heaps * (double)iters / (ms * 1e-3) : 0.0;
    return rc;
}

template <int W>
int run_heap(uint32_t n_heaps, const uint64_t* ops, uint32_t n_ops, uint32_t opl, uint32_t cap, int track,
             uint64_t* popped, uint32_t* popped_tag, uint64_t* heap_out, uint32_t* size_out, uint32_t* overflow_out) {
    const uint32_t stride = cap + W;  // logical entry j at physical j + W - 1
    const size_t total = (size_t)n_heaps * n_ops;
    DevBuf<HeapEntry> heaps;
    DevBuf<uint32_t> sizes, over, tags;
    DevBuf<uint64_t> dops, dpop;
    if (!heaps.alloc((size_t)n_heaps * stride) || !sizes.alloc(n_heaps) || !over.alloc(n_heaps) ||
        !tags.alloc(total) || !dops.alloc(total) || !dpop.alloc(total))
        return GA_E_NOMEM;
    int rc = hip_rc(hipMemset(heaps.p, 0, (size_t)n_heaps * stride * sizeof(HeapEntry)));
    if (rc == GA_OK) rc = hip_rc(hipMemset(sizes.p, 0, n_heaps * 4));
    if (rc == GA_OK) rc = hip_rc(hipMemset(over.p, 0, n_heaps * 4));
    if (rc == GA_OK) rc = hip_rc(hipMemset(dpop.p, 0xFF, total * 8));
    if (rc == GA_OK) rc = hip_rc(hipMemset(tags.p, 0xFF, total * 4));
    if (rc == GA_OK && total) rc = hip_rc(hipMemcpy(dops.p, ops, total * 8, hipMemcpyHostToDevice));
    if (rc != GA_OK) return rc;
    for (uint32_t b = 0; b < n_ops; b += opl) {
        const uint32_t e = n_ops - b < opl ? n_ops : b + opl;
        heap_kernel<W><<<1, 64>>>(heaps.p, sizes.p, stride, cap, n_heaps, dops.p, n_ops, b, e, track, dpop.p, tags.p,
                                  over.p);
        rc = hip_rc(hipGetLastError());
        if (rc != GA_OK) return rc;
    }
    rc = hip_rc(hipDeviceSynchronize());
    if (rc != GA_OK) return rc;
    std::vector<HeapEntry> host((size_t)n_heaps * stride);
    rc = hip_rc(hipMemcpy(host.data(), heaps.p, host.size() * sizeof(HeapEntry), hipMemcpyDeviceToHost));
    if (rc == GA_OK) rc = hip_rc(hipMemcpy(size_out, sizes.p, n_heaps * 4, hipMemcpyDeviceToHost));
    if (rc == GA_OK) rc = hip_rc(hipMemcpy(overflow_out, over.p, n_heaps * 4, hipMemcpyDeviceToHost));
    if (rc == GA_OK && total) rc = hip_rc(hipMemcpy(popped, dpop.p, total * 8, hipMemcpyDeviceToHost));
    if (rc == GA_OK && total) rc = hip_rc(hipMemcpy(popped_tag, tags.p, total * 4, hipMemcpyDeviceToHost));
    if (rc != GA_OK) return rc;
    for (uint32_t h = 0; h < n_heaps; ++h)
        for (uint32_t j = 0; j < cap; ++j)
            heap_out[(size_t)h * cap + j] = j < size_out[h] ? host[(size_t)h * stride + W - 1 + j].prio : ~0ull;
    return GA_OK;
}

}  // namespace

extern "C" {

int ga_debug_closed_set(int device, int layout, const uint32_t* keys, const uint32_t* g, const uint32_t* parent,
                        uint32_t n, uint32_t log2_slots, uint32_t max_probes, int serial, int32_t* outcome,
                        uint32_t* slot, uint64_t* tag_out, uint32_t* key_out, uint64_t* val_out) {
    if (layout < GA_SET_U64 || layout > GA_SET_WIDE12 || log2_slots < 1 || log2_slots > 24 || max_probes == 0 ||
        n > (1u << 24) || (serial != 0 && serial != 1) || (n && (!keys || !g || !parent || !outcome || !slot)))
        return GA_E_INVALID;
    // keys the layouts cannot hold: the empty key (64-bit keys), and boards
    // whose lo is all ones or whose hi has its top bit set (the layout's
    // empty and being-written marks; no 24-puzzle board has either)
    for (uint32_t i = 0; i < n; ++i) {
        if (layout == GA_SET_U64 && keys[2 * i] == ~0u && keys[2 * i + 1] == ~0u) return GA_E_INVALID;
        if (layout == GA_SET_BOARD && ((keys[4 * i] == ~0u && keys[4 * i + 1] == ~0u) || (keys[4 * i + 3] >> 31)))
            return GA_E_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) return GA_E_INVALID;
    switch (layout) {
        case GA_SET_U64:
            return run_closed_set<HashedU64Set>(keys, g, parent, n, log2_slots, max_probes, serial, outcome, slot,
                                                tag_out, key_out, val_out, sizeof(Slot));
        case GA_SET_BOARD:
            return run_closed_set<Board5KeySet>(keys, g, parent, n, log2_slots, max_probes, serial, outcome, slot,
                                                tag_out, key_out, val_out, sizeof(Board5KeySet::PairSlot));
        default:
            return run_closed_set<Wide12Set>(keys, g, parent, n, log2_slots, max_probes, serial, outcome, slot,
                                             tag_out, key_out, val_out, sizeof(Slot));
    }
}

int ga_debug_heap(int device, uint32_t arity, uint32_t n_heaps, const uint64_t* ops, uint32_t n_ops,
                  uint32_t ops_per_launch, uint32_t cap, int track_dirty, uint64_t* popped, uint32_t* popped_tag,
                  uint64_t* heap_out, uint32_t* size_out, uint32_t* overflow_out) {
    // the dirty mask has one bit per LDS top entry: 21 at arity 4, 25 at 8, 49 at 16.  Arity 64 (one
    // heap per wave) is refused: its first run here ran past a 120 s limit with no cause shown.
    if ((arity != 4 && arity != 8 && arity != 16) || n_heaps == 0 || n_heaps > 64 / arity ||
        cap == 0 || cap > (1u << 20) || n_ops > (1u << 20) || ops_per_launch == 0 ||
        (track_dirty != 0 && track_dirty != 1) || (track_dirty && arity > 8) || !heap_out ||
        !size_out || !overflow_out || (n_ops && (!ops || !popped || !popped_tag)))
        return GA_E_INVALID;
    for (size_t i = 0; i < (size_t)n_heaps * n_ops; ++i) {
        const uint64_t v = ops[i];
        if (v < 0xFFFFFFFF00000000ull || v == GA_HOP_POP || v == GA_HOP_FLUSH || v == GA_HOP_NOP || v == GA_HOP_GAP)
            continue;
        return GA_E_INVALID;  // a push whose f is the infinite cost
    }
    if (hipSetDevice(device) != hipSuccess) return GA_E_INVALID;
    switch (arity) {
        case 4:
            return run_heap<4>(n_heaps, ops, n_ops, ops_per_launch, cap, track_dirty, popped, popped_tag, heap_out,
                               size_out, overflow_out);
        case 8:
            return run_heap<8>(n_heaps, ops, n_ops, ops_per_launch, cap, track_dirty, popped, popped_tag, heap_out,
                               size_out, overflow_out);
        default:
            return run_heap<16>(n_heaps, ops, n_ops, ops_per_launch, cap, track_dirty, popped, popped_tag, heap_out,
                                size_out, overflow_out);
    }
}

int ga_heap_rate(int device, uint32_t arity, uint32_t waves, uint32_t fill, uint32_t iters, double* pops_per_sec) {
    if ((arity != 4 && arity != 8 && arity != 16) || waves == 0 || waves % 4 || waves > (1u << 16) ||
        fill < 2 || fill > (1u << 16) || iters == 0 || iters > (1u << 16) || !pops_per_sec)
        return GA_E_INVALID;
    // at most 16 GiB of heaps
    if ((uint64_t)(waves / 4) * (256 / arity) * (fill + 2 * arity) * sizeof(HeapEntry) > (16ull << 30))
        return GA_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return GA_E_INVALID;
    switch (arity) {
        case 4: return run_heap_rate<4>(waves, fill, iters, pops_per_sec);
        case 8: return run_heap_rate<8>(waves, fill, iters, pops_per_sec);
        default: return run_heap_rate<16>(waves, fill, iters, pops_per_sec);
    }
}

}  // extern "C"
