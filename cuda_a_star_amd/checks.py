"""Kernel-level checks of the device data structures through the C ABI.

`closed_set` runs the search step's closed-set update (find or claim the
state's slot, then atomicMin of g << 32 | parent; `closed_set_update` in
include/gastar/core.hpp) on caller keys, in each hashed layout, and returns
every outcome with the table after the last insert.  `heap` runs sequences of
batched pushes and pops on W-ary heaps whose top levels live in LDS for a
launch (`push_records`, `heap_pop`, `HeapRef`), and returns what every pop
returned and the heaps at the end; `heap_rate` times those operations over
the whole device at a given arity (4, 8 or 16).  The reference's counterparts are the mass
insert of fancy_hashmap.h:32-81,155-167 and the heap operations of
heaps.h:132-177.  Both need a GPU; their argument checks do not.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import GAStarError, lib

SET_U64, SET_BOARD, SET_WIDE12 = 0, 1, 2
KEY_WORDS = {SET_U64: 2, SET_BOARD: 4, SET_WIDE12: 3}
CS_REJECTED, CS_INSERTED, CS_UPDATED, CS_FULL = 0, 1, 2, 3
HOP_POP = 0xFFFFFFFFFFFFFFFF
HOP_FLUSH = 0xFFFFFFFFFFFFFFFE
HOP_NOP = 0xFFFFFFFFFFFFFFFD
HOP_GAP = 0xFFFFFFFFFFFFFFFC
NONE = 0xFFFFFFFF


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


def closed_set(layout: int, keys, g, parent, log2_slots: int, max_probes: int = 1 << 16, serial: bool = False,
               device: int = 0) -> dict:
    """keys: (n, KEY_WORDS[layout]) uint32 words per state; g, parent: n values.
    Returns outcome, slot (NONE for FULL) and the table: tag, key, val per slot."""
    kw = KEY_WORDS[layout]
    keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, kw)
    n = keys.shape[0]
    g = np.ascontiguousarray(g, dtype=np.uint32).reshape(n)
    parent = np.ascontiguousarray(parent, dtype=np.uint32).reshape(n)
    slots = 1 << log2_slots
    outcome = np.empty(n, dtype=np.int32)
    slot = np.empty(n, dtype=np.uint32)
    tag = np.empty(slots, dtype=np.uint64)
    key = np.empty((slots, kw), dtype=np.uint32)
    val = np.empty(slots, dtype=np.uint64)
    rc = lib().ga_debug_closed_set(device, layout, _ptr(keys), _ptr(g), _ptr(parent), n, log2_slots, max_probes,
                                   1 if serial else 0, _ptr(outcome), _ptr(slot), _ptr(tag), _ptr(key), _ptr(val))
    if rc != 0:
        raise GAStarError(rc, "ga_debug_closed_set")
    return {"outcome": outcome, "slot": slot, "tag": tag, "key": key, "val": val}


def heap(arity: int, ops, ops_per_launch: int, cap: int, track_dirty: bool = False, device: int = 0) -> dict:
    """ops: (n_heaps, n_ops) uint64, one row per heap (HOP_* codes or push priorities).
    Returns popped / popped_tag (n_heaps, n_ops), size, overflow and heap (list of the logical entries)."""
    ops = np.ascontiguousarray(ops, dtype=np.uint64)
    if ops.ndim == 1:
        ops = ops.reshape(1, -1)
    n_heaps, n_ops = ops.shape
    popped = np.empty((n_heaps, n_ops), dtype=np.uint64)
    popped_tag = np.empty((n_heaps, n_ops), dtype=np.uint32)
    heap_out = np.empty((n_heaps, cap), dtype=np.uint64)
    size = np.empty(n_heaps, dtype=np.uint32)
    overflow = np.empty(n_heaps, dtype=np.uint32)
    rc = lib().ga_debug_heap(device, arity, n_heaps, _ptr(ops), n_ops, ops_per_launch, cap, 1 if track_dirty else 0,
                             _ptr(popped), _ptr(popped_tag), _ptr(heap_out), _ptr(size), _ptr(overflow))
    if rc != 0:
        raise GAStarError(rc, "ga_debug_heap")
    return {"popped": popped, "popped_tag": popped_tag, "size": size, "overflow": overflow,
            "heap": [heap_out[h, :size[h]].copy() for h in range(n_heaps)]}


def heap_rate(arity: int, waves: int = 8192, fill: int = 4096, iters: int = 1024, device: int = 0) -> float:
    """Heap pops per second over the whole device (pushes run at the same rate): `waves`
    64-lane waves, each W-lane group popping two entries and pushing two from a heap of
    `fill` entries, `iters` times (ga_heap_rate)."""
    out = ctypes.c_double()
    rc = lib().ga_heap_rate(device, arity, waves, fill, iters, ctypes.byref(out))
    if rc != 0:
        raise GAStarError(rc, "ga_heap_rate")
    return out.value


if __name__ == "__main__":
    # python -m cuda_a_star_amd.checks: the heap pop rate per arity, the device full of waves
    import argparse

    ap = argparse.ArgumentParser(description="whole-device heap operation rate per arity (ga_heap_rate)")
    ap.add_argument("--waves", type=int, default=8192)
    ap.add_argument("--iters", type=int, default=256)
    ap.add_argument("--fill", type=int, nargs="+", default=[256, 4096])
    args = ap.parse_args()
    for fill in args.fill:
        for arity in (4, 8, 16):
            rate = heap_rate(arity, waves=args.waves, fill=fill, iters=args.iters)
            print(f"heap entries {fill} arity {arity}: {args.waves * 64 // arity} heaps, "
                  f"{rate / 1e9:.2f} G pops/s", flush=True)
