/*
 * gastar.h -- C ABI of the MI355X GA* engine (libgastar.so).
 *
 * The reference (michalborkowski96/cuda_a_star) exposes its search as a C++
 * template, a_star<T, max_neighbours, gen, h, bucket_depth>(start, target,
 * num_queues) -> {path, elapsed} (a_star.h:223-224), called by
 * SlidingPuzzle::solve (a_star_sliding_puzzle.h:165-171) and
 * Pathfinding::solve (a_star_pathfinding.h:223-238).  That template is kept
 * as a header drop-in in gastar/a_star.h.  This header is the thin C ABI over
 * the built-in problem instantiations: plain C types, no exceptions, integer
 * status codes.  Every entry point below names the reference interface it
 * replaces.
 *
 * States:
 *   sliding puzzle: uint64, nibble i (bits 4i..4i+3) = tile on cell i in
 *                   row-major order, 0 = blank; side 3 or 4.
 *   24-puzzle:      (side 5, the reference CLI's default, main.cu:72) a board
 *                   of 25 bytes, tile on each cell row-major, 0 = blank.  The
 *                   request carries start and goal boards in `cells` (50
 *                   bytes: start, then goal; `start`/`goal` are ignored), and
 *                   each path state is returned as 32 bytes (4 words): the 25
 *                   board bytes, then 7 zero bytes.
 *   grid:           cell index y*width + x into a row-major uint8 grid,
 *                   0 = wall, 1..255 = cost of entering the cell.
 * Paths are returned start..goal as states in these encodings.
 */
#ifndef GASTAR_H
#define GASTAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws std::runtime_error instead,
 *      helpers.h:46-50; "no path" is its empty vector, a_star.h:268-274) ---- */
#define GA_OK 0
#define GA_NO_PATH 1
#define GA_PENDING 2       /* not finished yet: every result holds it while a run is in progress,
                              and no result holds it once the run has returned */
#define GA_E_INVALID (-1)  /* malformed request */
#define GA_E_NOMEM (-2)    /* device memory exhausted */
#define GA_E_BUDGET (-3)   /* request can never fit the admission budget */
#define GA_E_HIP (-4)      /* HIP runtime error */
#define GA_E_TIMEOUT (-5)  /* iteration limit reached */
#define GA_E_OVERFLOW (-6) /* capacity exceeded after all retries */

#define GA_MAX_LANES 4 /* ga_config.lanes at most */

/* ---- problem kinds ---- */
#define GA_PUZZLE 1 /* sliding puzzle: a_star_sliding_puzzle.h */
#define GA_GRID 2   /* weighted grid:  a_star_pathfinding.h    */

typedef struct ga_engine ga_engine;

/* One search request (the arguments of a_star<>() plus the problem data that
 * the reference keeps in __constant__ globals, a_star_sliding_puzzle.h:125-142
 * and a_star_pathfinding.h:72-81). */
typedef struct ga_request {
    uint32_t kind;          /* GA_PUZZLE or GA_GRID */
    uint32_t side;          /* puzzle side, 3, 4 or 5 */
    uint64_t start;         /* start state */
    uint64_t goal;          /* target state */
    uint32_t width;         /* grid width  (x extent) */
    uint32_t height;        /* grid height (y extent) */
    uint32_t connectivity;  /* grid: 4 or 8 */
    uint32_t w_min;         /* grid heuristic scale; 0 = minimum open-cell weight */
    const uint8_t* cells;   /* grid cells in host memory, or NULL; side-5 puzzle: start and goal boards */
    const uint8_t* cells_device; /* grid cells already resident on the device, or NULL */
    uint32_t num_queues;    /* number of parallel heaps (a_star.h:224); 0 = planner's choice */
    uint32_t heap_capacity; /* entries per heap; 0 = planner's choice */
    uint64_t table_slots;   /* closed-set slots (puzzles); 0 = planner's choice */
    uint64_t user_tag;      /* copied into the result */
} ga_request;

/* Result of one request (the reference returns {path, elapsed ns},
 * a_star.h:224,281; the cost is not returned there and is added here). */
typedef struct ga_result {
    int32_t status;        /* GA_OK, GA_NO_PATH or an error */
    uint32_t cost;         /* optimal path cost */
    uint64_t expanded;     /* nodes expanded on the device, over every attempt (restarts included) */
    uint64_t stored;       /* distinct states inserted in the closed set */
    uint32_t iterations;   /* GA* iterations */
    uint32_t retries;      /* restarts with larger capacities (the search repeated from the start) */
    uint32_t path_len;     /* states in path */
    uint32_t grows;        /* in-place growths of the heaps or closed set (the search carried over) */
    uint64_t* path;        /* start..goal (side 5: 4 words per state); owned by the library, see ga_result_free */
    uint64_t search_ns;    /* admission to completion */
    uint64_t latency_ns;   /* arrival to result on the host (serving) */
    uint64_t user_tag;
} ga_result;

/* Engine configuration.  An invalid value (heap_arity other than 0, 4 or 8,
 * lanes above 4, block other than 0, 128, 256 or 512) makes ga_engine_create
 * return GA_E_INVALID. */
typedef struct ga_config {
    uint64_t budget_bytes;   /* admission budget for heap + closed-set bytes */
    uint64_t chunk_bytes;    /* device arena grows in chunks of this size */
    uint32_t max_instances;  /* instances resident at once */
    uint32_t chunk_iters;    /* GA* iterations launched between host checks */
    uint32_t use_graph;      /* replay each chunk from a captured hipGraph (turned off when the
                                process runs under rocprofv3, which sets ROCPROF_OUTPUT_PATH) */
    uint32_t max_retries;    /* capacity-overflow restarts per request */
    uint32_t max_iterations; /* per request; 0 = unlimited */
    uint32_t heap_arity;     /* lanes per heap = heap arity: 4 or 8 (0 = 4) */
    uint32_t max_load_pct;   /* closed-set load factor limit, percent */
    uint32_t pops_per_iter;  /* heap pops tried per heap per iteration (stale skip) */
    uint32_t profile;        /* 0 off; 1 count work and time every step with HIP events;
                                n > 1 sample every n-th chunk: time its steps, and count the
                                work marked (profile) below in those steps only */
    uint32_t rounds;         /* expansion rounds per heap per launch (local pushes between rounds) */
    uint32_t adaptive_start; /* initial heaps in use per request (doubles on demand); 0 = all num_queues */
    uint32_t grow_thresh;    /* double the heaps in use when open entries exceed this many per heap */
    uint32_t rounds_grid;    /* expansion rounds for grid requests; 0 = rounds */
    uint32_t lanes;          /* launch lanes (streams with their own iteration clock), 1..4 */
    uint64_t deep_cells;     /* grids with at least this many cells run on the last lane */
    uint32_t slack_puzzle;   /* 0 = off; k: expand a heap top only if f <= previous min f + k - 1 */
    uint32_t slack_grid;     /* the same for grids */
    uint32_t pdb24;          /* 1: 24-puzzles towards the ordered goal use additive pattern databases
                                (built on the device at the first such request, 48.8 MB), others
                                Manhattan; 0: Manhattan for every 24-puzzle */
    uint32_t grow_in_place;  /* 1: a search whose heaps or closed set fill up pauses at an iteration
                                boundary, the structure grows 4x with the search state carried over
                                (ga_result.grows), and it resumes; 0: it overflows and restarts from
                                scratch with larger structures (ga_result.retries) */
    uint32_t block;          /* threads per step-kernel workgroup: 128, 256 or 512 (0 = 256) */
    uint32_t step_flags;     /* step-kernel variants (each keeps results exact): 1 = write back only
                                the LDS heap-top entries a launch changed; 2 = grids read each entered
                                cell's stored g with its weight and skip the atomicMin that could not
                                lower it; 4 = a launch lane that runs only 4-connected grids pops two
                                nodes per round and issues both nodes' atomicMins before reading either
                                result (its own kernel instantiation; heap_arity 4 only, ignored at 8);
                                8 (with 4) = so do the 4-connected grids of a lane that also holds
                                other kinds */
    uint32_t persist;        /* ga_solve_batch only: 1 = solve the batch's 8- and 15-puzzles and its
                                grids of at most persist_cells cells in one launch per wave, one
                                workgroup per instance (persist_block / heap_arity heaps) running
                                every iteration of it; the other requests, and those that outgrow
                                their heaps or closed set, run on the launch lanes afterwards */
    uint32_t persist_block;  /* threads per workgroup of that launch: 256, 512 or 1024 (0 = 1024) */
    uint64_t persist_cells;  /* largest grid (cells) of the one-launch form (0 = puzzles only) */
    uint32_t pdb15;          /* 1: 15-puzzles towards the ordered goal (tile t on cell t-1, blank
                                last) use additive pattern databases (three 5-tile tables, 3 MiB,
                                built on the device at the first such request), others Manhattan;
                                0: Manhattan for every 15-puzzle.  Costs are the same either way */
} ga_config;

/* Engine statistics.  With cfg.profile = n, kernel time covers the timed
   chunks (every n-th chunk of a lane: timed_launches launches) and the counts
   marked (profile) the counted chunks (those half-way between: counted_launches
   launches); n = 1 times and counts every chunk; both are zero when
   cfg.profile == 0. */
typedef struct ga_engine_stats {
    uint64_t step_launches;  /* step kernels launched */
    double step_ms_total;    /* summed step-kernel time from HIP events (profile) */
    uint64_t expanded;       /* node expansions by all requests */
    uint64_t generated;      /* successors generated */
    uint64_t pushed;         /* heap pushes (profile) */
    uint64_t probes;         /* closed-set probes (profile) */
    uint64_t sift_down;      /* heap levels visited by pops (profile) */
    uint64_t sift_up;        /* heap levels visited by pushes (profile) */
    uint64_t pops;           /* heap pops, including stale ones (profile) */
    uint64_t completed;      /* requests completed */
    uint64_t peak_bytes;     /* peak admitted heap + closed-set bytes */
    uint64_t groups;         /* heap-group iterations of running instances (profile) */
    uint64_t inserted;       /* successors that lowered a stored g (profile) */
    uint64_t lookups;        /* closed-set find-or-insert calls (profile) */
    double algo_bytes;       /* algorithmic bytes of all step launches, the profiled terms scaled
                                from the sampled launches to all of them (per-unit figures: the
                                design notes' kernel table) */
    uint64_t timed_launches; /* step launches of the timed chunks: step_ms_total covers them */
    uint64_t heap_groups;    /* heap-group iterations launched in the counted launches (every mapped heap) */
    uint64_t work_groups;    /* heap-group iterations with work: pushes, or a pop within the f gate (profile) */
    uint64_t sift_down_hbm;  /* sift-down levels whose children were read from global memory (profile) */
    uint64_t sift_up_hbm;    /* sift-up levels whose parent was read from global memory (profile) */
    uint64_t push_in;        /* records pushed from the candidate buffer (profile) */
    uint64_t counted_launches; /* step launches of the counted chunks: the (profile) counts cover them */
    /* per launch lane: index l < lanes for the configured lanes, index lanes for the 24-puzzle lane */
    uint64_t lane_launches[GA_MAX_LANES + 1];        /* step launches */
    uint64_t lane_timed_launches[GA_MAX_LANES + 1];  /* step launches of timed chunks */
    uint64_t lane_counted_launches[GA_MAX_LANES + 1]; /* step launches of counted chunks */
    double lane_step_ms[GA_MAX_LANES + 1];           /* timed chunks' time from HIP events (one pair per chunk) */
    double lane_algo_bytes[GA_MAX_LANES + 1];        /* algorithmic bytes of all the lane's launches */
    uint64_t lane_atomics[GA_MAX_LANES + 1];         /* closed-set atomics of the lane's counted launches:
                                                        one per find-or-insert, plus one per hashed-set
                                                        probe (profile) */
    double lane_busy_ms[GA_MAX_LANES + 1];           /* time the lane's chunks occupied its stream, from one
                                                        HIP event pair around every chunk (profile) */
    /* host time of the run loop: admitting requests, handling chunk read-backs (finished requests, path
       walks launched, growth), uploading launch maps and launching chunks, finishing path walks */
    double host_admit_ms, host_snapshot_ms, host_launch_ms, host_path_ms;
    uint64_t persist_launches;  /* one-launch batch waves run (ga_config.persist) */
    uint64_t persist_instances; /* instances those waves ran */
    uint64_t persist_early_restarts; /* instances restarted on the lanes while their wave still ran (their
                                        workgroup's completion flag reported an overflow) */
} ga_engine_stats;

typedef struct ga_serve_stats {
    uint32_t completed;
    uint32_t failed;
    uint64_t wall_ns;        /* first arrival to last completion */
    double solves_per_sec;
    double p50_ms;
    double p99_ms;
    uint32_t max_inflight;   /* largest number of requests resident at once */
    uint32_t reserved;
} ga_serve_stats;

void ga_config_default(ga_config* cfg);

/* Engine lifetime.  An engine owns one HIP stream on `device`; it is not
 * thread-safe (the reference is not re-entrant either: a_star.h:205-212,
 * reduction.h:46-57). */
int ga_engine_create(int device, const ga_config* cfg, ga_engine** out);
void ga_engine_destroy(ga_engine* e);
int ga_engine_get_stats(ga_engine* e, ga_engine_stats* out);
void ga_engine_reset_stats(ga_engine* e);

/* Return codes of the solve and serve entry points.  A run ends in one of
 * two ways:
 *   - it completes: every request has its final status in its ga_result
 *     (GA_OK, GA_NO_PATH, or a per-request error such as GA_E_BUDGET), and
 *     ga_solve_batch / ga_serve_run return GA_OK;
 *   - it fails (a HIP error, device memory exhausted while the run holds
 *     requests): the entry point returns that negative code, every request
 *     that had not finished gets it as its status, and requests that had
 *     finished keep theirs.  After a HIP error the engine refuses further
 *     runs with GA_E_HIP; destroy it and create another.
 * ga_solve returns its request's status in both cases (the run's error when
 * the run failed). */

/* Plain GA*, one request at a time.  Replaces a_star<>() (a_star.h:223-282)
 * as called by SlidingPuzzle::solve (a_star_sliding_puzzle.h:165-171) and
 * Pathfinding::solve (a_star_pathfinding.h:223-238). */
int ga_solve(ga_engine* e, const ga_request* req, ga_result* out);
int ga_solve_puzzle(ga_engine* e, uint32_t side, uint64_t start, uint64_t goal, uint32_t num_queues,
                    ga_result* out);
int ga_solve_grid(ga_engine* e, const uint8_t* cells, uint32_t width, uint32_t height, uint64_t start,
                  uint64_t goal, uint32_t connectivity, uint32_t num_queues, ga_result* out);

/* Batched GA*: all requests share each iteration's launches, with heaps and
 * closed sets partitioned per request.  With cfg.persist the 8- and
 * 15-puzzles and the grids of at most cfg.persist_cells cells instead run
 * every iteration inside one launch, one workgroup per request, beside the
 * launch lanes.  No reference counterpart (the reference solves one instance
 * per a_star<>() call, a_star.h:223-224). */
int ga_solve_batch(ga_engine* e, const ga_request* reqs, uint32_t n, ga_result* outs);

/* Serving: replay an arrival stream (arrival_ns[i] relative to the call)
 * with continuous batching under cfg.budget_bytes.  No reference
 * counterpart. */
int ga_serve_run(ga_engine* e, const ga_request* reqs, const uint64_t* arrival_ns, uint32_t n,
                 ga_result* outs, ga_serve_stats* stats);

/* ---- online serving: requests submitted while the engine runs ----
 * No reference counterpart (the reference solves one instance per a_star<>()
 * call, a_star.h:223-224).  ga_serve_start hands the engine to a thread of
 * its own, which admits every submitted request as soon as a slot and the
 * admission budget (cfg.budget_bytes) allow, into the same launches as the
 * requests already resident (continuous batching); each result is the same
 * as the request's result alone.  While a server runs, the engine's other
 * entry points return GA_E_INVALID; stop the server before destroying the
 * engine.  The server functions may be called from any thread. */
typedef struct ga_server ga_server;
int ga_serve_start(ga_engine* e, ga_server** out);
/* Queues one request and returns its ticket.  Host data the request points
 * to (grid cells, side-5 boards) is copied; device cells (cells_device) must
 * stay valid until the request has finished.  Latency is measured from this
 * call.  After a failed run every call returns that run's error. */
int ga_serve_submit(ga_server* s, const ga_request* req, uint64_t* ticket);
/* GA_PENDING while the request runs; once it has finished, fills *out (the
 * request's own outcome is out->status; the path belongs to the caller, see
 * ga_result_free), forgets the ticket and returns GA_OK.  GA_E_INVALID for an
 * unknown ticket. */
int ga_serve_poll(ga_server* s, uint64_t ticket, ga_result* out);
/* Waits until every submitted request has finished; returns GA_OK, or the
 * error of a failed run (whose unfinished requests then carry it). */
int ga_serve_drain(ga_server* s);
/* Counts and latency percentiles of the requests finished so far. */
int ga_serve_status(ga_server* s, ga_serve_stats* out);
/* Waits for every submitted request, ends the thread, frees the results
 * never polled and the server; returns as ga_serve_drain. */
int ga_serve_stop(ga_server* s);

/* Plain GA* of one request with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, code that must not call hipMalloc).  The same search as ga_solve,
 * with every device buffer its kernels use (descriptor, state, launch map,
 * heaps, candidate records, closed set, a grid's cells, the path walk) taken
 * from one static array of 256 MiB per GPU, which the library's code object
 * holds on each GPU it runs on until the process ends; it calls no hipMalloc,
 * hipFree, hipHostMalloc, event or stream creation.  A request whose
 * structures (or path) do not fit 256 MiB gets GA_E_BUDGET.  The path is
 * written start..goal into the caller's host buffer path_buf (path_cap
 * states; 4 words per state for side-5 puzzles, else 1); out->path stays
 * NULL and out->path_len counts the states.  If the path is longer than
 * path_cap the status is GA_E_INVALID, out->cost holds the cost and
 * out->path_len the states needed.  Every call on a GPU shares that GPU's
 * array, which keeps the last call's contents (each call clears what it
 * needs): calls that overlap on one GPU race, from any engine or thread.
 * Returns the request's status. */
int ga_solve_noalloc(ga_engine* e, const ga_request* req, ga_result* out, uint64_t* path_buf, uint32_t path_cap);

void ga_result_free(ga_result* r);
const char* ga_status_str(int status);
/* Message of the last failed run or call on this engine ("" if none). */
const char* ga_engine_last_error(ga_engine* e);

/* The 24-puzzle pattern databases of the engine's device (ga_config.pdb24),
 * built first if need be, copied to `out` (bytes must be at least
 * 48,828,125: five tables of 25^5 bytes, the fifth using its first 25^4, the
 * rest 255).  No counterpart in the reference; for tests and inspection. */
int ga_pdb24_tables(ga_engine* e, uint8_t* out, size_t bytes);

/* The 15-puzzle pattern databases of the engine's device (ga_config.pdb15),
 * built first if need be, copied to `out` (bytes must be at least 3 * 16^5 =
 * 3,145,728: the tables of tiles {1,2,3,5,6}, {4,7,8,11,12}, {9,10,13,14,15},
 * indexed by the sum of cell(tile i) * 16^i, 255 where no placement is).  No
 * counterpart in the reference; for tests and inspection. */
int ga_pdb15_tables(ga_engine* e, uint8_t* out, size_t bytes);

/* Test support, no counterpart in the reference, no device call: applies the
 * request checks of ga_solve_batch to reqs, finishes the first finish_first
 * requests that pass them as completed (GA_OK, cost 0), then fails the run
 * with fail_code (negative) as an error inside a run would, and returns what
 * ga_solve_batch returns then. */
int ga_debug_fail_run(const ga_request* reqs, uint32_t n, ga_result* outs, uint32_t finish_first, int fail_code);

/* ---- kernel-level checks of the device data structures (tests) ----
 * Neither call has a counterpart in the reference.  Each runs the engine's
 * own device functions on caller data, on GPU `device`, in device memory it
 * allocates and frees itself, so that the closed sets and the heaps can be
 * checked against host models one operation at a time instead of only
 * through whole searches.  Both return GA_OK, GA_E_INVALID (arguments,
 * checked before any device call), GA_E_NOMEM or GA_E_HIP. */

/* Closed-set layouts: 64-bit keys (the 15- and 8-puzzles; 2 words per key,
 * low word first), 24-puzzle boards in 32-byte slots (4 words: lo, hi), and
 * 12-byte states behind a fingerprint tag (the header drop-in's layout for
 * states wider than 8 bytes, e.g. Position<3>; 3 words). */
enum { GA_SET_U64 = 0, GA_SET_BOARD = 1, GA_SET_WIDE12 = 2 };
/* Outcomes: the contract of the reference's mass insert, fancy_hashmap.h:4-8,
 * 56-61 (INSERTED: the state was new; UPDATED: it was present and its g was
 * lowered; REJECTED: neither), plus FULL: no slot found within max_probes. */
enum { GA_CS_REJECTED = 0, GA_CS_INSERTED = 1, GA_CS_UPDATED = 2, GA_CS_FULL = 3 };
/* Inserts n successors (key i = words [i*kw, (i+1)*kw), kw = 2, 4 or 3 by
 * layout; value g[i] << 32 | parent[i]) into an empty set of 2^log2_slots
 * slots (1 <= log2_slots <= 24), each as one step of the search does
 * (find or claim the slot, then atomicMin of the value): all n at once, one
 * device thread each (serial = 0), or one after another in index order on
 * one thread (serial = 1).  Writes outcome[i] and slot[i] (0xFFFFFFFF for
 * FULL) and, where not null, the table after the last insert:
 * tag_out[s] = the word a probe claims (the key; a board's lo; a 12-byte
 * state's tag: all ones for empty, else its hash >> 2), key_out[s*kw ..] =
 * the state stored in slot s, val_out[s] = its value (all ones if empty). */
int ga_debug_closed_set(int device, int layout, const uint32_t* keys, const uint32_t* g, const uint32_t* parent,
                        uint32_t n, uint32_t log2_slots, uint32_t max_probes, int serial, int32_t* outcome,
                        uint32_t* slot, uint64_t* tag_out, uint32_t* key_out, uint64_t* val_out);

/* Heap operations, one value per operation; any other value v, with
 * v < 0xFFFFFFFF00000000 (f below the infinite cost), is a push of a record
 * with priority v.  Pushes are collected, up to 2 * arity per batch, and
 * pushed as one batch of candidate records (the batch is flushed before a
 * pop, a FLUSH, at the end of a launch and when full); GAP leaves an empty
 * record position in the batch, as an expansion without a survivor does. */
#define GA_HOP_POP 0xFFFFFFFFFFFFFFFFull
#define GA_HOP_FLUSH 0xFFFFFFFFFFFFFFFEull
#define GA_HOP_NOP 0xFFFFFFFFFFFFFFFDull
#define GA_HOP_GAP 0xFFFFFFFFFFFFFFFCull
/* Runs n_heaps (1 <= n_heaps <= 64 / arity) heaps of `arity` (4, 8 or 16) and
 * capacity cap (1 <= cap <= 2^20) in one wave, heap h executing
 * ops[h*n_ops .. (h+1)*n_ops) (n_ops <= 2^20), ops_per_launch operations per
 * kernel launch, each launch copying the heaps' top levels into LDS and back
 * as a search step does (track_dirty, arity 4 and 8 only: write back only
 * the changed entries).
 * Writes, per heap and operation, popped[h*n_ops + i] (the priority a POP
 * returned; all ones for an empty heap or another operation) and
 * popped_tag[h*n_ops + i] (the index, within the heap's operations, of the
 * push that entry came from; 0xFFFFFFFE if it does not carry its own heap's
 * mark), then the heap after the last operation: size_out[h], heap_out
 * [h*cap + j] = the priority of logical entry j < size_out[h], and
 * overflow_out[h] = the number of batches refused because they would have
 * passed the capacity (such a batch pushes nothing). */
int ga_debug_heap(int device, uint32_t arity, uint32_t n_heaps, const uint64_t* ops, uint32_t n_ops,
                  uint32_t ops_per_launch, uint32_t cap, int track_dirty, uint64_t* popped, uint32_t* popped_tag,
                  uint64_t* heap_out, uint32_t* size_out, uint32_t* overflow_out);

/* Measurement helper, no counterpart in the reference: the device's rate
 * for heap operations at a given arity (4, 8 or 16).  waves (a multiple of 4, at most
 * 2^16) 64-lane waves run in 256-thread workgroups, each W-lane group of them
 * owning one heap (64 / arity heaps per wave) of `fill` entries (2..2^16) with
 * its top entries in LDS, as a search step keeps them; every group pops two
 * entries and pushes two successors of the first as one batch, `iters` times
 * (1..2^16), so its heap keeps its size.  Writes the pops per second of the
 * second of two launches (pushes run at the same rate) to *pops_per_sec. */
int ga_heap_rate(int device, uint32_t arity, uint32_t waves, uint32_t fill, uint32_t iters, double* pops_per_sec);

/* Device memory helpers for callers that keep grids resident in HBM. */
int ga_device_upload(ga_engine* e, const void* host, size_t bytes, void** device_out);
int ga_device_free(ga_engine* e, void* device);

/* Measurement helper, no counterpart in the reference: the device's rate for
 * the closed sets' operations, random 64-bit global atomics at agent scope
 * (kind 0: fetch-min, kind 1: CAS from the empty key), or random 64-bit plain
 * loads (kind 2: the rate at which the memory system serves scattered lines),
 * over a table of table_bytes (rounded down to a power of two), about `ops`
 * of them in one timed launch.  Writes operations per second to *ops_per_sec. */
int ga_atomic_rate(ga_engine* e, uint64_t table_bytes, uint64_t ops, int kind, double* ops_per_sec);

/* ---- seeded instance generators (host only; splitmix64 streams) ---- */
uint64_t ga_splitmix64(uint64_t* state);
/* Goal = tiles 1..n-1 in row-major order, blank last. */
uint64_t ga_puzzle_goal(uint32_t side);
/* Random walk of `depth` legal blank moves from `goal`, moves drawn uniformly
 * from the legal ones in the reference order (a_star_sliding_puzzle.h:88-105);
 * no_backtrack excludes the move that undoes the previous one. */
uint64_t ga_gen_puzzle(uint32_t side, uint64_t goal, uint32_t depth, int no_backtrack, uint64_t seed);
/* width*height cells: wall with probability wall_ppm/1e6, else weight uniform
 * in w_lo..w_hi; start and goal forced open; redrawn from the continuing
 * stream until goal is reachable from start (4- or 8-connected).  Returns the
 * number of draws (>= 1), or 0 on invalid arguments / no reachable draw in
 * max_attempts. */
uint32_t ga_gen_grid(uint32_t width, uint32_t height, uint32_t wall_ppm, uint32_t w_lo, uint32_t w_hi,
                     uint64_t start, uint64_t goal, uint32_t connectivity, uint64_t seed,
                     uint32_t max_attempts, uint8_t* cells_out);
/* 1 if goal is reachable from start by sliding moves (parity test). */
int ga_puzzle_solvable(uint32_t side, uint64_t start, uint64_t goal);
/* Boards of side 2..5 (side*side bytes, 0 = blank): the walk of ga_gen_puzzle
 * (same draws and moves, so sides 3 and 4 give the same boards) into out;
 * returns 1, or 0 on invalid arguments. */
int ga_gen_board(uint32_t side, const uint8_t* goal, uint32_t depth, int no_backtrack, uint64_t seed,
                 uint8_t* out);
/* ga_puzzle_solvable on boards of side 2..5. */
int ga_board_solvable(uint32_t side, const uint8_t* start, const uint8_t* goal);

#ifdef __cplusplus
}
#endif

#endif
