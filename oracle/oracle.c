/*
 * oracle.c -- sequential CPU oracle (test infrastructure only; see oracle.h).
 *
 * Reference citations are to michalborkowski96/cuda_a_star.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define NONE32 0xFFFFFFFFu
#define INF32 0xFFFFFFFFu

/* ------------------------------------------------------------------ */
/* priority queue: binary min-heap on (f asc, g desc)                  */
/* QueueValue::operator< in a_star.h:58-63: equal f -> larger g first. */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t prio; /* (f << 32) | (0xFFFFFFFF - g) */
    uint32_t idx;
    uint32_t g;
} pq_item;

typedef struct {
    pq_item* a;
    size_t n, cap;
} pq;

static uint64_t prio_of(uint32_t f, uint32_t g) {
    return ((uint64_t)f << 32) | (uint64_t)(0xFFFFFFFFu - g);
}

static int pq_push(pq* q, pq_item it) {
    if (q->n == q->cap) {
        size_t nc = q->cap ? q->cap * 2 : 1024;
        pq_item* na = (pq_item*)realloc(q->a, nc * sizeof(pq_item));
        if (!na) return -1;
        q->a = na;
        q->cap = nc;
    }
    size_t i = q->n++;
    while (i > 0) {
        size_t p = (i - 1) >> 1;
        if (it.prio < q->a[p].prio) {
            q->a[i] = q->a[p];
            i = p;
        } else {
            break;
        }
    }
    q->a[i] = it;
    return 0;
}

static pq_item pq_pop(pq* q) {
    pq_item top = q->a[0];
    pq_item last = q->a[--q->n];
    size_t n = q->n, i = 0;
    if (n) {
        for (;;) {
            size_t l = 2 * i + 1, r = l + 1, m = i;
            uint64_t mp = last.prio;
            if (l < n && q->a[l].prio < mp) { m = l; mp = q->a[l].prio; }
            if (r < n && q->a[r].prio < mp) { m = r; }
            if (m == i) break;
            q->a[i] = q->a[m];
            i = m;
        }
        q->a[i] = last;
    }
    return top;
}

/* ------------------------------------------------------------------ */
/* sliding puzzle                                                      */
/* ------------------------------------------------------------------ */
static inline uint32_t nib(uint64_t s, int c) { return (uint32_t)((s >> (4 * c)) & 0xF); }

static int blank_of(int side, uint64_t s) {
    int n = side * side;
    for (int c = 0; c < n; ++c)
        if (nib(s, c) == 0) return c;
    return -1;
}

uint32_t oracle_puzzle_manhattan(int side, uint64_t state, uint64_t goal) {
    int gr[16], gc[16];
    int n = side * side;
    for (int c = 0; c < n; ++c) {
        gr[nib(goal, c)] = c / side;
        gc[nib(goal, c)] = c % side;
    }
    uint32_t hsum = 0;
    for (int c = 0; c < n; ++c) {
        uint32_t t = nib(state, c);
        if (t == 0) continue; /* blank excluded: admissible (unlike a_star_sliding_puzzle.h:152-163) */
        int dr = c / side - gr[t], dc = c % side - gc[t];
        hsum += (uint32_t)((dr < 0 ? -dr : dr) + (dc < 0 ? -dc : dc));
    }
    return hsum;
}

/* Successor i of s in the reference order (a_star_sliding_puzzle.h:88-105):
 * 0: blank row+1, 1: blank row-1, 2: blank col+1, 3: blank col-1. */
static int puzzle_succ(int side, uint64_t s, int blank, int i, uint64_t* out) {
    int r = blank / side, c = blank % side;
    switch (i) {
        case 0: r += 1; break;
        case 1: r -= 1; break;
        case 2: c += 1; break;
        default: c -= 1; break;
    }
    if (r < 0 || r >= side || c < 0 || c >= side) return 0;
    int nb = r * side + c;
    uint64_t tile = (s >> (4 * nb)) & 0xF;
    uint64_t t = s & ~(0xFull << (4 * nb));
    t |= tile << (4 * blank);
    *out = t;
    return 1;
}

static int valid_puzzle(int side, uint64_t s) {
    int n = side * side;
    uint32_t seen = 0;
    for (int c = 0; c < n; ++c) {
        uint32_t t = nib(s, c);
        if (t >= (uint32_t)n || (seen >> t) & 1) return 0;
        seen |= 1u << t;
    }
    if (n < 16 && (s >> (4 * n)) != 0) return 0;
    return 1;
}

typedef struct {
    uint64_t* key;
    uint32_t* g;
    uint32_t* parent;
    size_t n, cap;
    uint32_t* slot; /* open addressing: node index + 1, 0 = empty */
    size_t slot_mask;
} pstore;

static uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static int pstore_grow_slots(pstore* st) {
    size_t ns = st->slot_mask ? (st->slot_mask + 1) * 2 : (1u << 16);
    uint32_t* sl = (uint32_t*)calloc(ns, sizeof(uint32_t));
    if (!sl) return -1;
    for (size_t i = 0; i < st->n; ++i) {
        size_t p = mix64(st->key[i]) & (ns - 1);
        while (sl[p]) p = (p + 1) & (ns - 1);
        sl[p] = (uint32_t)(i + 1);
    }
    free(st->slot);
    st->slot = sl;
    st->slot_mask = ns - 1;
    return 0;
}

/* Returns node index for key, inserting it with g = INF if absent; -1 on OOM. */
static int64_t pstore_get(pstore* st, uint64_t key) {
    if (2 * (st->n + 1) > st->slot_mask + 1)
        if (pstore_grow_slots(st)) return -1;
    size_t p = mix64(key) & st->slot_mask;
    while (st->slot[p]) {
        uint32_t i = st->slot[p] - 1;
        if (st->key[i] == key) return i;
        p = (p + 1) & st->slot_mask;
    }
    if (st->n == st->cap) {
        size_t nc = st->cap ? st->cap * 2 : 4096;
        uint64_t* k = (uint64_t*)realloc(st->key, nc * sizeof(uint64_t));
        if (!k) return -1;
        st->key = k;
        uint32_t* g = (uint32_t*)realloc(st->g, nc * sizeof(uint32_t));
        if (!g) return -1;
        st->g = g;
        uint32_t* pa = (uint32_t*)realloc(st->parent, nc * sizeof(uint32_t));
        if (!pa) return -1;
        st->parent = pa;
        st->cap = nc;
    }
    size_t i = st->n++;
    st->key[i] = key;
    st->g[i] = INF32;
    st->parent[i] = NONE32;
    st->slot[p] = (uint32_t)(i + 1);
    return (int64_t)i;
}

static void pstore_free(pstore* st) {
    free(st->key);
    free(st->g);
    free(st->parent);
    free(st->slot);
}

static uint32_t pdb15_h_state(const uint8_t* pdb, uint64_t s);

/* A* over 64-bit puzzle states; h = Manhattan without the blank, or (pdb, 4x4,
 * the ordered goal) the sum of the 15-puzzle pattern databases */
static int astar_puzzle_h(int side, uint64_t start, uint64_t goal, const uint8_t* pdb, uint64_t max_states,
                          oracle_result* res, uint64_t* path, uint32_t path_cap) {
    memset(res, 0, sizeof(*res));
    res->cost = -1;
    if ((side != 3 && side != 4) || !valid_puzzle(side, start) || !valid_puzzle(side, goal))
        return ORACLE_E_INVALID;
    pstore st;
    memset(&st, 0, sizeof(st));
    pq q;
    memset(&q, 0, sizeof(q));
    int rc = ORACLE_NO_PATH;
    int64_t s0 = pstore_get(&st, start);
    if (s0 < 0) { rc = ORACLE_E_NOMEM; goto out; }
    st.g[s0] = 0; /* the start is stored with g = 0 (the reference never stores it, a_star.h:226 vs :253) */
    uint32_t h0 = pdb ? pdb15_h_state(pdb, start) : oracle_puzzle_manhattan(side, start, goal);
    if (pq_push(&q, (pq_item){prio_of(h0, 0), (uint32_t)s0, 0})) { rc = ORACLE_E_NOMEM; goto out; }
    int64_t goal_idx = -1;
    while (q.n) {
        pq_item it = pq_pop(&q);
        if (it.g > st.g[it.idx]) continue; /* stale: a better copy was pushed later */
        uint64_t s = st.key[it.idx];
        if (s == goal) { goal_idx = it.idx; break; } /* goal test at pop: a_star.h:80-83 */
        res->expanded++;
        int b = blank_of(side, s);
        for (int i = 0; i < 4; ++i) {
            uint64_t t;
            if (!puzzle_succ(side, s, b, i, &t)) continue;
            res->generated++;
            uint32_t g2 = it.g + 1; /* unit cost: a_star_sliding_puzzle.h:84 */
            int64_t ti = pstore_get(&st, t);
            if (ti < 0) { rc = ORACLE_E_NOMEM; goto out; }
            if (max_states && st.n > max_states) { rc = ORACLE_E_LIMIT; goto out; }
            if (g2 >= st.g[ti]) continue; /* rejected: fancy_hashmap.h:56-61 */
            st.g[ti] = g2;
            st.parent[ti] = it.idx;
            uint32_t ht = pdb ? pdb15_h_state(pdb, t) : oracle_puzzle_manhattan(side, t, goal);
            if (pq_push(&q, (pq_item){prio_of(g2 + ht, g2), (uint32_t)ti, g2})) { rc = ORACLE_E_NOMEM; goto out; }
        }
    }
    res->stored = st.n;
    if (goal_idx >= 0) {
        rc = ORACLE_OK;
        res->cost = st.g[goal_idx];
        uint32_t len = 0;
        for (uint32_t i = (uint32_t)goal_idx; i != NONE32; i = st.parent[i]) ++len;
        res->path_len = len;
        if (path && path_cap >= len) {
            uint32_t k = len;
            for (uint32_t i = (uint32_t)goal_idx; i != NONE32; i = st.parent[i]) path[--k] = st.key[i];
        }
    }
out:
    if (rc != ORACLE_OK && rc != ORACLE_NO_PATH) res->stored = st.n;
    pstore_free(&st);
    free(q.a);
    return rc;
}

int oracle_astar_puzzle(int side, uint64_t start, uint64_t goal, uint64_t max_states,
                        oracle_result* res, uint64_t* path, uint32_t path_cap) {
    return astar_puzzle_h(side, start, goal, NULL, max_states, res, path, path_cap);
}

int oracle_astar_puzzle_pdb(uint64_t start, const uint8_t* pdb, uint64_t max_states, oracle_result* res,
                            uint64_t* path, uint32_t path_cap) {
    memset(res, 0, sizeof(*res));
    res->cost = -1;
    if (!pdb) return ORACLE_E_INVALID;
    return astar_puzzle_h(4, start, ORACLE_PUZZLE4_GOAL, pdb, max_states, res, path, path_cap);
}

/* ------------------------------------------------------------------ */
/* grids                                                               */
/* ------------------------------------------------------------------ */
/* Move order of a_star_pathfinding.h:150-182: x+1, x-1, y+1, y-1,
 * (+1,+1), (+1,-1), (-1,+1), (-1,-1). */
static const int DX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
static const int DY[8] = {0, 0, 1, -1, 1, -1, 1, -1};

static uint32_t grid_h(uint32_t x, uint32_t y, uint32_t gx, uint32_t gy, int conn, uint32_t w_min) {
    uint32_t dx = x > gx ? x - gx : gx - x;
    uint32_t dy = y > gy ? y - gy : gy - y;
    if (conn == 8) return (dx > dy ? dx : dy) * w_min;
    return (dx + dy) * w_min;
}

static uint32_t grid_wmin(const uint8_t* cells, size_t n) {
    uint32_t m = 256;
    for (size_t i = 0; i < n; ++i)
        if (cells[i] && cells[i] < m) m = cells[i];
    return m == 256 ? 1 : m;
}

int oracle_astar_grid(const uint8_t* cells, uint32_t w, uint32_t h, uint32_t start, uint32_t goal,
                      int connectivity, uint32_t w_min, oracle_result* res, uint32_t* path,
                      uint32_t path_cap) {
    memset(res, 0, sizeof(*res));
    res->cost = -1;
    size_t n = (size_t)w * h;
    if (!w || !h || start >= n || goal >= n || (connectivity != 4 && connectivity != 8) ||
        !cells[start] || !cells[goal])
        return ORACLE_E_INVALID;
    if (w_min == 0) w_min = grid_wmin(cells, n);
    uint32_t* g = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t* par = (uint32_t*)malloc(n * sizeof(uint32_t));
    if (!g || !par) {
        free(g);
        free(par);
        return ORACLE_E_NOMEM;
    }
    memset(g, 0xFF, n * sizeof(uint32_t));
    pq q;
    memset(&q, 0, sizeof(q));
    int rc = ORACLE_NO_PATH;
    uint32_t gx = goal % w, gy = goal / w;
    g[start] = 0;
    par[start] = NONE32;
    if (pq_push(&q, (pq_item){prio_of(grid_h(start % w, start / w, gx, gy, connectivity, w_min), 0), start, 0})) {
        rc = ORACLE_E_NOMEM;
        goto out;
    }
    int found = 0;
    uint64_t stored = 1;
    while (q.n) {
        pq_item it = pq_pop(&q);
        if (it.g > g[it.idx]) continue;
        if (it.idx == goal) { found = 1; break; }
        res->expanded++;
        uint32_t x = it.idx % w, y = it.idx / w;
        for (int i = 0; i < connectivity; ++i) {
            int64_t nx = (int64_t)x + DX[i], ny = (int64_t)y + DY[i];
            if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
            uint32_t c = (uint32_t)ny * w + (uint32_t)nx;
            if (!cells[c]) continue; /* wall */
            res->generated++;
            uint32_t g2 = it.g + cells[c]; /* cost = entered cell: a_star_pathfinding.h:203-208 */
            if (g2 >= g[c]) continue;
            if (g[c] == INF32) ++stored;
            g[c] = g2;
            par[c] = it.idx;
            uint32_t f = g2 + grid_h((uint32_t)nx, (uint32_t)ny, gx, gy, connectivity, w_min);
            if (pq_push(&q, (pq_item){prio_of(f, g2), c, g2})) { rc = ORACLE_E_NOMEM; goto out; }
        }
    }
    res->stored = stored;
    if (found) {
        rc = ORACLE_OK;
        res->cost = g[goal];
        uint32_t len = 0;
        for (uint32_t i = goal; i != NONE32; i = par[i]) ++len;
        res->path_len = len;
        if (path && path_cap >= len) {
            uint32_t k = len;
            for (uint32_t i = goal; i != NONE32; i = par[i]) path[--k] = i;
        }
    }
out:
    free(g);
    free(par);
    free(q.a);
    return rc;
}

/* Dial's algorithm: Dijkstra with a circular array of buckets, valid because
 * every move cost is in 1..255.  No heuristic, no binary heap: an independent
 * check of oracle_astar_grid. */
int64_t oracle_dial_grid(const uint8_t* cells, uint32_t w, uint32_t h, uint32_t start, uint32_t goal,
                         int connectivity) {
    size_t n = (size_t)w * h;
    if (!w || !h || start >= n || goal >= n || !cells[start] || !cells[goal]) return -1;
    enum { NB = 256 };
    uint32_t* dist = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t head[NB]; /* bucket d % NB: singly-linked list of pool entries */
    if (!dist) return -1;
    memset(dist, 0xFF, n * sizeof(uint32_t));
    for (int i = 0; i < NB; ++i) head[i] = NONE32;
    /* A cell may be queued several times; each queued copy is a pool entry
     * (cell, d) and is skipped when popped if dist[cell] has since dropped. */
    size_t pool_cap = n + 16, pool_n = 0;
    uint32_t* pool_cell = (uint32_t*)malloc(pool_cap * sizeof(uint32_t));
    uint32_t* pool_d = (uint32_t*)malloc(pool_cap * sizeof(uint32_t));
    uint32_t* pool_next = (uint32_t*)malloc(pool_cap * sizeof(uint32_t));
    int64_t ans = -1;
    if (!pool_cell || !pool_d || !pool_next) goto out;
    dist[start] = 0;
    pool_cell[0] = start;
    pool_d[0] = 0;
    pool_next[0] = NONE32;
    head[0] = 0;
    pool_n = 1;
    size_t live = 1;
    for (uint32_t d = 0; live; ++d) {
        uint32_t b = d % NB;
        while (head[b] != NONE32) {
            uint32_t e = head[b];
            head[b] = pool_next[e];
            --live;
            /* every move costs 1..255 < NB, so all live entries lie within one
             * lap of d and bucket d % NB holds only entries with pool_d == d */
            uint32_t c = pool_cell[e];
            if (dist[c] != d) continue;
            if (c == goal) { ans = d; goto out; }
            uint32_t x = c % w, y = c / w;
            for (int i = 0; i < connectivity; ++i) {
                int64_t nx = (int64_t)x + DX[i], ny = (int64_t)y + DY[i];
                if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                uint32_t t = (uint32_t)ny * w + (uint32_t)nx;
                if (!cells[t]) continue;
                uint32_t nd = d + cells[t];
                if (nd >= dist[t]) continue;
                dist[t] = nd;
                if (pool_n == pool_cap) {
                    size_t nc = pool_cap * 2;
                    uint32_t* a = (uint32_t*)realloc(pool_cell, nc * sizeof(uint32_t));
                    if (!a) goto out;
                    pool_cell = a;
                    a = (uint32_t*)realloc(pool_d, nc * sizeof(uint32_t));
                    if (!a) goto out;
                    pool_d = a;
                    a = (uint32_t*)realloc(pool_next, nc * sizeof(uint32_t));
                    if (!a) goto out;
                    pool_next = a;
                    pool_cap = nc;
                }
                size_t e2 = pool_n++;
                pool_cell[e2] = t;
                pool_d[e2] = nd;
                pool_next[e2] = head[nd % NB];
                head[nd % NB] = (uint32_t)e2;
                ++live;
            }
        }
    }
out:
    free(dist);
    free(pool_cell);
    free(pool_d);
    free(pool_next);
    return ans;
}

/* ------------------------------------------------------------------ */
/* 8-puzzle BFS                                                        */
/* ------------------------------------------------------------------ */
uint32_t oracle_perm_rank9(uint64_t state) {
    /* Lehmer code of the 9 nibbles, cell 0 most significant. */
    static const uint32_t fact[9] = {40320, 5040, 720, 120, 24, 6, 2, 1, 1};
    uint32_t used = 0, rank = 0;
    for (int c = 0; c < 9; ++c) {
        uint32_t t = nib(state, c);
        uint32_t smaller = (uint32_t)__builtin_popcount(((1u << t) - 1) & ~used);
        rank += smaller * fact[c];
        used |= 1u << t;
    }
    return rank;
}

uint32_t oracle_bfs_puzzle8(uint64_t goal, uint8_t* depth_by_rank, uint32_t* hist) {
    memset(depth_by_rank, 0xFF, 362880);
    memset(hist, 0, 64 * sizeof(uint32_t));
    uint64_t* queue = (uint64_t*)malloc(362880 * sizeof(uint64_t));
    if (!queue) return 0;
    size_t qh = 0, qt = 0;
    queue[qt++] = goal;
    depth_by_rank[oracle_perm_rank9(goal)] = 0;
    while (qh < qt) {
        uint64_t s = queue[qh++];
        uint8_t d = depth_by_rank[oracle_perm_rank9(s)];
        hist[d]++;
        int b = blank_of(3, s);
        for (int i = 0; i < 4; ++i) {
            uint64_t t;
            if (!puzzle_succ(3, s, b, i, &t)) continue;
            uint32_t r = oracle_perm_rank9(t);
            if (depth_by_rank[r] != 0xFF) continue;
            depth_by_rank[r] = (uint8_t)(d + 1);
            queue[qt++] = t;
        }
    }
    free(queue);
    return (uint32_t)qt;
}

/* ------------------------------------------------------------------ */
/* replay                                                              */
/* ------------------------------------------------------------------ */
int64_t oracle_replay_puzzle(int side, const uint64_t* path, uint32_t n, uint64_t start, uint64_t goal) {
    if (n == 0 || path[0] != start || path[n - 1] != goal) return -1;
    if (!valid_puzzle(side, start)) return -1;
    int64_t cost = 0;
    for (uint32_t k = 1; k < n; ++k) {
        uint64_t s = path[k - 1];
        int b = blank_of(side, s);
        int ok = 0;
        for (int i = 0; i < 4 && !ok; ++i) {
            uint64_t t;
            if (puzzle_succ(side, s, b, i, &t) && t == path[k]) ok = 1;
        }
        if (!ok) return -1;
        cost += 1;
    }
    return cost;
}

int64_t oracle_replay_grid(const uint8_t* cells, uint32_t w, uint32_t h, int connectivity,
                           const uint32_t* path, uint32_t n, uint32_t start, uint32_t goal) {
    size_t total = (size_t)w * h;
    if (n == 0 || path[0] != start || path[n - 1] != goal) return -1;
    if (start >= total || !cells[start]) return -1;
    int64_t cost = 0;
    for (uint32_t k = 1; k < n; ++k) {
        uint32_t a = path[k - 1], b = path[k];
        if (b >= total || !cells[b]) return -1;
        int64_t dx = (int64_t)(b % w) - (int64_t)(a % w);
        int64_t dy = (int64_t)(b / w) - (int64_t)(a / w);
        int ok = 0;
        for (int i = 0; i < connectivity; ++i)
            if (DX[i] == dx && DY[i] == dy) ok = 1;
        if (!ok) return -1;
        cost += cells[b];
    }
    return cost;
}

/* ------------------------------------------------------------------ */
/* boards of side <= 5 (up to the 24-puzzle): 5 bits per cell, 128 bits */
/* ------------------------------------------------------------------ */
/* Same rules as the 64-bit puzzle code above (move order
 * a_star_sliding_puzzle.h:88-105, unit cost, goal at pop, re-opening), on a
 * board given as side*side bytes (tile numbers, 0 = blank, row-major). */
typedef unsigned __int128 u128;

static inline uint32_t cell5(u128 s, int c) { return (uint32_t)(s >> (5 * c)) & 31u; }

static u128 board_pack(int side, const uint8_t* b) {
    u128 s = 0;
    for (int c = 0; c < side * side; ++c) s |= (u128)b[c] << (5 * c);
    return s;
}

static void board_unpack(int side, u128 s, uint8_t* b) {
    for (int c = 0; c < side * side; ++c) b[c] = (uint8_t)cell5(s, c);
}

static int board_valid(int side, const uint8_t* b) {
    if (side < 2 || side > 5) return 0;
    int n = side * side;
    uint32_t seen = 0;
    for (int c = 0; c < n; ++c) {
        if (b[c] >= n || (seen >> b[c]) & 1u) return 0;
        seen |= 1u << b[c];
    }
    return 1;
}

static int board_blank(int side, u128 s) {
    for (int c = 0; c < side * side; ++c)
        if (cell5(s, c) == 0) return c;
    return -1;
}

static int board_succ(int side, u128 s, int blank, int i, u128* out) {
    int r = blank / side, c = blank % side;
    switch (i) {
        case 0: r += 1; break;
        case 1: r -= 1; break;
        case 2: c += 1; break;
        default: c -= 1; break;
    }
    if (r < 0 || r >= side || c < 0 || c >= side) return 0;
    int nb = r * side + c;
    u128 tile = (u128)cell5(s, nb);
    u128 t = s & ~((u128)31 << (5 * nb));
    t |= tile << (5 * blank);
    *out = t;
    return 1;
}

static void goal_rc(int side, u128 goal, int* gr, int* gc) {
    for (int c = 0; c < side * side; ++c) {
        gr[cell5(goal, c)] = c / side;
        gc[cell5(goal, c)] = c % side;
    }
}

static uint32_t manh128(int side, u128 s, const int* gr, const int* gc) {
    uint32_t hsum = 0;
    for (int c = 0; c < side * side; ++c) {
        uint32_t t = cell5(s, c);
        if (t == 0) continue; /* blank excluded */
        int dr = c / side - gr[t], dc = c % side - gc[t];
        hsum += (uint32_t)((dr < 0 ? -dr : dr) + (dc < 0 ? -dc : dc));
    }
    return hsum;
}

uint32_t oracle_board_manhattan(int side, const uint8_t* state, const uint8_t* goal) {
    int gr[32], gc[32];
    goal_rc(side, board_pack(side, goal), gr, gc);
    return manh128(side, board_pack(side, state), gr, gc);
}

/* node store keyed by 128-bit states (open addressing over node indices) */
typedef struct {
    u128* key;
    uint32_t* g;
    uint32_t* parent;
    size_t n, cap;
    uint32_t* slot;
    size_t slot_mask;
} bstore;

static uint64_t hash128(u128 k) { return mix64((uint64_t)k ^ mix64((uint64_t)(k >> 64) + 0x9E3779B97F4A7C15ull)); }

static int bstore_grow_slots(bstore* st) {
    size_t ns = st->slot_mask ? (st->slot_mask + 1) * 2 : (1u << 16);
    uint32_t* sl = (uint32_t*)calloc(ns, sizeof(uint32_t));
    if (!sl) return -1;
    for (size_t i = 0; i < st->n; ++i) {
        size_t p = hash128(st->key[i]) & (ns - 1);
        while (sl[p]) p = (p + 1) & (ns - 1);
        sl[p] = (uint32_t)(i + 1);
    }
    free(st->slot);
    st->slot = sl;
    st->slot_mask = ns - 1;
    return 0;
}

static int64_t bstore_get(bstore* st, u128 key) {
    if (2 * (st->n + 1) > st->slot_mask + 1)
        if (bstore_grow_slots(st)) return -1;
    size_t p = hash128(key) & st->slot_mask;
    while (st->slot[p]) {
        uint32_t i = st->slot[p] - 1;
        if (st->key[i] == key) return i;
        p = (p + 1) & st->slot_mask;
    }
    if (st->n == st->cap) {
        size_t nc = st->cap ? st->cap * 2 : 4096;
        u128* k = (u128*)realloc(st->key, nc * sizeof(u128));
        if (!k) return -1;
        st->key = k;
        uint32_t* g = (uint32_t*)realloc(st->g, nc * sizeof(uint32_t));
        if (!g) return -1;
        st->g = g;
        uint32_t* pa = (uint32_t*)realloc(st->parent, nc * sizeof(uint32_t));
        if (!pa) return -1;
        st->parent = pa;
        st->cap = nc;
    }
    size_t i = st->n++;
    st->key[i] = key;
    st->g[i] = INF32;
    st->parent[i] = NONE32;
    st->slot[p] = (uint32_t)(i + 1);
    return (int64_t)i;
}

static void bstore_free(bstore* st) {
    free(st->key);
    free(st->g);
    free(st->parent);
    free(st->slot);
}

/* ---- additive pattern databases for 5x5 boards with the ordered goal ---- */
/* The 24 tiles are split into five disjoint patterns.  A pattern's table
 * holds, for every placement of its tiles, the fewest moves of those tiles
 * that bring them to their goal cells when a tile may step into any
 * neighbouring cell not held by another tile of the same pattern (the other
 * tiles and the blank are ignored).  That is a relaxation of the puzzle, and
 * every real move moves one tile of one pattern, so the sum over the patterns
 * never exceeds the true distance and changes by at most 1 per move.
 * Placement index: sum over the pattern's tiles i of cell(tile i) * 25^i.
 * Goal: tile t on cell t - 1, blank on cell 24. */
static const uint8_t PDB_TILES[ORACLE_PDB_PATTERNS][5] = {
    {1, 2, 3, 6, 7}, {4, 5, 8, 9, 10}, {11, 12, 16, 17, 21}, {13, 14, 18, 19, 22}, {15, 20, 23, 24, 0}};
static const int PDB_NT[ORACLE_PDB_PATTERNS] = {5, 5, 5, 5, 4};
static const uint32_t POW25[6] = {1u, 25u, 625u, 15625u, 390625u, 9765625u};

static size_t pdb_offset(int p) { return (size_t)p * POW25[5]; }

int oracle_pdb24_build(uint8_t* out) {
    if (!out) return ORACLE_E_INVALID;
    for (int p = 0; p < ORACLE_PDB_PATTERNS; ++p) {
        const int nt = PDB_NT[p];
        const uint32_t n = POW25[nt];
        uint8_t* dist = out + pdb_offset(p);
        memset(dist, 0xFF, POW25[5]);  /* the 4-tile table's unused tail too */
        uint32_t* queue = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
        if (!queue) return ORACLE_E_NOMEM;
        uint32_t goal = 0;
        for (int i = 0; i < nt; ++i) goal += (uint32_t)(PDB_TILES[p][i] - 1) * POW25[i];
        size_t head = 0, tail = 0;
        dist[goal] = 0;
        queue[tail++] = goal;
        while (head < tail) { /* FIFO breadth-first search from the goal placement */
            const uint32_t idx = queue[head++];
            int pos[5];
            for (int i = 0; i < nt; ++i) pos[i] = (int)(idx / POW25[i] % 25u);
            for (int i = 0; i < nt; ++i) {
                const int r = pos[i] / 5, c = pos[i] % 5;
                const int nr[4] = {r + 1, r - 1, r, r}, nc[4] = {c, c, c + 1, c - 1};
                for (int m = 0; m < 4; ++m) {
                    if (nr[m] < 0 || nr[m] > 4 || nc[m] < 0 || nc[m] > 4) continue;
                    const int np = nr[m] * 5 + nc[m];
                    int taken = 0;
                    for (int k = 0; k < nt; ++k) taken |= (k != i && pos[k] == np);
                    if (taken) continue;
                    const uint32_t nidx = idx + (uint32_t)(np - pos[i]) * POW25[i];
                    if (dist[nidx] != 0xFF) continue;
                    dist[nidx] = (uint8_t)(dist[idx] + 1);
                    queue[tail++] = nidx;
                }
            }
        }
        free(queue);
    }
    return ORACLE_OK;
}

static uint32_t pdb_h128(const uint8_t* pdb, u128 s) {
    uint32_t idx[ORACLE_PDB_PATTERNS] = {0, 0, 0, 0, 0};
    for (int c = 0; c < 25; ++c) {
        const uint32_t t = cell5(s, c);
        for (int p = 0; p < ORACLE_PDB_PATTERNS; ++p)
            for (int i = 0; i < PDB_NT[p]; ++i)
                if (PDB_TILES[p][i] == t && t) idx[p] += (uint32_t)c * POW25[i];
    }
    uint32_t h = 0;
    for (int p = 0; p < ORACLE_PDB_PATTERNS; ++p) h += pdb[pdb_offset(p) + idx[p]];
    return h;
}

uint32_t oracle_pdb24_h(const uint8_t* pdb, const uint8_t* board) { return pdb_h128(pdb, board_pack(5, board)); }

/* ---- additive pattern databases for 4x4 boards with the ordered goal ---- */
/* The same relaxation as the 5x5 tables above, for the 15-puzzle: its 15
 * tiles in three disjoint patterns of five; placement index: sum over the
 * pattern's tiles i of cell(tile i) * 16^i.  Goal: tile t on cell t - 1,
 * blank on cell 15. */
static const uint8_t PDB15_TILES[ORACLE_PDB15_PATTERNS][5] = {{1, 2, 3, 5, 6}, {4, 7, 8, 11, 12}, {9, 10, 13, 14, 15}};
#define PDB15_TABLE (1u << 20) /* 16^5 */

int oracle_pdb15_build(uint8_t* out) {
    if (!out) return ORACLE_E_INVALID;
    uint32_t* queue = (uint32_t*)malloc((size_t)PDB15_TABLE * sizeof(uint32_t));
    if (!queue) return ORACLE_E_NOMEM;
    for (int p = 0; p < ORACLE_PDB15_PATTERNS; ++p) {
        uint8_t* dist = out + (size_t)p * PDB15_TABLE;
        memset(dist, 0xFF, PDB15_TABLE);
        uint32_t goal = 0;
        for (int i = 0; i < 5; ++i) goal += (uint32_t)(PDB15_TILES[p][i] - 1) << (4 * i);
        size_t head = 0, tail = 0;
        dist[goal] = 0;
        queue[tail++] = goal;
        while (head < tail) { /* FIFO breadth-first search from the goal placement */
            const uint32_t idx = queue[head++];
            int pos[5];
            for (int i = 0; i < 5; ++i) pos[i] = (int)(idx >> (4 * i)) & 15;
            for (int i = 0; i < 5; ++i) {
                const int r = pos[i] / 4, c = pos[i] % 4;
                const int nr[4] = {r + 1, r - 1, r, r}, nc[4] = {c, c, c + 1, c - 1};
                for (int m = 0; m < 4; ++m) {
                    if (nr[m] < 0 || nr[m] > 3 || nc[m] < 0 || nc[m] > 3) continue;
                    const int np = nr[m] * 4 + nc[m];
                    int taken = 0;
                    for (int k = 0; k < 5; ++k) taken |= (k != i && pos[k] == np);
                    if (taken) continue;
                    const uint32_t nidx = (idx & ~(15u << (4 * i))) | ((uint32_t)np << (4 * i));
                    if (dist[nidx] != 0xFF) continue;
                    dist[nidx] = (uint8_t)(dist[idx] + 1);
                    queue[tail++] = nidx;
                }
            }
        }
    }
    free(queue);
    return ORACLE_OK;
}

static uint32_t pdb15_h_state(const uint8_t* pdb, uint64_t s) {
    uint8_t board[16];
    for (int c = 0; c < 16; ++c) board[c] = (uint8_t)((s >> (4 * c)) & 15u);
    return oracle_pdb15_h(pdb, board);
}

uint32_t oracle_pdb15_h(const uint8_t* pdb, const uint8_t* board) {
    uint32_t idx[ORACLE_PDB15_PATTERNS] = {0, 0, 0};
    for (int c = 0; c < 16; ++c)
        for (int p = 0; p < ORACLE_PDB15_PATTERNS; ++p)
            for (int i = 0; i < 5; ++i)
                if (PDB15_TILES[p][i] == board[c]) idx[p] += (uint32_t)c << (4 * i);
    uint32_t h = 0;
    for (int p = 0; p < ORACLE_PDB15_PATTERNS; ++p) h += pdb[(size_t)p * PDB15_TABLE + idx[p]];
    return h;
}

/* the heuristic of one board search: Manhattan without the blank, or the sum of the pattern databases */
typedef struct {
    int side;
    int gr[32], gc[32];
    const uint8_t* pdb;
} board_h;

static uint32_t board_h_of(const board_h* h, u128 s) {
    return h->pdb ? pdb_h128(h->pdb, s) : manh128(h->side, s, h->gr, h->gc);
}

static int astar_board_h(int side, const uint8_t* start_b, const uint8_t* goal_b, const uint8_t* pdb,
                         uint64_t max_states, oracle_result* res, uint8_t* path, uint32_t path_cap);

int oracle_astar_board(int side, const uint8_t* start_b, const uint8_t* goal_b, uint64_t max_states,
                       oracle_result* res, uint8_t* path, uint32_t path_cap) {
    return astar_board_h(side, start_b, goal_b, NULL, max_states, res, path, path_cap);
}

int oracle_astar_board_pdb(const uint8_t* start_b, const uint8_t* pdb, uint64_t max_states, oracle_result* res,
                           uint8_t* path, uint32_t path_cap) {
    uint8_t goal_b[25];
    for (int c = 0; c < 24; ++c) goal_b[c] = (uint8_t)(c + 1);
    goal_b[24] = 0;
    if (!pdb) return ORACLE_E_INVALID;
    return astar_board_h(5, start_b, goal_b, pdb, max_states, res, path, path_cap);
}

static int astar_board_h(int side, const uint8_t* start_b, const uint8_t* goal_b, const uint8_t* pdb,
                         uint64_t max_states, oracle_result* res, uint8_t* path, uint32_t path_cap) {
    memset(res, 0, sizeof(*res));
    res->cost = -1;
    if (!board_valid(side, start_b) || !board_valid(side, goal_b)) return ORACLE_E_INVALID;
    const u128 start = board_pack(side, start_b), goal = board_pack(side, goal_b);
    board_h H;
    H.side = side;
    H.pdb = pdb;
    goal_rc(side, goal, H.gr, H.gc);
    bstore st;
    memset(&st, 0, sizeof(st));
    pq q;
    memset(&q, 0, sizeof(q));
    int rc = ORACLE_NO_PATH;
    int64_t s0 = bstore_get(&st, start);
    if (s0 < 0) { rc = ORACLE_E_NOMEM; goto out; }
    st.g[s0] = 0;
    if (pq_push(&q, (pq_item){prio_of(board_h_of(&H, start), 0), (uint32_t)s0, 0})) { rc = ORACLE_E_NOMEM; goto out; }
    int64_t goal_idx = -1;
    while (q.n) {
        pq_item it = pq_pop(&q);
        if (it.g > st.g[it.idx]) continue; /* stale */
        u128 s = st.key[it.idx];
        if (s == goal) { goal_idx = it.idx; break; } /* goal test at pop: a_star.h:80-83 */
        res->expanded++;
        int b = board_blank(side, s);
        for (int i = 0; i < 4; ++i) {
            u128 t;
            if (!board_succ(side, s, b, i, &t)) continue;
            res->generated++;
            uint32_t g2 = it.g + 1;
            int64_t ti = bstore_get(&st, t);
            if (ti < 0) { rc = ORACLE_E_NOMEM; goto out; }
            if (max_states && st.n > max_states) { rc = ORACLE_E_LIMIT; goto out; }
            if (g2 >= st.g[ti]) continue; /* fancy_hashmap.h:56-61 */
            st.g[ti] = g2;
            st.parent[ti] = it.idx;
            if (pq_push(&q, (pq_item){prio_of(g2 + board_h_of(&H, t), g2), (uint32_t)ti, g2})) { rc = ORACLE_E_NOMEM; goto out; }
        }
    }
    res->stored = st.n;
    if (goal_idx >= 0) {
        rc = ORACLE_OK;
        res->cost = st.g[goal_idx];
        uint32_t len = 0;
        for (uint32_t i = (uint32_t)goal_idx; i != NONE32; i = st.parent[i]) ++len;
        res->path_len = len;
        if (path && path_cap >= len) {
            uint32_t k = len;
            for (uint32_t i = (uint32_t)goal_idx; i != NONE32; i = st.parent[i])
                board_unpack(side, st.key[i], path + (size_t)(--k) * side * side);
        }
    }
out:
    if (rc != ORACLE_OK && rc != ORACLE_NO_PATH) res->stored = st.n;
    bstore_free(&st);
    free(q.a);
    return rc;
}

int64_t oracle_bfs_board(int side, const uint8_t* start_b, const uint8_t* goal_b, uint32_t max_depth,
                         uint64_t max_states) {
    if (!board_valid(side, start_b) || !board_valid(side, goal_b)) return -2;
    const u128 start = board_pack(side, start_b), goal = board_pack(side, goal_b);
    if (start == goal) return 0;
    bstore st;
    memset(&st, 0, sizeof(st));
    int64_t result = -1;
    int64_t s0 = bstore_get(&st, start);
    if (s0 < 0) { result = -3; goto out; }
    st.g[s0] = 0;
    /* the store's insertion order is the BFS queue: node i is expanded after every node before it */
    for (size_t head = 0; head < st.n; ++head) {
        uint32_t d = st.g[head];
        if (d >= max_depth) break;
        u128 s = st.key[head];
        int b = board_blank(side, s);
        for (int i = 0; i < 4; ++i) {
            u128 t;
            if (!board_succ(side, s, b, i, &t)) continue;
            int64_t ti = bstore_get(&st, t);
            if (ti < 0) { result = -3; goto out; }
            if (st.g[ti] != INF32) continue;
            st.g[ti] = d + 1;
            if (t == goal) { result = d + 1; goto out; }
            if (max_states && st.n > max_states) { result = -3; goto out; }
        }
    }
out:
    bstore_free(&st);
    return result;
}

int64_t oracle_replay_board(int side, const uint8_t* path, uint32_t n, const uint8_t* start_b, const uint8_t* goal_b) {
    const size_t cells = (size_t)side * side;
    if (n == 0 || !board_valid(side, start_b) || !board_valid(side, goal_b)) return -1;
    if (memcmp(path, start_b, cells) || memcmp(path + (n - 1) * cells, goal_b, cells)) return -1;
    int64_t cost = 0;
    for (uint32_t k = 1; k < n; ++k) {
        if (!board_valid(side, path + k * cells)) return -1;
        u128 s = board_pack(side, path + (k - 1) * cells), want = board_pack(side, path + k * cells);
        int b = board_blank(side, s);
        int ok = 0;
        for (int i = 0; i < 4 && !ok; ++i) {
            u128 t;
            if (board_succ(side, s, b, i, &t) && t == want) ok = 1;
        }
        if (!ok) return -1;
        cost += 1;
    }
    return cost;
}
