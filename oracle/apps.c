/* ORACLE -- test infrastructure only (tests/, bench.py cpu_baseline).
 * CPU restatements of the heap consumers, used to check the GPU drivers:
 *   orc_dijkstra       textbook Dijkstra on a binary min-heap of
 *                      dist << 32 | vertex keys (the key of
 *                      applications/sssp_longfei_new/sssp.cu:6), stale keys
 *                      skipped; the reference ships no CPU SSSP
 *   orc_astar          sequential A* restated from
 *                      artifacts/astar/seq_astar.hpp:186-279 (open list of
 *                      f << 32 | node keys, closed list, stop when the goal
 *                      is popped) with the 10/14 costs and octile heuristic
 *   orc_knapsack_dp    textbook O(n * C) 0/1 knapsack dynamic program
 *   orc_knapsack_bnb   best-first branch-and-bound with the integer Dantzig
 *                      bound (artifacts/knapsack/knapsackKernel.cuh:4-91
 *                      uses a double bound), counting the nodes it pushes
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---- binary min-heap of u64 keys */
typedef struct {
    uint64_t* a;
    int64_t n, cap;
} orc_pq;

static int pq_push(orc_pq* q, uint64_t k) {
    if (q->n == q->cap) {
        int64_t nc = q->cap ? 2 * q->cap : 1024;
        uint64_t* na = (uint64_t*)realloc(q->a, (size_t)nc * sizeof(uint64_t));
        if (!na) return -1;
        q->a = na;
        q->cap = nc;
    }
    int64_t i = q->n++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (q->a[p] <= k) break;
        q->a[i] = q->a[p];
        i = p;
    }
    q->a[i] = k;
    return 0;
}

static uint64_t pq_pop(orc_pq* q) {
    uint64_t top = q->a[0];
    uint64_t last = q->a[--q->n];
    int64_t i = 0;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= q->n) break;
        if (c + 1 < q->n && q->a[c + 1] < q->a[c]) ++c;
        if (last <= q->a[c]) break;
        q->a[i] = q->a[c];
        i = c;
    }
    if (q->n > 0) q->a[i] = last;
    return top;
}

int orc_dijkstra(int32_t n, const int32_t* rowptr, const int32_t* col, const uint32_t* w, int32_t src,
                 uint32_t* dist, int64_t* pops) {
    orc_pq q = {0, 0, 0};
    for (int32_t v = 0; v < n; ++v) dist[v] = UINT32_MAX;
    dist[src] = 0;
    if (pq_push(&q, (uint64_t)src)) return -1;
    int64_t popped = 0;
    while (q.n > 0) {
        uint64_t k = pq_pop(&q);
        ++popped;
        uint32_t d = (uint32_t)(k >> 32), v = (uint32_t)k;
        if (d > dist[v]) continue; /* stale */
        for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            uint64_t nd = (uint64_t)d + w[e];
            uint32_t u = (uint32_t)col[e];
            if (nd >= UINT32_MAX) {
                free(q.a);
                return -2; /* distance does not fit */
            }
            if (nd < dist[u]) {
                dist[u] = (uint32_t)nd;
                if (pq_push(&q, nd << 32 | u)) return -1;
            }
        }
    }
    free(q.a);
    if (pops) *pops = popped;
    return 0;
}

static uint32_t octile(int32_t x, int32_t y, int32_t gx, int32_t gy) {
    uint32_t dx = (uint32_t)(x > gx ? x - gx : gx - x);
    uint32_t dy = (uint32_t)(y > gy ? y - gy : gy - y);
    uint32_t lo = dx < dy ? dx : dy, hi = dx < dy ? dy : dx;
    return 10u * hi + 4u * lo; /* 14*lo + 10*(hi-lo) */
}

int orc_astar(int32_t width, int32_t height, const uint8_t* free_cells, uint32_t* cost, int64_t* expanded) {
    const int64_t n = (int64_t)width * height;
    const int32_t gx = width - 1, gy = height - 1;
    const uint32_t goal = (uint32_t)(n - 1);
    uint32_t* g = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    uint8_t* closed = (uint8_t*)calloc((size_t)n, 1);
    if (!g || !closed) return -1;
    for (int64_t i = 0; i < n; ++i) g[i] = UINT32_MAX;
    orc_pq q = {0, 0, 0};
    g[0] = 0;
    pq_push(&q, (uint64_t)octile(0, 0, gx, gy) << 32 | 0u);
    int64_t exp = 0;
    *cost = UINT32_MAX;
    static const int dxs[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
    static const int dys[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    while (q.n > 0) {
        uint64_t k = pq_pop(&q);
        uint32_t node = (uint32_t)k;
        if (node == goal) {
            *cost = g[goal];
            break;
        }
        if (closed[node]) continue;
        closed[node] = 1;
        ++exp;
        const int32_t x = (int32_t)(node % (uint32_t)width), y = (int32_t)(node / (uint32_t)width);
        const uint32_t gn = g[node];
        for (int d = 0; d < 8; ++d) {
            const int32_t nx = x + dxs[d], ny = y + dys[d];
            if (nx < 0 || ny < 0 || nx >= width || ny >= height) continue;
            const uint32_t nb = (uint32_t)ny * (uint32_t)width + (uint32_t)nx;
            if (!free_cells[nb] || closed[nb]) continue;
            const uint32_t ng = gn + ((dxs[d] != 0 && dys[d] != 0) ? 14u : 10u);
            if (ng < g[nb]) {
                g[nb] = ng;
                pq_push(&q, (uint64_t)(ng + octile(nx, ny, gx, gy)) << 32 | nb);
            }
        }
    }
    if (expanded) *expanded = exp;
    free(q.a);
    free(g);
    free(closed);
    return 0;
}

int orc_knapsack_dp(int32_t n, const uint32_t* weight, const uint32_t* value, uint64_t capacity, uint64_t* best) {
    uint64_t* f = (uint64_t*)calloc((size_t)capacity + 1, sizeof(uint64_t));
    if (!f) return -1;
    for (int32_t i = 0; i < n; ++i) {
        const uint64_t wi = weight[i], vi = value[i];
        if (wi > capacity) continue;
        for (uint64_t c = capacity; c >= wi; --c) {
            const uint64_t t = f[c - wi] + vi;
            if (t > f[c]) f[c] = t;
            if (c == wi) break;
        }
    }
    *best = f[capacity];
    free(f);
    return 0;
}

/* items in order of value/weight, highest first (ties: lower index first) */
static void ratio_order(int32_t n, const uint32_t* w, const uint32_t* v, int32_t* ord) {
    for (int32_t i = 0; i < n; ++i) ord[i] = i;
    for (int32_t i = 1; i < n; ++i) { /* insertion sort: n <= 255 */
        int32_t x = ord[i], j = i - 1;
        while (j >= 0) {
            const int32_t y = ord[j];
            const uint64_t lhs = (uint64_t)v[x] * w[y], rhs = (uint64_t)v[y] * w[x];
            if (lhs > rhs || (lhs == rhs && x < y)) {
                ord[j + 1] = y;
                --j;
            } else {
                break;
            }
        }
        ord[j + 1] = x;
    }
}

/* Dantzig bound from item position `level` with room `room` left:
 * value + whole items while they fit + floor(room * v / w) of the next. */
static uint64_t dantzig(int32_t n, const uint32_t* w, const uint32_t* v, int32_t level, uint64_t room,
                        uint64_t value) {
    uint64_t ub = value;
    for (int32_t i = level; i < n; ++i) {
        if (w[i] <= room) {
            room -= w[i];
            ub += v[i];
        } else {
            ub += room * v[i] / w[i];
            break;
        }
    }
    return ub;
}

/* With `dominance`, a node is kept only if its value beats the best value
 * recorded for its (depth, used weight) cell, and a popped node whose cell
 * has since improved is stale: the search then visits at most n * (C + 1)
 * states, as the dynamic program does. */
int orc_knapsack_bnb(int32_t n, const uint32_t* weight, const uint32_t* value, uint64_t capacity,
                     int64_t max_nodes, int dominance, uint64_t* best, int64_t* nodes) {
    if (n < 1 || n > 255) return -1;
    int32_t ord[255];
    uint32_t w[255], v[255];
    ratio_order(n, weight, value, ord);
    for (int32_t i = 0; i < n; ++i) {
        w[i] = weight[ord[i]];
        v[i] = value[ord[i]];
    }
    typedef struct {
        uint64_t room, value;
        int32_t level;
    } node_t;
    int64_t cap = 1 << 16, cnt = 0;
    node_t* pool = (node_t*)malloc((size_t)cap * sizeof(node_t));
    orc_pq q = {0, 0, 0};
    uint64_t inc = 0;
    const uint64_t kUbTop = (1ull << 24) - 1;
    uint32_t* cell = NULL; /* best value + 1 per (depth, used weight); 0 = none */
    if (dominance) {
        cell = (uint32_t*)calloc((size_t)(n + 1) * (capacity + 1), sizeof(uint32_t));
        if (!cell) return -1;
        cell[0] = 1; /* the root: depth 0, nothing used, value 0 */
    }
    pool[cnt] = (node_t){capacity, 0, 0};
    uint64_t ub0 = dantzig(n, w, v, 0, capacity, 0);
    pq_push(&q, ((((kUbTop - ub0) << 8) | 255u) << 32) | (uint64_t)cnt);
    ++cnt;
    int rc = 0;
    while (q.n > 0) {
        uint64_t k = pq_pop(&q);
        node_t nd = pool[(uint32_t)k];
        uint64_t ub = kUbTop - (k >> 40);
        if (ub <= inc) continue; /* pruned */
        if (nd.level >= n) continue;
        if (cell && cell[(size_t)nd.level * (capacity + 1) + (capacity - nd.room)] != nd.value + 1) continue;
        for (int take = 1; take >= 0; --take) {
            node_t c = nd;
            c.level = nd.level + 1;
            if (take) {
                if (w[nd.level] > nd.room) continue;
                c.room -= w[nd.level];
                c.value += v[nd.level];
            }
            if (c.value > inc) inc = c.value;
            if (c.level >= n) continue;
            uint64_t cub = dantzig(n, w, v, c.level, c.room, c.value);
            if (cub <= inc) continue;
            if (cell) {
                uint32_t* slot = &cell[(size_t)c.level * (capacity + 1) + (capacity - c.room)];
                if (*slot >= c.value + 1) continue; /* dominated */
                *slot = (uint32_t)c.value + 1;
            }
            if (cnt == cap) {
                if (max_nodes > 0 && cnt >= max_nodes) {
                    rc = -3;
                    goto done;
                }
                cap *= 2;
                node_t* np = (node_t*)realloc(pool, (size_t)cap * sizeof(node_t));
                if (!np) {
                    rc = -1;
                    goto done;
                }
                pool = np;
            }
            pool[cnt] = c;
            pq_push(&q, ((((kUbTop - cub) << 8) | (uint64_t)(255 - c.level)) << 32) | (uint64_t)cnt);
            ++cnt;
        }
    }
done:
    free(cell);
    *best = inc;
    if (nodes) *nodes = cnt;
    free(pool);
    free(q.a);
    return rc;
}
