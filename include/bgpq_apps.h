/* Heap consumers on the GPU: single-source shortest paths, grid A* and 0/1
 * knapsack branch-and-bound, each one persistent kernel driving one
 * bgpq_heap through the workgroup-collective device API, plus the host
 * generators of their synthetic inputs.
 *
 * Counterparts in the reference (each is a host relaunch loop around a
 * persistent kernel there; here one launch runs each solve to completion):
 *   bgpq_sssp      applications/sssp_longfei_new/main.cu:24-113, sssp.cu:19-136
 *   bgpq_astar     artifacts/astar/astar_main.cu:12-196, pq_model.cuh:114-427
 *   bgpq_knapsack  artifacts/knapsack/main.cu:22-100, models_pq.cuh:50-256,
 *                  knapsackKernel.cuh:4-91
 *
 * The heap passed in must hold u64 keys, use a variant with a partial
 * buffer (BGPQ_BUITDD or BGPQ_TDITDD: pops and pushes are of any size), be
 * empty, and have a batch size in [256, 2048].  It is empty again when the
 * call returns successfully.  Every call is stream-ordered and returns when
 * the solve has finished.
 */
#ifndef BGPQ_APPS_H
#define BGPQ_APPS_H
#include <stdint.h>

#include "bgpq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bgpq_app_stats {
    int64_t popped;    /* keys taken out of the heap */
    int64_t expanded;  /* popped keys that were current and got expanded */
    int64_t stale;     /* popped keys superseded by a better label */
    int64_t pruned;    /* popped keys cut by the incumbent (A*, knapsack) */
    int64_t pushed;    /* keys inserted */
    int32_t error_flags;
    float ms;          /* device time of the solve (HIP events) */
} bgpq_app_stats;

/* ---- drivers */

/* Dijkstra with the heap as frontier: keys are dist << 32 | vertex (the
 * reference's MAKE_KEY, sssp.cu:6).  CSR with int32 indices, u32 weights;
 * d_dist[n_vertices] receives the distances (UINT32_MAX = unreachable).
 * grid is the number of workgroups (0: 64).  Fails with BGPQ_EHEAP if a distance would not
 * fit in 32 bits or the heap overflowed. */
int bgpq_sssp(bgpq_heap* h, const int32_t* d_rowptr, const int32_t* d_col, const uint32_t* d_w,
              int32_t n_vertices, int32_t source, uint32_t* d_dist, int grid, bgpq_app_stats* stats,
              void* stream);

/* A* on an 8-connected width x height grid from cell 0 to the last cell.
 * d_free[cell] != 0 marks free cells (row-major, cell = y * width + x);
 * moves cost 10 orthogonally and 14 diagonally into any free cell, the
 * heuristic is the octile distance.  Keys are f << 32 | cell.  Workgroups
 * pop only keys whose f is below the goal's label (the reference's
 * threshold deleteRoot, artifacts/astar/BGPQ_B/heap.cuh:260-287), and the
 * search ends by the heap's termination vote; the keys it leaves in the
 * heap are counted as pruned, and the heap is cleared.  d_g[width*height]
 * is scratch for the g labels; *cost receives the optimal cost
 * (UINT32_MAX if the goal is unreachable). */
int bgpq_astar(bgpq_heap* h, const uint8_t* d_free, int32_t width, int32_t height, uint32_t* d_g, int grid,
               uint32_t* cost, bgpq_app_stats* stats, void* stream);

/* Best-first 0/1 knapsack branch-and-bound with an integer Dantzig bound.
 * Items (host arrays, n <= 255) are ordered by value/weight internally.
 * Keys are (bound rank << 8 | depth rank) << 32 | node: the highest bound
 * first, deeper nodes first among equal bounds.  Nodes live in a pool of
 * pool_nodes entries (0 picks a default); *best receives the optimum.
 * Fails with BGPQ_EHEAP if the pool or the heap overflowed. */
int bgpq_knapsack(bgpq_heap* h, int32_t n, const uint32_t* weight, const uint32_t* value, uint64_t capacity,
                  int64_t pool_nodes, int grid, uint64_t* best, bgpq_app_stats* stats, void* stream);

/* bgpq_knapsack without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid it).  No counterpart in the reference.  Same arguments, checks and
 * results as bgpq_knapsack, but the control words, item arrays, prefix sums,
 * dominance cells and node pool live in one 16 MiB array the library keeps
 * per device for the life of the process (the heap is still the caller's):
 * an instance whose (n + 1) * (capacity + 1) dominance cells and node pool
 * do not fit is refused with BGPQ_EINVAL; pool_nodes = 0 takes what fits.
 * Calls that overlap on one device share that array and race; calls one
 * after another start from a cleared state.  No HIP events are used: stats->ms
 * is host time around the solve. */
int bgpq_knapsack_noalloc(bgpq_heap* h, int32_t n, const uint32_t* weight, const uint32_t* value, uint64_t capacity,
                          int64_t pool_nodes, int grid, uint64_t* best, bgpq_app_stats* stats, void* stream);

/* Summary of SSSP labels computed on the device, the same one the serving
 * path reports (bgpq_result.digest and .value): d_out2[0] = sum over v of
 * (dist[v] + 1) * (2v + 1) mod 2^64, d_out2[1] = vertices reached
 * (dist != UINT32_MAX).  d_out2 is a caller-owned device buffer of two
 * u64; the result is stream-ordered. */
int bgpq_label_digest(const uint32_t* d_labels, int64_t n, uint64_t* d_out2, void* stream);

/* ---- host generators of the synthetic inputs (deterministic for a seed) */

/* R-MAT graph with 2^scale vertices and degree * 2^scale directed edges,
 * quadrant probabilities a, b, c (d = 1 - a - b - c), weights uniform in
 * [wmin, wmax].  Writes CSR in host memory: rowptr[2^scale + 1], col[E],
 * w[E]; edges of a vertex keep generation order.  Edge e draws from its own
 * SplitMix64 stream seeded with seed + (e + 1) * 0x9E3779B97F4A7C15: per
 * level one 32-bit draw (high half first) against the thresholds
 * floor(p * 2^32), then one 64-bit draw for the weight, wmin + x % span. */
int bgpq_gen_rmat(int scale, int degree, double a, double b, double c, uint64_t seed, uint32_t wmin,
                  uint32_t wmax, int32_t* rowptr, int32_t* col, uint32_t* w);

/* Grid of width x height cells, each an obstacle with probability p
 * (cell i: the high 32 bits of the first SplitMix64 draw seeded with
 * seed + (i + 1) * 0x9E3779B97F4A7C15, below floor(p * 2^32)); the first
 * and the last cell are forced free.  free_cells[i] = 1 for free. */
int bgpq_gen_grid(int width, int height, double p, uint64_t seed, uint8_t* free_cells);

/* Knapsack instance: weights uniform in [wmin, wmax] (item i draws from the
 * stream of seed + (i + 1) * 0x9E3779B97F4A7C15), value = weight + value_add,
 * capacity = floor(sum of weights / 2). */
int bgpq_gen_knapsack(int n, uint64_t seed, uint32_t wmin, uint32_t wmax, uint32_t value_add, uint32_t* weight,
                      uint32_t* value, uint64_t* capacity);

#ifdef __cplusplus
}
#endif
#endif
