/* Batched and served heap solves: persistent kernels run many independent
 * queries (SSSP, A*, knapsack), each on its own device-resident heap,
 * solved by a team of workgroups.  New queries join while others run: the
 * host posts them into a ring in pinned host memory, which one idle
 * workgroup at a time reads for the others.  Device memory for every query (heap, labels,
 * search state) comes from one arena whose size is the HBM budget; a query
 * that does not fit waits (BGPQ_EAGAIN) until completions free memory.
 *
 * The reference has no counterpart: each of its applications is one heap
 * per launch inside a host relaunch loop (artifacts/astar/astar_main.cu:
 * 128-178, artifacts/knapsack/models_pq.cuh:199-246,
 * applications/sssp_longfei_new/main.cu:96-102).
 *
 * Queries use the buitdd heap variant with u64 keys and a batch size of
 * 256, 512 or 1024.  There are two serve kernels (tiers), each with its
 * own slot queue: 256-thread workgroups for B = 256 and 512, 512-thread
 * workgroups for B = 1024.
 */
#ifndef BGPQ_SERVE_H
#define BGPQ_SERVE_H
#include <stdint.h>

#include "bgpq.h"
#include "bgpq_apps.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BGPQ_EAGAIN 1 /* not an error: the query does not fit right now */

enum { BGPQ_Q_SSSP = 0, BGPQ_Q_ASTAR = 1, BGPQ_Q_KNAPSACK = 2 };

typedef struct bgpq_query {
    int32_t type;       /* BGPQ_Q_* */
    int32_t batch_size; /* 256, 512 or 1024 */
    int32_t team;       /* workgroups solving it together, 1..64 */
    int32_t reserved;
    int64_t heap_keys;  /* heap capacity in keys; 0: edges (SSSP), 4 * cells (A*), pool (knapsack) */
    /* SSSP: CSR graph in device memory */
    const int32_t* rowptr;
    const int32_t* col;
    const uint32_t* weight;
    int32_t n_vertices;
    int32_t source;
    int64_t n_edges;
    /* A*: obstacle map in device memory, start cell 0, goal the last cell */
    const uint8_t* free_cells;
    int32_t width;
    int32_t height;
    /* knapsack: items in host memory */
    const uint32_t* item_weight;
    const uint32_t* item_value;
    int32_t n_items;
    int32_t reserved2;
    uint64_t capacity;
    int64_t pool_nodes; /* 0: 2^20 */
    /* optional host buffer for the labels on completion: SSSP distances
     * (n_vertices) or A* g labels (width * height) */
    uint32_t* labels_out;
    uint64_t tag; /* returned with the result */
} bgpq_query;

typedef struct bgpq_result {
    uint64_t tag;
    int32_t type;
    int32_t status;    /* BGPQ_OK, or BGPQ_EHEAP when the solve raised error flags */
    uint64_t value;    /* A*: optimal cost; knapsack: optimum; SSSP: reachable vertices */
    uint64_t digest;   /* SSSP: sum over v of (dist[v] + 1) * (2v + 1) mod 2^64; else 0 */
    bgpq_app_stats stats;
    double t_submit;   /* bgpq_now() when submitted */
    double t_done;     /* bgpq_now() when bgpq_server_poll saw it complete */
} bgpq_result;

typedef struct bgpq_server bgpq_server;

/* Host clock of the server (monotonic seconds). */
double bgpq_now(void);

/* HBM footprint of a query in the arena (bytes), or a negative status. */
int64_t bgpq_query_bytes(const bgpq_query* q);

/* hbm_budget: arena bytes; workgroups: size of each tier's persistent
 * kernel (0: one per CU); queue_slots: most queries in flight (0: 4096).
 * Inputs the queries point to (CSR graphs, obstacle maps) must be in device
 * memory before bgpq_server_start: the kernel reads them with plain loads. */
int bgpq_server_create(bgpq_server** out, int64_t hbm_budget, int workgroups, int queue_slots);
void bgpq_server_destroy(bgpq_server* s);
/* Idle polls (each about 3.4 us) a workgroup of a persistent kernel makes
 * while nothing is posted before it leaves the kernel; 0 restores the
 * default (about a minute).  Set before bgpq_server_start. */
int bgpq_server_set_idle_limit(bgpq_server* s, int64_t polls);
/* Launch the persistent kernels (each tier on its own stream). */
int bgpq_server_start(bgpq_server* s);
/* Workgroups of the running kernels (both tiers) that have not left them. */
int bgpq_server_live(const bgpq_server* s);
/* Prepare the query's state in the arena and post it.  Returns BGPQ_OK,
 * BGPQ_EAGAIN when the arena or the queue is full now, or an error
 * (BGPQ_EHEAP once every workgroup of the query's tier has left at the
 * idle limit: nothing could take the query).  A workgroup takes a member
 * slot only once it is posted, so a posted query is solved while any
 * workgroup of its tier is left. */
int bgpq_server_submit(bgpq_server* s, const bgpq_query* q);
/* Collect up to max_results completed queries (their arena memory is
 * freed); returns how many were written, or BGPQ_EHEAP when none completed
 * and the queries still in flight never can (no workgroup is left). */
int bgpq_server_poll(bgpq_server* s, bgpq_result* out, int max_results);
/* Queries submitted and not yet collected. */
int bgpq_server_inflight(const bgpq_server* s);
/* Arena bytes in use. */
int64_t bgpq_server_used_bytes(const bgpq_server* s);
/* No more submissions: the kernel ends once every posted query is done.
 * Waits for it; collect the last results with bgpq_server_poll.  Fails
 * with BGPQ_EHEAP if a posted query never completed. */
int bgpq_server_close(bgpq_server* s);

/* Offline batch: every query of the list at once (they must all fit in
 * hbm_budget), prepared and solved by one captured hipGraph
 * [prepare kernels -> persistent kernel over a closed device-side ring],
 * replayed `repeats` times.  Each replay's answers are copied back and must
 * equal the first replay's (BGPQ_EHEAP otherwise); the last replay's
 * results are written to out[n].  ms_per_run receives the mean device time
 * of one replay. */
int bgpq_batch_run(const bgpq_query* q, int n, int64_t hbm_budget, int workgroups, int repeats, bgpq_result* out,
                   float* ms_per_run);

#ifdef __cplusplus
}
#endif
#endif
