/* ORACLE -- test infrastructure only (tests/, smoke(), bench.py cpu_baseline).
 * CPU restatements used to check the GPU product; never linked into it. */
#ifndef BGPQ_ORACLE_H
#define BGPQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BUITDD = 0, ORC_BUITDD_NP = 1, ORC_TDITDD = 2, ORC_TDITDD_NP = 3 };

typedef struct orc_heap {
    int variant, batch_num, B;
    uint64_t init;
    uint64_t* items; /* (batch_num+1)*B */
    int batch_count, partial, delete_count, error;
    uint64_t *x0, *x1, *x2, *tmp;
} orc_heap;

/* bgpq_model.c */
uint32_t orc_rev_idx(uint32_t x);
int orc_parent(int variant, int cur);
orc_heap* orc_create(int variant, int batch_num, int B, uint64_t init_limits);
void orc_reset(orc_heap* m);
void orc_destroy(orc_heap* m);
int orc_insert(orc_heap* m, const uint64_t* keys, int size);
int orc_delete_root(orc_heap* m, uint64_t* out, int* size);
void orc_delete_update(orc_heap* m);
int orc_take_last(orc_heap* m, uint64_t* cand);
int orc_root_out(orc_heap* m, uint64_t* out, uint64_t* cand, int check);
int orc_sift_step(orc_heap* m, int cur, uint64_t* cand);
int orc_check(const orc_heap* m, char* why, int why_len);
int orc_heapsort(int variant, int B, int batch_num, uint64_t init_limits, const uint64_t* keys, int64_t n,
                 int chunk, uint64_t* out);

/* keys.c: SplitMix64 and the key generators of the workloads */
uint64_t orc_splitmix64(uint64_t* state);
void orc_keys_u32(uint64_t seed, int64_t n, uint32_t* out);
void orc_keys_u64(uint64_t seed, int64_t n, uint64_t* out);

/* apps.c: CPU restatements of the heap consumers */
int orc_dijkstra(int32_t n, const int32_t* rowptr, const int32_t* col, const uint32_t* w, int32_t src,
                 uint32_t* dist, int64_t* pops);
int orc_astar(int32_t width, int32_t height, const uint8_t* free_cells, uint32_t* cost, int64_t* expanded);
int orc_knapsack_dp(int32_t n, const uint32_t* weight, const uint32_t* value, uint64_t capacity, uint64_t* best);
int orc_knapsack_bnb(int32_t n, const uint32_t* weight, const uint32_t* value, uint64_t capacity,
                     int64_t max_nodes, int dominance, uint64_t* best, int64_t* nodes);

#ifdef __cplusplus
}
#endif
#endif
