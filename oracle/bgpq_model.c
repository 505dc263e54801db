/* ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this code, and only as the
 * checker; the product never calls it.
 *
 * Sequential CPU restatement of the reference's batched heap (BGPQ), all
 * four variants, executing one operation at a time.  With a single workgroup
 * the GPU heap performs the same data movement, so its state after every
 * operation must equal this model's state bit for bit; with many workgroups
 * only the outputs (sorted) and the invariants are comparable.
 *
 * Followed reference files:
 *   heap/heaputil.cuh:23-47    ibitonicSort  -> sorted batch (any sort: result is the multiset, ascending)
 *   heap/heaputil.cuh:99-158   imergePath    -> merge2 (ties take a first; unobservable for scalar keys)
 *   heap/buitdd/heap.cuh:233-255  getReversedIdx / hostGetReversedIdx -> rev_idx
 *   heap/buitdd/heap.cuh:283-343  deleteRoot   (tditdd:267-341, *_np without the partial-only branch)
 *   heap/buitdd/heap.cuh:348-482  deleteUpdate (buitdd_np:288-423 drops the partial merge and stops on <=);
 *                                 split into orc_take_last / orc_root_out / orc_sift_step so that tests can
 *                                 script interleavings of concurrent deleters
 *   heap/tditdd/heap.cuh:345-534  deleteUpdate (natural indexing, stops on <=)
 *   heap/buitdd/heap.cuh:484-690  insertion, bottom-up (partial cases 1.1-2.2 at :522-613)
 *   heap/tditdd/heap.cuh:536-753  insertion, top-down toward the target slot
 *   heap/buitdd/heap.cuh:118-172  checkInsertHeap (tditdd:123-176)
 *
 * Keys are held as uint64_t; u32 heaps use init_limits = 0xFFFFFFFF.
 * Children beyond batch_num are treated as holding init_limits (the
 * reference instead over-allocates the array, test/heap_test.cu:97-100).
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

uint32_t orc_rev_idx(uint32_t x) {
    if (x <= 1) return x;
    int n = 0;
    for (uint32_t y = x; y; y >>= 1) ++n;           /* bit length */
    uint32_t r = 0;
    for (int i = 0; i < n - 1; ++i) r |= ((x >> i) & 1u) << (n - 2 - i);  /* reverse the low n-1 bits */
    return r | (1u << (n - 1));
}

static int is_bu(int v) { return v == ORC_BUITDD || v == ORC_BUITDD_NP; }
static int has_pb(int v) { return v == ORC_BUITDD || v == ORC_TDITDD; }

static uint64_t* node(orc_heap* m, int i) { return m->items + (size_t)i * m->B; }

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Merge sorted a[B], b[B]; smallest B -> small, largest B -> large (either may
 * be NULL; either may alias a or b). */
static void merge2(orc_heap* m, const uint64_t* a, const uint64_t* b, uint64_t* small, uint64_t* large) {
    const int B = m->B;
    uint64_t* t = m->tmp;
    int i = 0, j = 0;
    for (int k = 0; k < 2 * B; ++k) {
        if (j >= B || (i < B && a[i] <= b[j])) t[k] = a[i++];
        else t[k] = b[j++];
    }
    if (small) memcpy(small, t, sizeof(uint64_t) * B);
    if (large) memcpy(large, t + B, sizeof(uint64_t) * B);
}

static void fill(uint64_t* p, int n, uint64_t v) {
    for (int i = 0; i < n; ++i) p[i] = v;
}

orc_heap* orc_create(int variant, int batch_num, int B, uint64_t init_limits) {
    if (variant < 0 || variant > 3 || batch_num < 1 || B < 1) return NULL;
    orc_heap* m = (orc_heap*)calloc(1, sizeof(orc_heap));
    m->variant = variant;
    m->batch_num = batch_num;
    m->B = B;
    m->init = init_limits;
    m->items = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)B * (size_t)(batch_num + 1));
    m->x0 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)B * 4);
    m->x1 = m->x0 + B;
    m->x2 = m->x1 + B;
    m->tmp = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)B * 2);
    orc_reset(m);
    return m;
}

void orc_reset(orc_heap* m) {
    fill(m->items, m->B * (m->batch_num + 1), m->init);
    m->batch_count = 0;
    m->partial = 0;
    m->delete_count = 0;
    m->error = 0;
}

void orc_destroy(orc_heap* m) {
    if (!m) return;
    free(m->items);
    free(m->x0);
    free(m->tmp);
    free(m);
}

static int left_child(const orc_heap* m, int cur) {
    return is_bu(m->variant) ? (int)orc_rev_idx(orc_rev_idx((uint32_t)cur) << 1) : cur << 1;
}
static int right_child(const orc_heap* m, int cur) {
    return is_bu(m->variant) ? (int)orc_rev_idx(orc_rev_idx((uint32_t)cur) << 1 | 1u) : (cur << 1) | 1;
}
int orc_parent(int variant, int cur) {
    return (variant == ORC_BUITDD || variant == ORC_BUITDD_NP) ? (int)orc_rev_idx(orc_rev_idx((uint32_t)cur) >> 1)
                                                               : cur >> 1;
}

/* Child of cur on the path to target, natural indexing (tditdd:263-265). */
static int next_toward(int cur, int target) {
    int lc = 0, lt = 0;
    for (int y = cur; y; y >>= 1) ++lc;
    for (int y = target; y; y >>= 1) ++lt;
    return target >> (lt - lc - 1);
}

int orc_insert(orc_heap* m, const uint64_t* keys, int size) {
    const int B = m->B;
    if (size <= 0) return 0;
    if (size > B) {
        m->error |= 2;
        return -1;
    }
    uint64_t* X0 = m->x0;
    uint64_t* X1 = m->x1;
    for (int i = 0; i < B; ++i) X0[i] = i < size ? keys[i] : m->init;
    qsort(X0, (size_t)B, sizeof(uint64_t), cmp_u64);
    const int nb = m->batch_count, pbs = m->partial;
    /* A batch that goes into the tree (case 2.2, or any insertion without a
     * partial buffer) needs slot nb + 1.  The reference does not check (it
     * writes past its array); here an overflow changes nothing and sets
     * error 1, as the GPU heap does. */
    const int to_tree = !(has_pb(m->variant) && (size + pbs < B || (nb == 0 && size < B)));
    if (to_tree && nb + 1 > m->batch_num) {
        m->error |= 1;
        return -1;
    }
    if (has_pb(m->variant)) {
        if (nb == 0 && size < B) {
            if (pbs == 0) { /* case 1.1 */
                memcpy(node(m, 0), X0, sizeof(uint64_t) * B);
                m->partial = size;
                return 0;
            }
            memcpy(X1, node(m, 0), sizeof(uint64_t) * B);
            if (size + pbs < B) { /* case 1.2 */
                merge2(m, X0, X1, node(m, 0), NULL);
                m->partial = pbs + size;
                return 0;
            }
            /* case 1.3 */
            m->batch_count += 1;
            merge2(m, X0, X1, node(m, 1), node(m, 0));
            m->partial = pbs + size - B;
            return 0;
        }
        memcpy(X1, node(m, 0), sizeof(uint64_t) * B);
        if (size + pbs < B) { /* case 2.1 */
            merge2(m, X0, X1, X0, NULL);
            memcpy(X1, node(m, 1), sizeof(uint64_t) * B);
            merge2(m, X0, X1, node(m, 1), node(m, 0));
            m->partial = pbs + size;
            return 0;
        }
        /* case 2.2 */
        merge2(m, X0, X1, X0, node(m, 0));
        m->partial = pbs + size - B;
    }
    const int idx = ++m->batch_count;
    if (is_bu(m->variant)) {
        memcpy(node(m, idx), X0, sizeof(uint64_t) * B);
        int cur = idx;
        while (cur != 1) {
            const int par = orc_parent(m->variant, cur);
            /* merge(cur, parent): smaller half to the parent (buitdd:647-675) */
            memcpy(X0, node(m, cur), sizeof(uint64_t) * B);
            memcpy(X1, node(m, par), sizeof(uint64_t) * B);
            merge2(m, X0, X1, node(m, par), node(m, cur));
            cur = par;
        }
    } else {
        int cur = 1;
        while (cur != idx) {
            /* carried batch X0 vs node: smaller half stays (tditdd:698-717) */
            memcpy(X1, node(m, cur), sizeof(uint64_t) * B);
            merge2(m, X0, X1, node(m, cur), X0);
            cur = next_toward(cur, idx);
        }
        memcpy(node(m, idx), X0, sizeof(uint64_t) * B);
    }
    return 0;
}

int orc_delete_root(orc_heap* m, uint64_t* out, int* size) {
    const int B = m->B;
    const int off = m->delete_count;
    if (m->batch_count == 0 && m->partial == 0) {
        *size = 0;
        return 0;
    }
    if (m->batch_count == 0) {
        if (!has_pb(m->variant)) { /* unreachable: _np heaps never hold a partial buffer */
            *size = 0;
            return 0;
        }
        const int pbs = m->partial;
        for (int i = 0; i < pbs; ++i) {
            out[off + i] = node(m, 0)[i];
            node(m, 0)[i] = m->init;
        }
        m->delete_count += pbs;
        m->partial = 0;
        *size = pbs;
        return 0;
    }
    memcpy(out + off, node(m, 1), sizeof(uint64_t) * B);
    m->delete_count += B;
    *size = B;
    return 1;
}

/* One level step of the top-down sift (buitdd:405-480, tditdd:419-531): the
 * deleter holds node cur and carries candidate cand.  The two children are
 * merged, the larger half goes to the child with the larger maximum, and
 * the smaller half competes with the candidate.  Returns the node the sift
 * continues at, or 0 when it ends here. */
int orc_sift_step(orc_heap* m, int cur, uint64_t* cand) {
    const int B = m->B;
    uint64_t* X0 = cand;
    uint64_t* X1 = m->x1;
    uint64_t* X2 = m->x2;
    const int l = left_child(m, cur), r = right_child(m, cur);
    const int lin = l <= m->batch_num, rin = r <= m->batch_num;
    if (lin) memcpy(X1, node(m, l), sizeof(uint64_t) * B);
    else fill(X1, B, m->init);
    if (rin) memcpy(X2, node(m, r), sizeof(uint64_t) * B);
    else fill(X2, B, m->init);
    const int tr = X1[B - 1] < X2[B - 1];
    const int tgt = tr ? r : l, oth = tr ? l : r;
    const int tgt_in = tr ? rin : lin, oth_in = tr ? lin : rin;
    merge2(m, X1, X2, X1, tgt_in ? node(m, tgt) : NULL);
    const int stop = (m->variant == ORC_BUITDD) ? (X0[B - 1] < X1[0]) : (X0[B - 1] <= X1[0]);
    if (X0[0] >= X1[B - 1]) {
        memcpy(node(m, cur), X1, sizeof(uint64_t) * B);
    } else if (stop) {
        memcpy(node(m, cur), X0, sizeof(uint64_t) * B);
        if (oth_in) memcpy(node(m, oth), X1, sizeof(uint64_t) * B);
        return 0;
    } else {
        merge2(m, X0, X1, node(m, cur), X0);
    }
    if (!oth_in) {
        m->error |= 4;
        return 0;
    }
    return oth;
}

/* The last batch leaves as the deleter's candidate, with the partial
 * buffer folded in (its larger half stays there): buitdd:357-395.  Returns
 * the slot taken, 0 when the heap holds no full batch. */
int orc_take_last(orc_heap* m, uint64_t* cand) {
    const int B = m->B;
    const int last = m->batch_count;
    if (last < 1) return 0;
    memcpy(cand, node(m, last), sizeof(uint64_t) * B);
    fill(node(m, last), B, m->init);
    m->batch_count--;
    if (has_pb(m->variant)) merge2(m, cand, node(m, 0), cand, node(m, 0));
    return last;
}

/* The root batch leaves at the heap-sort cursor.  With a candidate taken
 * before the root (the GPU deleteRoot's early take of the last batch), the
 * root check hands out the B smallest of root and candidate instead, and
 * the candidate keeps the B largest; `check` = 0 leaves that check out, to
 * show what it repairs.  Returns 1 when the check merged. */
int orc_root_out(orc_heap* m, uint64_t* out, uint64_t* cand, int check) {
    const int B = m->B, off = m->delete_count;
    memcpy(out + off, node(m, 1), sizeof(uint64_t) * B);
    m->delete_count += B;
    if (cand && check && cand[0] < node(m, 1)[B - 1]) {
        merge2(m, cand, node(m, 1), out + off, cand);
        return 1;
    }
    return 0;
}

void orc_delete_update(orc_heap* m) {
    uint64_t* X0 = m->x0;
    if (m->batch_count == 1) {
        m->batch_count = 0;
        return;
    }
    orc_take_last(m, X0);
    int cur = 1;
    while ((cur = orc_sift_step(m, cur, X0)) > 0) {
    }
}

/* checkInsertHeap restated; returns 1 if every invariant holds. */
int orc_check(const orc_heap* m, char* why, int why_len) {
    const int B = m->B, nb = m->batch_count, pbs = m->partial;
    const uint64_t* it = m->items;
#define ORC_FAIL(...)                                       \
    do {                                                    \
        if (why && why_len > 0) snprintf(why, (size_t)why_len, __VA_ARGS__); \
        return 0;                                           \
    } while (0)
    if (m->error) ORC_FAIL("error flags 0x%x", m->error);
    if (nb < 0 || nb > m->batch_num) ORC_FAIL("batch count %d", nb);
    if (pbs < 0 || pbs >= B) ORC_FAIL("partial size %d", pbs);
    if (nb != 0 && pbs != 0) {
        if (it[(size_t)2 * B - 1] > it[0]) ORC_FAIL("partial buffer below the root's last key");
        for (int i = 1; i < pbs; ++i)
            if (it[i] < it[i - 1]) ORC_FAIL("partial buffer not ascending at %d", i);
    }
    for (int i = 1; i <= nb; ++i) {
        if (i > 1) {
            const int p = orc_parent(m->variant, i);
            if (it[(size_t)i * B] < it[(size_t)p * B + B - 1]) ORC_FAIL("batch %d below parent %d", i, p);
        }
        for (int j = 1; j < B; ++j)
            if (it[(size_t)i * B + j] < it[(size_t)i * B + j - 1]) ORC_FAIL("batch %d not ascending at %d", i, j);
    }
#undef ORC_FAIL
    return 1;
}

/* Whole heap sort through the model, mirroring the harness: insert n keys
 * in `chunk`-sized calls, then delete until empty.  out receives n keys. */
int orc_heapsort(int variant, int B, int batch_num, uint64_t init_limits, const uint64_t* keys, int64_t n,
                 int chunk, uint64_t* out) {
    orc_heap* m = orc_create(variant, batch_num, B, init_limits);
    if (!m) return -1;
    if (chunk <= 0) chunk = B;
    for (int64_t i = 0; i < n; i += chunk) {
        const int sz = (int)((n - i) < chunk ? (n - i) : chunk);
        if (orc_insert(m, keys + i, sz)) {
            orc_destroy(m);
            return -2;
        }
    }
    int size;
    while (1) {
        const int more = orc_delete_root(m, out, &size);
        if (size == 0) break;
        if (more) orc_delete_update(m);
    }
    const int err = m->error;
    const int64_t got = m->delete_count;
    orc_destroy(m);
    if (err) return -3;
    return got == n ? 0 : -4;
}
