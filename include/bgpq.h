/* C ABI of the MI355X batched GPU priority queue (BGPQ).
 *
 * The reference exposes its heap only as a header-only CUDA class template,
 * Heap<K>, one include directory per variant, plus harness kernels that drive
 * it.  This header is the boundary that replaces that path: every entry point
 * takes plain pointers and sizes, device buffers are raw device pointers, and
 * `stream` is a hipStream_t (NULL = the default stream).
 *
 * Status codes: 0 = OK, negative = error (bgpq_last_error() has the text).
 * Handles are not thread-safe; all operations on one handle are ordered on
 * the stream given to them.  A handle owns its device memory; callers own
 * key buffers.
 */
#ifndef BGPQ_H
#define BGPQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BGPQ_OK 0
#define BGPQ_EINVAL (-1)   /* bad argument */
#define BGPQ_EHIP (-2)     /* HIP runtime error */
#define BGPQ_EHEAP (-3)    /* the device flagged an error (overflow, bad size) */
#define BGPQ_ENOKERNEL (-4) /* no kernel built for this key/variant/batch size */

/* Variants: README.md:21-25 of the reference. */
#define BGPQ_BUITDD 0     /* heap/buitdd/heap.cuh    */
#define BGPQ_BUITDD_NP 1  /* heap/buitdd_np/heap.cuh */
#define BGPQ_TDITDD 2     /* heap/tditdd/heap.cuh    */
#define BGPQ_TDITDD_NP 3  /* heap/tditdd_np/heap.cuh */

#define BGPQ_KEY_U32 0
#define BGPQ_KEY_U64 1

typedef struct bgpq_heap bgpq_heap;

/* Heap(int batchNum, int batchSize, K init_limits) -- heap/buitdd/heap.cuh:34-68,
 * heap/tditdd/heap.cuh:33-67.  batch_size is a power of two in [64, 8192]; u64 keys stop at 4096
 * (the delete workspace, 3*batch_size*sizeof(key), must fit gfx950's 160 KB of LDS).
 * Every slot is filled with init_limits on the device; keys inserted later
 * must be strictly below init_limits.  The heap-sort output cursor
 * (deleteCount, the reference's -DHEAP_SORT mode) is always maintained. */
int bgpq_create(bgpq_heap** out, int variant, int key_type, int batch_num, int batch_size,
                uint64_t init_limits);
/* ~Heap() -- heap/buitdd/heap.cuh:94-116 */
void bgpq_destroy(bgpq_heap* h);
/* reset() -- heap/buitdd/heap.cuh:70-92 (all four variants here) */
int bgpq_reset(bgpq_heap* h, void* stream);

/* checkInsertHeap() -- heap/buitdd/heap.cuh:118-172, heap/tditdd/heap.cuh:123-176.
 * Returns 1 when every invariant holds, 0 on the first violation (its text
 * is in bgpq_last_error()), negative on error. */
int bgpq_check(bgpq_heap* h);
/* itemCount() / nodeCount() / isEmpty() -- heap/buitdd/heap.cuh:213-231 */
int bgpq_item_count(bgpq_heap* h, int64_t* out);
int bgpq_node_count(bgpq_heap* h, int64_t* out);
int bgpq_is_empty(bgpq_heap* h);

/* Bulk insertion: the harness insertKernel (test/heap_test.cu:41-55,
 * test/partial_heap_test.cu:45-58).  Inserts n device keys as ceil(n/chunk)
 * insertion() calls of at most `chunk` keys (0 = batch_size) spread over
 * `grid` workgroups (0 = default).  _np variants need chunk == batch_size
 * and n a multiple of it. */
int bgpq_insert(bgpq_heap* h, const void* d_keys, int64_t n, int chunk, int grid, void* stream);
/* Bulk deletion: the harness deleteKernel (test/heap_test.cu:57-74).  Runs
 * `calls` deleteRoot(+deleteUpdate) calls over `grid` workgroups; deleted
 * keys land in d_out at the heap-sort cursor (deleteCount), so d_out
 * receives the keys in ascending order.  The cursor persists across calls
 * (bgpq_set_cursor moves it); d_out holds out_keys keys, and a deletion that
 * would write past them is refused and flags error 16 instead. */
int bgpq_delete(bgpq_heap* h, void* d_out, int64_t out_keys, int64_t calls, int grid, void* stream);
/* Set the heap-sort output cursor (deleteCount). */
int bgpq_set_cursor(bgpq_heap* h, int64_t value);
/* concurrentKernel (test/heap_test.cu:13-38): the first half of the grid
 * inserts n_insert keys in chunks, the second half runs `calls` deletes. */
int bgpq_concurrent(bgpq_heap* h, const void* d_keys, int64_t n_insert, int chunk, void* d_out,
                    int64_t out_keys, int64_t calls, int grid, void* stream);
/* Heap sort: insert the n keys of d_in (chunk keys per insertion() call),
 * then delete until empty into d_out, which receives the keys ascending
 * (d_out may equal d_in).  The heap must be empty and hold batch_num >=
 * 2*ceil(n/batch_size) slots.  Optional outputs: the insert and delete phase
 * times in ms, from HIP events on `stream`. */
int bgpq_heapsort(bgpq_heap* h, const void* d_in, void* d_out, int64_t n, int chunk, int grid_insert,
                  int grid_delete, float* ms_insert, float* ms_delete, void* stream);

/* The heap's device storage, for kernels written outside this library
 * against include/bgpq_device.hpp (bgpq::view_from<K>(v) builds the
 * DevHeap's HeapView from it).  This replaces the reference's shallow
 * device copy of the Heap<K> object (test/heap_test.cu:128-129): the
 * pointers stay owned by the handle.  Status words and counters sit
 * word_stride ints apart. */
typedef struct bgpq_view_t {
    void* items;     /* (batch_num + 1) * batch_size keys */
    int* status;     /* (batch_num + 1) * word_stride ints */
    int* ctr;        /* 8 * word_stride ints */
    uint64_t init_limits;
    int32_t batch_num;
    int32_t batch_size;
    int32_t variant;   /* BGPQ_BUITDD .. BGPQ_TDITDD_NP */
    int32_t key_type;  /* BGPQ_KEY_U32 / BGPQ_KEY_U64 */
    int32_t word_stride;
    int32_t reserved;
} bgpq_view_t;
int bgpq_view(bgpq_heap* h, bgpq_view_t* out);

/* Termination vote of persistent kernels (the reference's PBS_MODEL
 * tbstate[] / terminate, heap/tditdd/heap.cuh:282-289 and
 * artifacts/astar/BGPQ_B/heap.cuh:273-283): a workgroup is busy from a
 * deleteRoot that hands it keys until a deleteRoot finds nothing for it;
 * the vote ends when no workgroup is busy.  bgpq_vote_reset starts a vote
 * with `busy` workgroups counted busy (those that call voteBusy() because
 * they seed the heap, as the reference presets tbstate[0] = 1).
 * bgpq_terminated returns 1 once the vote has ended, 0 before. */
int bgpq_vote_reset(bgpq_heap* h, int busy, void* stream);
int bgpq_terminated(bgpq_heap* h);

/* Empty the heap like bgpq_reset, touching only the slots in use (the
 * partial buffer and batches 1..batchCount, their status words, and the
 * counters): for a heap that a threshold search left holding keys it
 * will never pop.  Stream-ordered. */
int bgpq_clear(bgpq_heap* h, void* stream);

/* Host copies of the whole device state, for parity tests: items
 * ((batch_num+1)*batch_size keys), status (batch_num+1 ints), counters (8 ints:
 * batchCount, partialBufferSize, deleteCount, error flags).  Any may be NULL. */
int bgpq_snapshot(bgpq_heap* h, void* h_items, int* h_status, int* h_counters);
/* Device error flags (1 overflow, 2 bad insert size, 4 keys lost, 8 spin limit,
 * 16 output buffer full). */
int bgpq_error_flags(bgpq_heap* h, int* out);

/* Parameters of the kernels built for a batch size. */
int bgpq_threads(int batch_size);                      /* workgroup size */
int bgpq_lds_bytes(int key_type, int batch_size);      /* dynamic LDS per workgroup */
int bgpq_default_grid(int batch_size);

/* Test hooks for the batch primitives, one workgroup per batch:
 * bgpq_test_sort: the insert path's pad-and-sort (ibitonicSort, heap/heaputil.cuh:23-47):
 *   batch i = d_in[i*B .. i*B+d_sizes[i]) padded with `pad`, sorted into d_out[i*B ..].
 * bgpq_test_merge: merge of sorted d_a[i], d_b[i] (imergePath, heap/heaputil.cuh:99-158):
 *   the B smallest keys to d_small[i], the B largest to d_large[i]; algo 0 is
 *   the register bitonic merge, 1 the merge path with a 4-ary split search,
 *   2 the group merge, 3 the serial branch-free merge path (the heap picks
 *   one of them by batch size), 4 the two halves computed one after the
 *   other from the same inputs, as the heap's sift steps do, 5 the
 *   low-register merge path of the delete's rare root step, 6 the two
 *   halves and 7 the group merge with copy windows (a window that takes all
 *   its keys from one input is copied), as an insertion takes them. */
int bgpq_test_sort(int key_type, int batch_size, const void* d_in, const int* d_sizes, void* d_out,
                   uint64_t pad, int batches);
int bgpq_test_merge(int key_type, int batch_size, const void* d_a, const void* d_b, void* d_small, void* d_large,
                    int batches, int algo);

/* Synthetic heap-sort keys on the device: key i is built from the (i+1)-th
 * SplitMix64 output for `seed`; u32 keys lie in [0, 2^32-2], u64 keys are
 * (u32 key << 32) | i.  Ordered on `stream`. */
int bgpq_gen_keys(int key_type, uint64_t seed, int64_t n, void* d_out, void* stream);

const char* bgpq_last_error(void);

/* Device memory helpers so hosts without a GPU framework can drive the API. */
int bgpq_dev_malloc(void** p, size_t bytes);
int bgpq_dev_free(void* p);
int bgpq_h2d(void* dst, const void* src, size_t bytes);
int bgpq_d2h(void* dst, const void* src, size_t bytes);
int bgpq_sync(void);
int bgpq_device_name(char* buf, int len);

#ifdef __cplusplus
}
#endif
#endif
