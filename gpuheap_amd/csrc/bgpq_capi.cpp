// Host side of the C ABI declared in include/bgpq.h.
#include "bgpq.h"

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bgpq_handle.hpp"
#include "bgpq_kernels.hpp"

namespace bgpq {
const LaunchRow* rows_u32_v0();
const LaunchRow* rows_u32_v1();
const LaunchRow* rows_u32_v2();
const LaunchRow* rows_u32_v3();
const LaunchRow* rows_u64_v0();
const LaunchRow* rows_u64_v1();
const LaunchRow* rows_u64_v2();
const LaunchRow* rows_u64_v3();
}  // namespace bgpq

using namespace bgpq;

static thread_local std::string g_err;

int bgpq_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define fail bgpq_fail
#define HIPCHK BGPQ_HIPCHK

static size_t key_bytes(int key_type) { return key_type == BGPQ_KEY_U64 ? 8 : 4; }

static const LaunchRow* rows_for(int key_type, int variant) {
    if (key_type == BGPQ_KEY_U32) {
        switch (variant) {
            case 0: return rows_u32_v0();
            case 1: return rows_u32_v1();
            case 2: return rows_u32_v2();
            case 3: return rows_u32_v3();
        }
    } else if (key_type == BGPQ_KEY_U64) {
        switch (variant) {
            case 0: return rows_u64_v0();
            case 1: return rows_u64_v1();
            case 2: return rows_u64_v2();
            case 3: return rows_u64_v3();
        }
    }
    return nullptr;
}

// Status words and counters sit kWordStride ints apart (one 128-byte line each).
static size_t status_bytes(int batch_num) { return sizeof(int) * kWordStride * (size_t)(batch_num + 1); }
static size_t ctr_bytes() { return sizeof(int) * kWordStride * CTR_WORDS; }

static HeapArgs args_of(const bgpq_heap* h, long long out_cap = 0) {
    return HeapArgs{h->items, h->status, h->ctr, h->init_limits, h->batch_num, 1, out_cap};
}


static int fill_items(bgpq_heap* h, hipStream_t s) {
    const long long n = (long long)(h->batch_num + 1) * h->batch_size;
    hipError_t e = h->key_type == BGPQ_KEY_U64 ? fill_u64(h->items, n, h->init_limits, s)
                                               : fill_u32(h->items, n, (uint32_t)h->init_limits, s);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "fill: %s", hipGetErrorString(e));
    HIPCHK(hipMemsetAsync(h->status, 0, status_bytes(h->batch_num), s));
    HIPCHK(hipMemsetAsync(h->ctr, 0, ctr_bytes(), s));
    return BGPQ_OK;
}

static int counters(bgpq_heap* h, int* c) {
    std::vector<int> raw((size_t)kWordStride * CTR_WORDS);
    HIPCHK(hipMemcpy(raw.data(), h->ctr, ctr_bytes(), hipMemcpyDeviceToHost));
    for (int i = 0; i < CTR_WORDS; ++i) c[i] = raw[(size_t)i * kWordStride];
    return BGPQ_OK;
}

int bgpq_handle_counters(bgpq_heap* h, int* c) { return counters(h, c); }

HeapArgs bgpq_handle_args(const bgpq_heap* h, int heap_sort) {
    HeapArgs a = args_of(h, 0);
    a.heap_sort = heap_sort;
    return a;
}

// status words 0..count-1 gathered from the padded array
static int status_words(bgpq_heap* h, int* out, int count) {
    std::vector<int> raw((size_t)kWordStride * count);
    HIPCHK(hipMemcpy(raw.data(), h->status, sizeof(int) * raw.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < count; ++i) out[i] = raw[(size_t)i * kWordStride];
    return BGPQ_OK;
}

// checkInsertHeap restated (heap/buitdd/heap.cuh:118-172, heap/tditdd/heap.cuh:123-176):
// every status word AVAIL, each batch ascending, each child's first key >=
// its parent's last key, and, when both exist, the partial buffer ascending
// and not below the root's last key.  The first violation is reported in
// the same words, and found in the same order, as the oracle's orc_check.
template <typename K>
static int check_impl(bgpq_heap* h) {
    int c[CTR_WORDS];
    int rc = counters(h, c);
    if (rc) return rc;
    const int nb = c[CTR_BATCH_COUNT], pbs = c[CTR_PARTIAL];
    const int B = h->batch_size;
    if (c[CTR_ERROR]) return fail(0, "error flags 0x%x", c[CTR_ERROR]);
    if (nb < 0 || nb > h->batch_num) return fail(0, "batch count %d", nb);
    if (pbs < 0 || pbs >= B) return fail(0, "partial size %d", pbs);
    std::vector<K> items((size_t)(nb + 1) * B);
    std::vector<int> status(nb + 1);
    HIPCHK(hipMemcpy(items.data(), h->items, sizeof(K) * items.size(), hipMemcpyDeviceToHost));
    rc = status_words(h, status.data(), nb + 1);
    if (rc) return rc;
    const bool bu = h->variant == BGPQ_BUITDD || h->variant == BGPQ_BUITDD_NP;
    if (status[0] != 0) return fail(0, "partial batch state %d, expected AVAIL", status[0]);
    if (nb != 0 && pbs != 0) {
        if (items[(size_t)2 * B - 1] > items[0])
            return fail(0, "partial buffer below the root's last key");
        for (int i = 1; i < pbs; ++i)
            if (items[i] < items[i - 1]) return fail(0, "partial buffer not ascending at %d", i);
    }
    for (int i = 1; i <= nb; ++i) {
        if (status[i] != 0) return fail(0, "state of batch %d is %d, expected AVAIL", i, status[i]);
        if (i > 1) {
            const int p = bu ? (int)rev_idx(rev_idx((uint32_t)i) >> 1) : i / 2;
            if (items[(size_t)i * B] < items[(size_t)p * B + B - 1])
                return fail(0, "batch %d below parent %d", i, p);
        }
        for (int j = 1; j < B; ++j)
            if (items[(size_t)i * B + j] < items[(size_t)i * B + j - 1])
                return fail(0, "batch %d not ascending at %d", i, j);
    }
    return 1;
}

extern "C" {

const char* bgpq_last_error(void) { return g_err.c_str(); }

int bgpq_threads(int batch_size) {
    if (batch_index(batch_size) < 0) return fail(BGPQ_EINVAL, "batch size %d not built", batch_size);
    return threads_for(batch_size);
}

int bgpq_lds_bytes(int key_type, int batch_size) {
    if (batch_index(batch_size) < 0) return fail(BGPQ_EINVAL, "batch size %d not built", batch_size);
    if (key_type != BGPQ_KEY_U32 && key_type != BGPQ_KEY_U64) return fail(BGPQ_EINVAL, "key type %d", key_type);
    return 3 * batch_size * (int)key_bytes(key_type);
}

int bgpq_default_grid(int batch_size) {
    (void)batch_size;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 32;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 32;
    return cus;
}

int bgpq_create(bgpq_heap** out, int variant, int key_type, int batch_num, int batch_size,
                uint64_t init_limits) {
    if (!out) return fail(BGPQ_EINVAL, "null handle pointer");
    *out = nullptr;
    if (variant < 0 || variant > 3) return fail(BGPQ_EINVAL, "variant %d", variant);
    if (key_type != BGPQ_KEY_U32 && key_type != BGPQ_KEY_U64) return fail(BGPQ_EINVAL, "key type %d", key_type);
    const int bi = batch_index(batch_size);
    if (bi < 0) return fail(BGPQ_EINVAL, "batch size %d (need a power of two in [64, 8192])", batch_size);
    if (rows_for(key_type, variant)[bi].batch_size == 0)
        return fail(BGPQ_EINVAL, "batch size %d: %d bytes of LDS per workgroup exceed %d", batch_size,
                    3 * batch_size * (int)key_bytes(key_type), kMaxLdsBytes);
    if (batch_num < 1) return fail(BGPQ_EINVAL, "batch_num %d", batch_num);
    if (key_type == BGPQ_KEY_U32 && init_limits > 0xFFFFFFFFull)
        return fail(BGPQ_EINVAL, "init_limits does not fit u32");
    bgpq_heap* h = new bgpq_heap();
    h->app_ctl = nullptr;
    h->variant = variant;
    h->key_type = key_type;
    h->batch_num = batch_num;
    h->batch_size = batch_size;
    h->init_limits = init_limits;
    h->row = &rows_for(key_type, variant)[bi];
    const size_t bytes = key_bytes(key_type) * (size_t)batch_size * (size_t)(batch_num + 1);
    auto alloc = [](void** p, size_t n) { return hipMalloc(p, n); };
    hipError_t e = alloc(&h->items, bytes);
    if (e == hipSuccess) e = alloc((void**)&h->status, status_bytes(batch_num));
    if (e == hipSuccess) e = alloc((void**)&h->ctr, ctr_bytes());
    if (e != hipSuccess) {
        bgpq_destroy(h);
        return fail(BGPQ_EHIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    int rc = fill_items(h, nullptr);
    if (rc == BGPQ_OK && hipStreamSynchronize(nullptr) != hipSuccess) rc = fail(BGPQ_EHIP, "sync after fill");
    if (rc != BGPQ_OK) {
        bgpq_destroy(h);
        return rc;
    }
    *out = h;
    return BGPQ_OK;
}

void bgpq_destroy(bgpq_heap* h) {
    if (!h) return;
    if (h->items) (void)hipFree(h->items);
    if (h->status) (void)hipFree(h->status);
    if (h->ctr) (void)hipFree(h->ctr);
    if (h->app_ctl) (void)hipFree(h->app_ctl);
    delete h;
}

int bgpq_reset(bgpq_heap* h, void* stream) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    return fill_items(h, (hipStream_t)stream);
}

int bgpq_view(bgpq_heap* h, bgpq_view_t* out) {
    if (!h || !out) return fail(BGPQ_EINVAL, "null argument");
    out->items = h->items;
    out->status = h->status;
    out->ctr = h->ctr;
    out->init_limits = h->init_limits;
    out->batch_num = h->batch_num;
    out->batch_size = h->batch_size;
    out->variant = h->variant;
    out->key_type = h->key_type;
    out->word_stride = kWordStride;
    out->reserved = 0;
    return BGPQ_OK;
}

int bgpq_vote_reset(bgpq_heap* h, int busy, void* stream) {
    if (!h || busy < 0) return fail(BGPQ_EINVAL, "bad vote reset");
    const int w[2] = {busy, 0};
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipMemcpyAsync(h->ctr + CTR_BUSY * kWordStride, &w[0], sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(h->ctr + CTR_TERM * kWordStride, &w[1], sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));  // the words above are on this stack
    return BGPQ_OK;
}

int bgpq_terminated(bgpq_heap* h) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    int c[CTR_WORDS];
    const int rc = counters(h, c);
    if (rc) return rc;
    return c[CTR_TERM] != 0 ? 1 : 0;
}

int bgpq_clear(bgpq_heap* h, void* stream) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    hipStream_t s = (hipStream_t)stream;
    const hipError_t e =
        h->key_type == BGPQ_KEY_U64
            ? clear_u64(h->items, h->status, h->ctr, h->batch_num, h->batch_size, h->init_limits, s)
            : clear_u32(h->items, h->status, h->ctr, h->batch_num, h->batch_size, (uint32_t)h->init_limits, s);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "clear: %s", hipGetErrorString(e));
    return BGPQ_OK;
}

int bgpq_set_cursor(bgpq_heap* h, int64_t value) {
    if (!h || value < 0 || value > 0x7FFFFFFF) return fail(BGPQ_EINVAL, "bad cursor");
    const int v = (int)value;
    HIPCHK(hipMemcpy(h->ctr + CTR_DELETE_COUNT * kWordStride, &v, sizeof v, hipMemcpyHostToDevice));
    return BGPQ_OK;
}

int bgpq_error_flags(bgpq_heap* h, int* out) {
    if (!h || !out) return fail(BGPQ_EINVAL, "null argument");
    int c[CTR_WORDS];
    const int rc = counters(h, c);
    if (rc) return rc;
    *out = c[CTR_ERROR];
    return BGPQ_OK;
}


int bgpq_item_count(bgpq_heap* h, int64_t* out) {
    if (!h || !out) return fail(BGPQ_EINVAL, "null argument");
    int c[CTR_WORDS];
    int rc = counters(h, c);
    if (rc) return rc;
    *out = (int64_t)c[CTR_PARTIAL] + (int64_t)c[CTR_BATCH_COUNT] * h->batch_size;
    return BGPQ_OK;
}

int bgpq_node_count(bgpq_heap* h, int64_t* out) {
    if (!h || !out) return fail(BGPQ_EINVAL, "null argument");
    int c[CTR_WORDS];
    int rc = counters(h, c);
    if (rc) return rc;
    *out = c[CTR_BATCH_COUNT];
    return BGPQ_OK;
}

int bgpq_is_empty(bgpq_heap* h) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    int c[CTR_WORDS];
    int rc = counters(h, c);
    if (rc) return rc;
    return (c[CTR_PARTIAL] == 0 && c[CTR_BATCH_COUNT] == 0) ? 1 : 0;
}

int bgpq_snapshot(bgpq_heap* h, void* h_items, int* h_status, int* h_counters) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    HIPCHK(hipDeviceSynchronize());
    if (h_items)
        HIPCHK(hipMemcpy(h_items, h->items,
                         key_bytes(h->key_type) * (size_t)h->batch_size * (size_t)(h->batch_num + 1),
                         hipMemcpyDeviceToHost));
    if (h_status) {
        const int rc = status_words(h, h_status, h->batch_num + 1);
        if (rc) return rc;
    }
    if (h_counters) {
        int c[CTR_WORDS];
        const int rc = counters(h, c);
        if (rc) return rc;
        for (int i = 0; i < kPublicCounters; ++i) h_counters[i] = c[i];
    }
    return BGPQ_OK;
}

int bgpq_check(bgpq_heap* h) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    HIPCHK(hipDeviceSynchronize());
    return h->key_type == BGPQ_KEY_U64 ? check_impl<uint64_t>(h) : check_impl<uint32_t>(h);
}

static int device_error(bgpq_heap* h) {
    int c[CTR_WORDS];
    int rc = counters(h, c);
    if (rc) return rc;
    if (c[CTR_ERROR]) return fail(BGPQ_EHEAP, "device error flags 0x%x", c[CTR_ERROR]);
    return BGPQ_OK;
}

static int resolve_grid(bgpq_heap* h, int grid) { return grid > 0 ? grid : bgpq_default_grid(h->batch_size); }

static bool np_variant(const bgpq_heap* h) { return h->variant == BGPQ_BUITDD_NP || h->variant == BGPQ_TDITDD_NP; }

int bgpq_insert(bgpq_heap* h, const void* d_keys, int64_t n, int chunk, int grid, void* stream) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    if (n < 0) return fail(BGPQ_EINVAL, "n %lld", (long long)n);
    if (n == 0) return BGPQ_OK;
    if (!d_keys) return fail(BGPQ_EINVAL, "null keys");
    if (chunk <= 0) chunk = h->batch_size;
    if (chunk > h->batch_size) return fail(BGPQ_EINVAL, "chunk %d > batch size %d", chunk, h->batch_size);
    if (np_variant(h) && (chunk != h->batch_size || n % h->batch_size))
        return fail(BGPQ_EINVAL, "_np variants insert full batches only (chunk %d, n %lld)", chunk, (long long)n);
    hipError_t e = h->row->insert(args_of(h), d_keys, n, chunk, resolve_grid(h, grid), (hipStream_t)stream);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "insert launch: %s", hipGetErrorString(e));
    return BGPQ_OK;
}

int bgpq_delete(bgpq_heap* h, void* d_out, int64_t out_keys, int64_t calls, int grid, void* stream) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    if (calls < 0 || out_keys < 0) return fail(BGPQ_EINVAL, "calls %lld, out_keys %lld", (long long)calls,
                                               (long long)out_keys);
    if (calls == 0) return BGPQ_OK;
    if (!d_out) return fail(BGPQ_EINVAL, "null output");
    hipError_t e = h->row->del(args_of(h, out_keys), d_out, calls, resolve_grid(h, grid), (hipStream_t)stream);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "delete launch: %s", hipGetErrorString(e));
    return BGPQ_OK;
}

int bgpq_concurrent(bgpq_heap* h, const void* d_keys, int64_t n_insert, int chunk, void* d_out, int64_t out_keys,
                    int64_t calls, int grid, void* stream) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    if (n_insert < 0 || calls < 0) return fail(BGPQ_EINVAL, "negative size");
    if (chunk <= 0) chunk = h->batch_size;
    if (chunk > h->batch_size) return fail(BGPQ_EINVAL, "chunk %d > batch size %d", chunk, h->batch_size);
    if (np_variant(h) && (chunk != h->batch_size || n_insert % h->batch_size))
        return fail(BGPQ_EINVAL, "_np variants insert full batches only");
    grid = resolve_grid(h, grid);
    if (grid < 2) return fail(BGPQ_EINVAL, "concurrent run needs a grid of at least 2");
    if ((n_insert && !d_keys) || (calls && !d_out)) return fail(BGPQ_EINVAL, "null buffer");
    if (out_keys < 0) return fail(BGPQ_EINVAL, "out_keys %lld", (long long)out_keys);
    hipError_t e = h->row->conc(args_of(h, out_keys), d_keys, n_insert, chunk, d_out, calls, grid, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "concurrent launch: %s", hipGetErrorString(e));
    return BGPQ_OK;
}

int bgpq_heapsort(bgpq_heap* h, const void* d_in, void* d_out, int64_t n, int chunk, int grid_insert,
                  int grid_delete, float* ms_insert, float* ms_delete, void* stream) {
    if (!h) return fail(BGPQ_EINVAL, "null handle");
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return fail(BGPQ_EINVAL, "bad buffers");
    if (n > 0x7FFFFFFF) return fail(BGPQ_EINVAL, "the heap-sort cursor is 32-bit: n <= 2^31-1");
    hipStream_t s = (hipStream_t)stream;
    if (chunk <= 0) chunk = h->batch_size;
    int64_t items = 0;
    int rc = bgpq_item_count(h, &items);
    if (rc) return rc;
    if (items != 0) return fail(BGPQ_EINVAL, "heap sort needs an empty heap (%lld items)", (long long)items);
    const int64_t B = h->batch_size;
    if ((n + B - 1) / B > (int64_t)h->batch_num / 2)
        return fail(BGPQ_EINVAL, "capacity: %lld keys need batch_num >= %lld", (long long)n,
                    (long long)(2 * ((n + B - 1) / B)));
    HIPCHK(hipMemsetAsync(h->ctr + CTR_DELETE_COUNT * kWordStride, 0, sizeof(int), s));
    hipEvent_t e0, e1, e2;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    HIPCHK(hipEventRecord(e0, s));
    rc = bgpq_insert(h, d_in, n, chunk, grid_insert, stream);
    if (!rc) {
        HIPCHK(hipEventRecord(e1, s));
        // enough calls to drain: every call removes a full batch or the partial buffer
        rc = bgpq_delete(h, d_out, n, (n + B - 1) / B + 1, grid_delete, stream);
    }
    if (!rc) {
        HIPCHK(hipEventRecord(e2, s));
        HIPCHK(hipEventSynchronize(e2));
        float a = 0, b = 0;
        HIPCHK(hipEventElapsedTime(&a, e0, e1));
        HIPCHK(hipEventElapsedTime(&b, e1, e2));
        if (ms_insert) *ms_insert = a;
        if (ms_delete) *ms_delete = b;
        rc = device_error(h);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    return rc;
}

int bgpq_test_sort(int key_type, int batch_size, const void* d_in, const int* d_sizes, void* d_out,
                   uint64_t pad, int batches) {
    const int bi = batch_index(batch_size);
    const LaunchRow* rows = rows_for(key_type, 0);
    if (bi < 0 || !rows || batches < 1 || rows[bi].batch_size == 0)
        return fail(BGPQ_EINVAL, "bad primitive test arguments");
    hipError_t e = rows[bi].sort(d_in, d_sizes, d_out, pad, batches, nullptr);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "sort launch: %s", hipGetErrorString(e));
    HIPCHK(hipDeviceSynchronize());
    return BGPQ_OK;
}

int bgpq_test_merge(int key_type, int batch_size, const void* d_a, const void* d_b, void* d_small, void* d_large,
                    int batches, int algo) {
    const int bi = batch_index(batch_size);
    const LaunchRow* rows = rows_for(key_type, 0);
    if (bi < 0 || !rows || batches < 1 || rows[bi].batch_size == 0 || algo < 0 || algo > 7)
        return fail(BGPQ_EINVAL, "bad primitive test arguments");
    hipError_t e = rows[bi].merge(d_a, d_b, d_small, d_large, batches, algo, nullptr);
    if (e != hipSuccess) return fail(BGPQ_EHIP, "merge launch: %s", hipGetErrorString(e));
    HIPCHK(hipDeviceSynchronize());
    return BGPQ_OK;
}

int bgpq_dev_malloc(void** p, size_t bytes) {
    if (!p) return fail(BGPQ_EINVAL, "null pointer");
    HIPCHK(hipMalloc(p, bytes));
    return BGPQ_OK;
}
int bgpq_dev_free(void* p) {
    HIPCHK(hipFree(p));
    return BGPQ_OK;
}
int bgpq_h2d(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BGPQ_OK;
}
int bgpq_d2h(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BGPQ_OK;
}
int bgpq_sync(void) {
    HIPCHK(hipDeviceSynchronize());
    return BGPQ_OK;
}
int bgpq_device_name(char* buf, int len) {
    if (!buf || len <= 0) return fail(BGPQ_EINVAL, "buffer");
    hipDeviceProp_t p;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return BGPQ_OK;
}

}  // extern "C"
