// Driver kernels and their C ABI (include/bgpq_apps.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <numeric>
#include <vector>

#include "bgpq_apps.h"
#include "bgpq_apps.hpp"
#include "bgpq_handle.hpp"

namespace bgpq {

template <int VAR, int B>
__global__ __launch_bounds__(threads_for(B)) void k_sssp(HeapView<uint64_t> h, SsspApp app, long long* ctl,
                                                         uint64_t first) {
    drive<VAR, B, threads_for(B)>(h, app, ctl, first, blockIdx.x == 0);
}
template <int VAR, int B>
__global__ __launch_bounds__(threads_for(B)) void k_astar(HeapView<uint64_t> h, AstarApp app, long long* ctl,
                                                          uint64_t first) {
    drive<VAR, B, threads_for(B)>(h, app, ctl, first, blockIdx.x == 0);
}
template <int VAR, int B>
__global__ __launch_bounds__(threads_for(B)) void k_knapsack(HeapView<uint64_t> h, KnapsackApp app,
                                                             long long* ctl, uint64_t first) {
    drive<VAR, B, threads_for(B)>(h, app, ctl, first, blockIdx.x == 0);
}

// Control words to their start state, plus one 32-bit label set (the
// source's distance, the start cell's g), and the heap's termination vote
// with `busy` workgroups counted busy (the seeding one, for a threshold
// search).
__global__ void k_app_setup(long long* ctl, long long pool, uint32_t* label, uint32_t value, int* heap_ctr,
                            int busy) {
    const int i = threadIdx.x;
    if (i < APP_WORDS) ctl[(size_t)i * kCtlStride] = i == APP_OUTSTANDING ? 1 : (i == APP_POOL ? pool : 0);
    if (i == 0 && label) *label = value;
    if (i == 0) {
        heap_ctr[CTR_BUSY * kWordStride] = busy;
        heap_ctr[CTR_TERM * kWordStride] = 0;
    }
}

struct AppLaunch {
    int batch_size;
    hipError_t (*sssp)(const HeapArgs&, const SsspApp&, long long*, uint64_t, int, hipStream_t);
    hipError_t (*astar)(const HeapArgs&, const AstarApp&, long long*, uint64_t, int, hipStream_t);
    hipError_t (*knapsack)(const HeapArgs&, const KnapsackApp&, long long*, uint64_t, int, hipStream_t);
};

template <int VAR, int B>
struct AppRows {
    static constexpr size_t lds = (size_t)app_lds_bytes<B>();
    static_assert(lds <= (size_t)kMaxLdsBytes, "driver LDS");
    template <class Kern>
    static hipError_t prep(Kern k) {
        if (lds <= 64 * 1024) return hipSuccess;
        return hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds);
    }
    static hipError_t sssp(const HeapArgs& a, const SsspApp& app, long long* ctl, uint64_t first, int grid,
                           hipStream_t s) {
        static const hipError_t ok = prep(&k_sssp<VAR, B>);
        if (ok != hipSuccess) return ok;
        (void)hipGetLastError();
        hipLaunchKernelGGL((k_sssp<VAR, B>), dim3(grid), dim3(threads_for(B)), lds, s, view_of<uint64_t>(a), app,
                           ctl, first);
        return hipGetLastError();
    }
    static hipError_t astar(const HeapArgs& a, const AstarApp& app, long long* ctl, uint64_t first, int grid,
                            hipStream_t s) {
        static const hipError_t ok = prep(&k_astar<VAR, B>);
        if (ok != hipSuccess) return ok;
        (void)hipGetLastError();
        hipLaunchKernelGGL((k_astar<VAR, B>), dim3(grid), dim3(threads_for(B)), lds, s, view_of<uint64_t>(a), app,
                           ctl, first);
        return hipGetLastError();
    }
    static hipError_t knapsack(const HeapArgs& a, const KnapsackApp& app, long long* ctl, uint64_t first, int grid,
                               hipStream_t s) {
        static const hipError_t ok = prep(&k_knapsack<VAR, B>);
        if (ok != hipSuccess) return ok;
        (void)hipGetLastError();
        hipLaunchKernelGGL((k_knapsack<VAR, B>), dim3(grid), dim3(threads_for(B)), lds, s, view_of<uint64_t>(a),
                           app, ctl, first);
        return hipGetLastError();
    }
    static AppLaunch row() { return AppLaunch{B, &sssp, &astar, &knapsack}; }
};

// rows for the partial-buffer variants (BGPQ_BUITDD, BGPQ_TDITDD) and
// batch sizes 256 .. 2048
static const AppLaunch* app_row(int variant, int B) {
    static const AppLaunch bu[4] = {AppRows<VAR_BUITDD, 256>::row(), AppRows<VAR_BUITDD, 512>::row(),
                                    AppRows<VAR_BUITDD, 1024>::row(), AppRows<VAR_BUITDD, 2048>::row()};
    static const AppLaunch td[4] = {AppRows<VAR_TDITDD, 256>::row(), AppRows<VAR_TDITDD, 512>::row(),
                                    AppRows<VAR_TDITDD, 1024>::row(), AppRows<VAR_TDITDD, 2048>::row()};
    const int i = B == 256 ? 0 : B == 512 ? 1 : B == 1024 ? 2 : B == 2048 ? 3 : -1;
    if (i < 0) return nullptr;
    if (variant == VAR_BUITDD) return &bu[i];
    if (variant == VAR_TDITDD) return &td[i];
    return nullptr;
}

hipError_t fill_u32(void* p, long long n, uint32_t v, hipStream_t s);

// bgpq_label_digest: per-lane partial sums, a wave reduction, one atomic per
// wave and word (additions mod 2^64).
__global__ void k_digest_zero(unsigned long long* out) {
    if (threadIdx.x < 2) out[threadIdx.x] = 0;
}
__global__ void k_digest(const uint32_t* labels, long long n, unsigned long long* out) {
    unsigned long long dig = 0, reach = 0;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        const uint32_t x = labels[v];
        dig += ((unsigned long long)x + 1) * (2 * (unsigned long long)v + 1);
        reach += x != 0xFFFFFFFFu;
    }
    for (int o = 32; o > 0; o >>= 1) {
        dig += __shfl_xor(dig, o, 64);
        reach += __shfl_xor(reach, o, 64);
    }
    if (__lane_id() == 0) {
        atomicAdd(out, dig);
        atomicAdd(out + 1, reach);
    }
}

}  // namespace bgpq

using namespace bgpq;

namespace {

// Checks shared by the drivers: a u64 heap with a partial buffer, a batch
// size with a driver instantiation, empty and without error flags.  With
// alloc_ctl the handle's control words are allocated on first use.
int check_heap(bgpq_heap* h, const AppLaunch** row, bool alloc_ctl = true) {
    if (!h) return bgpq_fail(BGPQ_EINVAL, "null heap");
    if (h->key_type != BGPQ_KEY_U64) return bgpq_fail(BGPQ_EINVAL, "drivers need a u64 heap");
    *row = app_row(h->variant, h->batch_size);
    if (!*row)
        return bgpq_fail(BGPQ_EINVAL, "drivers need variant buitdd or tditdd and batch size 256..2048 (got %d, %d)",
                         h->variant, h->batch_size);
    int c[CTR_WORDS];
    int rc = bgpq_handle_counters(h, c);
    if (rc) return rc;
    if (c[CTR_BATCH_COUNT] || c[CTR_PARTIAL]) return bgpq_fail(BGPQ_EINVAL, "heap is not empty");
    if (c[CTR_ERROR] || c[CTR_ABORT]) return bgpq_fail(BGPQ_EINVAL, "heap carries error flags 0x%x", c[CTR_ERROR]);
    if (alloc_ctl && !h->app_ctl) BGPQ_HIPCHK(hipMalloc(&h->app_ctl, sizeof(long long) * APP_WORDS * kCtlStride));
    return BGPQ_OK;
}

// Default workgroups per solve: every pop and push of a solve meets at the
// root and status[0], and past about 64 workgroups more of them only add
// contention (R-MAT scale 22 SSSP: 157 ms at 64, 266 ms at 256).
int resolve_grid(int grid) { return grid > 0 ? grid : 64; }

// Solve launch (timed with HIP events, or with the host clock around the
// stream's completion when events=false), then the control words `ctl` and
// the heap counters are read back.  A threshold search (A*) ends with its
// pruned keys still in the heap: there must be exactly `outstanding` of
// them, and the heap is cleared.
template <class F>
int run_solve(bgpq_heap* h, long long* ctl, hipStream_t s, bgpq_app_stats* stats, F launch, bool threshold = false,
              bool events = true) {
    float ms = 0.f;
    if (events) {
        hipEvent_t e0, e1;
        BGPQ_HIPCHK(hipEventCreate(&e0));
        BGPQ_HIPCHK(hipEventCreate(&e1));
        BGPQ_HIPCHK(hipEventRecord(e0, s));
        hipError_t e = launch();
        if (e != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return bgpq_fail(BGPQ_EHIP, "driver launch: %s", hipGetErrorString(e));
        }
        BGPQ_HIPCHK(hipEventRecord(e1, s));
        BGPQ_HIPCHK(hipEventSynchronize(e1));
        BGPQ_HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = launch();
        if (e != hipSuccess) return bgpq_fail(BGPQ_EHIP, "driver launch: %s", hipGetErrorString(e));
        BGPQ_HIPCHK(hipStreamSynchronize(s));
        ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    long long raw[APP_WORDS * kCtlStride];
    BGPQ_HIPCHK(hipMemcpy(raw, ctl, sizeof raw, hipMemcpyDeviceToHost));
    auto word = [&](int i) { return raw[(size_t)i * kCtlStride]; };
    int c[CTR_WORDS];
    int rc = bgpq_handle_counters(h, c);
    if (rc) return rc;
    const long long left = threshold ? (long long)c[CTR_BATCH_COUNT] * h->batch_size + c[CTR_PARTIAL] : 0;
    if (stats) {
        stats->popped = word(APP_POPPED);
        stats->expanded = word(APP_EXPANDED);
        stats->stale = word(APP_STALE);
        stats->pruned = word(APP_PRUNED) + left;
        stats->pushed = word(APP_PUSHED);
        stats->error_flags = (int32_t)word(APP_ERROR) | (c[CTR_ERROR] << 8);
        stats->ms = ms;
    }
    if (word(APP_ERROR) || c[CTR_ERROR])
        return bgpq_fail(BGPQ_EHEAP, "driver error flags 0x%llx, heap error flags 0x%x", word(APP_ERROR), c[CTR_ERROR]);
    if (threshold) {
        if (word(APP_OUTSTANDING) != left || !c[CTR_TERM])
            return bgpq_fail(BGPQ_EHEAP, "threshold search ended with %lld keys outstanding, %lld in the heap",
                             word(APP_OUTSTANDING), left);
        rc = bgpq_clear(h, s);
        if (rc) return rc;
        BGPQ_HIPCHK(hipStreamSynchronize(s));
        return BGPQ_OK;
    }
    if (word(APP_OUTSTANDING) != 0 || c[CTR_BATCH_COUNT] || c[CTR_PARTIAL])
        return bgpq_fail(BGPQ_EHEAP, "driver ended with %lld keys outstanding", word(APP_OUTSTANDING));
    return BGPQ_OK;
}

}  // namespace

extern "C" int bgpq_sssp(bgpq_heap* h, const int32_t* d_rowptr, const int32_t* d_col, const uint32_t* d_w,
                         int32_t n_vertices, int32_t source, uint32_t* d_dist, int grid, bgpq_app_stats* stats,
                         void* stream) {
    const AppLaunch* row = nullptr;
    int rc = check_heap(h, &row);
    if (rc) return rc;
    if (n_vertices < 1 || source < 0 || source >= n_vertices || !d_rowptr || !d_col || !d_w || !d_dist)
        return bgpq_fail(BGPQ_EINVAL, "sssp: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    BGPQ_HIPCHK(fill_u32(d_dist, n_vertices, 0xFFFFFFFFu, s));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_app_setup, dim3(1), dim3(64), 0, s, h->app_ctl, 0ll, d_dist + source, 0u, h->ctr, 0);
    BGPQ_HIPCHK(hipGetLastError());
    const SsspApp app{d_rowptr, d_col, d_w, d_dist, h->app_ctl};
    const HeapArgs a = bgpq_handle_args(h, 0);
    const int g = resolve_grid(grid);
    return run_solve(h, h->app_ctl, s, stats,
                     [&] { return row->sssp(a, app, h->app_ctl, (uint64_t)(uint32_t)source, g, s); });
}

extern "C" int bgpq_astar(bgpq_heap* h, const uint8_t* d_free, int32_t width, int32_t height, uint32_t* d_g, int grid,
                          uint32_t* cost, bgpq_app_stats* stats, void* stream) {
    const AppLaunch* row = nullptr;
    int rc = check_heap(h, &row);
    if (rc) return rc;
    if (width < 1 || height < 1 || (int64_t)width * height > INT32_MAX || !d_free || !d_g || !cost)
        return bgpq_fail(BGPQ_EINVAL, "astar: bad arguments");
    // the largest f: 14 per step along a path through every cell
    if ((uint64_t)width * height * 14ull + 14ull * (width + height) >= 0xFFFFFFFFull)
        return bgpq_fail(BGPQ_EINVAL, "astar: grid too large for 32-bit costs");
    hipStream_t s = (hipStream_t)stream;
    const long long cells = (long long)width * height;
    BGPQ_HIPCHK(fill_u32(d_g, cells, 0xFFFFFFFFu, s));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_app_setup, dim3(1), dim3(64), 0, s, h->app_ctl, 0ll, d_g, 0u, h->ctr, 1);
    BGPQ_HIPCHK(hipGetLastError());
    const AstarApp app{d_free, width, height, d_g, h->app_ctl};
    const uint32_t dx = (uint32_t)(width - 1), dy = (uint32_t)(height - 1);
    const uint32_t h0 = 10u * std::max(dx, dy) + 4u * std::min(dx, dy);
    const HeapArgs a = bgpq_handle_args(h, 0);
    const int g = resolve_grid(grid);
    rc = run_solve(h, h->app_ctl, s, stats, [&] { return row->astar(a, app, h->app_ctl, (uint64_t)h0 << 32, g, s); },
                   true);
    if (rc) return rc;
    BGPQ_HIPCHK(hipMemcpy(cost, d_g + (cells - 1), sizeof(uint32_t), hipMemcpyDeviceToHost));
    return BGPQ_OK;
}

extern "C" int bgpq_knapsack(bgpq_heap* h, int32_t n, const uint32_t* weight, const uint32_t* value,
                             uint64_t capacity, int64_t pool_nodes, int grid, uint64_t* best,
                             bgpq_app_stats* stats, void* stream) {
    const AppLaunch* row = nullptr;
    int rc = check_heap(h, &row);
    if (rc) return rc;
    if (n < 1 || n > 255 || !weight || !value || !best) return bgpq_fail(BGPQ_EINVAL, "knapsack: bad arguments");
    uint64_t vsum = 0;
    for (int i = 0; i < n; ++i) {
        if (weight[i] == 0) return bgpq_fail(BGPQ_EINVAL, "knapsack: zero weight");
        vsum += value[i];
    }
    if (vsum >= kUbTop) return bgpq_fail(BGPQ_EINVAL, "knapsack: total value %llu exceeds 2^24 - 2", vsum);
    const uint64_t cells = (uint64_t)(n + 1) * (capacity + 1);
    if (capacity >= 0xFFFFFFFFull || cells > (1ull << 34))
        return bgpq_fail(BGPQ_EINVAL, "knapsack: capacity %llu too large", capacity);
    // items by value/weight, highest first, ties by index
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        return (uint64_t)value[x] * weight[y] > (uint64_t)value[y] * weight[x];
    });
    std::vector<uint32_t> w(n), v(n);
    std::vector<uint64_t> pw(n + 1, 0), pv(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        w[i] = weight[ord[i]];
        v[i] = value[ord[i]];
        pw[i + 1] = pw[i] + w[i];
        pv[i + 1] = pv[i] + v[i];
    }
    if (pool_nodes <= 0) pool_nodes = (int64_t)std::min<uint64_t>(std::max<uint64_t>(cells, 1u << 16), 1ull << 26);
    if (pool_nodes > (int64_t)0xFFFFFFFFll) return bgpq_fail(BGPQ_EINVAL, "knapsack: pool over 2^32 nodes");
    hipStream_t s = (hipStream_t)stream;
    // one allocation: item arrays, prefix sums, dominance cells, node pool
    const size_t o_w = 0, o_v = o_w + 4 * (size_t)n, o_pw = (o_v + 4 * (size_t)n + 15) & ~(size_t)15;
    const size_t o_pv = o_pw + 8 * (size_t)(n + 1), o_cell = (o_pv + 8 * (size_t)(n + 1) + 15) & ~(size_t)15;
    const size_t o_pool = (o_cell + 4 * cells + 15) & ~(size_t)15;
    const size_t bytes = o_pool + sizeof(KnapNode) * (size_t)pool_nodes;
    char* mem = nullptr;
    BGPQ_HIPCHK(hipMalloc((void**)&mem, bytes));
    auto release = [&](int code) {
        (void)hipFree(mem);
        return code;
    };
    const KnapNode root{0, 0, 0, 0};
    const uint32_t one = 1;
    if (hipMemcpyAsync(mem + o_w, w.data(), 4 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(mem + o_v, v.data(), 4 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(mem + o_pw, pw.data(), 8 * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(mem + o_pv, pv.data(), 8 * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(mem + o_cell, 0, 4 * cells, s) != hipSuccess ||
        hipMemcpyAsync(mem + o_cell, &one, 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(mem + o_pool, &root, sizeof root, hipMemcpyHostToDevice, s) != hipSuccess)
        return release(bgpq_fail(BGPQ_EHIP, "knapsack: upload failed"));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_app_setup, dim3(1), dim3(64), 0, s, h->app_ctl, 1ll, nullptr, 0u, h->ctr, 0);
    if (hipGetLastError() != hipSuccess) return release(bgpq_fail(BGPQ_EHIP, "knapsack: setup launch"));
    const KnapsackApp app{n,
                          capacity,
                          (const uint32_t*)(mem + o_w),
                          (const uint32_t*)(mem + o_v),
                          (const uint64_t*)(mem + o_pw),
                          (const uint64_t*)(mem + o_pv),
                          (uint32_t*)(mem + o_cell),
                          (KnapNode*)(mem + o_pool),
                          (long long)pool_nodes,
                          h->app_ctl};
    // root bound, as the kernel computes it
    uint64_t ub0 = 0, room = capacity;
    for (int i = 0; i < n; ++i) {
        if (w[i] <= room) {
            room -= w[i];
            ub0 += v[i];
        } else {
            ub0 += room * v[i] / w[i];
            break;
        }
    }
    const uint64_t first = ((((kUbTop - ub0) << 8) | 255ull) << 32) | 0ull;
    const HeapArgs a = bgpq_handle_args(h, 0);
    const int g = resolve_grid(grid);
    rc = run_solve(h, h->app_ctl, s, stats, [&] { return row->knapsack(a, app, h->app_ctl, first, g, s); });
    if (rc) return release(rc);
    long long b = 0;
    if (hipMemcpy(&b, h->app_ctl + (size_t)APP_BEST * kCtlStride, sizeof b, hipMemcpyDeviceToHost) != hipSuccess)
        return release(bgpq_fail(BGPQ_EHIP, "knapsack: result copy"));
    *best = (uint64_t)b;
    return release(BGPQ_OK);
}

extern "C" int bgpq_label_digest(const uint32_t* d_labels, int64_t n, uint64_t* d_out2, void* stream) {
    if (n < 0 || !d_out2 || (n > 0 && !d_labels)) return bgpq_fail(BGPQ_EINVAL, "label digest: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    long long blocks = (n + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_digest_zero, dim3(1), dim3(64), 0, s, (unsigned long long*)d_out2);
    hipLaunchKernelGGL(k_digest, dim3((unsigned)blocks), dim3(256), 0, s, d_labels, (long long)n,
                       (unsigned long long*)d_out2);
    BGPQ_HIPCHK(hipGetLastError());
    return BGPQ_OK;
}

// ------------------------------------------------------- no allocation
// bgpq_knapsack_noalloc keeps all that bgpq_knapsack allocates (the
// control words, the item arrays, the prefix sums, the dominance cells and
// the node pool) in this array instead.  Each device the library runs on
// holds its own copy until the process ends.  In host code `scratch` only
// names the symbol: kernels get the address hipGetSymbolAddress returns.
namespace {

constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_at[kMaxDevices];  // per device, found on the first call there

// The current device's copy of `scratch`, looked up once per device.
int scratch_for_current_device(unsigned char** out) {
    int dev = 0;
    BGPQ_HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return bgpq_fail(BGPQ_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return bgpq_fail(BGPQ_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_at[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_at[dev];
    return BGPQ_OK;
}

size_t up16(size_t x) { return (x + 15) & ~(size_t)15; }

}  // namespace

extern "C" int bgpq_knapsack_noalloc(bgpq_heap* h, int32_t n, const uint32_t* weight, const uint32_t* value,
                                     uint64_t capacity, int64_t pool_nodes, int grid, uint64_t* best,
                                     bgpq_app_stats* stats, void* stream) {
    const AppLaunch* row = nullptr;
    int rc = check_heap(h, &row, false);
    if (rc) return rc;
    if (n < 1 || n > 255 || !weight || !value || !best) return bgpq_fail(BGPQ_EINVAL, "knapsack: bad arguments");
    uint64_t vsum = 0;
    for (int i = 0; i < n; ++i) {
        if (weight[i] == 0) return bgpq_fail(BGPQ_EINVAL, "knapsack: zero weight");
        vsum += value[i];
    }
    if (vsum >= kUbTop) return bgpq_fail(BGPQ_EINVAL, "knapsack: total value %llu exceeds 2^24 - 2", vsum);
    if (capacity >= 0xFFFFFFFFull) return bgpq_fail(BGPQ_EINVAL, "knapsack: capacity %llu too large", capacity);
    // the layout of bgpq_knapsack's allocation, after the control words
    const uint64_t cells = (uint64_t)(n + 1) * (capacity + 1);
    const size_t o_w = (size_t)APP_WORDS * kCtlStride * sizeof(long long), o_v = o_w + 4 * (size_t)n;
    const size_t o_pw = up16(o_v + 4 * (size_t)n), o_pv = o_pw + 8 * (size_t)(n + 1);
    const size_t o_cell = up16(o_pv + 8 * (size_t)(n + 1));
    if (cells > kScratchBytes / 4 || up16(o_cell + 4 * cells) + sizeof(KnapNode) > kScratchBytes)
        return bgpq_fail(BGPQ_EINVAL,
                         "knapsack_noalloc: %llu dominance cells ((n + 1) * (capacity + 1)) and a node pool do not fit "
                         "the %zu-byte scratch array; use bgpq_knapsack",
                         (unsigned long long)cells, kScratchBytes);
    const size_t o_pool = up16(o_cell + 4 * cells);
    const int64_t pool_max = (int64_t)((kScratchBytes - o_pool) / sizeof(KnapNode));
    if (pool_nodes <= 0) pool_nodes = pool_max;
    if (pool_nodes > pool_max)
        return bgpq_fail(BGPQ_EINVAL, "knapsack_noalloc: %lld pool nodes exceed the %lld that fit the scratch array",
                         (long long)pool_nodes, (long long)pool_max);
    unsigned char* base = nullptr;
    rc = scratch_for_current_device(&base);
    if (rc) return rc;
    // items by value/weight, highest first, ties by index (as bgpq_knapsack;
    // an insertion sort on the stack)
    int ord[255];
    uint32_t w[255], v[255];
    uint64_t pw[256], pv[256];
    for (int i = 0; i < n; ++i) {
        int j = i;
        while (j > 0 && (uint64_t)value[i] * weight[ord[j - 1]] > (uint64_t)value[ord[j - 1]] * weight[i]) {
            ord[j] = ord[j - 1];
            --j;
        }
        ord[j] = i;
    }
    pw[0] = pv[0] = 0;
    for (int i = 0; i < n; ++i) {
        w[i] = weight[ord[i]];
        v[i] = value[ord[i]];
        pw[i + 1] = pw[i] + w[i];
        pv[i + 1] = pv[i] + v[i];
    }
    hipStream_t s = (hipStream_t)stream;
    long long* ctl = reinterpret_cast<long long*>(base);
    const KnapNode root{0, 0, 0, 0};
    const uint32_t one = 1;
    // every call starts from zeroed dominance cells; the control words are
    // set by k_app_setup, the pool's root below, the rest is written before
    // it is read
    BGPQ_HIPCHK(fill_u32(base + o_cell, (long long)cells, 0u, s));
    if (hipMemcpyAsync(base + o_w, w, 4 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + o_v, v, 4 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + o_pw, pw, 8 * (size_t)(n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + o_pv, pv, 8 * (size_t)(n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + o_cell, &one, 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + o_pool, &root, sizeof root, hipMemcpyHostToDevice, s) != hipSuccess) {
        (void)hipStreamSynchronize(s);  // no copy may still read the arrays on this stack
        return bgpq_fail(BGPQ_EHIP, "knapsack_noalloc: upload failed");
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_app_setup, dim3(1), dim3(64), 0, s, ctl, 1ll, nullptr, 0u, h->ctr, 0);
    if (hipGetLastError() != hipSuccess) {
        (void)hipStreamSynchronize(s);
        return bgpq_fail(BGPQ_EHIP, "knapsack_noalloc: setup launch");
    }
    const KnapsackApp app{n,
                          capacity,
                          (const uint32_t*)(base + o_w),
                          (const uint32_t*)(base + o_v),
                          (const uint64_t*)(base + o_pw),
                          (const uint64_t*)(base + o_pv),
                          (uint32_t*)(base + o_cell),
                          (KnapNode*)(base + o_pool),
                          (long long)pool_nodes,
                          ctl};
    uint64_t ub0 = 0, room = capacity;
    for (int i = 0; i < n; ++i) {
        if (w[i] <= room) {
            room -= w[i];
            ub0 += v[i];
        } else {
            ub0 += room * v[i] / w[i];
            break;
        }
    }
    const uint64_t first = ((((kUbTop - ub0) << 8) | 255ull) << 32) | 0ull;
    const HeapArgs a = bgpq_handle_args(h, 0);
    const int g = resolve_grid(grid);
    // the host arrays above stay alive: run_solve waits for the stream
    rc = run_solve(h, ctl, s, stats, [&] { return row->knapsack(a, app, ctl, first, g, s); }, false, false);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    long long b = 0;
    BGPQ_HIPCHK(hipMemcpy(&b, ctl + (size_t)APP_BEST * kCtlStride, sizeof b, hipMemcpyDeviceToHost));
    *best = (uint64_t)b;
    return BGPQ_OK;
}
