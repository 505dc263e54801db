// A heap consumer written outside the library: 0/1 knapsack best-first
// branch-and-bound in its own persistent kernel, built only against the
// public headers (include/bgpq.h, include/bgpq_device.hpp) and linked with
// libbgpq.so.  Its loop has the shape of the reference's
// oneHeapApplication (artifacts/knapsack/models_pq.cuh:50-156): every
// workgroup deletes a batch, expands it (one thread per popped node, as
// appKernel does, knapsackKernel.cuh:35-91) and inserts the children in
// batch-sized insertion() calls, until the heap's termination vote ends
// (PBS_MODEL: tbstate / terminate).  Pops use the threshold deleteRoot of
// the A* forks (artifacts/astar/BGPQ_B/heap.cuh:260-287): only nodes whose
// bound beats the incumbent leave the heap.
//
//   ext_knapsack <instance file> [workgroups] [batch_num]
//
// The instance file has the reference's format (artifacts/knapsack/main.cu:
// 53-63): "n capacity", then n lines "benefit weight".  Prints one JSON
// line with the optimum.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "bgpq.h"
#include "bgpq_device.hpp"

namespace {

constexpr int kB = 256;
constexpr int kNT = bgpq::threads_for(kB);
using K = uint64_t;

// A node is its own key: (0xFFFF - bound) << 48 | depth << 40 | weight used
// << 20 | value, so the smallest key is the node with the highest bound.
constexpr uint64_t kTop = 0xFFFF;
__host__ __device__ inline K node_key(uint32_t ub, uint32_t level, uint32_t used, uint32_t value) {
    return ((kTop - ub) << 48) | ((uint64_t)level << 40) | ((uint64_t)used << 20) | value;
}
// pop only keys whose bound is above `best`
__device__ inline K threshold_of(uint32_t best) { return (kTop - best) << 48; }

struct Items {
    const uint32_t* w;  // by value/weight, highest first
    const uint32_t* v;
    int n;
    uint32_t cap;
};

// value + whole items from `level` while they fit + the fractional part of
// the first that does not (Dantzig bound, knapsackKernel.cuh:4-33 in integers)
__host__ __device__ inline uint32_t bound(const uint32_t* w, const uint32_t* v, int n, uint32_t cap, int level,
                                          uint32_t used, uint32_t value) {
    uint32_t room = cap - used, ub = value;
    for (int i = level; i < n; ++i) {
        if (w[i] <= room) {
            room -= w[i];
            ub += v[i];
        } else {
            ub += (uint32_t)((uint64_t)room * v[i] / w[i]);
            break;
        }
    }
    return ub;
}

// Dynamic LDS: popped[B], children[2B], then the heap's workspace.
constexpr int kHeapOffset = 3 * kB * (int)sizeof(K) / 4;  // ints, a multiple of 4
constexpr int kLdsBytes = 3 * kB * (int)sizeof(K) + bgpq::delete_lds_bytes<K, kB>();
// Idle rounds (each a deleteRoot and a sleep) before a workgroup gives up:
// far above any real wait, it only bounds a fault.
constexpr long long kMaxIdle = 1ll << 22;

__global__ __launch_bounds__(kNT) void ext_knapsack(bgpq_view_t view, Items it, uint32_t* best,
                                                    unsigned long long* popped, int* failed) {
    extern __shared__ __align__(16) int smem[];
    K* del = reinterpret_cast<K*>(smem);
    K* ins = del + kB;
    __shared__ int s_ins;
    __shared__ K s_thr;
    const int t = threadIdx.x;
    bgpq::DevHeap<K, bgpq::VAR_BUITDD, kB, kNT> heap{bgpq::view_from<K>(view)};
    if (blockIdx.x == 0) {
        // the seed: the root node, counted busy by the host's vote reset
        if (t == 0) del[0] = node_key(bound(it.w, it.v, it.n, it.cap, 0, 0, 0), 0, 0, 0);
        __syncthreads();
        heap.voteBusy();
        heap.insertion(del, 1, kHeapOffset);
    }
    long long idle = 0;
    unsigned long long mine = 0;
    while (true) {
        if (t == 0) s_thr = threshold_of(__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        __syncthreads();
        const K thr = s_thr;
        int size = 0;
        if (heap.deleteRoot(del, size, thr, kHeapOffset)) heap.deleteUpdate(kHeapOffset);
        if (heap.aborted()) {
            if (t == 0) atomicOr(failed, 1);
            break;
        }
        if (size == 0) {
            if (heap.ifTerminate()) break;
            if (++idle > kMaxIdle) {
                if (t == 0) atomicOr(failed, 2);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        idle = 0;
        mine += (unsigned long long)size;
        if (t == 0) s_ins = 0;
        __syncthreads();
        // one thread per popped node: the two children (take / skip the next item)
        for (int i = t; i < size; i += kNT) {
            const K key = del[i];
            const uint32_t ub = (uint32_t)(kTop - (key >> 48));
            const int level = (int)((key >> 40) & 0xFF);
            const uint32_t used = (uint32_t)((key >> 20) & 0xFFFFF), value = (uint32_t)(key & 0xFFFFF);
            if (ub <= __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || level >= it.n) continue;
            for (int take = 1; take >= 0; --take) {
                uint32_t u = used, val = value;
                if (take) {
                    if (used + it.w[level] > it.cap) continue;
                    u += it.w[level];
                    val += it.v[level];
                    atomicMax(best, val);
                }
                const uint32_t cb = bound(it.w, it.v, it.n, it.cap, level + 1, u, val);
                if (level + 1 < it.n && cb > __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    ins[atomicAdd(&s_ins, 1)] = node_key(cb, level + 1, u, val);
            }
        }
        __syncthreads();
        const int n = s_ins;
        for (int off = 0; off < n; off += kB) heap.insertion(ins + off, n - off < kB ? n - off : kB, kHeapOffset);
    }
    if (t == 0) atomicAdd(popped, mine);
}

#define CHECK(expr)                                                                                  \
    do {                                                                                             \
        const int rc_ = (expr);                                                                      \
        if (rc_ < 0) {                                                                               \
            std::fprintf(stderr, "%s failed (%d): %s\n", #expr, rc_, bgpq_last_error());             \
            std::exit(1);                                                                            \
        }                                                                                            \
    } while (0)
#define HIP(expr)                                                                                    \
    do {                                                                                             \
        const hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                      \
            std::fprintf(stderr, "%s failed: %s\n", #expr, hipGetErrorString(e_));                   \
            std::exit(1);                                                                            \
        }                                                                                            \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <instance> [workgroups] [batch_num]\n", argv[0]);
        return 2;
    }
    const int grid = argc > 2 ? std::atoi(argv[2]) : 8;
    const int batch_num = argc > 3 ? std::atoi(argv[3]) : (1 << 14);
    FILE* f = std::fopen(argv[1], "r");
    if (!f) return 2;
    int n = 0;
    unsigned cap = 0;
    if (std::fscanf(f, "%d %u", &n, &cap) != 2 || n < 1 || n > 255) return 2;
    std::vector<uint32_t> w(n), v(n);
    for (int i = 0; i < n; ++i)
        if (std::fscanf(f, "%u %u", &v[i], &w[i]) != 2 || w[i] == 0) return 2;
    std::fclose(f);
    // the key layout: bounds below 2^16, weights and values below 2^20
    if ((uint64_t)std::accumulate(v.begin(), v.end(), 0ull) >= kTop || cap >= (1u << 20)) {
        std::fprintf(stderr, "instance outside the key layout\n");
        return 2;
    }
    // by value/weight, highest first (artifacts/knapsack/main.cu:73-85)
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return (uint64_t)v[a] * w[b] > (uint64_t)v[b] * w[a]; });
    std::vector<uint32_t> ws(n), vs(n);
    for (int i = 0; i < n; ++i) {
        ws[i] = w[ord[i]];
        vs[i] = v[ord[i]];
    }

    bgpq_heap* h = nullptr;
    CHECK(bgpq_create(&h, BGPQ_BUITDD, BGPQ_KEY_U64, batch_num, kB, ~0ull));
    bgpq_view_t view;
    CHECK(bgpq_view(h, &view));
    if (view.batch_size != kB || view.variant != BGPQ_BUITDD || view.key_type != BGPQ_KEY_U64 ||
        view.word_stride != bgpq::kWordStride) {
        std::fprintf(stderr, "heap view does not match the kernel\n");
        return 1;
    }
    CHECK(bgpq_vote_reset(h, 1, nullptr));  // workgroup 0 seeds the heap

    uint32_t *d_w = nullptr, *d_v = nullptr, *d_best = nullptr;
    unsigned long long* d_popped = nullptr;
    int* d_failed = nullptr;
    HIP(hipMalloc(&d_w, 4 * n));
    HIP(hipMalloc(&d_v, 4 * n));
    HIP(hipMalloc(&d_best, 4));
    HIP(hipMalloc(&d_popped, 8));
    HIP(hipMalloc(&d_failed, 4));
    HIP(hipMemcpy(d_w, ws.data(), 4 * n, hipMemcpyHostToDevice));
    HIP(hipMemcpy(d_v, vs.data(), 4 * n, hipMemcpyHostToDevice));
    HIP(hipMemset(d_best, 0, 4));
    HIP(hipMemset(d_popped, 0, 8));
    HIP(hipMemset(d_failed, 0, 4));
    const Items items{d_w, d_v, n, cap};
    hipEvent_t e0, e1;
    HIP(hipEventCreate(&e0));
    HIP(hipEventCreate(&e1));
    HIP(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(ext_knapsack, dim3(grid), dim3(kNT), kLdsBytes, nullptr, view, items, d_best, d_popped,
                       d_failed);
    HIP(hipGetLastError());
    HIP(hipEventRecord(e1, nullptr));
    HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP(hipEventElapsedTime(&ms, e0, e1));
    uint32_t best = 0;
    unsigned long long popped = 0;
    int failed = 0, flags = 0;
    HIP(hipMemcpy(&best, d_best, 4, hipMemcpyDeviceToHost));
    HIP(hipMemcpy(&popped, d_popped, 8, hipMemcpyDeviceToHost));
    HIP(hipMemcpy(&failed, d_failed, 4, hipMemcpyDeviceToHost));
    CHECK(bgpq_error_flags(h, &flags));
    const int term = bgpq_terminated(h);
    int64_t left = 0;
    CHECK(bgpq_item_count(h, &left));
    CHECK(bgpq_clear(h, nullptr));
    CHECK(bgpq_sync());
    const int empty = bgpq_is_empty(h);
    std::printf("{\"best\": %u, \"popped\": %llu, \"left_pruned\": %lld, \"terminated\": %d, \"failed\": %d, "
                "\"heap_flags\": %d, \"empty_after_clear\": %d, \"ms\": %.3f}\n",
                best, popped, (long long)left, term, failed, flags, empty, ms);
    bgpq_destroy(h);
    (void)hipFree(d_w);
    (void)hipFree(d_v);
    (void)hipFree(d_best);
    (void)hipFree(d_popped);
    (void)hipFree(d_failed);
    return (failed || flags || term != 1 || empty != 1) ? 1 : 0;
}
