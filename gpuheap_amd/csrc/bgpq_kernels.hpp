// Bulk kernels that drive one shared heap from every workgroup of a grid,
// and the launch table the C ABI dispatches through.
//
// Counterparts: insertKernel / deleteKernel / concurrentKernel of
// test/heap_test.cu:13-74 and test/partial_heap_test.cu:13-77.
#pragma once
#include <hip/hip_runtime.h>
#include "bgpq_device.hpp"

namespace bgpq {

// gfx950 gives a workgroup up to 160 KB of LDS.  The delete path needs
// 3*B*sizeof(K) of it: every u32 batch size up to 8192 fits, u64 up to 4096.
constexpr int kMaxLdsBytes = 160 * 1024;
template <typename K, int B>
constexpr bool batch_fits() { return delete_lds_bytes<K, B>() <= kMaxLdsBytes; }

template <typename K, int VAR, int B>
__global__ __launch_bounds__(threads_for(B)) void k_insert(HeapView<K> h, const K* keys, long long n,
                                                           int chunk) {
    constexpr int NT = threads_for(B);
    DevHeap<K, VAR, B, NT> hp{h};
    const long long calls = (n + chunk - 1) / chunk;
    for (long long i = blockIdx.x; i < calls; i += gridDim.x) {
        const long long rem = n - i * chunk;
        hp.insertion(keys + i * chunk, (int)(rem < chunk ? rem : chunk), 0);
    }
}

template <typename K, int VAR, int B>
__global__ __launch_bounds__(threads_for(B)) void k_delete(HeapView<K> h, K* out, long long calls) {
    constexpr int NT = threads_for(B);
    DevHeap<K, VAR, B, NT> hp{h};
    for (long long i = blockIdx.x; i < calls; i += gridDim.x) {
        int size = 0;
        if (hp.deleteRoot(out, size, 0)) hp.deleteUpdate(0);
    }
}

template <typename K, int VAR, int B>
__global__ __launch_bounds__(threads_for(B)) void k_concurrent(HeapView<K> h, const K* keys, long long n,
                                                               int chunk, K* out, long long calls) {
    constexpr int NT = threads_for(B);
    DevHeap<K, VAR, B, NT> hp{h};
    const int half = (int)gridDim.x / 2;
    const int id = (int)blockIdx.x;
    if (id < half) {
        const long long ins = (n + chunk - 1) / chunk;
        for (long long i = id; i < ins; i += half) {
            const long long rem = n - i * chunk;
            hp.insertion(keys + i * chunk, (int)(rem < chunk ? rem : chunk), 0);
        }
    } else {
        const int nd = (int)gridDim.x - half;
        for (long long i = id - half; i < calls; i += nd) {
            int size = 0;
            if (hp.deleteRoot(out, size, 0)) hp.deleteUpdate(0);
        }
    }
}

template <typename K, int B>
__global__ __launch_bounds__(threads_for(B)) void k_prim_sort(const K* in, const int* sizes, K* out, K pad) {
    constexpr int NT = threads_for(B);
    extern __shared__ __align__(16) int lds[];
    K* X0 = reinterpret_cast<K*>(lds);
    wg_sort_batch<K, B, NT>(X0, in + (size_t)blockIdx.x * B, sizes[blockIdx.x], pad);
    l2g_batch<K, B, NT>(out + (size_t)blockIdx.x * B, X0);
}

// algo 0: the register bitonic merge, 1: the LDS merge path, 2: the group
// merge, 3: the serial merge path, 4: the two halves one after the other,
// 5: the lean merge path of the delete's rare root step, 6: the two halves
// with copy windows (the insertion's), 7: the group merge with copy windows
template <typename K, int B, int ALGO>
__global__ __launch_bounds__(threads_for(B)) void k_prim_merge(const K* a, const K* b, K* small, K* large) {
    constexpr int NT = threads_for(B);
    extern __shared__ __align__(16) int lds[];
    K* X0 = reinterpret_cast<K*>(lds);
    K* X1 = X0 + B;
    const size_t o = (size_t)blockIdx.x * B;
    g2l_batch<K, B, NT>(X0, a + o);
    g2l_batch<K, B, NT>(X1, b + o);
    __syncthreads();
    if constexpr (ALGO == 0) wg_merge_bitonic<K, B, NT>(X0, X1, ToGlobal<K>{small + o}, ToGlobal<K>{large + o});
    else if constexpr (ALGO == 1) wg_merge_mp<K, B, NT>(X0, X1, ToGlobal<K>{small + o}, ToGlobal<K>{large + o});
    else if constexpr (ALGO == 2) wg_merge_grp<K, B, NT>(X0, X1, ToGlobal<K>{small + o}, ToGlobal<K>{large + o});
    else if constexpr (ALGO == 3) wg_merge_ser<K, B, NT>(X0, X1, ToGlobal<K>{small + o}, ToGlobal<K>{large + o});
    else if constexpr (ALGO == 5) wg_merge_lean<K, B, NT>(X0, X1, small + o, large + o);
    else if constexpr (ALGO == 7) wg_merge_grp<K, B, NT, true>(X0, X1, ToGlobal<K>{small + o}, ToGlobal<K>{large + o});
    else {
        // the two halves one after the other, as the heap's sift steps take them
        constexpr bool CP = ALGO == 6;
        HalfMerge<K, B, NT, false, CP> lo;
        lo.run(X0, X1);
        lo.put(ToGlobal<K>{small + o});
        HalfMerge<K, B, NT, true, CP> hi;
        hi.run(X0, X1);
        hi.put(ToGlobal<K>{large + o});
    }
}

template <typename K>
__global__ void k_fill(K* p, long long n, K v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// bgpq_clear: every slot in use (the partial buffer and batches
// 1..batchCount) back to `v` and their status words to 0.  Launched when no
// kernel of the heap runs, so batchCount is read with a plain load; the
// counters are zeroed by k_clear_ctr afterwards.
template <typename K>
__global__ void k_clear(K* items, int* status, const int* ctr, int batch_num, int batch_size, K v) {
    int nb = ctr[CTR_BATCH_COUNT * kWordStride];
    nb = nb < 0 ? 0 : (nb > batch_num ? batch_num : nb);
    const long long n = (long long)(nb + 1) * batch_size;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (long long i = i0; i < n; i += stride) items[i] = v;
    for (long long i = i0; i <= nb; i += stride) status[i * kWordStride] = 0;
}
template <typename K>
__global__ void k_clear_ctr(int* ctr) {
    for (int i = threadIdx.x; i < CTR_WORDS * kWordStride; i += blockDim.x) ctr[i] = 0;
}

// Type-erased launchers, one row per (key type, variant, batch size).
struct HeapArgs {
    void* items;
    int* status;
    int* ctr;
    uint64_t init_limits;
    int batch_num;
    int heap_sort;
    long long out_cap;
};

struct LaunchRow {
    int batch_size;
    hipError_t (*sort)(const void* in, const int* sizes, void* out, uint64_t pad, int batches, hipStream_t s);
    hipError_t (*merge)(const void* a, const void* b, void* small, void* large, int batches, int algo, hipStream_t s);
    hipError_t (*insert)(const HeapArgs&, const void* keys, long long n, int chunk, int grid, hipStream_t s);
    hipError_t (*del)(const HeapArgs&, void* out, long long calls, int grid, hipStream_t s);
    hipError_t (*conc)(const HeapArgs&, const void* keys, long long n, int chunk, void* out, long long calls,
                       int grid, hipStream_t s);
};

template <typename K>
inline HeapView<K> view_of(const HeapArgs& a) {
    return HeapView<K>{(K*)a.items, a.status, a.ctr, (K)a.init_limits, a.batch_num, a.heap_sort, a.out_cap};
}

// hipGetLastError reports the most recent failure of any runtime call on the
// thread; it is cleared before each launch so that a launch reports its own.
inline void clear_last_error() { (void)hipGetLastError(); }

template <typename K, int VAR, int B>
struct Launch {
    static constexpr size_t lds = (size_t)delete_lds_bytes<K, B>();
    static_assert(batch_fits<K, B>(), "dynamic LDS per workgroup");
    // Dynamic LDS above 64 KB must be allowed per kernel before a launch.
    static hipError_t allow_lds(const void* fn, size_t bytes) {
        if (bytes <= 64 * 1024) return hipSuccess;
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    }
    static hipError_t insert(const HeapArgs& a, const void* keys, long long n, int chunk, int grid,
                             hipStream_t s) {
        static const hipError_t ok = allow_lds(reinterpret_cast<const void*>(&k_insert<K, VAR, B>), lds);
        if (ok != hipSuccess) return ok;
        clear_last_error();
        hipLaunchKernelGGL((k_insert<K, VAR, B>), dim3(grid), dim3(threads_for(B)), lds, s, view_of<K>(a),
                           (const K*)keys, n, chunk);
        return hipGetLastError();
    }
    static hipError_t del(const HeapArgs& a, void* out, long long calls, int grid, hipStream_t s) {
        static const hipError_t ok = allow_lds(reinterpret_cast<const void*>(&k_delete<K, VAR, B>), lds);
        if (ok != hipSuccess) return ok;
        clear_last_error();
        hipLaunchKernelGGL((k_delete<K, VAR, B>), dim3(grid), dim3(threads_for(B)), lds, s, view_of<K>(a),
                           (K*)out, calls);
        return hipGetLastError();
    }
    static hipError_t conc(const HeapArgs& a, const void* keys, long long n, int chunk, void* out,
                           long long calls, int grid, hipStream_t s) {
        static const hipError_t ok = allow_lds(reinterpret_cast<const void*>(&k_concurrent<K, VAR, B>), lds);
        if (ok != hipSuccess) return ok;
        clear_last_error();
        hipLaunchKernelGGL((k_concurrent<K, VAR, B>), dim3(grid), dim3(threads_for(B)), lds, s,
                           view_of<K>(a), (const K*)keys, n, chunk, (K*)out, calls);
        return hipGetLastError();
    }
    static hipError_t sort(const void* in, const int* sizes, void* out, uint64_t pad, int batches, hipStream_t s) {
        clear_last_error();
        constexpr int lds = sort_xch_bytes<K, B, threads_for(B)>();
        hipLaunchKernelGGL((k_prim_sort<K, B>), dim3(batches), dim3(threads_for(B)), lds, s, (const K*)in, sizes,
                           (K*)out, (K)pad);
        return hipGetLastError();
    }
    template <int ALGO>
    static hipError_t merge_as(const void* a, const void* b, void* small, void* large, int batches, hipStream_t s) {
        static const hipError_t ok =
            allow_lds(reinterpret_cast<const void*>(&k_prim_merge<K, B, ALGO>), 2 * B * sizeof(K));
        if (ok != hipSuccess) return ok;
        clear_last_error();
        hipLaunchKernelGGL((k_prim_merge<K, B, ALGO>), dim3(batches), dim3(threads_for(B)), 2 * B * sizeof(K), s,
                           (const K*)a, (const K*)b, (K*)small, (K*)large);
        return hipGetLastError();
    }
    static hipError_t merge(const void* a, const void* b, void* small, void* large, int batches, int algo,
                            hipStream_t s) {
        switch (algo) {
            case 0: return merge_as<0>(a, b, small, large, batches, s);
            case 1: return merge_as<1>(a, b, small, large, batches, s);
            case 2: return merge_as<2>(a, b, small, large, batches, s);
            case 3: return merge_as<3>(a, b, small, large, batches, s);
            case 4: return merge_as<4>(a, b, small, large, batches, s);
            case 5: return merge_as<5>(a, b, small, large, batches, s);
            case 6: return merge_as<6>(a, b, small, large, batches, s);
            default: return merge_as<7>(a, b, small, large, batches, s);
        }
    }
    static LaunchRow row() { return LaunchRow{B, &sort, &merge, &insert, &del, &conc}; }
};

// A batch size whose delete workspace does not fit in LDS: an empty row
// (batch_size 0), which bgpq_create rejects.
template <typename K, int VAR, int B>
LaunchRow launch_row() {
    if constexpr (batch_fits<K, B>()) return Launch<K, VAR, B>::row();
    else return LaunchRow{0, nullptr, nullptr, nullptr, nullptr, nullptr};
}

constexpr int kNumBatchSizes = 8;  // 64 .. 8192
inline int batch_index(int B) {
    switch (B) {
        case 64: return 0;
        case 128: return 1;
        case 256: return 2;
        case 512: return 3;
        case 1024: return 4;
        case 2048: return 5;
        case 4096: return 6;
        case 8192: return 7;
        default: return -1;
    }
}

// Defined in the per-(key, variant) instantiation units.
const LaunchRow* rows_u32(int var);
const LaunchRow* rows_u64(int var);
hipError_t fill_u32(void* p, long long n, uint32_t v, hipStream_t s);
hipError_t fill_u64(void* p, long long n, uint64_t v, hipStream_t s);
hipError_t clear_u32(void* items, int* status, int* ctr, int batch_num, int batch_size, uint32_t v, hipStream_t s);
hipError_t clear_u64(void* items, int* status, int* ctr, int batch_num, int batch_size, uint64_t v, hipStream_t s);

}  // namespace bgpq
