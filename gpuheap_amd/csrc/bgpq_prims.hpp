// Workgroup-collective batch primitives for the batched GPU priority queue,
// written for CDNA4 (gfx950, wave64).
//
// Counterpart of the reference's heap/heaputil.cuh:
//   batchCopy   (heap/heaputil.cuh:4-12)   -> load_batch / store_batch / fill_batch
//   ibitonicSort(heap/heaputil.cuh:23-47)  -> wg_sort_batch
//   imergePath  (heap/heaputil.cuh:99-158) -> wg_merge_path (+ the store helpers)
//
// Layout conventions
//   * A batch is B keys, ascending, B a power of two.
//   * NT threads per workgroup, NT a multiple of 64 and NT <= B.
//   * For sorting, thread t holds the P = B/NT consecutive keys [t*P, t*P+P)
//     in registers ("blocked" layout): the compare-exchange stages with
//     stride < P run inside a lane, strides < 64*P across lanes of one wave
//     (ds_swizzle / DPP / bpermute), and larger strides through LDS.
//   * For merging, thread t produces the L = 2B/NT consecutive outputs
//     [t*L, t*L+L) of the 2B-key merge; threads t < NT/2 own the small half,
//     the others the large half, so every store is a contiguous L-key run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bgpq {

// ---------------------------------------------------------------- memory model
// Cross-workgroup hand-off on gfx950: a workgroup that publishes batch data
// runs an agent-scope release fence in every wave (buffer_wbl2 sc1 +
// s_waitcnt vmcnt(0)), then a barrier, then lane 0 stores the status word.
// The consumer's lane 0 spins on the status word, runs an acquire fence,
// then a barrier, then every wave runs its own acquire fence (buffer_inv sc1)
// before reading batch data.  __syncthreads() alone does not wait for other
// waves' global stores on gfx950, hence the per-wave fences.
__device__ __forceinline__ void fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ int a_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void a_store(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void a_store_rel(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool a_cas(int* p, int expect, int desired) {
    return __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int a_cas_val(int* p, int expect, int desired) {
    __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return expect;  // value found in memory
}
__device__ __forceinline__ int a_add(int* p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int a_max(int* p, int v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void backoff() { __builtin_amdgcn_s_sleep(1); }

// ---------------------------------------------------------------- key helpers
template <typename K> __device__ __forceinline__ K kmin(K a, K b) { return b < a ? b : a; }
template <typename K> __device__ __forceinline__ K kmax(K a, K b) { return b < a ? a : b; }

// xor-lane exchange inside a wave64.  Distances 1..16 use ds_swizzle in
// bit mode (and=0x1f, xor=m), 32 uses ds_bpermute.
template <int M> __device__ __forceinline__ uint32_t xlane(uint32_t v) {
    static_assert(M >= 1 && M <= 32 && (M & (M - 1)) == 0, "xor distance");
    if constexpr (M == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (M == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (M <= 16) {
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (M << 10) | 0x1F);
    } else {
        int lane = (int)__lane_id();
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, (int)v);
    }
}
template <int M> __device__ __forceinline__ uint64_t xlane(uint64_t v) {
    uint32_t lo = xlane<M>((uint32_t)v);
    uint32_t hi = xlane<M>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// ---------------------------------------------------------------- vector moves
// Copy B keys global -> LDS or LDS -> global with 16-byte lanes.
template <typename K, int B, int NT>
__device__ __forceinline__ void copy_batch(K* __restrict__ dst, const K* __restrict__ src) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int c = threadIdx.x; c < NV; c += NT) d[c] = s[c];
}

// Move B keys global -> LDS and reset the source to `fill` in the same
// pass: each thread resets exactly the 16-byte chunk it loaded, so no other
// thread can overwrite a chunk before it is read.
template <typename K, int B, int NT>
__device__ __forceinline__ void take_batch(K* __restrict__ dst, K* __restrict__ src, K fill) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    uint4 f;
    if constexpr (sizeof(K) == 4) {
        f.x = f.y = f.z = f.w = (uint32_t)fill;
    } else {
        f.x = f.z = (uint32_t)fill;
        f.y = f.w = (uint32_t)((uint64_t)fill >> 32);
    }
    uint4* s = reinterpret_cast<uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int c = threadIdx.x; c < NV; c += NT) {
        const uint4 v = s[c];
        d[c] = v;
        s[c] = f;
    }
}

template <typename K, int B, int NT>
__device__ __forceinline__ void fill_batch(K* dst, K v) {
#pragma unroll
    for (int i = threadIdx.x; i < B; i += NT) dst[i] = v;
}

// Store a thread's run of N consecutive keys (N*sizeof(K) a multiple of 16
// when possible).
template <typename K, int N>
__device__ __forceinline__ void store_run(K* dst, const K (&v)[N]) {
    if constexpr ((N * sizeof(K)) % 16 == 0) {
        constexpr int W = 16 / sizeof(K);
#pragma unroll
        for (int c = 0; c < N / W; ++c) {
            uint4 q;
            if constexpr (sizeof(K) == 4) {
                q.x = (uint32_t)v[c * 4 + 0]; q.y = (uint32_t)v[c * 4 + 1];
                q.z = (uint32_t)v[c * 4 + 2]; q.w = (uint32_t)v[c * 4 + 3];
            } else {
                q.x = (uint32_t)v[c * 2 + 0]; q.y = (uint32_t)((uint64_t)v[c * 2 + 0] >> 32);
                q.z = (uint32_t)v[c * 2 + 1]; q.w = (uint32_t)((uint64_t)v[c * 2 + 1] >> 32);
            }
            reinterpret_cast<uint4*>(dst)[c] = q;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = v[i];
    }
}

template <typename K, int N>
__device__ __forceinline__ void load_run(K (&v)[N], const K* src) {
    if constexpr ((N * sizeof(K)) % 16 == 0) {
#pragma unroll
        for (int c = 0; c < N * (int)sizeof(K) / 16; ++c) {
            uint4 q = reinterpret_cast<const uint4*>(src)[c];
            if constexpr (sizeof(K) == 4) {
                v[c * 4 + 0] = (K)q.x; v[c * 4 + 1] = (K)q.y;
                v[c * 4 + 2] = (K)q.z; v[c * 4 + 3] = (K)q.w;
            } else {
                v[c * 2 + 0] = (K)(((uint64_t)q.y << 32) | q.x);
                v[c * 2 + 1] = (K)(((uint64_t)q.w << 32) | q.z);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = src[i];
    }
}

// ---------------------------------------------------------------- bitonic sort
// Ascending sort of the B = P*NT keys held blocked in registers.  `xch` is an
// LDS scratch of B keys used for the cross-wave strides (may alias nothing
// that is live).  Semantics match ibitonicSort (heap/heaputil.cuh:23-47):
// the result is the sorted multiset.
template <typename K, int P, int NT>
__device__ __forceinline__ void wg_sort_regs(K (&v)[P], K* xch) {
    constexpr int B = P * NT;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 2; k <= B; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j < P) {
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    if ((i & j) == 0) {
                        const int g = t * P + i;
                        const bool asc = (g & k) == 0;
                        K a = v[i], b = v[i | j];
                        const bool sw = asc ? (b < a) : (a < b);
                        v[i] = sw ? b : a;
                        v[i | j] = sw ? a : b;
                    }
                }
            } else {
                const int m = j / P;
                const bool lower = (t & m) == 0;
                const bool asc = ((t * P) & k) == 0;
                const bool take_min = (lower == asc);
                if (j < 64 * P) {
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        K w;
                        switch (m) {
                            case 1: w = xlane<1>(v[i]); break;
                            case 2: w = xlane<2>(v[i]); break;
                            case 4: w = xlane<4>(v[i]); break;
                            case 8: w = xlane<8>(v[i]); break;
                            case 16: w = xlane<16>(v[i]); break;
                            default: w = xlane<32>(v[i]); break;
                        }
                        v[i] = take_min ? kmin(v[i], w) : kmax(v[i], w);
                    }
                } else {
                    store_run<K, P>(xch + t * P, v);
                    __syncthreads();
                    K w[P];
                    load_run<K, P>(w, xch + (t ^ m) * P);
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < P; ++i) v[i] = take_min ? kmin(v[i], w[i]) : kmax(v[i], w[i]);
                }
            }
        }
    }
}

// Load `size` keys (size <= B) from `src`, pad the rest with `pad`, sort,
// and leave the sorted batch in LDS `dst` (B keys).  `src` is global memory.
template <typename K, int B, int NT>
__device__ __forceinline__ void wg_sort_batch(K* dst, const K* src, int size, K pad) {
    constexpr int P = B / NT;
    K v[P];
    const int t = threadIdx.x;
    if (size == B && ((uintptr_t)src % 16) == 0) {
        load_run<K, P>(v, src + t * P);
    } else {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int g = t * P + i;
            v[i] = g < size ? src[g] : pad;
        }
    }
    wg_sort_regs<K, P, NT>(v, dst);
    store_run<K, P>(dst + t * P, v);
    __syncthreads();
}

// ---------------------------------------------------------------- merge path
// Merge the sorted LDS batches a[B] and b[B]; thread t returns outputs
// [t*L, t*L+L) of the merged 2B keys, L = 2B/NT.  Ties take `a` first, as in
// imergePath (heap/heaputil.cuh:122-145).  Each thread finds its diagonal
// split with a binary search over LDS, then merges serially with one LDS read
// per output.
template <typename K, int B, int NT>
__device__ __forceinline__ void wg_merge_path(const K* a, const K* b, K (&out)[2 * B / NT]) {
    constexpr int L = 2 * B / NT;
    const int d = threadIdx.x * L;
    int lo = d > B ? d - B : 0;
    int hi = d < B ? d : B;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(b[d - mid - 1] < a[mid])) lo = mid + 1;  // a[mid] <= b[d-mid-1]
        else hi = mid;
    }
    int ai = lo, bi = d - lo;
    K av = a[ai < B ? ai : B - 1];
    K bv = b[bi < B ? bi : B - 1];
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const bool ta = (bi >= B) || (ai < B && !(bv < av));
        out[s] = ta ? av : bv;
        if (ta) {
            ++ai;
            av = a[ai < B ? ai : B - 1];
        } else {
            ++bi;
            bv = b[bi < B ? bi : B - 1];
        }
    }
}

// Destination tags for the two halves of a merge.
struct ToNone {};
template <typename K> struct ToPtr { K* p; };   // LDS or global, B keys

template <typename K, int B, int NT, typename D>
__device__ __forceinline__ void put_half(const D& dst, const K (&out)[2 * B / NT], int run_off) {
    if constexpr (!std::is_same<D, ToNone>::value) {
        constexpr int L = 2 * B / NT;
        store_run<K, L>(dst.p + run_off, out);
    }
}

// Merge a and b (LDS) and write the B smallest keys to `small`, the B largest
// to `large`.  Destinations may alias a or b: all reads finish before the
// first write.  Ends with a barrier.
template <typename K, int B, int NT, typename DS, typename DL>
__device__ __forceinline__ void wg_merge(const K* a, const K* b, const DS& small, const DL& large) {
    constexpr int L = 2 * B / NT;
    K out[L];
    wg_merge_path<K, B, NT>(a, b, out);
    __syncthreads();
    const int d = threadIdx.x * L;
    if (d < B) put_half<K, B, NT>(small, out, d);
    else put_half<K, B, NT>(large, out, d - B);
    __syncthreads();
}

}  // namespace bgpq
