// Workgroup-collective batch primitives for the batched GPU priority queue,
// written for CDNA4 (gfx950, wave64).
//
// Counterpart of the reference's heap/heaputil.cuh:
//   batchCopy   (heap/heaputil.cuh:4-12)   -> g2l_batch / l2g_batch / take_batch
//   ibitonicSort(heap/heaputil.cuh:23-47)  -> wg_sort_batch
//   imergePath  (heap/heaputil.cuh:99-158) -> wg_merge (one of wg_merge_bitonic,
//                                             wg_merge_mp, wg_merge_grp; wg_merge_ser
//                                             is a tested alternative)
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
#include <climits>
#include <stdint.h>

#include <type_traits>

namespace bgpq {

// ---------------------------------------------------------------- memory model
// Cross-workgroup hand-off on gfx950.  Every access to shared heap data is
// device-coherent: 16-byte buffer loads and stores with the sc1 cache policy
// (the encoding LLVM uses for agent-scope atomic loads and stores), which
// neither hit stale lines in another XCD's L2 nor leave dirty lines behind.
// A publisher therefore only has to wait until its own stores are acknowledged
// (s_waitcnt vmcnt(0) in every wave), pass a barrier, and let lane 0 store the
// status word; a consumer's lane 0 spins on the status word, then a barrier
// orders every wave's data loads after it.  No whole-L2 writeback
// (buffer_wbl2) or invalidate (buffer_inv) is needed, which is what the plain
// acquire/release fences cost (about 0.7 us and 0.3 us per wave on MI355X).
constexpr int kSC1 = 16;  // CPol::SC1 on gfx940+: device-coherent

__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
    // raw buffer over [base, base + 2 GiB); 0x00020000 = the gfx9 word-3 format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7FFFFFFF, 0x00020000);
}
__device__ __forceinline__ uint4 gload16(__amdgpu_buffer_rsrc_t r, int byte_off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, kSC1);
    return make_uint4(v.x, v.y, v.z, v.w);
}
// One key, device-coherent, at lane_off (per lane) + row_off (a constant)
// bytes from the resource's base.
template <typename K>
__device__ __forceinline__ void gstore_row(__amdgpu_buffer_rsrc_t r, int lane_off, int row_off, K v) {
    if constexpr (sizeof(K) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, r, lane_off, row_off, kSC1);
    } else {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        u32x2 q;
        q.x = (uint32_t)v;
        q.y = (uint32_t)((uint64_t)v >> 32);
        __builtin_amdgcn_raw_buffer_store_b64(q, r, lane_off, row_off, kSC1);
    }
}
__device__ __forceinline__ void gstore16(__amdgpu_buffer_rsrc_t r, int byte_off, uint4 q) {
    u32x4 v;
    v.x = q.x; v.y = q.y; v.z = q.z; v.w = q.w;
    __builtin_amdgcn_raw_buffer_store_b128(v, r, byte_off, 0, kSC1);
}

// single-key coherent access (relaxed agent-scope atomics compile to sc1)
template <typename K>
__device__ __forceinline__ K gload_key(const K* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename K>
__device__ __forceinline__ void gstore_key(K* p, K v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }

// ---------------------------------------------------------------- key helpers
template <typename K> __device__ __forceinline__ K kmin(K a, K b) { return b < a ? b : a; }
template <typename K> __device__ __forceinline__ K kmax(K a, K b) { return b < a ? a : b; }

// One side of a compare-exchange whose side is a per-lane predicate:
// take_min ? min(a, b) : max(a, b).  For 32-bit keys this is
// med3(a, b, take_min ? 0 : ~0), which the compiler emits as a single
// v_med3_u32 (the networks here are VALU-bound, and the select form costs
// three operations); 64-bit keys keep the select form.
template <typename K> __device__ __forceinline__ K kminmax(K a, K b, bool take_min) {
    if constexpr (sizeof(K) == 4) {
        const K k = take_min ? K(0) : ~K(0);
        return kmax(kmin(a, b), kmin(kmax(a, b), k));
    } else {
        return take_min ? kmin(a, b) : kmax(a, b);
    }
}

// xor-lane exchange inside a wave64.  Distances 1 and 2 use DPP quad_perm,
// 4..16 ds_swizzle in bit mode (and=0x1f, xor=m), 32 ds_bpermute.  With
// 1024-thread workgroups the sort and merge networks are VALU-bound (a
// wave64 op occupies a 16-lane SIMD for 4 cycles), and these one-instruction
// forms measured faster there than the DPP row-rotate and v_permlane*_swap
// sequences, which need a select per exchange.
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
// Heap nodes live in global memory and are moved with coherent 16-byte
// accesses; LDS copies are plain.  Thread t moves chunks c = t, t+NT, ...
template <typename K, int B, int NT>
__device__ __forceinline__ void g2l_batch(K* __restrict__ lds, const K* __restrict__ g) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(g);
    uint4* d = reinterpret_cast<uint4*>(lds);
#pragma unroll
    for (int k = 0; k < (NV + NT - 1) / NT; ++k) {
        const int c = threadIdx.x + k * NT;
        if (NV % NT == 0 || c < NV) d[c] = gload16(r, c * 16);
    }
}
// g2l_batch in two steps: issue() starts the coherent loads into registers
// and returns at once; commit() writes them to LDS (waiting for them there).
// Other work placed between the two hides the load latency.
template <typename K, int B, int NT>
struct BatchLoad {
    static constexpr int NV = B * (int)sizeof(K) / 16;
    static constexpr int PER = (NV + NT - 1) / NT;
    uint4 q[PER];
    __device__ __forceinline__ void issue(const K* g) {
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(g);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = threadIdx.x + k * NT;
            if (NV % NT == 0 || c < NV) q[k] = gload16(r, c * 16);
        }
    }
    __device__ __forceinline__ void commit(K* lds) const {
        uint4* d = reinterpret_cast<uint4*>(lds);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = threadIdx.x + k * NT;
            if (NV % NT == 0 || c < NV) d[c] = q[k];
        }
    }
};

template <typename K, int B, int NT>
__device__ __forceinline__ void l2g_batch(K* __restrict__ g, const K* __restrict__ lds) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(g);
    const uint4* s = reinterpret_cast<const uint4*>(lds);
#pragma unroll
    for (int k = 0; k < (NV + NT - 1) / NT; ++k) {
        const int c = threadIdx.x + k * NT;
        if (NV % NT == 0 || c < NV) gstore16(r, c * 16, s[c]);
    }
}
template <typename K, int B, int NT>
__device__ __forceinline__ void l2l_batch(K* __restrict__ dst, const K* __restrict__ src) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int k = 0; k < (NV + NT - 1) / NT; ++k) {
        const int c = threadIdx.x + k * NT;
        if (NV % NT == 0 || c < NV) d[c] = s[c];
    }
}
// global -> global (heap node to caller buffer), coherent loads
// Global-to-global copy of a batch (a coherent read of a heap node into the
// caller's buffer) that also leaves the batch's last key in *last (LDS)
template <typename K, int B, int NT>
__device__ __forceinline__ void g2g_batch_last(K* __restrict__ dst, const K* __restrict__ g, K* last) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(g);
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int k = 0; k < (NV + NT - 1) / NT; ++k) {
        const int c = threadIdx.x + k * NT;
        if (NV % NT == 0 || c < NV) {
            const uint4 v = gload16(r, c * 16);
            d[c] = v;
            if (c == NV - 1) {
                if constexpr (sizeof(K) == 4) *last = (K)v.w;
                else *last = (K)(((unsigned long long)v.w << 32) | v.z);
            }
        }
    }
}

template <typename K>
__device__ __forceinline__ uint4 splat16(K v) {
    uint4 f;
    if constexpr (sizeof(K) == 4) {
        f.x = f.y = f.z = f.w = (uint32_t)v;
    } else {
        f.x = f.z = (uint32_t)v;
        f.y = f.w = (uint32_t)((uint64_t)v >> 32);
    }
    return f;
}

// Move B keys global -> LDS and reset the source to `fill` in the same
// pass: each thread resets exactly the 16-byte chunk it loaded, so no other
// thread can overwrite a chunk before it is read.
template <typename K, int B, int NT>
__device__ __forceinline__ void take_batch(K* __restrict__ lds, K* __restrict__ g, K fill) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(g);
    const uint4 f = splat16(fill);
    uint4* d = reinterpret_cast<uint4*>(lds);
#pragma unroll
    for (int k = 0; k < (NV + NT - 1) / NT; ++k) {
        const int c = threadIdx.x + k * NT;
        if (NV % NT == 0 || c < NV) {
            d[c] = gload16(r, c * 16);
            gstore16(r, c * 16, f);
        }
    }
}

template <typename K, int B, int NT>
__device__ __forceinline__ void fill_lds(K* lds, K v) {
    constexpr int NV = B * (int)sizeof(K) / 16;
    const uint4 f = splat16(v);
    uint4* d = reinterpret_cast<uint4*>(lds);
#pragma unroll
    for (int k = 0; k < (NV + NT - 1) / NT; ++k) {
        const int c = threadIdx.x + k * NT;
        if (NV % NT == 0 || c < NV) d[c] = f;
    }
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
    } else if constexpr (N * sizeof(K) == 8 && sizeof(K) == 4) {
        // one 8-byte access: two 4-byte ones two words apart would put lanes
        // t and t + 16 of an LDS pass on the same bank
        reinterpret_cast<uint2*>(dst)[0] = make_uint2((uint32_t)v[0], (uint32_t)v[1]);
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
    } else if constexpr (N * sizeof(K) == 8 && sizeof(K) == 4) {
        const uint2 q = reinterpret_cast<const uint2*>(src)[0];
        v[0] = (K)q.x;
        v[1] = (K)q.y;
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = src[i];
    }
}

// A 32-byte run as two 16-byte words and back, in registers (a run cast
// to uint4 in a local array, indexed by a per-lane value, went through
// scratch memory).
template <typename K, int N>
__device__ __forceinline__ void run_to_words(const K (&v)[N], uint4& q0, uint4& q1) {
    static_assert(N * sizeof(K) == 32, "32-byte run");
    if constexpr (sizeof(K) == 4) {
        q0 = make_uint4(v[0], v[1], v[2], v[3]);
        q1 = make_uint4(v[4], v[5], v[6], v[7]);
    } else {
        q0 = make_uint4((uint32_t)v[0], (uint32_t)(v[0] >> 32), (uint32_t)v[1], (uint32_t)(v[1] >> 32));
        q1 = make_uint4((uint32_t)v[2], (uint32_t)(v[2] >> 32), (uint32_t)v[3], (uint32_t)(v[3] >> 32));
    }
}
template <typename K, int N>
__device__ __forceinline__ void words_to_run(K (&v)[N], const uint4& q0, const uint4& q1) {
    static_assert(N * sizeof(K) == 32, "32-byte run");
    if constexpr (sizeof(K) == 4) {
        v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w;
        v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
    } else {
        v[0] = (K)q0.x | ((K)q0.y << 32);
        v[1] = (K)q0.z | ((K)q0.w << 32);
        v[2] = (K)q1.x | ((K)q1.y << 32);
        v[3] = (K)q1.z | ((K)q1.w << 32);
    }
}
__device__ __forceinline__ uint4 sel_words(bool c, const uint4& x, const uint4& y) {
    return make_uint4(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z, c ? x.w : y.w);
}

// The same moves for runs in LDS.  A 32-byte run goes as two 16-byte
// accesses, which the LDS serves 8 lanes at a time; runs 8 words apart put
// lanes t and t + 4 on the same banks, so lanes of odd quads move their
// second half first and every access instruction covers all 32 banks.
template <typename K, int N>
__device__ __forceinline__ void store_run_lds(K* dst, const K (&v)[N]) {
    if constexpr (N * sizeof(K) == 32) {
        uint4 q0, q1;
        run_to_words<K, N>(v, q0, q1);
        const int sw = ((int)threadIdx.x >> 2) & 1;
        uint4* d = reinterpret_cast<uint4*>(dst);
        d[sw] = sel_words(sw, q1, q0);
        d[sw ^ 1] = sel_words(sw, q0, q1);
    } else {
        store_run<K, N>(dst, v);
    }
}
template <typename K, int N>
__device__ __forceinline__ void load_run_lds(K (&v)[N], const K* src) {
    if constexpr (N * sizeof(K) == 32) {
        const int sw = ((int)threadIdx.x >> 2) & 1;
        const uint4* s = reinterpret_cast<const uint4*>(src);
        const uint4 x = s[sw], y = s[sw ^ 1];
        words_to_run<K, N>(v, sel_words(sw, y, x), sel_words(sw, x, y));
    } else {
        load_run<K, N>(v, src);
    }
}

// ---------------------------------------------------------------- bitonic sort
// Position of thread t's run of P keys in the sort's LDS exchange.  LDS has
// 32 banks of 4 bytes.  A run of 32 bytes (P = 8 u32 or 4 u64 keys) moves as
// two 16-byte accesses, which the LDS serves 8 lanes at a time; runs 8 words
// apart would put lanes t and t + 4 on the same banks (a 2-way conflict), so
// 16 bytes of padding follow every 4 runs and the 8 lanes cover all 32
// banks.  Shorter runs are conflict-free unpadded.  The exchange then needs
// B*sizeof(K) + NT*4 bytes.
template <typename K, int P>
__device__ __forceinline__ int xch_pos(int t) {
    if constexpr (P * sizeof(K) == 32) return t * P + (t >> 2) * (16 / (int)sizeof(K));
    else return t * P;
}
template <typename K, int B, int NT>
constexpr int sort_xch_bytes() {
    return (B / NT) * (int)sizeof(K) == 32 ? B * (int)sizeof(K) + NT * 4 : B * (int)sizeof(K);
}

// Ascending sort of the B = P*NT keys held blocked in registers.  `xch` is an
// LDS scratch of sort_xch_bytes() used for the cross-wave strides (may alias
// nothing that is live).  Semantics match ibitonicSort (heap/heaputil.cuh:
// 23-47): the result is the sorted multiset.
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
                        v[i] = kminmax(v[i], w, take_min);
                    }
                } else {
                    store_run<K, P>(xch + xch_pos<K, P>(t), v);
                    __syncthreads();
                    K w[P];
                    load_run<K, P>(w, xch + xch_pos<K, P>(t ^ m));
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < P; ++i) v[i] = kminmax(v[i], w[i], take_min);
                }
            }
        }
    }
}

// Load `size` keys (size <= B) from `src`, pad the rest with `pad`, sort,
// and leave the sorted batch in LDS `dst` (B keys; the sort's exchange uses
// sort_xch_bytes() from dst).  `src` is global memory or LDS.
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
    store_run_lds<K, P>(dst + t * P, v);
    __syncthreads();
}

// ---------------------------------------------------------------- merge path
// The merges below all produce the sorted union of a[B] and b[B] split at
// B.  imergePath (heap/heaputil.cuh:122-145) takes `a` first on ties; for
// scalar keys the tie rule leaves the output values unchanged.
//
// Merge path with a short dependent chain.  The diagonal split is found
// with a 4-ary search (three independent LDS probes per round, so about
// log4(B) rounds instead of log2(B)), then the thread loads the L keys that
// follow the split in each input at once and keeps the L smallest of those
// 2L keys: c[s] = min(x[s], y[L-1-s]) is bitonic and holds them, and
// log2(L) in-register half-cleaning stages sort it.
template <typename K>
__device__ __forceinline__ K key_max() { return ~K(0); }

template <typename K, int B, int NT>
__device__ __forceinline__ void wg_merge_path_fast(const K* a, const K* b, K (&out)[2 * B / NT]);

// Outputs d .. d+L-1 of the merge of a[B] and b[B] (0 <= d, d+L <= 2B).
template <typename K, int B, int L>
__device__ __forceinline__ void merge_path_at(const K* a, const K* b, int d, K (&out)[L]) {
    int lo = d > B ? d - B : 0;
    int hi = d < B ? d : B;
    // f(i) = a[i] is among the first d outputs = !(b[d-1-i] < a[i]); true then false
    while (hi - lo > 3) {
        const int span = hi - lo;
        const int q1 = lo + (span >> 2), q2 = lo + (span >> 1), q3 = lo + ((3 * span) >> 2);
        const K a1 = a[q1], a2 = a[q2], a3 = a[q3];
        const K b1 = b[d - 1 - q1], b2 = b[d - 1 - q2], b3 = b[d - 1 - q3];
        if (!(b3 < a3)) {
            lo = q3 + 1;
        } else if (!(b2 < a2)) {
            lo = q2 + 1;
            hi = q3;
        } else if (!(b1 < a1)) {
            lo = q1 + 1;
            hi = q2;
        } else {
            hi = q1;
        }
    }
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(b[d - 1 - mid] < a[mid])) lo = mid + 1;
        else hi = mid;
    }
    const int ai = lo, bi = d - lo;
    if (ai + L <= B && bi + L <= B) {
        // both runs inside the inputs: plain loads at fixed offsets
        K x[L], y[L];
#pragma unroll
        for (int s = 0; s < L; ++s) {
            x[s] = a[ai + s];
            y[s] = b[bi + s];
        }
#pragma unroll
        for (int s = 0; s < L; ++s) out[s] = kmin(x[s], y[L - 1 - s]);
    } else {
        // unconditional loads at clamped indexes, then a select: a guarded
        // load would become a branch with its own wait per key
        K x[L], y[L];
#pragma unroll
        for (int s = 0; s < L; ++s) {
            x[s] = a[kmin(ai + s, B - 1)];
            y[s] = b[kmin(bi + s, B - 1)];
        }
#pragma unroll
        for (int s = 0; s < L; ++s) {
            x[s] = (ai + s < B) ? x[s] : key_max<K>();
            y[s] = (bi + s < B) ? y[s] : key_max<K>();
        }
#pragma unroll
        for (int s = 0; s < L; ++s) out[s] = kmin(x[s], y[L - 1 - s]);
    }
#pragma unroll
    for (int j = L / 2; j > 0; j >>= 1) {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            if ((i & j) == 0) {
                const K u = out[i], v = out[i | j];
                out[i] = kmin(u, v);
                out[i | j] = kmax(u, v);
            }
        }
    }
}

template <typename K, int B, int NT>
__device__ __forceinline__ void wg_merge_path_fast(const K* a, const K* b, K (&out)[2 * B / NT]) {
    constexpr int L = 2 * B / NT;
    merge_path_at<K, B, L>(a, b, (int)threadIdx.x * L, out);
}

// Destination tags for the halves of a merge: LDS, global heap node
// (coherent stores), or dropped.
// Merge of two sorted LDS batches with few registers: each thread finds the
// co-rank of its own run of 2B/NT outputs by a binary search and merges that
// run serially (a first among equal keys).  Output i < B goes to small[i],
// the others to large[i - B], by plain stores; neither may alias a or b.
// No barrier inside: the caller orders the inputs' stores before the call
// and the outputs' use after it.
template <typename K, int B, int NT>
__device__ void wg_merge_lean(const K* a, const K* b, K* small, K* large) {
    constexpr int PT = 2 * B / NT;
    const int d0 = (int)threadIdx.x * PT;
    int lo = d0 > B ? d0 - B : 0, hi = d0 < B ? d0 : B;
    while (lo < hi) {  // first i with b[d0 - i - 1] < a[i]
        const int mid = (lo + hi) >> 1;
        if (!(b[d0 - mid - 1] < a[mid])) lo = mid + 1;
        else hi = mid;
    }
    int i = lo, j = d0 - lo;
    for (int k = 0; k < PT; ++k) {
        const K x = a[i < B ? i : B - 1], y = b[j < B ? j : B - 1];
        const bool ta = i < B && (j >= B || !(y < x));
        const K v = ta ? x : y;
        i += ta;
        j += !ta;
        const int d = d0 + k;
        if (d < B) small[d] = v;
        else large[d - B] = v;
    }
}

struct ToNone {};
template <typename K> struct ToLds { K* p; };
template <typename K> struct ToGlobal { K* p; };

// Store a thread's run of N keys at key offset `off` of a destination.
template <typename K, int N, typename D>
__device__ __forceinline__ void put_run(const D& dst, int off, const K (&v)[N]) {
    if constexpr (std::is_same<D, ToNone>::value) {
        return;
    } else if constexpr (std::is_same<D, ToLds<K>>::value) {
        store_run<K, N>(dst.p + off, v);
    } else if constexpr ((N * sizeof(K)) % 16 != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) gstore_key(dst.p + off + i, v[i]);
    } else {
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(dst.p);
        constexpr int W = 16 / (int)sizeof(K);
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
            gstore16(r, (off + c * W) * (int)sizeof(K), q);
        }
    }
}

// Merge path with a serial, branch-free merge: a binary search finds the
// thread's diagonal split, then each of its L outputs costs one compare,
// two selects and one LDS load at a selected address (the next key of the
// side just consumed).  About 8 VALU operations per output, against the
// log2(2L)-stage cleaner of wg_merge_path_fast; the dependent chain is
// longer (one LDS round trip per output), which the other waves hide.
template <typename K, int B, int NT>
__device__ __forceinline__ void wg_merge_path_serial(const K* a, const K* b, K (&out)[2 * B / NT]) {
    constexpr int L = 2 * B / NT;
    const int d = threadIdx.x * L;
    int lo = d > B ? d - B : 0;
    int hi = d < B ? d : B;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(b[d - 1 - mid] < a[mid])) lo = mid + 1;
        else hi = mid;
    }
    int ai = lo, bi = d - lo;
    K av = a[kmin(ai, B - 1)];
    K bv = b[kmin(bi, B - 1)];
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const bool ta = bi >= B || (ai < B && !(bv < av));
        out[s] = ta ? av : bv;
        ai += ta ? 1 : 0;
        bi += ta ? 0 : 1;
        const K nv = ta ? a[kmin(ai, B - 1)] : b[kmin(bi, B - 1)];
        av = ta ? nv : av;
        bv = ta ? bv : nv;
    }
}

template <typename K, int B, int NT, typename DS, typename DL>
__device__ __forceinline__ void wg_merge_ser(const K* a, const K* b, const DS& small, const DL& large) {
    constexpr int L = 2 * B / NT;
    K out[L];
    wg_merge_path_serial<K, B, NT>(a, b, out);
    __syncthreads();
    const int d = threadIdx.x * L;
    if (d < B) put_run<K, L>(small, d, out);
    else put_run<K, L>(large, d - B, out);
    __syncthreads();
}

// Merge a and b (LDS) with the merge-path kernel above and write the B
// smallest keys to `small`, the B largest to `large`.  Destinations may
// alias a or b.  Ends with a barrier.
template <typename K, int B, int NT, typename DS, typename DL>
__device__ __forceinline__ void wg_merge_mp(const K* a, const K* b, const DS& small, const DL& large) {
    constexpr int L = 2 * B / NT;
    K out[L];
    wg_merge_path_fast<K, B, NT>(a, b, out);
    __syncthreads();
    const int d = threadIdx.x * L;
    if (d < B) put_run<K, L>(small, d, out);
    else put_run<K, L>(large, d - B, out);
    __syncthreads();
}

// Register bitonic merge.  For ascending a and b, c[i] = min(a[i], b[B-1-i])
// holds the B smallest keys of a and b and d[i] = max(a[i], b[B-1-i]) the B
// largest, each as a bitonic sequence (Batcher's half-cleaner on a ++
// reverse(b)); log2(B) half-cleaning stages then sort c and d together.
// Thread t keeps c[tP..tP+P) and d[tP..tP+P) in registers, P = B/NT: strides
// below P stay in the lane, strides below 64P use DPP / ds_swizzle /
// ds_bpermute, and larger ones go through LDS, reusing a and b as scratch.
// No step depends on the key values, so the latency is fixed.  Ends with a
// barrier.
template <typename K, int P, int NT>
__device__ __forceinline__ void xlane_minmax(K (&c)[P], K (&d)[P], int m, bool lower) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
        K wc, wd;
        switch (m) {
            case 1: wc = xlane<1>(c[i]); wd = xlane<1>(d[i]); break;
            case 2: wc = xlane<2>(c[i]); wd = xlane<2>(d[i]); break;
            case 4: wc = xlane<4>(c[i]); wd = xlane<4>(d[i]); break;
            case 8: wc = xlane<8>(c[i]); wd = xlane<8>(d[i]); break;
            case 16: wc = xlane<16>(c[i]); wd = xlane<16>(d[i]); break;
            default: wc = xlane<32>(c[i]); wd = xlane<32>(d[i]); break;
        }
        c[i] = kminmax(c[i], wc, lower);
        d[i] = kminmax(d[i], wd, lower);
    }
}

template <typename K, int B, int NT, typename DS, typename DL>
__device__ __forceinline__ void wg_merge_bitonic(K* a, K* b, const DS& small, const DL& large) {
    constexpr int P = B / NT;
    const int t = threadIdx.x;
    K c[P], d[P], av[P], bv[P];
    load_run_lds<K, P>(av, a + t * P);
    load_run_lds<K, P>(bv, b + (B - P - t * P));
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const K x = av[i], y = bv[P - 1 - i];
        c[i] = kmin(x, y);
        d[i] = kmax(x, y);
    }
    constexpr bool kCrossWave = (B / 2) >= 64 * P;
    if constexpr (kCrossWave) __syncthreads();  // a and b become exchange scratch
    constexpr int kStages = ilog2(B);
#pragma unroll
    for (int s = 0; s < kStages; ++s) {
        const int j = (B / 2) >> s;
        if (j < P) {
#pragma unroll
            for (int i = 0; i < P; ++i) {
                if ((i & j) == 0) {
                    const K c0 = c[i], c1 = c[i | j], d0 = d[i], d1 = d[i | j];
                    c[i] = kmin(c0, c1); c[i | j] = kmax(c0, c1);
                    d[i] = kmin(d0, d1); d[i | j] = kmax(d0, d1);
                }
            }
        } else {
            const int m = j / P;
            const bool lower = (t & m) == 0;
            if (j < 64 * P) {
                xlane_minmax<K, P, NT>(c, d, m, lower);
            } else {
                store_run_lds<K, P>(a + t * P, c);
                store_run_lds<K, P>(b + t * P, d);
                __syncthreads();
                K wc[P], wd[P];
                load_run_lds<K, P>(wc, a + (t ^ m) * P);
                load_run_lds<K, P>(wd, b + (t ^ m) * P);
                __syncthreads();
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    c[i] = kminmax(c[i], wc[i], lower);
                    d[i] = kminmax(d[i], wd[i], lower);
                }
            }
        }
    }
    if constexpr (!kCrossWave) __syncthreads();  // all reads of a and b done
    put_run<K, P>(small, t * P, c);
    put_run<K, P>(large, t * P, d);
    __syncthreads();
}

// Group merge.  The 2B outputs are cut into windows of W = G*L keys, one
// per group of G consecutive lanes (merge_group: 128 bytes of keys; L =
// 2B/NT keys per lane):
//  1. The group finds the merge-path split of its window's first output with
//     a cooperative (G+1)-ary search: each round, lane j probes position
//     lo + (j+1)*n/(G+1) and a ballot counts the probes that still lie
//     before the split (2-4 rounds), with no barrier.
//  2. Past the split, the window holds the W smallest of the next W keys of
//     each input.  c[p] = min(a[ai+p], b[bi+W-1-p]) (p < W) is bitonic and
//     holds exactly them (Batcher); log2(W) half-cleaning stages sort it.
// Lane j of a group owns window positions p = s*G + j (s < L): every LDS
// read and write of the group is G consecutive words, free of bank
// conflicts (the split search's probes aside); strides >= G of the cleaning
// stages stay in the lane, smaller ones use DPP / ds_swizzle.  No step depends on a previous output, so the
// dependent chain is the search plus one LDS round trip.
// The group spans one LDS pass: the LDS serves 128 bytes (32 banks of 4
// bytes) per cycle, so a wave's 4-byte accesses go 32 lanes at a time and
// its 8-byte accesses 16 lanes at a time.  A group of that many lanes reads
// 128 consecutive bytes, which hit 32 distinct banks wherever its window
// starts: G = 32 for u32 keys, 16 for u64.  (Two 16-lane groups of
// unrelated offsets in one 32-lane pass of u32 reads met on the same banks:
// 31-35% of the merges' LDS-active cycles.  For u64 keys 32-lane groups
// only add a cleaning stage: a served single-workgroup SSSP solve took 45%
// longer with them.)
template <typename K>
constexpr int merge_group() { return 128 / (int)sizeof(K); }

template <typename K, int B, int NT, int GS = merge_group<K>()>
struct GroupMerge {
    static constexpr int G = GS;
    static constexpr int L = 2 * B / NT;
    static constexpr int W = G * L;
    static_assert(NT % G == 0 && B % W == 0, "group merge shape");

    // Outputs of lane threadIdx.x: window base (first output index of its
    // group) and the lane's column j; out[s] is output base + s*G + j.
    template <bool CP = false>
    __device__ __forceinline__ static void run(const K* a, const K* b, K (&out)[L], int& base, int& j) {
        const int t = threadIdx.x;
        j = t % G;
        base = (t / G) * W;
        window<L, CP>(a, b, base, j, out);
    }

    // One group's window of G*LL outputs starting at output D; lane j gets
    // outputs D + s*G + j (s < LL).  D is the same for the G lanes.
    // CP: look for copy windows (below) first.
    template <int LL, bool CP = false>
    __device__ __forceinline__ static void window(const K* a, const K* b, int D, int j, K (&out)[LL]) {
        constexpr int W = G * LL;
        int lo = D > B ? D - B : 0;
        int hi = D < B ? D : B;
        const int shift = ((int)__lane_id() / G) * G;
        // split = first q in [lo, hi] with b[D-1-q] < a[q] (or hi).  Lane j
        // probes q_j = lo + (j+1)*s - 1 with an odd stride s >= n/(G+1)
        // (probes at or past hi count as not before the split), so the G
        // probes of a group fall on G distinct banks in both inputs and
        // each round leaves an interval of at most s; with n <= G + 1 the
        // probes are lo, lo+1, ... and at most one more round follows.
        // Branch-free rounds: both probe loads at the clamped index every
        // time (a short-circuit would make them a branch with its own exec
        // mask), the group's count from its 32-bit half of the ballot.
        while (lo < hi) {
            const int n = hi - lo;
            const int st = ((n + G) / (G + 1)) | 1;
            const int q = lo + (j + 1) * st - 1;
            const int qc = kmin(q, hi - 1);
            const bool before = (q < hi) & !(b[D - 1 - qc] < a[qc]);
            const uint64_t m = __builtin_amdgcn_ballot_w64(before);
            int c;
            if constexpr (G == 32) {
                const int c_lo = __popc((uint32_t)m), c_hi = __popc((uint32_t)(m >> 32));
                c = shift ? c_hi : c_lo;
            } else if constexpr (G == 64) {
                c = __popcll(m);
            } else {
                c = __popcll((m >> shift) & ((1ull << G) - 1));
            }
            const int nlo = lo + c * st;
            const int nhi = c < G ? kmin(nlo + st - 1, hi) : hi;
            lo = nlo;
            hi = nhi;
        }
        const int ai = lo, bi = D - lo;
        if constexpr (CP) {
            // A window whose W outputs all come from one input (its next W
            // keys there are no larger than the other input's next key, or
            // the other input is used up) is a copy, already sorted: no
            // compares.  Where a batch of spread keys meets a node near the
            // root, which holds the smallest keys, that is nearly every
            // window of both halves.  The four probes are the same for the
            // whole group (broadcast loads); they cost one LDS round trip,
            // which the merges of the sift, whose windows are mixed, do not
            // pay (the delete measured 2-6% slower with them at B <= 4096).
            const K a_first = a[kmin(ai, B - 1)], a_last = a[kmin(ai + W - 1, B - 1)];
            const K b_first = b[kmin(bi, B - 1)], b_last = b[kmin(bi + W - 1, B - 1)];
            const bool pa = (ai + W <= B) & ((bi >= B) | !(b_first < a_last));
            const bool pb = (bi + W <= B) & ((ai >= B) | !(a_first < b_last));
            if (pa | pb) {
                const K* src = (pa ? a + ai : b + bi) + j;
#pragma unroll
                for (int s = 0; s < LL; ++s) out[s] = src[s * G];
                return;
            }
        }
        if (ai + W <= B && bi + W <= B) {
            // the window lies inside both inputs (every group but a few near
            // the ends): loads at fixed offsets from two base addresses
            const K* xa = a + ai + j;
            const K* yb = b + bi + (W - 1) - j - (LL - 1) * G;
            K x[LL], y[LL];
#pragma unroll
            for (int s = 0; s < LL; ++s) {
                x[s] = xa[s * G];
                y[s] = yb[(LL - 1 - s) * G];
            }
#pragma unroll
            for (int s = 0; s < LL; ++s) out[s] = kmin(x[s], y[s]);
        } else {
            // unconditional loads at clamped indexes, then a select (a guarded
            // load would become a branch with its own wait per key)
            K x[LL], y[LL];
#pragma unroll
            for (int s = 0; s < LL; ++s) {
                const int p = s * G + j;
                x[s] = a[kmin(ai + p, B - 1)];
                y[s] = b[kmin(bi + W - 1 - p, B - 1)];
            }
#pragma unroll
            for (int s = 0; s < LL; ++s) {
                const int p = s * G + j;
                const K xs = ai + p < B ? x[s] : key_max<K>();
                const K ys = bi + W - 1 - p < B ? y[s] : key_max<K>();
                out[s] = kmin(xs, ys);
            }
        }
        // strides W/2 .. G: partner position p ^ stride is column j, row s ^ (stride/G)
#pragma unroll
        for (int r = LL / 2; r > 0; r >>= 1) {
#pragma unroll
            for (int s = 0; s < LL; ++s) {
                if ((s & r) == 0) {
                    const K u = out[s], v = out[s | r];
                    out[s] = kmin(u, v);
                    out[s | r] = kmax(u, v);
                }
            }
        }
        // strides G/2 .. 1: partner lane j ^ m
        xclean<G / 2, LL>(out, j);
    }

    template <int M, int LL>
    __device__ __forceinline__ static void xclean(K (&out)[LL], int j) {
        if constexpr (M >= 1) {
            const bool lower = (j & M) == 0;
#pragma unroll
            for (int s = 0; s < LL; ++s) {
                const K w = xlane<M>(out[s]);
                out[s] = kminmax(out[s], w, lower);
            }
            xclean<M / 2, LL>(out, j);
        }
    }
};

// Store the group layout (lane j of a G-lane group holds outputs off + s*G + j).
// A row s of a group is one or two 128-byte lines (G keys), so plain
// key-sized stores, one per row, already write whole lines; the row's offset
// is a constant in the instruction.  (A 4x4 transpose inside each quad of
// lanes into 16-byte runs cost about 32 VALU operations per lane per half
// merge; without it the heap sort of 2^26 u32 keys measured 8% faster at
// B = 8192 and 5% at B = 4096, and the same at B <= 2048.)
template <typename K, int L, int G, typename D>
__device__ __forceinline__ void put_strided(const D& dst, int off, int j, const K (&v)[L]) {
    if constexpr (std::is_same<D, ToNone>::value) {
        return;
    } else if constexpr (std::is_same<D, ToLds<K>>::value) {
#pragma unroll
        for (int s = 0; s < L; ++s) dst.p[off + s * G + j] = v[s];
    } else {
        static_assert(G * sizeof(K) % 128 == 0, "a group row is whole 128-byte lines");
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(dst.p);
        const int lane_off = (off + j) * (int)sizeof(K);
#pragma unroll
        for (int s = 0; s < L; ++s) gstore_row<K>(rs, lane_off, s * G * (int)sizeof(K), v[s]);
    }
}

// Merge a and b (LDS) with the group merge and write the B smallest keys to
// `small`, the B largest to `large`.  Destinations may alias a or b.  Ends
// with a barrier.
template <typename K, int B, int NT, bool CP = false, typename DS, typename DL, int GS = merge_group<K>()>
__device__ __forceinline__ void wg_merge_grp(const K* a, const K* b, const DS& small, const DL& large) {
    using GM = GroupMerge<K, B, NT, GS>;
    K out[GM::L];
    int base, j;
    GM::template run<CP>(a, b, out, base, j);
    __syncthreads();
    if (base < B) put_strided<K, GM::L, GM::G>(small, base, j, out);
    else put_strided<K, GM::L, GM::G>(large, base - B, j, out);
    __syncthreads();
}

// One half of a merge: the B smallest (HI = false) or the B largest
// (HI = true) keys of the union of sorted a[B] and b[B], sorted, N = B/NT of
// them in each thread's registers.  a and b are only read, so both halves
// can be taken from the same inputs one after the other, and the first can
// be stored (and its node released) before the second is computed.  Thread
// 0's v[0] is the half's smallest key and thread NT-1's v[N-1] its largest.
// B >= 256 uses the group merge's windows (one workgroup, u32, lower half:
// 1.14 / 1.30 / 1.85 us at B = 2048 / 4096 / 8192 against 1.55 / 1.82 /
// 2.51 us for the merge path), smaller B the merge path.
// CP: look for copy windows first (GroupMerge::window), for the merges of
// an insertion, whose two batches seldom overlap near the root.
// GS: lanes per merge group, 0 for the default: the whole wave for u32 keys
// up to B = 4096, whose merges are latency-bound (the split search is then
// uniform and takes a round fewer, for one more cleaner stage through
// ds_bpermute; heap sort of 2^26 keys: +6% at B = 1024, +5% at 2048 and
// 4096), 128 bytes of keys otherwise (at B = 8192, 8 keys per lane, the
// extra stage cost more than the search saves: -4.7%).
template <typename K, int B, int NT, bool HI, bool CP = false, int GS = 0, bool GRP = (B >= 256)>
struct HalfMerge {
    static constexpr int N = B / NT;
    static constexpr bool kGroup = GRP;
    static constexpr int G = GS ? GS : (sizeof(K) == 4 && B <= 4096) ? 64 : merge_group<K>();
    static_assert(!kGroup || NT % G == 0, "half merge shape");
    K v[N];

    __device__ __forceinline__ void run(const K* a, const K* b) {
        const int t = threadIdx.x;
        if constexpr (kGroup) {
            GroupMerge<K, B, NT, G>::template window<N, CP>(a, b, (HI ? B : 0) + (t / G) * (G * N), t % G, v);
        } else {
            merge_path_at<K, B, N>(a, b, (HI ? B : 0) + t * N, v);
        }
    }
    // Store the half at positions [0, B) of dst.
    template <typename D>
    __device__ __forceinline__ void put(const D& dst) const {
        const int t = threadIdx.x;
        if constexpr (kGroup) put_strided<K, N, G>(dst, (t / G) * (G * N), t % G, v);
        else put_run<K, N>(dst, t * N, v);
    }
    __device__ __forceinline__ K last() const { return v[N - 1]; }
};

// The merge used by the heap, by measurement on MI355X (one workgroup,
// LDS inputs, u32 keys, small half to LDS and large half to global memory):
//   B <= 1024: the register bitonic merge (B = 1024, 512 threads: 0.80 us
//     against 0.93 for the merge path);
//   B = 2048, 4096: the group merge (1.35 / 1.91 us against 1.59 / 2.24 for
//     the merge path);
//   B = 8192: the two group-window halves one after the other (3.55 us
//     against 4.52 for the one-pass group merge, whose 16 keys per lane
//     crowd the registers).
// CP: copy windows (GroupMerge::window), for B >= 2048.
template <typename K, int B, int NT, bool CP = false, typename DS, typename DL>
__device__ __forceinline__ void wg_merge(K* a, K* b, const DS& small, const DL& large) {
    if constexpr (B >= 8192) {
        HalfMerge<K, B, NT, false, CP> lo;
        lo.run(a, b);
        HalfMerge<K, B, NT, true, CP> hi;
        hi.run(a, b);
        __syncthreads();  // destinations may alias a or b
        lo.put(small);
        hi.put(large);
        __syncthreads();
    } else if constexpr (B >= 2048) {
        wg_merge_grp<K, B, NT, CP>(a, b, small, large);
    } else {
        wg_merge_bitonic<K, B, NT>(a, b, small, large);
    }
}

}  // namespace bgpq
