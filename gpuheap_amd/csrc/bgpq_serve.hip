// Batched and served solves (include/bgpq_serve.h): the persistent serve
// kernel, the prepare kernels, the HBM arena and the host side of the ring.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "bgpq_apps.hpp"
#include "bgpq_handle.hpp"
#include "bgpq_serve.h"

namespace bgpq {

// Serve tiers: one persistent kernel per workgroup size, so that every
// batch size runs at the thread count its heap kernels are built for.
// Tier 0 (256 threads) takes B = 256 and 512, tier 1 (512 threads) B = 1024:
// at 256 threads a B = 1024 solve holds twice the keys per lane and spilled
// registers to scratch.  Each tier has its own slot queue.
constexpr int kTiers = 2;
constexpr int kTierNT[kTiers] = {256, 512};
constexpr int tier_of(int B) { return B == 1024 ? 1 : 0; }
// Polls (each ~3.4 us of sleep plus a host-memory read) a workgroup waits
// for the host to post or close before it gives up: about a minute.
constexpr long long kMaxRingPolls = 1ll << 24;

// What a query's arena block needs before its solve: fills (heap with
// init_limits, status, counters, control words, labels), copies of small
// host arrays (knapsack items), and single words set after them.
constexpr int kMaxFills = 8, kMaxPokes = 4, kMaxCopies = 4;
struct FillJobs {
    void* p[kMaxFills];
    long long n[kMaxFills];
    unsigned long long v[kMaxFills];
    int esz[kMaxFills];
    int count;
    void* pp[kMaxPokes];
    unsigned long long pv[kMaxPokes];
    int pesz[kMaxPokes];
    int pokes;
    const void* cs[kMaxCopies];  // pinned host memory
    void* cd[kMaxCopies];
    long long cb[kMaxCopies];    // bytes, a multiple of 4
    int copies;
};
// A served query is prepared by its own team, in chunks of this many bytes
// of fill (and one chunk per copy).
constexpr long long kSetupChunk = 256 << 10;

// What a workgroup needs to take part in a query, in 8-byte words so that it
// can be copied from host memory one word per lane.
struct QueryDesc {
    int32_t type, B, team, n;         // n: vertices (SSSP), width (A*), items (knapsack)
    int32_t source, height, pad0, pad1;
    uint64_t* items;
    int* status;
    int* ctr;
    long long batch_num;
    long long* ctl;
    const int* rowptr;
    const int* col;
    const uint32_t* w;
    uint32_t* labels;
    long long n_labels;
    const uint8_t* free_cells;
    const uint32_t* kw;
    const uint32_t* kv;
    const uint64_t* pw;
    const uint64_t* pv;
    uint32_t* cell;
    KnapNode* pool;
    long long pool_cap;
    uint64_t cap;
    uint64_t first_key;
    // served queries: the team prepares the block (setup_chunks > 0), with
    // the chunk counter and the chunks-done count at sync[0] and sync[1]
    // (device words of the query index, zero between queries)
    FillJobs setup;
    int32_t setup_chunks, pad2;
    unsigned* sync;
    // served queries that want their labels back: pinned host memory the
    // last member copies them to (the host copies on from there)
    uint32_t* labels_host;
};
static_assert(sizeof(QueryDesc) % 8 == 0 && sizeof(QueryDesc) / 8 <= 256, "descriptor words");

// Written by the last team member to leave, read by the host.
struct Completion {
    uint64_t value, digest;
    long long popped, expanded, stale, pruned, pushed;
    int32_t error;
    int32_t done;
};

struct ServeRing {
    const QueryDesc* desc;  // qmask + 1 entries
    const uint32_t* slots;  // smask + 1 entries: query index << 8 | member
    const long long* tail;  // member slots posted
    const int* closed;
    Completion* done;       // per query index
    long long* claim;       // device: member slots taken (only posted ones are)
    // device: the tail as one workgroup last read it from the ring, with
    // kClosedBit once the ring was seen closed (then the tail is final)
    unsigned long long* mirror;
    int* poller;            // device: 1 while a workgroup is the one reading the ring
    int* fault;             // set when a workgroup left after max_polls idle polls
    int* live;              // workgroups still in the kernel
    long long max_polls;
    unsigned qmask, smask;
};
constexpr unsigned long long kClosedBit = 1ull << 62;

// Ring words are read with relaxed system-scope loads, which bypass the
// caches without invalidating them (an acquire would invalidate the L2 on
// every poll, under the workgroups that are solving).  The descriptor words
// are loaded the same way once the tail says they are posted.
__device__ __forceinline__ long long sys_load(const long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int sys_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// One solve, inlined into its tier's kernel: as calls (noinline) the
// solves took 216-246 VGPRs at 512 threads, inlined the kernel takes 162,
// so both tiers' workgroups fit on a CU together.
template <int B, int NT>
__device__ __forceinline__ void solve_sssp(const QueryDesc& d, bool first) {
    const HeapView<uint64_t> hv{d.items, d.status, d.ctr, ~0ull, (int)d.batch_num, 0, 0};
    const SsspApp app{d.rowptr, d.col, d.w, d.labels, d.ctl};
    drive<VAR_BUITDD, B, NT>(hv, app, d.ctl, d.first_key, first);
}
template <int B, int NT>
__device__ __forceinline__ void solve_astar(const QueryDesc& d, bool first) {
    const HeapView<uint64_t> hv{d.items, d.status, d.ctr, ~0ull, (int)d.batch_num, 0, 0};
    const AstarApp app{d.free_cells, d.n, d.height, d.labels, d.ctl};
    drive<VAR_BUITDD, B, NT>(hv, app, d.ctl, d.first_key, first);
}
template <int B, int NT>
__device__ __forceinline__ void solve_knapsack(const QueryDesc& d, bool first) {
    const HeapView<uint64_t> hv{d.items, d.status, d.ctr, ~0ull, (int)d.batch_num, 0, 0};
    const KnapsackApp app{d.n, d.cap, d.kw, d.kv, d.pw, d.pv, d.cell, d.pool, d.pool_cap, d.ctl};
    drive<VAR_BUITDD, B, NT>(hv, app, d.ctl, d.first_key, first);
}
template <int B, int NT>
__device__ __forceinline__ void solve_b(const QueryDesc& d, bool first) {
    if (d.type == BGPQ_Q_SSSP) solve_sssp<B, NT>(d, first);
    else if (d.type == BGPQ_Q_ASTAR) solve_astar<B, NT>(d, first);
    else solve_knapsack<B, NT>(d, first);
}

// ---- a served query's setup, done by its team (no kernel launch of its
// own: with the serve kernels resident, a prepare kernel or a copy kernel
// of the runtime can wait for a free CU until idle workgroups leave).
// Chunk c of the setup: a kSetupChunk slice of one fill, or one copy.  The
// stores are device-coherent (sc1, or agent-scope word stores), as every
// access of the solve is.
template <int NT>
__device__ void setup_chunk(const FillJobs& j, unsigned c) {
    for (int k = 0; k < j.count; ++k) {
        const long long bytes = j.n[k] * j.esz[k];
        const unsigned nc = (unsigned)((bytes + kSetupChunk - 1) / kSetupChunk);
        if (c >= nc) {
            c -= nc;
            continue;
        }
        const long long begin = (long long)c * kSetupChunk;
        const int len = (int)(bytes - begin < kSetupChunk ? bytes - begin : kSetupChunk);
        char* base = (char*)j.p[k] + begin;
        const uint32_t lo = (uint32_t)j.v[k], hi = j.esz[k] == 8 ? (uint32_t)(j.v[k] >> 32) : lo;
        const uint4 q = make_uint4(lo, hi, lo, hi);
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(base);
        const int n16 = len / 16;
        for (int i = threadIdx.x; i < n16; i += NT) gstore16(r, i * 16, q);
        // the rest, in words (begin is a multiple of 16, so word parity
        // within the slice picks the half of an 8-byte pattern)
        for (int o = n16 * 16 + 4 * threadIdx.x; o < len; o += 4 * NT)
            __hip_atomic_store((uint32_t*)(base + o), (o & 4) ? hi : lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if ((int)c < j.copies) {
        const uint32_t* src = (const uint32_t*)j.cs[c];
        uint32_t* dst = (uint32_t*)j.cd[c];
        for (long long i = threadIdx.x; i < j.cb[c] / 4; i += NT)
            __hip_atomic_store(dst + i, __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Every member takes chunks until none are left; the member that completes
// the last one stores the single words and then marks the setup done
// (sync[1] = chunks + 1), which every member waits for.  A chunk is only
// ever taken by a running workgroup, so the wait always ends; it is
// bounded anyway (false: the workgroup gives up).
template <int NT>
__device__ bool setup_query(const QueryDesc& d, long long max_polls) {
    __shared__ unsigned s_c;
    __shared__ int s_fin;
    unsigned* next = d.sync;
    unsigned* done = d.sync + 1;
    const unsigned total = (unsigned)d.setup_chunks;
    while (true) {
        if (threadIdx.x == 0) s_c = __hip_atomic_fetch_add(next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const unsigned c = s_c;
        __syncthreads();  // every lane has read s_c
        if (c >= total) break;
        setup_chunk<NT>(d.setup, c);
        wait_stores();
        __syncthreads();
        if (threadIdx.x == 0)
            s_fin = __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == total;
        __syncthreads();
        if (s_fin) {
            const int k = threadIdx.x;
            if (k < d.setup.pokes) {
                if (d.setup.pesz[k] == 8)
                    __hip_atomic_store((unsigned long long*)d.setup.pp[k], d.setup.pv[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                else
                    __hip_atomic_store((uint32_t*)d.setup.pp[k], (uint32_t)d.setup.pv[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            wait_stores();
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(done, total + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();  // every lane has read s_fin
    }
    if (threadIdx.x == 0) {
        int ok = 1;
        long long n = 0;
        while (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != total + 1) {
            if (++n > max_polls) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        s_fin = ok;
    }
    __syncthreads();
    const bool ok = s_fin != 0;
    __syncthreads();
    return ok;
}

// The last member's epilogue: result, digest and statistics to the host.
template <int NT>
__device__ void finish(const QueryDesc& d, Completion* c) {
    __shared__ unsigned long long red[2][NT / 64];
    const int t = threadIdx.x;
    uint64_t dig = 0, reach = 0;
    if (d.type == BGPQ_Q_SSSP) {
        for (long long v = t; v < d.n_labels; v += NT) {
            const uint32_t x = u32_load(d.labels + v);
            dig += ((uint64_t)x + 1) * (2 * (uint64_t)v + 1);
            reach += x != 0xFFFFFFFFu;
        }
        for (int o = 32; o > 0; o >>= 1) {
            dig += __shfl_xor(dig, o, 64);
            reach += __shfl_xor(reach, o, 64);
        }
        if (__lane_id() == 0) {
            red[0][t / 64] = dig;
            red[1][t / 64] = reach;
        }
    }
    if (d.labels_host) {
        for (long long v = t; v < d.n_labels; v += NT) d.labels_host[v] = u32_load(d.labels + v);
        __threadfence_system();  // before the completion below is released
    }
    __syncthreads();
    if (t == 0) {
        uint64_t value = 0;
        dig = 0;
        if (d.type == BGPQ_Q_SSSP) {
            for (int i = 0; i < NT / 64; ++i) {
                dig += red[0][i];
                value += red[1][i];
            }
        } else if (d.type == BGPQ_Q_ASTAR) {
            value = u32_load(d.labels + d.n_labels - 1);
        } else {
            value = (uint64_t)ctl_load(d.ctl, APP_BEST);
        }
        const int heap_err = a_load(d.ctr + CTR_ERROR * kWordStride);
        const long long left = ctl_load(d.ctl, APP_OUTSTANDING);
        int err = (int)ctl_load(d.ctl, APP_ERROR) | (heap_err << 8);
        // A* ends by the termination vote with its pruned keys in the heap:
        // exactly the outstanding ones (the block is refilled before reuse)
        long long in_heap = 0;
        if (d.type == BGPQ_Q_ASTAR)
            in_heap = (long long)a_load(d.ctr + CTR_BATCH_COUNT * kWordStride) * d.B +
                      a_load(d.ctr + CTR_PARTIAL * kWordStride);
        if (left != in_heap) err |= APPERR_HEAP;
        // every member is past its setup wait: the query index's words are
        // cleared for its next query (ordered before the release of `done`)
        if (d.sync) {
            __hip_atomic_store(d.sync, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(d.sync + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        auto put = [](auto* p, auto v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); };
        put(&c->value, value);
        put(&c->digest, dig);
        put(&c->popped, ctl_load(d.ctl, APP_POPPED));
        put(&c->expanded, ctl_load(d.ctl, APP_EXPANDED));
        put(&c->stale, ctl_load(d.ctl, APP_STALE));
        put(&c->pruned, ctl_load(d.ctl, APP_PRUNED) + in_heap);
        put(&c->pushed, ctl_load(d.ctl, APP_PUSHED));
        put(&c->error, err);
        __hip_atomic_store(&c->done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Persistent serve kernel.  Each workgroup takes the next member slot once
// the host has posted it (a CAS on the claim counter, only while it is
// below the tail, so an idle workgroup holds no slot), joins that query's
// team and runs its solve; the last member to leave publishes the
// completion.  A workgroup ends when the ring is closed and every posted
// slot is taken, or after max_polls unanswered polls (then `fault` is set);
// either way it leaves the `live` count, which the host reads.
// Idle workgroups poll device words only.  One of them at a time (the
// holder of `poller`) reads the host's tail and closed flag and raises
// `mirror` to what it saw: reads of pinned host memory by every idle
// workgroup of the chip (hundreds, each every few microseconds) slowed the
// workgroups that were solving about fivefold.
template <int NT>
__global__ __launch_bounds__(NT) void k_serve(ServeRing r) {
    __shared__ long long s_slot;
    __shared__ int s_go, s_last;
    __shared__ uint32_t s_entry;
    __shared__ __align__(16) QueryDesc s_d;
    const int t = threadIdx.x;
    while (true) {
        if (t == 0) {
            int go = 0;
            long long s = -1, polls = 0;
            bool poller = false;
            while (true) {
                const unsigned long long m = __hip_atomic_load(r.mirror, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const long long c = __hip_atomic_load(r.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (c < (long long)(m & (kClosedBit - 1))) {
                    long long e = c;
                    if (__hip_atomic_compare_exchange_strong(r.claim, &e, c + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        s = c;
                        go = 1;
                        break;
                    }
                    continue;  // another workgroup took slot c
                }
                if (m & kClosedBit) break;  // the tail is final and every slot is taken
                if (!poller && __hip_atomic_load(r.poller, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                    int z = 0;
                    poller = __hip_atomic_compare_exchange_strong(r.poller, &z, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
                }
                if (poller) {
                    // the host posts every slot before it closes the ring: a
                    // tail read after the closed flag was seen is final
                    const int closed = sys_load(r.closed);
                    if (closed) __atomic_thread_fence(__ATOMIC_ACQUIRE);
                    const unsigned long long v = (unsigned long long)sys_load(r.tail) | (closed ? kClosedBit : 0);
                    if (v > m) {
                        __hip_atomic_fetch_max(r.mirror, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        continue;  // act on it at once
                    }
                }
                if (++polls > r.max_polls) {
                    __hip_atomic_fetch_or(r.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                __builtin_amdgcn_s_sleep(127);
            }
            if (poller) __hip_atomic_store(r.poller, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_go = go;
            s_slot = s;
            if (go)
                s_entry = __hip_atomic_load(r.slots + (s_slot & r.smask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            else
                __hip_atomic_fetch_add(r.live, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
        if (!s_go) break;
        const uint32_t qi = s_entry >> 8;
        const int member = (int)(s_entry & 0xFF);
        if (t < (int)(sizeof(QueryDesc) / 8)) {
            const unsigned long long* src = reinterpret_cast<const unsigned long long*>(r.desc + (qi & r.qmask));
            reinterpret_cast<unsigned long long*>(&s_d)[t] =
                __hip_atomic_load(src + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
        if (s_d.setup_chunks > 0 && !setup_query<NT>(s_d, r.max_polls)) {
            // a member of this query never finished its chunk: leave
            if (t == 0) {
                __hip_atomic_fetch_or(r.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_fetch_add(r.live, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            break;
        }
        const bool first = member == 0;
        if constexpr (NT == kTierNT[1]) {
            solve_b<1024, NT>(s_d, first);
        } else {
            if (s_d.B == 256) solve_b<256, NT>(s_d, first);
            else solve_b<512, NT>(s_d, first);
        }
        if (t == 0) {
            wait_stores();
            s_last = __hip_atomic_fetch_add(ctl_word(s_d.ctl, APP_DONE), 1ll, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) == s_d.team - 1;
        }
        __syncthreads();
        if (s_last) finish<NT>(s_d, r.done + (qi & r.qmask));
        __syncthreads();
    }
}

__global__ void k_prepare(FillJobs j) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = 0; k < j.count; ++k) {
        if (j.esz[k] == 8) {
            unsigned long long* p = (unsigned long long*)j.p[k];
            for (long long i = i0; i < j.n[k]; i += stride) p[i] = j.v[k];
        } else {
            uint32_t* p = (uint32_t*)j.p[k];
            for (long long i = i0; i < j.n[k]; i += stride) p[i] = (uint32_t)j.v[k];
        }
    }
}
__global__ void k_poke(FillJobs j) {
    const int k = threadIdx.x;
    if (k < j.pokes) {
        if (j.pesz[k] == 8) *(unsigned long long*)j.pp[k] = j.pv[k];
        else *(uint32_t*)j.pp[k] = (uint32_t)j.pv[k];
    }
}
__global__ void k_ring_reset(long long* claim, unsigned long long* mirror, int* poller, int* fault, int* live,
                             int workgroups) {
    __hip_atomic_store(claim, 0ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(mirror, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(poller, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(fault, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(live, workgroups, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace bgpq

using namespace bgpq;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// First-fit allocator over the arena, 256-byte granules, coalescing frees.
struct Arena {
    size_t size = 0;
    size_t used = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
    void init(size_t s) {
        size = s;
        used = 0;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t roff = it->first + bytes;
                free_.erase(it);
                if (rest) free_[roff] = rest;
                used += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes, 256);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Byte layout of one query's block in the arena.
struct Layout {
    size_t ctl, ctr, status, items, labels, kw, kv, pw, pv, cell, pool, total;
    long long batch_num, n_labels, pool_nodes, cells;
};

int check_query(const bgpq_query* q) {
    if (!q) return bgpq_fail(BGPQ_EINVAL, "null query");
    if (q->batch_size != 256 && q->batch_size != 512 && q->batch_size != 1024)
        return bgpq_fail(BGPQ_EINVAL, "query batch size %d (need 256, 512 or 1024)", q->batch_size);
    if (q->team < 1 || q->team > 64) return bgpq_fail(BGPQ_EINVAL, "query team %d (need 1..64)", q->team);
    switch (q->type) {
        case BGPQ_Q_SSSP:
            if (!q->rowptr || !q->col || !q->weight || q->n_vertices < 1 || q->source < 0 ||
                q->source >= q->n_vertices || q->n_edges < 0)
                return bgpq_fail(BGPQ_EINVAL, "sssp query: bad arguments");
            break;
        case BGPQ_Q_ASTAR:
            if (!q->free_cells || q->width < 1 || q->height < 1 || (int64_t)q->width * q->height > INT32_MAX ||
                (uint64_t)q->width * q->height * 14ull + 14ull * (q->width + q->height) >= 0xFFFFFFFFull)
                return bgpq_fail(BGPQ_EINVAL, "astar query: bad arguments");
            break;
        case BGPQ_Q_KNAPSACK: {
            if (!q->item_weight || !q->item_value || q->n_items < 1 || q->n_items > 255 ||
                q->capacity >= 0xFFFFFFFFull)
                return bgpq_fail(BGPQ_EINVAL, "knapsack query: bad arguments");
            uint64_t vs = 0;
            for (int i = 0; i < q->n_items; ++i) {
                if (q->item_weight[i] == 0) return bgpq_fail(BGPQ_EINVAL, "knapsack query: zero weight");
                vs += q->item_value[i];
            }
            if (vs >= kUbTop) return bgpq_fail(BGPQ_EINVAL, "knapsack query: total value too large");
            break;
        }
        default:
            return bgpq_fail(BGPQ_EINVAL, "query type %d", q->type);
    }
    return BGPQ_OK;
}

Layout layout_of(const bgpq_query* q) {
    Layout L{};
    const long long B = q->batch_size;
    long long keys = q->heap_keys;
    L.pool_nodes = q->pool_nodes > 0 ? q->pool_nodes : (1ll << 20);
    if (keys <= 0) {
        if (q->type == BGPQ_Q_SSSP) keys = q->n_edges + 1;
        else if (q->type == BGPQ_Q_ASTAR) keys = 4ll * q->width * q->height;
        else keys = L.pool_nodes;
    }
    // full batches for every key, the workgroups' in-flight pushes and slack
    L.batch_num = (keys + B - 1) / B + 2ll * q->team + 8;
    L.n_labels = q->type == BGPQ_Q_SSSP ? q->n_vertices : q->type == BGPQ_Q_ASTAR ? (long long)q->width * q->height : 0;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.ctl = take(sizeof(long long) * APP_WORDS * kCtlStride);
    L.ctr = take(sizeof(int) * CTR_WORDS * kWordStride);
    L.status = take(sizeof(int) * (size_t)(L.batch_num + 1) * kWordStride);
    L.items = take(8 * (size_t)(L.batch_num + 1) * B);
    L.labels = take(4 * (size_t)L.n_labels);
    if (q->type == BGPQ_Q_KNAPSACK) {
        const size_t n = (size_t)q->n_items;
        L.cells = (long long)(n + 1) * (long long)(q->capacity + 1);
        L.kw = take(4 * n);
        L.kv = take(4 * n);
        L.pw = take(8 * (n + 1));
        L.pv = take(8 * (n + 1));
        L.cell = take(4 * (size_t)L.cells);
        L.pool = take(sizeof(KnapNode) * (size_t)L.pool_nodes);
    }
    L.total = o;
    return L;
}

// Pinned host bytes a knapsack query's item arrays take while it is
// prepared (w, v: 4 bytes per item; pw, pv: 8 bytes per item + 8).
constexpr size_t kStageBytes = 8192;
static_assert(4 * 255 * 2 + 8 * 256 * 2 <= kStageBytes, "knapsack staging");

// Descriptor and fills of one query placed at `base`.  Batch form: knapsack items are uploaded here (on stream s, finished before
// returning) and the fills are left to launch_fills.  Served form: the
// items are written to `stage` (pinned host memory) and everything becomes
// the descriptor's setup, which the query's team runs (setup_query).
int build_query(const bgpq_query* q, char* base, const Layout& L, hipStream_t s, QueryDesc* d, FillJobs* jobs,
                bool served = false, char* stage = nullptr) {
    std::memset(d, 0, sizeof *d);
    d->type = q->type;
    d->B = q->batch_size;
    d->team = q->team;
    d->items = (uint64_t*)(base + L.items);
    d->status = (int*)(base + L.status);
    d->ctr = (int*)(base + L.ctr);
    d->batch_num = L.batch_num;
    d->ctl = (long long*)(base + L.ctl);
    d->labels = (uint32_t*)(base + L.labels);
    d->n_labels = L.n_labels;
    FillJobs j{};
    auto fill = [&](void* p, long long n, unsigned long long v, int esz) {
        j.p[j.count] = p;
        j.n[j.count] = n;
        j.v[j.count] = v;
        j.esz[j.count] = esz;
        ++j.count;
    };
    auto poke = [&](void* p, unsigned long long v, int esz) {
        j.pp[j.pokes] = p;
        j.pv[j.pokes] = v;
        j.pesz[j.pokes] = esz;
        ++j.pokes;
    };
    fill(d->items, (L.batch_num + 1) * q->batch_size, ~0ull, 8);
    fill(d->status, (L.batch_num + 1) * kWordStride, 0, 4);
    fill(d->ctr, CTR_WORDS * kWordStride, 0, 4);
    fill(d->ctl, APP_WORDS * kCtlStride, 0, 8);
    poke(d->ctl + (size_t)APP_OUTSTANDING * kCtlStride, 1, 8);
    if (q->type == BGPQ_Q_SSSP) {
        d->n = q->n_vertices;
        d->source = q->source;
        d->rowptr = q->rowptr;
        d->col = q->col;
        d->w = q->weight;
        fill(d->labels, L.n_labels, 0xFFFFFFFFull, 4);
        poke(d->labels + q->source, 0, 4);
        d->first_key = (uint64_t)(uint32_t)q->source;
    } else if (q->type == BGPQ_Q_ASTAR) {
        d->n = q->width;
        d->height = q->height;
        d->free_cells = q->free_cells;
        fill(d->labels, L.n_labels, 0xFFFFFFFFull, 4);
        poke(d->labels, 0, 4);
        poke(d->ctr + CTR_BUSY * kWordStride, 1, 4);  // the seeding member, for the termination vote
        const uint32_t dx = (uint32_t)(q->width - 1), dy = (uint32_t)(q->height - 1);
        d->first_key = (uint64_t)(10u * std::max(dx, dy) + 4u * std::min(dx, dy)) << 32;
    } else {
        const int n = q->n_items;
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
            return (uint64_t)q->item_value[x] * q->item_weight[y] > (uint64_t)q->item_value[y] * q->item_weight[x];
        });
        std::vector<uint32_t> w(n), v(n);
        std::vector<uint64_t> pw(n + 1, 0), pv(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            w[i] = q->item_weight[ord[i]];
            v[i] = q->item_value[ord[i]];
            pw[i + 1] = pw[i] + w[i];
            pv[i + 1] = pv[i] + v[i];
        }
        uint64_t ub0 = 0, room = q->capacity;
        for (int i = 0; i < n; ++i) {
            if (w[i] <= room) {
                room -= w[i];
                ub0 += v[i];
            } else {
                ub0 += room * v[i] / w[i];
                break;
            }
        }
        d->n = n;
        d->cap = q->capacity;
        d->kw = (const uint32_t*)(base + L.kw);
        d->kv = (const uint32_t*)(base + L.kv);
        d->pw = (const uint64_t*)(base + L.pw);
        d->pv = (const uint64_t*)(base + L.pv);
        d->cell = (uint32_t*)(base + L.cell);
        d->pool = (KnapNode*)(base + L.pool);
        d->pool_cap = L.pool_nodes;
        d->first_key = ((((kUbTop - ub0) << 8) | 255ull) << 32) | 0ull;
        const void* src[4] = {w.data(), v.data(), pw.data(), pv.data()};
        void* dst[4] = {(void*)d->kw, (void*)d->kv, (void*)d->pw, (void*)d->pv};
        const size_t len[4] = {4 * (size_t)n, 4 * (size_t)n, 8 * (size_t)(n + 1), 8 * (size_t)(n + 1)};
        if (served) {
            if (!stage) return bgpq_fail(BGPQ_EINVAL, "knapsack query: no staging memory");
            size_t o = 0;
            for (int k = 0; k < 4; ++k) {
                std::memcpy(stage + o, src[k], len[k]);
                j.cs[k] = stage + o;
                j.cd[k] = dst[k];
                j.cb[k] = (long long)len[k];
                o += align_up(len[k], 16);
            }
            j.copies = 4;
        } else {
            for (int k = 0; k < 4; ++k) BGPQ_HIPCHK(hipMemcpyAsync(dst[k], src[k], len[k], hipMemcpyHostToDevice, s));
        }
        fill(d->cell, L.cells, 0, 4);
        fill(d->pool, 4, 0, 4);  // the root: depth 0, nothing used, value 0
        poke(d->cell, 1, 4);
        poke(d->ctl + (size_t)APP_POOL * kCtlStride, 1, 8);
        if (!served) BGPQ_HIPCHK(hipStreamSynchronize(s));  // the host vectors above go out of scope
    }
    *jobs = j;
    if (served) {
        d->setup = j;
        long long chunks = j.copies;
        for (int k = 0; k < j.count; ++k) chunks += (j.n[k] * j.esz[k] + kSetupChunk - 1) / kSetupChunk;
        d->setup_chunks = (int32_t)chunks;
    }
    return BGPQ_OK;
}

// The fills, then the single words (a second launch: a word inside a
// filled range must be set after the fill).
void launch_fills(const FillJobs& j, hipStream_t s, int blocks) {
    hipLaunchKernelGGL(k_prepare, dim3(blocks), dim3(256), 0, s, j);
    hipLaunchKernelGGL(k_poke, dim3(1), dim3(64), 0, s, j);
}

// Dynamic LDS of a tier's kernel: the workspace of its largest batch size.
int tier_lds(int t) { return t == 1 ? app_lds_bytes<1024>() : app_lds_bytes<512>(); }

hipError_t allow_serve_lds() {
    static const hipError_t e = [] {
        hipError_t r = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_serve<kTierNT[0]>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, tier_lds(0));
        if (r == hipSuccess)
            r = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_serve<kTierNT[1]>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, tier_lds(1));
        return r;
    }();
    return e;
}

// The ring words of one tier, reset, then that tier's serve kernel.
void launch_tier(int t, int wgs, hipStream_t st, const ServeRing& r) {
    hipLaunchKernelGGL(k_ring_reset, dim3(1), dim3(1), 0, st, r.claim, r.mirror, r.poller, r.fault, r.live, wgs);
    if (t == 0) hipLaunchKernelGGL(k_serve<kTierNT[0]>, dim3(wgs), dim3(kTierNT[0]), tier_lds(0), st, r);
    else hipLaunchKernelGGL(k_serve<kTierNT[1]>, dim3(wgs), dim3(kTierNT[1]), tier_lds(1), st, r);
}

// One serve workgroup per CU and tier by default: the two tiers' workgroups
// fit on a CU together (registers, LDS), and the prepare kernels of new
// queries must still find room on the CUs while the serve kernels run.
int default_workgroups(int wg) {
    if (wg > 0) return wg;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus;
}

void to_result(const Completion& c, uint64_t tag, int type, double t0, double t1, bgpq_result* r) {
    r->tag = tag;
    r->type = type;
    r->status = c.error ? BGPQ_EHEAP : BGPQ_OK;
    r->value = c.value;
    r->digest = c.digest;
    r->stats.popped = c.popped;
    r->stats.expanded = c.expanded;
    r->stats.stale = c.stale;
    r->stats.pruned = c.pruned;
    r->stats.pushed = c.pushed;
    r->stats.error_flags = c.error;
    r->stats.ms = 0.f;
    r->t_submit = t0;
    r->t_done = t1;
}

}  // namespace

struct bgpq_server {
    int workgroups = 0;  // per tier
    unsigned qcap = 0, scap = 0;
    char* arena_base = nullptr;
    Arena arena;
    // pinned host memory, one allocation: descriptors and completions (by
    // query index), each tier's slot queue and words, the closed flag
    void* host = nullptr;
    QueryDesc* desc = nullptr;
    Completion* done = nullptr;
    int* closed = nullptr;
    long long max_polls = kMaxRingPolls;
    // device words of every tier: claim counter, ring mirror, poller flag
    char* dev = nullptr;
    // device: a query index's setup words (chunk counter, chunks done),
    // 16 bytes apart, zero between queries
    unsigned* qsync = nullptr;
    // pinned host slots for knapsack item arrays until their setup is done
    char* stage = nullptr;
    std::vector<int> stage_free;
    // pinned host buffers the kernels write returned labels to (bytes,
    // pointer): kept for reuse until the server is destroyed
    std::vector<std::pair<size_t, void*>> lab_free;
    std::vector<void*> lab_all;
    struct Tier {
        uint32_t* slots = nullptr;
        long long* tail = nullptr;
        int* fault = nullptr;    // written by the kernel
        int* live_wg = nullptr;  // workgroups in the kernel, written by it
        long long posted = 0, slots_live = 0;
        hipStream_t run = nullptr;
    } tier[kTiers];
    hipStream_t prep = nullptr;
    bool running = false;
    struct Live {
        uint64_t tag;
        int type, team, tier, stage;
        size_t off, bytes;
        unsigned qi;
        uint32_t* labels;
        long long n_labels;
        uint32_t* labels_out;
        void* lab_host;  // pinned copy of the labels, or null
        size_t lab_bytes;
        double t_submit;
    };
    std::vector<Live> live;
    std::vector<char> qi_busy;
    unsigned next_qi = 0;

    ServeRing ring(int t) const {
        char* d = dev + 512 * t;
        return ServeRing{desc,
                         tier[t].slots,
                         tier[t].tail,
                         closed,
                         done,
                         reinterpret_cast<long long*>(d),
                         reinterpret_cast<unsigned long long*>(d + 128),
                         reinterpret_cast<int*>(d + 256),
                         tier[t].fault,
                         tier[t].live_wg,
                         max_polls,
                         qcap - 1,
                         scap - 1};
    }
    int live_of(int t) const { return __atomic_load_n(tier[t].live_wg, __ATOMIC_ACQUIRE); }
};

extern "C" double bgpq_now(void) { return now_s(); }

extern "C" int64_t bgpq_query_bytes(const bgpq_query* q) {
    const int rc = check_query(q);
    if (rc) return rc;
    return (int64_t)layout_of(q).total;
}

extern "C" int bgpq_server_create(bgpq_server** out, int64_t hbm_budget, int workgroups, int queue_slots) {
    if (!out) return bgpq_fail(BGPQ_EINVAL, "null server pointer");
    *out = nullptr;
    if (hbm_budget < (1 << 20)) return bgpq_fail(BGPQ_EINVAL, "HBM budget %lld too small", (long long)hbm_budget);
    if (queue_slots <= 0) queue_slots = 4096;
    unsigned qcap = 1;
    while ((int)qcap < queue_slots) qcap <<= 1;
    if (qcap > (1u << 20)) return bgpq_fail(BGPQ_EINVAL, "queue_slots %d too large", queue_slots);
    bgpq_server* s = new bgpq_server();
    s->workgroups = default_workgroups(workgroups);
    s->qcap = qcap;
    s->scap = qcap * 16;
    // per tier: the slots, then a 512-byte block (tail, live count, fault:
    // a 128-byte line each); the closed flag at the end
    const size_t o_done = sizeof(QueryDesc) * qcap, o_tiers = o_done + sizeof(Completion) * qcap;
    const size_t tier_bytes = 4ull * s->scap + 512, hbytes = o_tiers + kTiers * tier_bytes + 128;
    const int stage_slots = (int)std::min<unsigned>(qcap, 1024);
    hipError_t e = hipHostMalloc(&s->host, hbytes, hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&s->stage, kStageBytes * stage_slots, hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) e = hipMalloc((void**)&s->arena_base, (size_t)hbm_budget);
    if (e == hipSuccess) e = hipMalloc((void**)&s->dev, 512 * kTiers);
    if (e == hipSuccess) e = hipMalloc((void**)&s->qsync, 16 * (size_t)qcap);
    for (int t = 0; t < kTiers && e == hipSuccess; ++t) e = hipStreamCreateWithFlags(&s->tier[t].run, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->prep, hipStreamNonBlocking);
    if (e == hipSuccess) e = allow_serve_lds();
    if (e != hipSuccess) {
        bgpq_server_destroy(s);
        return bgpq_fail(BGPQ_EHIP, "server allocation: %s", hipGetErrorString(e));
    }
    std::memset(s->host, 0, hbytes);
    char* h = (char*)s->host;
    s->desc = (QueryDesc*)h;
    s->done = (Completion*)(h + o_done);
    for (int t = 0; t < kTiers; ++t) {
        char* b = h + o_tiers + t * tier_bytes;
        s->tier[t].slots = (uint32_t*)b;
        s->tier[t].tail = (long long*)(b + 4ull * s->scap);
        s->tier[t].live_wg = (int*)(b + 4ull * s->scap + 128);
        s->tier[t].fault = (int*)(b + 4ull * s->scap + 256);
    }
    s->closed = (int*)(h + hbytes - 128);
    for (int i = stage_slots - 1; i >= 0; --i) s->stage_free.push_back(i);
    s->arena.init((size_t)hbm_budget);
    s->qi_busy.assign(qcap, 0);
    *out = s;
    return BGPQ_OK;
}

extern "C" int bgpq_server_start(bgpq_server* s) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (s->running) return bgpq_fail(BGPQ_EINVAL, "server already running");
    __atomic_store_n(s->closed, 0, __ATOMIC_RELEASE);
    (void)hipGetLastError();
    // before any serve kernel runs: setup words of every query index zero
    BGPQ_HIPCHK(hipMemset(s->qsync, 0, 16 * (size_t)s->qcap));
    for (int t = 0; t < kTiers; ++t) {
        bgpq_server::Tier& T = s->tier[t];
        __atomic_store_n(T.tail, 0ll, __ATOMIC_RELEASE);
        T.posted = 0;
        __atomic_store_n(T.live_wg, s->workgroups, __ATOMIC_RELEASE);
        __atomic_store_n(T.fault, 0, __ATOMIC_RELEASE);
        launch_tier(t, s->workgroups, T.run, s->ring(t));
    }
    BGPQ_HIPCHK(hipGetLastError());
    s->running = true;
    return BGPQ_OK;
}

extern "C" int bgpq_server_set_idle_limit(bgpq_server* s, int64_t polls) {
    if (!s || polls < 0) return bgpq_fail(BGPQ_EINVAL, "bad idle limit");
    if (s->running) return bgpq_fail(BGPQ_EINVAL, "set the idle limit before bgpq_server_start");
    s->max_polls = polls ? (long long)polls : kMaxRingPolls;
    return BGPQ_OK;
}

extern "C" int bgpq_server_live(const bgpq_server* s) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (!s->running) return 0;
    int n = 0;
    for (int t = 0; t < kTiers; ++t) n += s->live_of(t);
    return n;
}

extern "C" int bgpq_server_submit(bgpq_server* s, const bgpq_query* q) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (!s->running) return bgpq_fail(BGPQ_EINVAL, "server not running");
    int rc = check_query(q);
    if (rc) return rc;
    const int t = tier_of(q->batch_size);
    bgpq_server::Tier& T = s->tier[t];
    // no workgroup of the query's tier is left to take it: refuse it rather
    // than queue it
    if (s->live_of(t) <= 0)
        return bgpq_fail(BGPQ_EHEAP, "the serve kernel has no workgroup left (idle limit reached)");
    if (s->live.size() >= s->qcap || T.slots_live + q->team > (long long)s->scap) return BGPQ_EAGAIN;
    const Layout L = layout_of(q);
    if (L.total > s->arena.size) return bgpq_fail(BGPQ_EINVAL, "query needs %zu bytes, budget is %zu", L.total, s->arena.size);
    const bool knap = q->type == BGPQ_Q_KNAPSACK;
    if (knap && s->stage_free.empty()) return BGPQ_EAGAIN;
    size_t off = 0;
    if (!s->arena.alloc(L.total, &off)) return BGPQ_EAGAIN;
    unsigned qi = s->next_qi;
    while (s->qi_busy[qi & (s->qcap - 1)]) ++qi;
    qi &= s->qcap - 1;
    s->next_qi = qi + 1;
    const int stage = knap ? s->stage_free.back() : -1;
    // no GPU call here: the query's team prepares its block (setup_query)
    QueryDesc d;
    FillJobs jobs;
    rc = build_query(q, s->arena_base + off, L, s->prep, &d, &jobs, true,
                     knap ? s->stage + kStageBytes * (size_t)stage : nullptr);
    if (rc) {
        s->arena.release(off, L.total);
        return rc;
    }
    // labels requested: a pinned buffer the last member writes them to (the
    // runtime's device-to-host copy can wait for a free CU, which the
    // resident serve kernels may not leave)
    void* lab = nullptr;
    size_t lab_bytes = 0;
    if (q->labels_out && L.n_labels > 0) {
        lab_bytes = 4 * (size_t)L.n_labels;
        for (size_t i = 0; i < s->lab_free.size(); ++i) {
            if (s->lab_free[i].first >= lab_bytes) {
                lab = s->lab_free[i].second;
                lab_bytes = s->lab_free[i].first;
                s->lab_free.erase(s->lab_free.begin() + (long)i);
                break;
            }
        }
        if (!lab) {
            if (hipHostMalloc(&lab, lab_bytes, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
                s->arena.release(off, L.total);
                return bgpq_fail(BGPQ_EHIP, "pinned memory for %zu label bytes", lab_bytes);
            }
            s->lab_all.push_back(lab);
        }
    }
    if (knap) s->stage_free.pop_back();
    d.sync = s->qsync + 4 * (size_t)qi;
    d.labels_host = (uint32_t*)lab;
    s->qi_busy[qi] = 1;
    std::memcpy(&s->desc[qi], &d, sizeof d);
    s->done[qi].done = 0;
    for (int m = 0; m < q->team; ++m) T.slots[(T.posted + m) & (s->scap - 1)] = (qi << 8) | (uint32_t)m;
    std::atomic_thread_fence(std::memory_order_release);
    T.posted += q->team;
    __atomic_store_n(T.tail, T.posted, __ATOMIC_RELEASE);
    T.slots_live += q->team;
    s->live.push_back(bgpq_server::Live{q->tag, q->type, q->team, t, stage, off, L.total, qi, d.labels, L.n_labels,
                                        q->labels_out, lab, lab_bytes, now_s()});
    return BGPQ_OK;
}

extern "C" int bgpq_server_poll(bgpq_server* s, bgpq_result* out, int max_results) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    // read before the done flags: with no workgroup of its tier left, a
    // query still not done below can never complete
    bool gone[kTiers];
    for (int t = 0; t < kTiers; ++t) gone[t] = s->running && s->live_of(t) <= 0;
    int n = 0;
    size_t stuck = 0;
    for (size_t i = 0; i < s->live.size() && n < max_results;) {
        bgpq_server::Live& L = s->live[i];
        if (!__atomic_load_n(&s->done[L.qi].done, __ATOMIC_ACQUIRE)) {
            stuck += gone[L.tier];
            ++i;
            continue;
        }
        const double t1 = now_s();
        Completion c;
        std::memcpy(&c, &s->done[L.qi], sizeof c);
        if (L.lab_host) {
            std::memcpy(L.labels_out, L.lab_host, 4 * (size_t)L.n_labels);
            s->lab_free.emplace_back(L.lab_bytes, L.lab_host);
        }
        if (out) to_result(c, L.tag, L.type, L.t_submit, t1, &out[n]);
        ++n;
        s->arena.release(L.off, L.bytes);
        if (L.stage >= 0) s->stage_free.push_back(L.stage);
        s->qi_busy[L.qi] = 0;
        s->tier[L.tier].slots_live -= L.team;
        s->live[i] = s->live.back();
        s->live.pop_back();
    }
    if (n == 0 && stuck)
        return bgpq_fail(BGPQ_EHEAP, "%zu queries cannot complete: the serve kernel has no workgroup left", stuck);
    return n;
}

extern "C" int bgpq_server_inflight(const bgpq_server* s) { return s ? (int)s->live.size() : 0; }

extern "C" int64_t bgpq_server_used_bytes(const bgpq_server* s) { return s ? (int64_t)s->arena.used : 0; }

extern "C" int bgpq_server_close(bgpq_server* s) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (!s->running) return BGPQ_OK;
    __atomic_store_n(s->closed, 1, __ATOMIC_RELEASE);
    hipError_t e = hipSuccess;
    for (int t = 0; t < kTiers; ++t) {
        const hipError_t et = hipStreamSynchronize(s->tier[t].run);
        if (e == hipSuccess) e = et;
    }
    s->running = false;
    if (e != hipSuccess) return bgpq_fail(BGPQ_EHIP, "serve kernel: %s", hipGetErrorString(e));
    // Workgroups that left at the idle limit are no fault; a posted query
    // that the ended kernels never completed is.
    size_t lost = 0;
    for (const auto& L : s->live)
        if (!__atomic_load_n(&s->done[L.qi].done, __ATOMIC_ACQUIRE)) ++lost;
    if (lost) return bgpq_fail(BGPQ_EHEAP, "%zu posted queries never completed", lost);
    return BGPQ_OK;
}

extern "C" void bgpq_server_destroy(bgpq_server* s) {
    if (!s) return;
    if (s->running) (void)bgpq_server_close(s);
    for (int t = 0; t < kTiers; ++t)
        if (s->tier[t].run) (void)hipStreamDestroy(s->tier[t].run);
    if (s->prep) (void)hipStreamDestroy(s->prep);
    if (s->dev) (void)hipFree(s->dev);
    if (s->qsync) (void)hipFree(s->qsync);
    if (s->stage) (void)hipHostFree(s->stage);
    for (void* p : s->lab_all) (void)hipHostFree(p);
    if (s->arena_base) (void)hipFree(s->arena_base);
    if (s->host) (void)hipHostFree(s->host);
    delete s;
}

// ---------------------------------------------------------------- batch
// All queries at once: the arena holds every query, each tier's ring lives
// in device memory fully posted and closed, and one captured graph runs
// [fills of every query -> (ring reset -> serve kernel) of each tier, the
// tiers as parallel branches]; replays start from the same state.
extern "C" int bgpq_batch_run(const bgpq_query* q, int n, int64_t hbm_budget, int workgroups, int repeats,
                              bgpq_result* out, float* ms_per_run) {
    if (!q || n < 1 || repeats < 1) return bgpq_fail(BGPQ_EINVAL, "batch: bad arguments");
    std::vector<Layout> lay(n);
    size_t total = 0;
    long long slots[kTiers] = {};
    for (int i = 0; i < n; ++i) {
        const int rc = check_query(&q[i]);
        if (rc) return rc;
        lay[i] = layout_of(&q[i]);
        total += align_up(lay[i].total, 256);
        slots[tier_of(q[i].batch_size)] += q[i].team;
    }
    if ((int64_t)total > hbm_budget)
        return bgpq_fail(BGPQ_EINVAL, "batch needs %zu bytes, budget %lld", total, (long long)hbm_budget);
    unsigned qcap = 1, scap[kTiers];
    while ((int)qcap < n) qcap <<= 1;
    if (qcap > (1u << 24)) return bgpq_fail(BGPQ_EINVAL, "batch too large");
    // ring: descriptors, completions, then per tier its slots and a
    // 1024-byte block of words (a 128-byte line each)
    const size_t o_done = sizeof(QueryDesc) * qcap;
    size_t o_slots[kTiers], o_words[kTiers], rbytes = o_done + sizeof(Completion) * qcap;
    for (int t = 0; t < kTiers; ++t) {
        scap[t] = 1;
        while ((long long)scap[t] < slots[t]) scap[t] <<= 1;
        o_slots[t] = align_up(rbytes, 256);
        o_words[t] = align_up(o_slots[t] + 4ull * scap[t], 256);
        rbytes = o_words[t] + 1024;
    }
    enum { W_TAIL = 0, W_CLOSED = 128, W_CLAIM = 256, W_MIRROR = 384, W_POLLER = 512, W_LIVE = 640, W_FAULT = 768 };
    char* arena = nullptr;
    char* ring = nullptr;
    hipStream_t s = nullptr, s1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, fork = nullptr, join = nullptr;
    std::vector<QueryDesc> desc(qcap);
    std::vector<FillJobs> jobs(n);
    std::vector<uint32_t> slot[kTiers];
    for (int t = 0; t < kTiers; ++t) slot[t].assign(scap[t], 0);
    std::vector<Completion> comp(qcap);
    int rc = BGPQ_OK;
    if (hipMalloc((void**)&arena, total) != hipSuccess || hipMalloc((void**)&ring, rbytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s1, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess || allow_serve_lds() != hipSuccess)
        rc = bgpq_fail(BGPQ_EHIP, "batch allocation");
    size_t off = 0;
    long long posted[kTiers] = {};
    for (int i = 0; i < n && rc == BGPQ_OK; ++i) {
        rc = build_query(&q[i], arena + off, lay[i], s, &desc[i], &jobs[i]);
        const int t = tier_of(q[i].batch_size);
        for (int m = 0; m < q[i].team; ++m) slot[t][posted[t]++] = ((uint32_t)i << 8) | (uint32_t)m;
        off += align_up(lay[i].total, 256);
    }
    const int closed = 1;
    if (rc == BGPQ_OK &&
        hipMemcpyAsync(ring, desc.data(), sizeof(QueryDesc) * qcap, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = bgpq_fail(BGPQ_EHIP, "batch ring upload");
    for (int t = 0; t < kTiers && rc == BGPQ_OK; ++t) {
        if (hipMemcpyAsync(ring + o_slots[t], slot[t].data(), 4ull * scap[t], hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(ring + o_words[t] + W_TAIL, &posted[t], sizeof(long long), hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(ring + o_words[t] + W_CLOSED, &closed, sizeof closed, hipMemcpyHostToDevice, s) != hipSuccess)
            rc = bgpq_fail(BGPQ_EHIP, "batch ring upload");
    }
    if (rc == BGPQ_OK && hipStreamSynchronize(s) != hipSuccess) rc = bgpq_fail(BGPQ_EHIP, "batch ring upload");
    auto ring_of = [&](int t) {
        char* w = ring + o_words[t];
        return ServeRing{(const QueryDesc*)ring,
                         (const uint32_t*)(ring + o_slots[t]),
                         (const long long*)(w + W_TAIL),
                         (const int*)(w + W_CLOSED),
                         (Completion*)(ring + o_done),
                         (long long*)(w + W_CLAIM),
                         (unsigned long long*)(w + W_MIRROR),
                         (int*)(w + W_POLLER),
                         (int*)(w + W_FAULT),
                         (int*)(w + W_LIVE),
                         kMaxRingPolls,
                         qcap - 1,
                         scap[t] - 1};
    };
    if (rc == BGPQ_OK) {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = bgpq_fail(BGPQ_EHIP, "begin capture");
        } else {
            for (int i = 0; i < n; ++i) launch_fills(jobs[i], s, 64);
            const int wgs = default_workgroups(workgroups);
            // tier 1 on a forked branch, joined before the capture ends
            const bool both = posted[0] > 0 && posted[1] > 0;
            if (both) {
                (void)hipEventRecord(fork, s);
                (void)hipStreamWaitEvent(s1, fork, 0);
            }
            if (posted[0] > 0) launch_tier(0, wgs, s, ring_of(0));
            if (posted[1] > 0) launch_tier(1, wgs, both ? s1 : s, ring_of(1));
            if (both) {
                (void)hipEventRecord(join, s1);
                (void)hipStreamWaitEvent(s, join, 0);
            }
            if (hipStreamEndCapture(s, &graph) != hipSuccess || !graph) rc = bgpq_fail(BGPQ_EHIP, "end capture");
        }
    }
    if (rc == BGPQ_OK && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        rc = bgpq_fail(BGPQ_EHIP, "graph instantiate");
    // Every replay is read back (an eager copy between replays) and must
    // give the first replay's answers: a replay starts from the state the
    // graph's own fill kernels write, never from what a previous one left.
    float total_ms = 0.f;
    std::vector<Completion> first;
    for (int r = 0; r < repeats && rc == BGPQ_OK; ++r) {
        float ms = 0.f;
        int fault = 0;
        if (hipEventRecord(e0, s) != hipSuccess || hipGraphLaunch(exec, s) != hipSuccess ||
            hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = bgpq_fail(BGPQ_EHIP, "graph launch: %s", hipGetErrorString(hipGetLastError()));
            break;
        }
        total_ms += ms;
        if (hipMemcpy(comp.data(), ring + o_done, sizeof(Completion) * qcap, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = bgpq_fail(BGPQ_EHIP, "batch results");
            break;
        }
        for (int t = 0; t < kTiers && rc == BGPQ_OK; ++t) {
            int f = 0;
            if (posted[t] > 0 && hipMemcpy(&f, ring + o_words[t] + W_FAULT, sizeof f, hipMemcpyDeviceToHost) != hipSuccess)
                rc = bgpq_fail(BGPQ_EHIP, "batch results");
            fault |= f;
        }
        if (rc != BGPQ_OK) break;
        if (fault) {
            rc = bgpq_fail(BGPQ_EHEAP, "batch serve kernel fault");
            break;
        }
        if (r == 0) {
            first = comp;
        } else {
            for (int i = 0; i < n; ++i) {
                if (comp[i].value != first[i].value || comp[i].digest != first[i].digest ||
                    comp[i].error != first[i].error) {
                    rc = bgpq_fail(BGPQ_EHEAP, "batch replay %d differs from the first at query %d", r, i);
                    break;
                }
            }
        }
    }
    if (rc == BGPQ_OK) {
        if (ms_per_run) *ms_per_run = total_ms / repeats;
        for (int i = 0; i < n && rc == BGPQ_OK; ++i) {
            if (!comp[i].done) rc = bgpq_fail(BGPQ_EHEAP, "batch query %d did not complete", i);
            if (out) to_result(comp[i], q[i].tag, q[i].type, 0.0, 0.0, &out[i]);
            if (rc == BGPQ_OK && q[i].labels_out && lay[i].n_labels > 0 &&
                hipMemcpy(q[i].labels_out, desc[i].labels, 4 * (size_t)lay[i].n_labels, hipMemcpyDeviceToHost) !=
                    hipSuccess)
                rc = bgpq_fail(BGPQ_EHIP, "batch labels");
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (s1) (void)hipStreamDestroy(s1);
    if (s) (void)hipStreamDestroy(s);
    if (ring) (void)hipFree(ring);
    if (arena) (void)hipFree(arena);
    return rc;
}
