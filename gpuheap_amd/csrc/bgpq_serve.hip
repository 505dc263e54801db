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

constexpr int kServeNT = 256;
// Polls (each ~3.4 us of sleep plus a host-memory read) a workgroup waits
// for the host to post or close before it gives up: about a minute.
constexpr long long kMaxRingPolls = 1ll << 24;

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
};
static_assert(sizeof(QueryDesc) % 8 == 0 && sizeof(QueryDesc) / 8 <= kServeNT, "descriptor words");

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
    int* fault;             // set when a workgroup left after max_polls idle polls
    int* live;              // workgroups still in the kernel
    long long max_polls;
    unsigned qmask, smask;
};

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

// One solve at batch size B: a call of its own, so that the register
// allocation of the three batch sizes is not merged into one function.
template <int B>
__device__ __attribute__((noinline)) void solve_b(const QueryDesc& d, bool first) {
    const HeapView<uint64_t> hv{d.items, d.status, d.ctr, ~0ull, (int)d.batch_num, 0, 0};
    if (d.type == BGPQ_Q_SSSP) {
        const SsspApp app{d.rowptr, d.col, d.w, d.labels, d.ctl};
        drive<VAR_BUITDD, B, kServeNT>(hv, app, d.ctl, d.first_key, first);
    } else if (d.type == BGPQ_Q_ASTAR) {
        const AstarApp app{d.free_cells, d.n, d.height, d.labels, d.ctl};
        drive<VAR_BUITDD, B, kServeNT>(hv, app, d.ctl, d.first_key, first);
    } else {
        const KnapsackApp app{d.n, d.cap, d.kw, d.kv, d.pw, d.pv, d.cell, d.pool, d.pool_cap, d.ctl};
        drive<VAR_BUITDD, B, kServeNT>(hv, app, d.ctl, d.first_key, first);
    }
}

// The last member's epilogue: result, digest and statistics to the host.
__device__ void finish(const QueryDesc& d, Completion* c) {
    __shared__ unsigned long long red[2][kServeNT / 64];
    const int t = threadIdx.x;
    uint64_t dig = 0, reach = 0;
    if (d.type == BGPQ_Q_SSSP) {
        for (long long v = t; v < d.n_labels; v += kServeNT) {
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
    __syncthreads();
    if (t == 0) {
        uint64_t value = 0;
        dig = 0;
        if (d.type == BGPQ_Q_SSSP) {
            for (int i = 0; i < kServeNT / 64; ++i) {
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
__global__ __launch_bounds__(kServeNT) void k_serve(ServeRing r) {
    __shared__ long long s_slot;
    __shared__ int s_go, s_last;
    __shared__ uint32_t s_entry;
    __shared__ __align__(16) QueryDesc s_d;
    const int t = threadIdx.x;
    while (true) {
        if (t == 0) {
            int go = 0;
            long long s = -1, polls = 0;
            while (true) {
                const long long c = __hip_atomic_load(r.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (c < sys_load(r.tail)) {
                    long long e = c;
                    if (__hip_atomic_compare_exchange_strong(r.claim, &e, c + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        s = c;
                        go = 1;
                        break;
                    }
                    continue;  // another workgroup took slot c
                }
                // the host posts every slot before it closes the ring, and
                // this load of the tail issues after the closed flag is seen
                if (sys_load(r.closed)) {
                    if (__hip_atomic_load(r.claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sys_load(r.tail))
                        continue;
                    break;
                }
                if (++polls > r.max_polls) {
                    __hip_atomic_fetch_or(r.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                __builtin_amdgcn_s_sleep(127);
            }
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
        const bool first = member == 0;
        if (s_d.B == 256) solve_b<256>(s_d, first);
        else if (s_d.B == 512) solve_b<512>(s_d, first);
        else solve_b<1024>(s_d, first);
        if (t == 0) {
            wait_stores();
            s_last = __hip_atomic_fetch_add(ctl_word(s_d.ctl, APP_DONE), 1ll, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) == s_d.team - 1;
        }
        __syncthreads();
        if (s_last) finish(s_d, r.done + (qi & r.qmask));
        __syncthreads();
    }
}

// Fills of a query's arena block (heap with init_limits, status, counters,
// control words, labels) and the single words set after them.
constexpr int kMaxFills = 8, kMaxPokes = 4;
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
};

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
__global__ void k_ring_reset(long long* claim, int* fault, int* live, int workgroups) {
    __hip_atomic_store(claim, 0ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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

// Descriptor and fills of one query placed at `base`; knapsack items are
// uploaded here (on stream s, finished before returning), the fills are
// left to launch_fills.
int build_query(const bgpq_query* q, char* base, const Layout& L, hipStream_t s, QueryDesc* d, FillJobs* jobs) {
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
        BGPQ_HIPCHK(hipMemcpyAsync((void*)d->kw, w.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
        BGPQ_HIPCHK(hipMemcpyAsync((void*)d->kv, v.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
        BGPQ_HIPCHK(hipMemcpyAsync((void*)d->pw, pw.data(), 8 * (size_t)(n + 1), hipMemcpyHostToDevice, s));
        BGPQ_HIPCHK(hipMemcpyAsync((void*)d->pv, pv.data(), 8 * (size_t)(n + 1), hipMemcpyHostToDevice, s));
        fill(d->cell, L.cells, 0, 4);
        fill(d->pool, 4, 0, 4);  // the root: depth 0, nothing used, value 0
        poke(d->cell, 1, 4);
        poke(d->ctl + (size_t)APP_POOL * kCtlStride, 1, 8);
        BGPQ_HIPCHK(hipStreamSynchronize(s));  // the host vectors above go out of scope
    }
    *jobs = j;
    return BGPQ_OK;
}

// The fills, then the single words (a second launch: a word inside a
// filled range must be set after the fill).
void launch_fills(const FillJobs& j, hipStream_t s, int blocks) {
    hipLaunchKernelGGL(k_prepare, dim3(blocks), dim3(256), 0, s, j);
    hipLaunchKernelGGL(k_poke, dim3(1), dim3(64), 0, s, j);
}

hipError_t allow_serve_lds() {
    static const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_serve),
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    app_lds_bytes<1024>());
    return e;
}

// One serve workgroup per CU by default: a workgroup's solve calls keep up
// to ~250 VGPRs per lane, and the prepare kernels of new queries must still
// find room on the CUs while the serve kernel runs.
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
    int workgroups = 0;
    unsigned qcap = 0, scap = 0;
    char* arena_base = nullptr;
    Arena arena;
    // pinned host memory, one allocation
    void* host = nullptr;
    QueryDesc* desc = nullptr;
    uint32_t* slots = nullptr;
    long long* tail = nullptr;
    int* closed = nullptr;
    int* fault = nullptr;  // written by the kernel
    int* live_wg = nullptr;  // workgroups in the kernel, written by it
    Completion* done = nullptr;
    long long max_polls = kMaxRingPolls;
    // device word
    long long* claim = nullptr;
    hipStream_t run = nullptr, prep = nullptr;
    bool running = false;
    long long posted = 0;
    long long slots_live = 0;
    struct Live {
        uint64_t tag;
        int type, team;
        size_t off, bytes;
        unsigned qi;
        uint32_t* labels;
        long long n_labels;
        uint32_t* labels_out;
        double t_submit;
    };
    std::vector<Live> live;
    std::vector<char> qi_busy;
    unsigned next_qi = 0;
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
    const size_t hbytes = sizeof(QueryDesc) * qcap + sizeof(Completion) * qcap + 4ull * s->scap + 512;
    hipError_t e = hipHostMalloc(&s->host, hbytes, hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) e = hipMalloc((void**)&s->arena_base, (size_t)hbm_budget);
    if (e == hipSuccess) e = hipMalloc((void**)&s->claim, 256);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->run, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->prep, hipStreamNonBlocking);
    if (e == hipSuccess) e = allow_serve_lds();
    if (e != hipSuccess) {
        bgpq_server_destroy(s);
        return bgpq_fail(BGPQ_EHIP, "server allocation: %s", hipGetErrorString(e));
    }
    std::memset(s->host, 0, hbytes);
    char* h = (char*)s->host;
    s->desc = (QueryDesc*)h;
    s->done = (Completion*)(h + sizeof(QueryDesc) * qcap);
    s->slots = (uint32_t*)(h + sizeof(QueryDesc) * qcap + sizeof(Completion) * qcap);
    s->tail = (long long*)(h + hbytes - 128);
    s->closed = (int*)(h + hbytes - 64);
    s->live_wg = (int*)(h + hbytes - 192);
    s->fault = (int*)(h + hbytes - 256);
    s->arena.init((size_t)hbm_budget);
    s->qi_busy.assign(qcap, 0);
    *out = s;
    return BGPQ_OK;
}

extern "C" int bgpq_server_start(bgpq_server* s) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (s->running) return bgpq_fail(BGPQ_EINVAL, "server already running");
    __atomic_store_n(s->tail, 0ll, __ATOMIC_RELEASE);
    __atomic_store_n(s->closed, 0, __ATOMIC_RELEASE);
    s->posted = 0;
    __atomic_store_n(s->live_wg, s->workgroups, __ATOMIC_RELEASE);
    __atomic_store_n(s->fault, 0, __ATOMIC_RELEASE);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_ring_reset, dim3(1), dim3(1), 0, s->run, s->claim, s->fault, s->live_wg, s->workgroups);
    const ServeRing r{s->desc, s->slots,   s->tail,      s->closed,    s->done,
                      s->claim, s->fault, s->live_wg, s->max_polls, s->qcap - 1, s->scap - 1};
    hipLaunchKernelGGL(k_serve, dim3(s->workgroups), dim3(kServeNT), app_lds_bytes<1024>(), s->run, r);
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
    return s->running ? __atomic_load_n(s->live_wg, __ATOMIC_ACQUIRE) : 0;
}

extern "C" int bgpq_server_submit(bgpq_server* s, const bgpq_query* q) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (!s->running) return bgpq_fail(BGPQ_EINVAL, "server not running");
    // no workgroup is left to take the query: refuse it rather than queue it
    if (__atomic_load_n(s->live_wg, __ATOMIC_ACQUIRE) <= 0)
        return bgpq_fail(BGPQ_EHEAP, "the serve kernel has no workgroup left (idle limit reached)");
    int rc = check_query(q);
    if (rc) return rc;
    if (s->live.size() >= s->qcap || s->slots_live + q->team > (long long)s->scap) return BGPQ_EAGAIN;
    const Layout L = layout_of(q);
    if (L.total > s->arena.size) return bgpq_fail(BGPQ_EINVAL, "query needs %zu bytes, budget is %zu", L.total, s->arena.size);
    size_t off = 0;
    if (!s->arena.alloc(L.total, &off)) return BGPQ_EAGAIN;
    unsigned qi = s->next_qi;
    while (s->qi_busy[qi & (s->qcap - 1)]) ++qi;
    qi &= s->qcap - 1;
    s->next_qi = qi + 1;
    QueryDesc d;
    FillJobs jobs;
    rc = build_query(q, s->arena_base + off, L, s->prep, &d, &jobs);
    if (rc == BGPQ_OK) {
        (void)hipGetLastError();
        launch_fills(jobs, s->prep, 256);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s->prep) != hipSuccess)
            rc = bgpq_fail(BGPQ_EHIP, "prepare");
    }
    if (rc) {
        s->arena.release(off, L.total);
        return rc;
    }
    s->qi_busy[qi] = 1;
    std::memcpy(&s->desc[qi], &d, sizeof d);
    s->done[qi].done = 0;
    for (int m = 0; m < q->team; ++m) s->slots[(s->posted + m) & (s->scap - 1)] = (qi << 8) | (uint32_t)m;
    std::atomic_thread_fence(std::memory_order_release);
    s->posted += q->team;
    __atomic_store_n(s->tail, s->posted, __ATOMIC_RELEASE);
    s->slots_live += q->team;
    s->live.push_back(bgpq_server::Live{q->tag, q->type, q->team, off, L.total, qi, d.labels, L.n_labels,
                                        q->labels_out, now_s()});
    return BGPQ_OK;
}

extern "C" int bgpq_server_poll(bgpq_server* s, bgpq_result* out, int max_results) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    // read before the done flags: with no workgroup left, a query still
    // not done below can never complete
    const bool gone = s->running && __atomic_load_n(s->live_wg, __ATOMIC_ACQUIRE) <= 0;
    int n = 0;
    for (size_t i = 0; i < s->live.size() && n < max_results;) {
        bgpq_server::Live& L = s->live[i];
        if (!__atomic_load_n(&s->done[L.qi].done, __ATOMIC_ACQUIRE)) {
            ++i;
            continue;
        }
        const double t1 = now_s();
        Completion c;
        std::memcpy(&c, &s->done[L.qi], sizeof c);
        if (L.labels_out && L.n_labels > 0) {
            BGPQ_HIPCHK(hipMemcpyAsync(L.labels_out, L.labels, 4 * (size_t)L.n_labels, hipMemcpyDeviceToHost, s->prep));
            BGPQ_HIPCHK(hipStreamSynchronize(s->prep));
        }
        if (out) to_result(c, L.tag, L.type, L.t_submit, t1, &out[n]);
        ++n;
        s->arena.release(L.off, L.bytes);
        s->qi_busy[L.qi] = 0;
        s->slots_live -= L.team;
        s->live[i] = s->live.back();
        s->live.pop_back();
    }
    if (n == 0 && gone && !s->live.empty())
        return bgpq_fail(BGPQ_EHEAP, "%zu queries cannot complete: the serve kernel has no workgroup left",
                         s->live.size());
    return n;
}

extern "C" int bgpq_server_inflight(const bgpq_server* s) { return s ? (int)s->live.size() : 0; }

extern "C" int64_t bgpq_server_used_bytes(const bgpq_server* s) { return s ? (int64_t)s->arena.used : 0; }

extern "C" int bgpq_server_close(bgpq_server* s) {
    if (!s) return bgpq_fail(BGPQ_EINVAL, "null server");
    if (!s->running) return BGPQ_OK;
    __atomic_store_n(s->closed, 1, __ATOMIC_RELEASE);
    hipError_t e = hipStreamSynchronize(s->run);
    s->running = false;
    if (e != hipSuccess) return bgpq_fail(BGPQ_EHIP, "serve kernel: %s", hipGetErrorString(e));
    // Workgroups that left at the idle limit are no fault; a posted query
    // that the ended kernel never completed is.
    size_t lost = 0;
    for (const auto& L : s->live)
        if (!__atomic_load_n(&s->done[L.qi].done, __ATOMIC_ACQUIRE)) ++lost;
    if (lost) return bgpq_fail(BGPQ_EHEAP, "%zu posted queries never completed", lost);
    return BGPQ_OK;
}

extern "C" void bgpq_server_destroy(bgpq_server* s) {
    if (!s) return;
    if (s->running) (void)bgpq_server_close(s);
    if (s->run) (void)hipStreamDestroy(s->run);
    if (s->prep) (void)hipStreamDestroy(s->prep);
    if (s->claim) (void)hipFree(s->claim);
    if (s->arena_base) (void)hipFree(s->arena_base);
    if (s->host) (void)hipHostFree(s->host);
    delete s;
}

// ---------------------------------------------------------------- batch
// All queries at once: the arena holds every query, the ring lives in
// device memory fully posted and closed, and one captured graph runs
// [fills of every query -> ring reset -> serve kernel]; replays start from
// the same state.
extern "C" int bgpq_batch_run(const bgpq_query* q, int n, int64_t hbm_budget, int workgroups, int repeats,
                              bgpq_result* out, float* ms_per_run) {
    if (!q || n < 1 || repeats < 1) return bgpq_fail(BGPQ_EINVAL, "batch: bad arguments");
    std::vector<Layout> lay(n);
    size_t total = 0;
    long long slots = 0;
    for (int i = 0; i < n; ++i) {
        const int rc = check_query(&q[i]);
        if (rc) return rc;
        lay[i] = layout_of(&q[i]);
        total += align_up(lay[i].total, 256);
        slots += q[i].team;
    }
    if ((int64_t)total > hbm_budget)
        return bgpq_fail(BGPQ_EINVAL, "batch needs %zu bytes, budget %lld", total, (long long)hbm_budget);
    unsigned qcap = 1, scap = 1;
    while ((int)qcap < n) qcap <<= 1;
    while ((long long)scap < slots) scap <<= 1;
    if (qcap > (1u << 24)) return bgpq_fail(BGPQ_EINVAL, "batch too large");
    const size_t o_done = sizeof(QueryDesc) * qcap, o_slots = o_done + sizeof(Completion) * qcap;
    const size_t o_tail = align_up(o_slots + 4ull * scap, 256), o_closed = o_tail + 64, o_claim = o_tail + 128,
                 o_live = o_tail + 160, o_fault = o_tail + 192, rbytes = o_tail + 256;
    char* arena = nullptr;
    char* ring = nullptr;
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<QueryDesc> desc(qcap);
    std::vector<FillJobs> jobs(n);
    std::vector<uint32_t> slot(scap, 0);
    std::vector<Completion> comp(qcap);
    int rc = BGPQ_OK;
    if (hipMalloc((void**)&arena, total) != hipSuccess || hipMalloc((void**)&ring, rbytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || allow_serve_lds() != hipSuccess)
        rc = bgpq_fail(BGPQ_EHIP, "batch allocation");
    size_t off = 0;
    long long posted = 0;
    for (int i = 0; i < n && rc == BGPQ_OK; ++i) {
        rc = build_query(&q[i], arena + off, lay[i], s, &desc[i], &jobs[i]);
        for (int m = 0; m < q[i].team; ++m) slot[posted++] = ((uint32_t)i << 8) | (uint32_t)m;
        off += align_up(lay[i].total, 256);
    }
    const long long tail = posted;
    const int closed = 1;
    if (rc == BGPQ_OK &&
        (hipMemcpyAsync(ring, desc.data(), sizeof(QueryDesc) * qcap, hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(ring + o_slots, slot.data(), 4ull * scap, hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(ring + o_tail, &tail, sizeof tail, hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(ring + o_closed, &closed, sizeof closed, hipMemcpyHostToDevice, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = bgpq_fail(BGPQ_EHIP, "batch ring upload");
    if (rc == BGPQ_OK) {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = bgpq_fail(BGPQ_EHIP, "begin capture");
        } else {
            for (int i = 0; i < n; ++i) launch_fills(jobs[i], s, 64);
            const int wgs = default_workgroups(workgroups);
            hipLaunchKernelGGL(k_ring_reset, dim3(1), dim3(1), 0, s, (long long*)(ring + o_claim),
                               (int*)(ring + o_fault), (int*)(ring + o_live), wgs);
            const ServeRing r{(const QueryDesc*)ring, (const uint32_t*)(ring + o_slots),
                              (const long long*)(ring + o_tail), (const int*)(ring + o_closed),
                              (Completion*)(ring + o_done), (long long*)(ring + o_claim), (int*)(ring + o_fault),
                              (int*)(ring + o_live), kMaxRingPolls, qcap - 1, scap - 1};
            hipLaunchKernelGGL(k_serve, dim3(wgs), dim3(kServeNT), app_lds_bytes<1024>(), s, r);
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
        if (hipMemcpy(comp.data(), ring + o_done, sizeof(Completion) * qcap, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&fault, ring + o_fault, sizeof fault, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = bgpq_fail(BGPQ_EHIP, "batch results");
            break;
        }
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
    if (s) (void)hipStreamDestroy(s);
    if (ring) (void)hipFree(ring);
    if (arena) (void)hipFree(arena);
    return rc;
}
