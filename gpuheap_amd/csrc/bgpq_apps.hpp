// Heap consumers on gfx950: one persistent kernel per solve, every
// workgroup running the loop below against one shared u64 heap.
//
// The reference runs each application as a host relaunch loop around a
// persistent kernel (applications/sssp_longfei_new/main.cu:96-102,
// artifacts/astar/astar_main.cu:128-178, artifacts/knapsack/models_pq.cuh:
// 199-246) that polls heap counters between launches.  Here one launch runs
// the solve to completion: an exact count of outstanding keys (inserted and
// not yet fully processed) tells every workgroup when to leave.
//
// Loop of one workgroup:
//   1. deleteRoot/deleteUpdate pops up to B keys into LDS (a cheap look at
//      the heap's counters first keeps idle workgroups off status[0]); an
//      App with kThreshold (A*) pops only below App::threshold(), the A*
//      forks' deleteRoot (artifacts/astar/BGPQ_B/heap.cuh:260-287);
//   2. App::degree gives each key its number of children (0 when the key is
//      stale or pruned); a workgroup prefix scan turns the degrees into
//      offsets;
//   3. the children are expanded edge-parallel, NT at a time
//      (App::expand); improving ones are appended to an LDS buffer, which
//      is inserted B keys at a time (the rest at the end of the batch);
//   4. outstanding += children before they are inserted, -= popped after.
// Keys are never lost: outstanding reaches 0 only once the heap is empty
// and no workgroup holds keys.  A threshold search ends instead with the
// heap's termination vote (PBS_MODEL, tditdd:282-289): every workgroup has
// found nothing below the threshold since its last pop.  The keys left in
// the heap then number exactly `outstanding`, all pruned.
#pragma once
#include "bgpq_device.hpp"
#include "bgpq_kernels.hpp"

namespace bgpq {

// Control words of a solve, each alone in a 128-byte line.
enum {
    APP_OUTSTANDING = 0,
    APP_POPPED = 1,
    APP_EXPANDED = 2,
    APP_STALE = 3,
    APP_PRUNED = 4,
    APP_PUSHED = 5,
    APP_ERROR = 6,
    APP_BEST = 7,  // knapsack incumbent
    APP_POOL = 8,  // knapsack node-pool cursor
    APP_DONE = 9,  // team members that have left the solve
    APP_WORDS = 10
};
constexpr int kCtlStride = 16;  // long longs per control word
enum { APPERR_RANGE = 1, APPERR_POOL = 2, APPERR_IDLE = 4, APPERR_HEAP = 8 };

// Polls of an idle workgroup (each about a microsecond) before it gives up
// and flags APPERR_IDLE: far above any real wait, it only bounds a fault.
constexpr long long kMaxIdlePolls = 1ll << 23;

__device__ __forceinline__ long long* ctl_word(long long* ctl, int i) { return ctl + (size_t)i * kCtlStride; }
__device__ __forceinline__ long long ctl_load(long long* ctl, int i) {
    return __hip_atomic_load(ctl_word(ctl, i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ctl_add(long long* ctl, int i, long long v) {
    __hip_atomic_fetch_add(ctl_word(ctl, i), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ctl_or(long long* ctl, int i, long long v) {
    __hip_atomic_fetch_or(ctl_word(ctl, i), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t u32_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t u32_min(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t u32_max(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t u64_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix sum over the workgroup; `tmp` holds NT/64 + 1 ints.
template <int NT>
__device__ __forceinline__ int wg_exclusive_scan(int v, int* tmp, int& total) {
    const int lane = (int)__lane_id(), wid = (int)threadIdx.x / 64;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < NT / 64; ++w) {
            const int c = tmp[w];
            tmp[w] = acc;
            acc += c;
        }
        tmp[NT / 64] = acc;
    }
    __syncthreads();
    const int r = x - v + tmp[wid];
    total = tmp[NT / 64];
    __syncthreads();
    return r;
}

// Dynamic LDS of a driver kernel, in ints: popped[B] and emit[2B] u64 keys,
// off[B + 4] ints, then the heap workspace (3B keys) at app_ws_offset.
template <int B>
constexpr int app_ws_offset() { return 2 * B + 4 * B + B + 4; }
template <int B>
constexpr int app_lds_bytes() { return (app_ws_offset<B>() + 6 * B) * 4; }

// Per-workgroup statistics gathered in LDS, added to the control words once.
struct AppStats {
    int popped, expanded, stale, pruned, pushed;
};

// One workgroup's part in a solve.  Any number of workgroups may run it on
// the same heap and control words (a team); the one with `first` set
// inserts the first key, which the outstanding count already includes.
template <int VAR, int B, int NT, class App>
__device__ void drive(HeapView<uint64_t> h, const App& app, long long* ctl, uint64_t first_key, bool first) {
    using K = uint64_t;
    static_assert(NT <= B && B % NT == 0, "workgroup shape");
    constexpr int P = B / NT;
    constexpr int kWs = app_ws_offset<B>();
    static_assert(kWs % 4 == 0, "heap workspace offset must be 16-byte aligned");
    extern __shared__ __align__(16) int app_lds[];
    K* popped = reinterpret_cast<K*>(app_lds);
    K* emit = popped + B;
    int* off = reinterpret_cast<int*>(emit + 2 * B);
    __shared__ int s_n, s_word, s_quit;
    __shared__ int s_scan[NT / 64 + 1];
    __shared__ AppStats st;
    DevHeap<K, VAR, B, NT> hp{h};
    constexpr bool kThr = App::kThreshold;
    const int t = threadIdx.x;
    if (t == 0) {
        st = AppStats{0, 0, 0, 0, 0};
        s_quit = 0;
        bcast_slots()[BC_ABORT] = 0;
    }
    __syncthreads();
    if (first) {
        if (t == 0) popped[0] = first_key;
        __syncthreads();
        hp.insertion(popped, 1, kWs);
        if (t == 0) st.pushed = 1;
        if (kThr) hp.voteBusy();  // counted busy by the solve's setup
    }
    long long idle = 0;
    while (!hp.aborted()) {
        // a busy workgroup must reach deleteRoot to cast its idle vote
        if (t == 0)
            s_word = (a_load(h.ctr + CTR_BATCH_COUNT * kWordStride) | a_load(h.ctr + CTR_PARTIAL * kWordStride)) != 0 ||
                     (kThr && hp.busy_);
        __syncthreads();
        int size = 0;
        if (s_word) {
            if constexpr (kThr) {
                if (hp.deleteRoot(popped, size, app.threshold(), kWs)) hp.deleteUpdate(kWs);
            } else {
                if (hp.deleteRoot(popped, size, kWs)) hp.deleteUpdate(kWs);
            }
            if (hp.aborted()) break;
        }
        if (size == 0) {
            if (t == 0) {
                s_word = kThr ? a_load(h.ctr + CTR_TERM * kWordStride) != 0 : ctl_load(ctl, APP_OUTSTANDING) == 0;
                if (!s_word && ++idle > kMaxIdlePolls) {
                    ctl_or(ctl, APP_ERROR, APPERR_IDLE);
                    s_word = 1;
                }
            }
            __syncthreads();
            const bool done = s_word != 0;
            __syncthreads();
            if (done) break;
            backoff();
            continue;
        }
        idle = 0;
        // degrees and their offsets
        int deg[P];
        int dsum = 0;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int i = t * P + p;
            deg[p] = i < size ? app.degree(popped[i], st) : 0;
            dsum += deg[p];
        }
        int total = 0;
        int pre = wg_exclusive_scan<NT>(dsum, s_scan, total);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int i = t * P + p;
            if (i < size) off[i] = pre;
            pre += deg[p];
        }
        if (t == 0) {
            off[size] = total;
            s_n = 0;
            st.popped += size;
        }
        __syncthreads();
        for (int e0 = 0; e0 < total; e0 += NT) {
            const int e = e0 + t;
            if (e < total) {
                int lo = 0, hi = size - 1;  // last i with off[i] <= e
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (off[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                K child;
                if (app.expand(popped[lo], e - off[lo], child, st)) emit[atomicAdd(&s_n, 1)] = child;
            }
            __syncthreads();
            const int n = s_n;
            if (n >= B) {
                if (t == 0) {
                    ctl_add(ctl, APP_OUTSTANDING, B);
                    wait_stores();
                    st.pushed += B;
                }
                hp.insertion(emit, B, kWs);
                if (hp.aborted()) {
                    if (t == 0) s_quit = 1;
                    break;
                }
                // n - B < NT: one key per thread moves down
                const K rest = t < n - B ? emit[B + t] : K(0);
                __syncthreads();
                if (t < n - B) emit[t] = rest;
                if (t == 0) s_n = n - B;
                __syncthreads();
            }
        }
        if (s_quit) break;
        const int n = s_n;
        if (n > 0) {
            if (t == 0) {
                ctl_add(ctl, APP_OUTSTANDING, n);
                wait_stores();
                st.pushed += n;
            }
            hp.insertion(emit, n, kWs);
            if (hp.aborted()) break;
        }
        if (t == 0) {
            wait_stores();
            ctl_add(ctl, APP_OUTSTANDING, -size);
        }
    }
    __syncthreads();
    if (t == 0) {
        if (hp.aborted()) ctl_or(ctl, APP_ERROR, APPERR_HEAP);
        ctl_add(ctl, APP_POPPED, st.popped);
        ctl_add(ctl, APP_EXPANDED, st.expanded);
        ctl_add(ctl, APP_STALE, st.stale);
        ctl_add(ctl, APP_PRUNED, st.pruned);
        ctl_add(ctl, APP_PUSHED, st.pushed);
    }
}

// ------------------------------------------------------------------ SSSP
// Keys dist << 32 | vertex (applications/sssp_longfei_new/sssp.cu:6).  A key
// whose dist is above the vertex's current label is stale (sssp.cu:48-57);
// relaxations use atomicMin on the labels (sssp.cu:103-120).
struct SsspApp {
    static constexpr bool kThreshold = false;
    const int* rowptr;
    const int* col;
    const uint32_t* w;
    uint32_t* dist;
    long long* ctl;
    __device__ int degree(uint64_t key, AppStats& st) const {
        const uint32_t d = (uint32_t)(key >> 32), v = (uint32_t)key;
        if (d > u32_load(dist + v)) {
            atomicAdd(&st.stale, 1);
            return 0;
        }
        atomicAdd(&st.expanded, 1);
        return rowptr[v + 1] - rowptr[v];
    }
    __device__ bool expand(uint64_t key, int j, uint64_t& child, AppStats&) const {
        const uint32_t d = (uint32_t)(key >> 32), v = (uint32_t)key;
        const int idx = rowptr[v] + j;
        const uint32_t u = (uint32_t)col[idx];
        const uint64_t nd = (uint64_t)d + w[idx];
        if (nd >= 0xFFFFFFFFull) {
            ctl_or(ctl, APP_ERROR, APPERR_RANGE);
            return false;
        }
        if (nd >= u32_load(dist + u)) return false;
        if (nd >= u32_min(dist + u, (uint32_t)nd)) return false;
        child = nd << 32 | u;
        return true;
    }
};

// -------------------------------------------------------------------- A*
// Keys f << 32 | cell on an 8-connected grid; moves cost 10 / 14 into any
// free cell (artifacts/astar/pq_model.cuh:114-184 checks only the target
// cell), octile heuristic.  Pops stop at the goal's label: deleteRoot with
// the threshold g[goal] << 32 hands out only keys whose f is below it (the
// reference passes dist[target], pq_model.cuh:287-296, to
// artifacts/astar/BGPQ_B/heap.cuh:273-287).  A popped key whose g is above
// the cell's label is stale; one whose f has reached the goal's label since
// it was popped is pruned.
struct AstarApp {
    static constexpr bool kThreshold = true;
    const uint8_t* free_cells;
    int width, height;
    uint32_t* g;
    long long* ctl;
    __device__ uint32_t h(uint32_t cell) const {
        const int x = (int)(cell % (uint32_t)width), y = (int)(cell / (uint32_t)width);
        const uint32_t dx = (uint32_t)(width - 1 - x), dy = (uint32_t)(height - 1 - y);
        const uint32_t lo = dx < dy ? dx : dy, hi = dx < dy ? dy : dx;
        return 10u * hi + 4u * lo;
    }
    __device__ uint32_t goal() const { return (uint32_t)width * (uint32_t)height - 1u; }
    __device__ uint64_t threshold() const { return (uint64_t)u32_load(g + goal()) << 32; }
    __device__ int degree(uint64_t key, AppStats& st) const {
        const uint32_t f = (uint32_t)(key >> 32), cell = (uint32_t)key;
        const uint32_t gc = f - h(cell);
        if (gc > u32_load(g + cell)) {
            atomicAdd(&st.stale, 1);
            return 0;
        }
        if (cell == goal() || f >= u32_load(g + goal())) {
            atomicAdd(&st.pruned, 1);
            return 0;
        }
        atomicAdd(&st.expanded, 1);
        return 8;
    }
    __device__ bool expand(uint64_t key, int j, uint64_t& child, AppStats&) const {
        const uint32_t f = (uint32_t)(key >> 32), cell = (uint32_t)key;
        const uint32_t gc = f - h(cell);
        const int k = j < 4 ? j : j + 1;  // 3x3 neighbourhood without the centre
        const int dx = k % 3 - 1, dy = k / 3 - 1;
        const int x = (int)(cell % (uint32_t)width) + dx, y = (int)(cell / (uint32_t)width) + dy;
        if (x < 0 || y < 0 || x >= width || y >= height) return false;
        const uint32_t nb = (uint32_t)y * (uint32_t)width + (uint32_t)x;
        if (!free_cells[nb]) return false;
        const uint32_t ng = gc + ((dx != 0 && dy != 0) ? 14u : 10u);
        const uint32_t nf = ng + h(nb);
        if (nf >= u32_load(g + goal())) return false;
        if (ng >= u32_load(g + nb)) return false;
        if (ng >= u32_min(g + nb, ng)) return false;
        child = (uint64_t)nf << 32 | nb;
        return true;
    }
};

// -------------------------------------------------------------- knapsack
// Best-first 0/1 branch-and-bound (artifacts/knapsack/knapsackKernel.cuh:
// 4-91: branch on the next item in value/weight order, bound by the
// fractional relaxation) with an integer Dantzig bound and dominance: a
// node survives only if its value beats the best recorded for its
// (depth, used weight) cell, and a popped node whose cell has since
// improved is stale.  Nodes are 16-byte records in a pool; the key's low
// word is the record index.  Every array of the solve is read with
// device-coherent loads: a long-running serving kernel may meet them in
// memory it has cached under an earlier query.
struct KnapNode {
    uint32_t level, used, value, pad;
};
constexpr uint64_t kUbTop = (1ull << 24) - 1;

__device__ __forceinline__ uint64_t knap_key(uint64_t ub, uint32_t level, uint32_t idx) {
    return ((((kUbTop - ub) << 8) | (uint64_t)(255u - level)) << 32) | idx;
}

struct KnapsackApp {
    static constexpr bool kThreshold = false;
    int n;
    uint64_t cap;
    const uint32_t* w;      // item weights in value/weight order
    const uint32_t* v;
    const uint64_t* pw;     // prefix sums of w and v, n + 1 entries
    const uint64_t* pv;
    uint32_t* cell;         // (n + 1) * (cap + 1): best value + 1
    KnapNode* pool;
    long long pool_cap;
    long long* ctl;

    __device__ KnapNode load_node(uint32_t idx) const {
        const uint4 q = gload16(buf_rsrc(pool + idx), 0);
        return KnapNode{q.x, q.y, q.z, q.w};
    }
    // value + whole items from `level` while they fit + floor(room * v / w)
    // of the first that does not (binary search on the prefix sums)
    __device__ uint64_t bound(uint32_t level, uint64_t room, uint64_t value) const {
        const uint64_t base_w = u64_load(pw + level);
        int lo = (int)level, hi = n;  // largest m with pw[m] - pw[level] <= room
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (u64_load(pw + mid) - base_w <= room) lo = mid;
            else hi = mid - 1;
        }
        uint64_t ub = value + (u64_load(pv + lo) - u64_load(pv + level));
        if (lo < n) ub += (room - (u64_load(pw + lo) - base_w)) * u32_load(v + lo) / u32_load(w + lo);
        return ub;
    }
    __device__ uint32_t* slot(uint32_t level, uint64_t used) const {
        return cell + (size_t)level * (cap + 1) + used;
    }
    __device__ int degree(uint64_t key, AppStats& st) const {
        const KnapNode nd = load_node((uint32_t)key);
        const uint64_t ub = kUbTop - (key >> 40);
        if (ub <= (uint64_t)ctl_load(ctl, APP_BEST) || (int)nd.level >= n) {
            atomicAdd(&st.pruned, 1);
            return 0;
        }
        if (u32_load(slot(nd.level, nd.used)) != nd.value + 1) {
            atomicAdd(&st.stale, 1);
            return 0;
        }
        atomicAdd(&st.expanded, 1);
        return 2;
    }
    __device__ bool expand(uint64_t key, int j, uint64_t& child, AppStats&) const {
        const KnapNode nd = load_node((uint32_t)key);
        const uint32_t lv = nd.level;
        uint64_t used = nd.used, value = nd.value;
        if (j == 0) {  // take item lv
            const uint32_t wl = u32_load(w + lv);
            if (used + wl > cap) return false;
            used += wl;
            value += u32_load(v + lv);
        }
        const uint32_t cl = lv + 1;
        if ((long long)value > ctl_load(ctl, APP_BEST))
            __hip_atomic_fetch_max(ctl_word(ctl, APP_BEST), (long long)value, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        if ((int)cl >= n) return false;
        const uint64_t ub = bound(cl, cap - used, value);
        if (ub <= (uint64_t)ctl_load(ctl, APP_BEST)) return false;
        if ((uint32_t)value + 1 <= u32_max(slot(cl, used), (uint32_t)value + 1)) return false;  // dominated
        const long long idx = __hip_atomic_fetch_add(ctl_word(ctl, APP_POOL), 1ll, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        if (idx >= pool_cap) {
            ctl_or(ctl, APP_ERROR, APPERR_POOL);
            return false;
        }
        gstore16(buf_rsrc(pool + idx), 0, make_uint4(cl, (uint32_t)used, (uint32_t)value, 0u));
        child = knap_key(ub, cl, (uint32_t)idx);
        return true;
    }
};

}  // namespace bgpq
