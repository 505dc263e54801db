// Device API of the batched GPU priority queue (BGPQ) for gfx950.
//
// This is the workgroup-collective interface that the reference exposes as
// the Heap<K> class template (heap/{buitdd,buitdd_np,tditdd,tditdd_np}/heap.cuh):
//   bool deleteRoot(K* items, int& size)               (buitdd:283-343, tditdd:267-341)
//   bool deleteRoot(K* items, int& size, K threshold)  (artifacts/astar/BGPQ_B/heap.cuh:260-329)
//   void deleteUpdate(int smemOffset)                   (buitdd:348-482, tditdd:345-534)
//   void insertion(const K* items, int size, int smemOffset)   (buitdd:484-690, tditdd:536-753)
//   int  getItemCount()                                 (buitdd:206-211)
//   bool ifTerminate()                                  (tditdd:117-120, PBS_MODEL)
// Every thread of the workgroup must call them.  deleteUpdate must follow a
// deleteRoot that returned true, on the same DevHeap object (it carries the
// state deleteRoot found under status[0]: the batch taken early, the partial
// buffer's size) and with the same smemOffset.  `smemOffset` counts ints
// into the kernel's dynamic LDS (`extern __shared__ int[]`), exactly as in
// the reference; it must be a multiple of 4.  LDS needed past the offset:
// insertion 2*B*sizeof(K) (3*B*sizeof(K) for B >= 1024), deleteUpdate
// 3*B*sizeof(K).
//
// A kernel written outside this library gets the heap's storage from the
// host handle with bgpq_view() (include/bgpq.h) and builds its DevHeap with
// view_from<K>(): the counterpart of the reference's shallow device copy of
// the Heap<K> object (test/heap_test.cu:128-129).  DevHeap's VAR and B must
// match the view's variant and batch_size.
//
// Storage (HBM): items[(batch_num+1)*B]: slot 0 is the partial buffer,
// slots 1..batchCount the full sorted batches, every other slot holds
// init_limits.  One status word per slot and the counters (batchCount,
// partialBufferSize, deleteCount = heap-sort output cursor, error flags,
// abort word, the termination vote) each sit alone in a 128-byte line
// (kWordStride ints apart), so spinning workgroups do not share a line
// with the words the lock holder updates.
//
// Variants (template parameter VAR):
//   VAR_BUITDD    bottom-up insertion, top-down deletion, partial buffer
//   VAR_BUITDD_NP same without the partial buffer (inserts must be full batches)
//   VAR_TDITDD    top-down insertion toward a target slot, partial buffer
//   VAR_TDITDD_NP same without the partial buffer
// bu variants address slot s as logical tree node rev(s) (bit reversal below
// the leading one), so consecutive new slots take disjoint insert paths.
#pragma once
#include "bgpq.h"
#include "bgpq_prims.hpp"

// Phase hook for timeline tracing builds: such a build defines
// BGPQ_EVENT(id) before including this header.  By default it is empty.
#ifndef BGPQ_EVENT
#define BGPQ_EVENT(id) ((void)0)
#endif

namespace bgpq {

enum { VAR_BUITDD = 0, VAR_BUITDD_NP = 1, VAR_TDITDD = 2, VAR_TDITDD_NP = 3 };
// Counter words.  CTR_BUSY and CTR_TERM are the termination vote of the
// reference's PBS_MODEL (tbstate[] and terminate, tditdd:282-289), kept as
// a count of busy workgroups instead of one flag per workgroup.
enum {
    CTR_BATCH_COUNT = 0,
    CTR_PARTIAL = 1,
    CTR_DELETE_COUNT = 2,
    CTR_ERROR = 3,
    CTR_ABORT = 4,
    CTR_BUSY = 5,
    CTR_TERM = 6,
    CTR_TICKET = 8,  // status[0]'s ticket lock: 64 bits, [serving : 32 | next ticket : 32]
    CTR_WORDS = 9
};
// Counters a host snapshot reports (the ticket words are internal).
constexpr int kPublicCounters = 8;
enum { ERR_OVERFLOW = 1, ERR_BAD_SIZE = 2, ERR_LOST_KEYS = 4, ERR_SPIN_LIMIT = 8, ERR_OUTPUT_FULL = 16 };

// Ints between consecutive status words / counters: one 128-byte line each.
constexpr int kWordStride = 32;

// A lock wait longer than this many polls (each at least one device-scope
// load plus a sleep, seconds in total) trips the heap's abort word: every
// workgroup then leaves its current operation at the next hand-off point,
// so a protocol fault ends the kernel with ERR_SPIN_LIMIT instead of
// hanging.  status[0], which every operation takes, is a fair (FIFO) ticket
// lock, so no waiter there can lose indefinitely to the others.
constexpr int kSpinLimit = 1 << 22;

template <typename K>
struct HeapView {
    K* items;
    int* status;  // (batch_num + 1) * kWordStride ints
    int* ctr;     // CTR_WORDS * kWordStride ints
    K init_limits;
    int batch_num;
    int heap_sort;     // nonzero: deleteRoot writes at items + deleteCount (HEAP_SORT mode)
    long long out_cap; // heap-sort mode: keys the caller's output buffer holds
};

// The device view of a heap described by bgpq_view() on the host.  Popped
// batches go to the caller's buffer (no heap-sort cursor).
template <typename K>
__host__ __device__ inline HeapView<K> view_from(const bgpq_view_t& v) {
    return HeapView<K>{reinterpret_cast<K*>(v.items), v.status, v.ctr, (K)v.init_limits, v.batch_num, 0, 0};
}

// Workgroup size the library's own kernels use for a batch size: every
// thread holds B/NT keys while sorting and produces 2B/NT keys per merge.
constexpr int threads_for(int B) { return B <= 128 ? 64 : (B / 2 < 1024 ? B / 2 : 1024); }

// Slot <-> logical node map of the bu variants.  rev(1) is defined as 1: the
// reference's device form (buitdd:233-236) shifts by 32 there, its host form
// (buitdd:238-239) special-cases it.
__device__ __host__ __forceinline__ uint32_t rev_idx(uint32_t x) {
    if (x <= 1) return x;
    const int n = 32 - __builtin_clz(x);  // bit length
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t r = __builtin_bitreverse32(x);
#else
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
#endif
    return (r >> (33 - n)) | (1u << (n - 1));
}

// Child of `cur` on the path to `target` in natural indexing (tditdd:263-265).
__device__ __forceinline__ int next_toward(int cur, int target) {
    return target >> (__builtin_clz((uint32_t)cur) - __builtin_clz((uint32_t)target) - 1);
}

// LDS words through which lane 0 broadcasts protocol decisions.
enum { BC_A = 0, BC_B = 1, BC_C = 2, BC_D = 3, BC_E = 4, BC_F = 5, BC_TAG = 6, BC_ABORT = 7 };
__device__ __forceinline__ int* bcast_slots() {
    __shared__ int slots[8];
    return slots;
}

// LDS word through which one lane broadcasts a key.
template <typename K>
__device__ __forceinline__ K* key_slot() {
    __shared__ K slot;
    return &slot;
}

template <typename K, int B>
constexpr int insert_lds_bytes() { return (B >= 1024 ? 3 : 2) * B * (int)sizeof(K); }
template <typename K, int B>
constexpr int delete_lds_bytes() { return 3 * B * (int)sizeof(K); }

template <typename K, int VAR, int B, int NT>
struct DevHeap {
    static_assert((B & (B - 1)) == 0 && B >= 64, "batch size must be a power of two >= 64");
    static_assert(NT % 64 == 0 && NT <= B && NT <= 1024, "threads per workgroup");
    static_assert(sort_xch_bytes<K, B, NT>() <= insert_lds_bytes<K, B>(), "the sort's exchange fits the workspace");
    static constexpr bool BU = (VAR == VAR_BUITDD || VAR == VAR_BUITDD_NP);
    static constexpr bool PB = (VAR == VAR_BUITDD || VAR == VAR_TDITDD);
    // state words (buitdd:6-9, tditdd:6-9).  bu words carry a tag: an
    // inserter marks the node it is sifting up as (tag << 2) | INSHOLD with a
    // tag no other live insertion uses, and INUSE is INT_MAX so the atomicMax
    // lock idiom still works.  Finding its own tag again proves that no
    // deleter touched the node in between (see the early stop in insertion).
    static constexpr int AVAIL = 0;
    static constexpr int INUSE = BU ? 0x7FFFFFFF : 1;
    static constexpr int INSHOLD = 1, DELMOD = 2;  // bu, low two bits
    static constexpr int TARGET = 2, MARKED = 3;   // td
    static constexpr int kTagMask = (1 << 29) - 1;
    __device__ __forceinline__ static bool is_inshold(int w) { return w != INUSE && (w & 3) == INSHOLD; }

    HeapView<K> h;
    // Set by deleteRoot when it already took the last batch into LDS (and
    // folded the partial buffer in): the slot it took, 0 otherwise.
    mutable int pre_last_ = 0;
    // Set by deleteRoot: the partial buffer's size it found under status[0]
    // (still held by deleteUpdate).  With an empty partial buffer (all
    // init_limits) the fold into the candidate is the identity and is skipped.
    mutable int pbs_ = 0;
    // Lane 0: this workgroup's vote, busy since its last pop (counted in
    // CTR_BUSY) until a deleteRoot finds nothing to hand out.
    mutable int busy_ = 0;
    // Lane 0: the ticket this workgroup holds status[0] with.
    mutable unsigned t0_ = 0;

    __device__ __forceinline__ K* node(int i) const { return h.items + (size_t)i * B; }
    __device__ __forceinline__ int* st(int i) const { return h.status + (size_t)i * kWordStride; }
    __device__ __forceinline__ int* ctr(int i) const { return h.ctr + i * kWordStride; }
    __device__ __forceinline__ K* lds_ws(int smemOffset) const {
        extern __shared__ __align__(16) int bgpq_dyn_lds[];
        return reinterpret_cast<K*>(bgpq_dyn_lds + smemOffset);
    }

    // ---- lane-level lock primitives (called by one lane each).  Every spin
    // reads before it writes (test-and-test-and-set) and returns false / -1
    // when the heap's abort word is set or its own wait exceeds kSpinLimit.
    __device__ __forceinline__ bool poll_abort(int& n) const {
        if ((++n & 63) != 0) return false;
        if (n > kSpinLimit || a_load(ctr(CTR_ABORT)) != 0) {
            trip();
            return true;
        }
        return false;
    }
    __device__ void trip() const {
        a_store(ctr(CTR_ABORT), 1);
        __hip_atomic_fetch_or(ctr(CTR_ERROR), (int)ERR_SPIN_LIMIT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bcast_slots()[BC_ABORT] = 1;
    }
    // The first attempt goes straight to the atomic (an uncontended lock costs
    // one round trip); only a failed attempt falls back to polling.
    __device__ __forceinline__ bool lock_cas(int i, int from, int to) const {
        if (a_cas(st(i), from, to)) return true;
        int n = 0;
        while (!(a_load(st(i)) == from && a_cas(st(i), from, to))) {
            if (poll_abort(n)) return false;
            backoff();
        }
        return true;
    }
    // bu "lock whatever state, remember it" idiom (buitdd:372-377,409-418);
    // returns the previous state, or -1 on abort
    __device__ __forceinline__ int lock_max(int i) const {
        int p = a_max(st(i), INUSE);
        if (p != INUSE) return p;
        int n = 0;
        while (true) {
            if (a_load(st(i)) != INUSE) {
                p = a_max(st(i), INUSE);
                if (p != INUSE) return p;
            }
            if (poll_abort(n)) return -1;
            backoff();
        }
    }
    // lock_max for a deleter's children, where at most a workgroup or two
    // wait for a node: the waiter retries the atomic itself, so the attempt
    // that follows the holder's release takes the lock one round trip after
    // it (polling first would cost two).
    __device__ __forceinline__ int lock_max_eager(int i) const {
        int n = 0;
        while (true) {
            const int p = a_max(st(i), INUSE);
            if (p != INUSE) return p;
            if (poll_abort(n)) return -1;
        }
    }
    // lock_cas that retries the CAS itself instead of polling first: the
    // attempt after the holder's release takes the lock one round trip after
    // it.  For locks with one waiter at most: the delete's root lock, which
    // only the status[0] holder waits for (2.5% faster this way, td).
    __device__ __forceinline__ bool lock_cas_eager(int i, int from, int to) const {
        int n = 0;
        while (!a_cas(st(i), from, to)) {
            if (poll_abort(n)) return false;
        }
        return true;
    }
    // The insertion walk's parent lock, which many inserters may wait for
    // (every walk that reaches the root takes it): after a failed CAS the
    // waiter polls until the word shows `from` and only then tries again.
    // Measured on the heap sort of 2^26 u32 keys at B = 8192 with 256
    // inserting workgroups against the CAS retried at once: the insert's
    // atomics fell from 3,500 to 33 per insertion, its time rose from 44.0
    // to 46.1 ms.
    __device__ __forceinline__ bool lock_cas_polled(int i, int from, int to) const {
        int n = 0;
        while (!a_cas(st(i), from, to)) {
            while (a_load(st(i)) != from) {
                if (poll_abort(n)) return false;
            }
        }
        return true;
    }
    // td child lock: AVAIL -> INUSE, waiting while INUSE; returns the state
    // found (AVAIL means locked), or -1 on abort
    __device__ __forceinline__ int lock_child_td(int i) const {
        int s = a_cas_val(st(i), AVAIL, INUSE);
        if (s != INUSE) return s;
        int n = 0;
        while (true) {
            s = a_load(st(i));
            if (s != INUSE) {
                if (s != AVAIL) return s;
                if (a_cas(st(i), AVAIL, INUSE)) return AVAIL;
            }
            if (poll_abort(n)) return -1;
            backoff();
        }
    }
    // status[0], the lock every operation takes first (the partial buffer,
    // the counters and the order of roots), is a ticket lock: one atomic per
    // acquisition, waiters served in arrival order, and each waiter polls
    // the serving word with a sleep that grows with its place in the queue
    // (the one next in line polls at once).  The status word itself stays
    // AVAIL.  Lane 0; false on abort.
    // The ticket and the serving count share one 64-bit word (serving in the
    // low half), so one fetch-add takes a ticket and shows whether the lock
    // is free: an uncontended acquisition costs one round trip.
    __device__ __forceinline__ bool lock0() const {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(ctr(CTR_TICKET));
        const unsigned long long old =
            __hip_atomic_fetch_add(w, 1ull << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned t = (unsigned)(old >> 32);
        t0_ = t;
        unsigned d = t - (unsigned)old;
        int n = 0;
        while (d != 0) {
            if (poll_abort(n)) return false;
            for (unsigned k = 1; k < d && k < 32; ++k) __builtin_amdgcn_s_sleep(8);
            backoff();
            d = t - (unsigned)a_load(ctr(CTR_TICKET));
        }
        return true;
    }
    // Lane 0: this lane's earlier stores (counters, the partial buffer's
    // words) are acknowledged first, then the next ticket is served (an
    // increment of the low half only: no carry reaches the tickets).
    __device__ __forceinline__ void unlock0() const {
        wait_stores();
        a_add(ctr(CTR_TICKET), 1);
    }
    // bu, lane 0 under status[0]: batch slot nb + 1 becomes the last one and
    // is locked for the new batch (its index), or -1 when the heap is full,
    // -2 on abort.
    __device__ __forceinline__ int claim_slot(int nb) const {
        const int idx = nb + 1;
        if (idx > h.batch_num) return -1;
        a_store(ctr(CTR_BATCH_COUNT), idx);
        // the new slot is nearly always free: try the CAS first
        return (a_cas(st(idx), AVAIL, INUSE) || lock_cas(idx, AVAIL, INUSE)) ? idx : -2;
    }
    // Release of the node above `cur` in a td insertion walk (status[0]
    // above the root).
    __device__ __forceinline__ void release_above(int cur) const {
        if (cur >> 1) release_st(cur >> 1, AVAIL);
        else unlock0();
    }

    // Uniform abort test after a hand-off barrier.
    __device__ __forceinline__ bool aborted() const { return bcast_slots()[BC_ABORT] != 0; }
    // Start of every operation: thread 0 samples the abort word, all threads
    // agree on it through LDS.
    __device__ __forceinline__ bool op_begin() const {
        if (threadIdx.x == 0) bcast_slots()[BC_ABORT] = a_load(ctr(CTR_ABORT)) != 0;
        __syncthreads();
        return !aborted();
    }
    // Status-word release by one lane: its own earlier stores (counters
    // written under status[0]) are acknowledged first.
    __device__ __forceinline__ void release_st(int i, int v) const {
        wait_stores();
        a_store(st(i), v);
    }
    // Every wave's stores to heap memory are acknowledged, then a barrier:
    // after this, lane 0 may publish with a status store.
    __device__ __forceinline__ void publish() const {
        wait_stores();
        __syncthreads();
    }
    __device__ __forceinline__ void flag_error(int e) const {
        if (threadIdx.x == 0) __hip_atomic_fetch_or(ctr(CTR_ERROR), e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    __device__ __forceinline__ int left_child(int cur) const {
        if constexpr (BU) return (int)rev_idx(rev_idx((uint32_t)cur) << 1);
        else return cur << 1;
    }
    __device__ __forceinline__ int right_child(int cur) const {
        if constexpr (BU) return (int)rev_idx(rev_idx((uint32_t)cur) << 1 | 1u);
        else return (cur << 1) | 1;
    }
    __device__ __forceinline__ int parent(int cur) const {
        if constexpr (BU) return (int)rev_idx(rev_idx((uint32_t)cur) >> 1);
        else return cur >> 1;
    }

    // ------------------------------------------------------------- deleteRoot
    // Lock the partial buffer (status[0]); if the heap holds full batches,
    // also lock the root and hand the root batch to the caller.  Returns true
    // when deleteUpdate must follow (both locks stay held until then).
    //
    // With smemOffset >= 0 (bu variants), the deleter also takes the last
    // batch as its candidate, and folds the partial buffer into it, *before*
    // it waits for the root: that work then overlaps the previous deleter's
    // sift instead of following it.  status[0] is still held until the root
    // is locked, so roots are taken in heap-sort cursor order.  deleteUpdate
    // must then be called with the same smemOffset.
    __device__ bool deleteRoot(K* items, int& size, int smemOffset = -1) const {
        return delete_root(items, size, smemOffset, false, K(0));
    }
    // The A* forks' deleteRoot(K*, uint32_t&, uint32_t app_t_val)
    // (artifacts/astar/BGPQ_B/heap.cuh:260-287, BGPQ_T/heap.cuh:275-289):
    // when the heap holds full batches and the root's smallest key is not
    // below `threshold`, nothing is handed out (size 0, false) and the
    // workgroup votes idle, exactly as for an empty heap.  Pass the threshold
    // as a K (an int third argument is smemOffset).  The batch is not taken
    // early in this form: the root is checked under its lock first.
    __device__ bool deleteRoot(K* items, int& size, K threshold, int smemOffset = -1) const {
        return delete_root(items, size, smemOffset, true, threshold);
    }

    // ---- termination vote (PBS_MODEL: tbstate[] / terminate, tditdd:282-289).
    // A workgroup is busy from the pop that hands it keys until a deleteRoot
    // finds nothing for it (the heap empty, or the root at or above the
    // threshold); that deleteRoot, under status[0], ends the vote when no
    // workgroup is busy any more.  Workgroups that hold keys before their
    // first deleteRoot (the ones that seed the heap) call voteBusy() first,
    // and the host counts them in bgpq_vote_reset().
    __device__ void voteBusy() const {
        if (threadIdx.x == 0) busy_ = 1;
    }
    // Collective: the vote has ended (reference: ifTerminate, tditdd:117-120).
    __device__ bool ifTerminate() const {
        int* bc = bcast_slots();
        __syncthreads();
        if (threadIdx.x == 0) bc[BC_F] = a_load(ctr(CTR_TERM));
        __syncthreads();
        const bool t = bc[BC_F] != 0;
        __syncthreads();
        return t;
    }
    // Lane 0, under status[0]: record this workgroup's vote after a
    // deleteRoot that handed out keys (got) or nothing.
    __device__ __forceinline__ void vote(bool got) const {
        if (got) {
            if (!busy_) {
                busy_ = 1;
                a_add(ctr(CTR_BUSY), 1);
            }
            return;
        }
        int now = a_load(ctr(CTR_BUSY));
        if (busy_) {
            busy_ = 0;
            now = a_add(ctr(CTR_BUSY), -1) - 1;
        }
        if (now <= 0) a_store(ctr(CTR_TERM), 1);
    }

    __device__ bool delete_root(K* items, int& size, int smemOffset, bool use_thr, K thr) const {
        int* bc = bcast_slots();
        size = 0;
        pre_last_ = 0;
        if (!op_begin()) return false;
        const bool prefetch = BU && smemOffset >= 0 && !use_thr;
        BGPQ_EVENT(1);
        if (threadIdx.x == 0) {
            do {
                if (!lock0()) break;
                BGPQ_EVENT(2);
                const int nb = a_load(ctr(CTR_BATCH_COUNT));
                const int pbs = a_load(ctr(CTR_PARTIAL));
                const int off = h.heap_sort ? a_load(ctr(CTR_DELETE_COUNT)) : 0;
                int mode;
                if (nb == 0 && pbs == 0) mode = 0;
                else if (nb == 0) mode = PB ? 1 : 0;
                else mode = 2;
                // the heap-sort cursor never writes past the caller's buffer:
                // refuse the deletion and flag it instead
                const long long need = mode == 2 ? B : (mode == 1 ? pbs : 0);
                if (h.heap_sort && mode != 0 && (long long)off + need > h.out_cap) {
                    __hip_atomic_fetch_or(ctr(CTR_ERROR), (int)ERR_OUTPUT_FULL, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                    mode = 0;
                }
                // threshold form: the root is read under its own lock (the
                // deleter ahead may still be writing it), and a root at or
                // above the threshold is left in place
                bool root_held = false;
                if (mode == 2 && use_thr) {
                    if (!lock_cas_eager(1, AVAIL, INUSE)) break;
                    if (!(gload_key(node(1)) < thr)) {
                        release_st(1, AVAIL);
                        mode = 0;
                    } else {
                        root_held = true;
                    }
                }
                vote(mode != 0);
                // Only a batch below the root's children may leave early (the
                // previous deleter may still be refilling the root from slots
                // 2 and 3).  The batch taken usually holds keys >= everything
                // the new root receives, but not always: a deleter further
                // ahead may still be refilling its ancestor from its children,
                // and once the batch has left, that refill draws on the
                // sibling alone.  The heap without the batch stays consistent
                // (the slot reads as empty), so the root is then the smallest
                // batch of the rest, and the root step below merges the two.
                int last = 0;
                if (mode == 2 && prefetch && nb >= 4) {
                    last = nb;
                    a_store(ctr(CTR_BATCH_COUNT), nb - 1);  // every change of it happens under status[0]
                    if (lock_max(last) < 0) break;
                    BGPQ_EVENT(3);
                } else if (mode == 2) {
                    if (!root_held && !lock_cas_eager(1, AVAIL, INUSE)) break;  // one waiter, as below
                    if (h.heap_sort) a_store(ctr(CTR_DELETE_COUNT), off + B);
                }
                bc[BC_A] = mode;
                bc[BC_B] = pbs;
                bc[BC_C] = off;
                bc[BC_D] = last;
            } while (0);
        }
        __syncthreads();
        if (aborted()) return false;
        const int mode = bc[BC_A], pbs = bc[BC_B], off = bc[BC_C], last = bc[BC_D];
        pbs_ = pbs;
        __syncthreads();
        if (mode == 0) {
            if (threadIdx.x == 0) unlock0();
            return false;
        }
        if (mode == 1) {
            // only the partial buffer holds keys: hand all of it out (buitdd:305-320)
            K* pb = node(0);
            for (int i = threadIdx.x; i < pbs; i += NT) {
                items[off + i] = gload_key(pb + i);
                gstore_key(pb + i, h.init_limits);
            }
            publish();
            if (threadIdx.x == 0) {
                if (h.heap_sort) a_store(ctr(CTR_DELETE_COUNT), off + pbs);
                a_store(ctr(CTR_PARTIAL), 0);
                unlock0();
            }
            size = pbs;
            return false;
        }
        if (last > 0) {
            // candidate = the last batch (+ the partial buffer's smaller half),
            // taken while the previous deleter still holds the root
            K* X0 = lds_ws(smemOffset);
            K* X1 = X0 + B;
            take_batch<K, B, NT>(X0, node(last), h.init_limits);
            publish();
            BGPQ_EVENT(4);
            if (threadIdx.x == 0) release_st(last, AVAIL);
            if constexpr (PB) {
                if (pbs > 0) {
                    g2l_batch<K, B, NT>(X1, node(0));
                    __syncthreads();
                    if constexpr (kSplitFold) fold_large(X0, X1);
                    else wg_merge<K, B, NT>(X0, X1, ToLds<K>{X0}, ToGlobal<K>{node(0)});
                    publish();
                }
            }
            if (threadIdx.x == 0) {
                lock_cas_eager(1, AVAIL, INUSE);  // only the status[0] holder waits here
                BGPQ_EVENT(5);
                if (h.heap_sort) a_store(ctr(CTR_DELETE_COUNT), off + B);
                unlock0();
            }
            __syncthreads();
            if (aborted()) return false;
            if constexpr (PB && kSplitFold) {
                if (pbs > 0) fold_small(X0, X1);
            }
            pre_last_ = last;
        }
        // the root leaves; its largest key is kept for the check below
        K* dst = items + off;
        if (((uintptr_t)dst % 16) == 0) {
            g2g_batch_last<K, B, NT>(dst, node(1), key_slot<K>());
        } else {
            for (int i = threadIdx.x; i < B; i += NT) {
                const K v = gload_key(node(1) + i);
                dst[i] = v;
                if (i == B - 1) *key_slot<K>() = v;
            }
        }
        __syncthreads();
        if (last > 0 && X0_min(smemOffset) < *key_slot<K>()) {
            // the rare case: the B smallest of root and candidate leave
            K* X0 = lds_ws(smemOffset);
            K* X1 = X0 + B;
            wait_stores();  // the root's copy lands before the merge overwrites it
            g2l_batch<K, B, NT>(X1, node(1));
            merge_root_rare(X0, X1, X1 + B, dst);
        }
        size = B;
        BGPQ_EVENT(6);
        return true;
    }

    // The rare root step: dst <- the B smallest of the candidate X0 and the
    // root X1, X0 <- the B largest.  A merge with few registers on purpose:
    // it shares the kernel with the sift's merges.
    __device__ __forceinline__ K X0_min(int smemOffset) const { return lds_ws(smemOffset)[0]; }
    __device__ void merge_root_rare(K* X0, const K* X1, K* X2, K* dst) const {
        for (int i = threadIdx.x; i < B; i += NT) X2[i] = X0[i];
        __syncthreads();
        wg_merge_lean<K, B, NT>(X2, X1, dst, X0);
        __syncthreads();
    }

    // ----------------------------------------------------------- deleteUpdate
    __device__ void deleteUpdate(int smemOffset) const {
        K* X0 = lds_ws(smemOffset);  // candidate carried down the tree
        K* X1 = X0 + B;              // smaller half of the children
        K* X2 = X1 + B;
        int* bc = bcast_slots();
        if (!op_begin()) return;
        if (pre_last_ > 0) {
            // deleteRoot already took the candidate and released status[0]
            pre_last_ = 0;
            sift_down(X0, X1, X2);
            return;
        }

        if (threadIdx.x == 0) {
            const int last = a_add(ctr(CTR_BATCH_COUNT), -1);
            if (last == 1) {
                release_st(1, AVAIL);
                unlock0();
            }
            bc[BC_A] = last;
        }
        __syncthreads();
        const int last = bc[BC_A];
        __syncthreads();
        if (last == 1) return;

        // ---- take the last batch as the candidate
        if constexpr (BU) {
            if (threadIdx.x == 0) lock_max(last);
            __syncthreads();
            if (aborted()) return;
            take_batch<K, B, NT>(X0, node(last), h.init_limits);
            publish();
            if (threadIdx.x == 0) release_st(last, AVAIL);
        } else {
            // td: the last slot may be an insert target still in flight; then
            // mark it and let the inserter hand its batch over through the
            // root slot (tditdd:371-393, 727-744).
            if (threadIdx.x == 0) {
                int type = -1, n = 0;
                while (true) {
                    const int s = a_load(st(last));
                    if (s == AVAIL && a_cas(st(last), AVAIL, INUSE)) { type = 0; break; }
                    if (s == TARGET && a_cas(st(last), TARGET, MARKED)) { type = 1; break; }
                    if (poll_abort(n)) break;
                    backoff();
                }
                if (type == 1 && !lock_cas(last, TARGET, AVAIL)) type = -1;
                bc[BC_B] = type;
            }
            __syncthreads();
            if (aborted()) return;
            const int type = bc[BC_B];
            if (type == 1) {
                g2l_batch<K, B, NT>(X0, node(1));
            } else {
                take_batch<K, B, NT>(X0, node(last), h.init_limits);
                publish();
                if (threadIdx.x == 0) release_st(last, AVAIL);
            }
        }
        __syncthreads();

        // ---- fold the partial buffer in: the larger half stays there
        const bool fold = PB && pbs_ > 0;
        if constexpr (PB) {
            if (fold) {
                g2l_batch<K, B, NT>(X1, node(0));
                __syncthreads();
                if constexpr (kSplitFold) fold_large(X0, X1);
                else wg_merge<K, B, NT>(X0, X1, ToLds<K>{X0}, ToGlobal<K>{node(0)});
                wait_stores();
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) unlock0();
        if constexpr (PB && kSplitFold) {
            if (fold) fold_small(X0, X1);
        }
        sift_down(X0, X1, X2);
    }

    // Smallest batch size whose level steps take the two merge halves one
    // after the other, with the rest of the pipeline below.  Measured heap
    // sort of 2^28 u32 keys (buitdd_np): B = 1024 1.36e8 keys/s split against
    // 1.26e8 with one bitonic merge per step.  With the half merges on
    // 64-lane group windows, splitting also pays for u32 keys at B = 512 (the
    // delete of 2^24 keys 12-16% faster for bu, 4% for td) and B = 256
    // (7-11%).  u64 keys split from B = 1024 only: at B = 256 the A* and
    // knapsack solves were 4-6% slower split, and at B = 512 (SSSP 8% faster)
    // the split sift raises the 256-thread serve kernel from 159 to 179
    // registers, so it no longer fits on a CU beside the 512-thread one,
    // which the serve kernels need (each CU holds one workgroup of each).
    static constexpr int kSplitFrom = sizeof(K) == 4 ? 256 : 1024;
    // bu insertion walk for B >= 1024: two half merges per step, with copy
    // windows, the node's batch stays in LDS between steps (a third buffer)
    // and the parent's load overlaps the node's lock.  At B = 1024 this
    // measured 2% faster than one bitonic merge per step and the plain walk
    // once the copy windows came in (4% slower without them).
    static constexpr bool kSplitWalk = BU && B >= 1024;
    // The walk's merge groups: the half merges' default, except that the
    // partial-buffer insertion at B = 4096, whose kernel already holds 128
    // registers, keeps 32-lane groups (64-lane ones made it spill and its
    // insert 12% slower).
    static constexpr int kWalkGroup = (PB && B == 4096) ? merge_group<K>() : 0;

    // bu: lanes 0 and 1 request the locks of the two children of `nxt`
    // without waiting; the state word found (INUSE if another workgroup holds
    // it), or kNoPre.  The next level's lock phase takes the result.  (td
    // variants measured no faster with it: their insert targets make a
    // request fail more often.)
    static constexpr int kNoPre = INT_MIN;
    __device__ __forceinline__ int lock_ahead(int nxt) const {
        if (threadIdx.x < 2) {
            const int c = threadIdx.x == 0 ? left_child(nxt) : right_child(nxt);
            if (c <= h.batch_num) return a_max(st(c), INUSE);
        }
        return kNoPre;
    }
    __device__ __forceinline__ static bool held_ahead(int pre) {
        return pre != kNoPre && pre != INUSE;
    }

    // bu, B >= kSplitFrom: the lock phase of a level step whose children are
    // not in LDS yet.  Lanes 0 and 1 request the two locks (or take the ones
    // requested ahead) until at least one is held; the workgroup then starts
    // loading that child while lane 0 or 1 waits for the other, which is
    // loaded after a second barrier.  The deleter ahead releases the larger
    // child about one publish before the node it descends from, so usually
    // only the last child's load is left after the wait.  lp, rp: the state
    // words found.  Returns false on abort.
    __device__ bool lock_load_split(int l, int r, bool lin, bool rin, int pre, K* X1, K* X2, int& lp,
                                    int& rp) const {
        int* bc = bcast_slots();
        bool need = false;  // lanes 0, 1: the lock is still to be taken
        const int c = threadIdx.x == 0 ? l : r;
        if (threadIdx.x < 64) {
            int word = AVAIL;
            if (threadIdx.x < 2 && (threadIdx.x == 0 ? lin : rin)) {
                if (held_ahead(pre)) word = pre;
                else need = true;
            }
            int n = 0;
            while (true) {
                if (need) {
                    const int q = a_max(st(c), INUSE);
                    if (q != INUSE) {
                        word = q;
                        need = false;
                    } else if (poll_abort(n)) {
                        word = -1;
                        need = false;
                    }
                }
                if (__ballot(threadIdx.x < 2 && !need) != 0) break;
            }
            if (threadIdx.x < 2) bc[BC_C + threadIdx.x] = need ? kNoPre : (is_inshold(word) ? DELMOD : word);
        }
        __syncthreads();
        if (aborted()) return false;
        const int w0 = bc[BC_C], w1 = bc[BC_D];
        const bool h0 = w0 != kNoPre, h1 = w1 != kNoPre;
        BatchLoad<K, B, NT> f0, f1;
        if (h0 && lin) f0.issue(node(l));
        if (h1 && rin) f1.issue(node(r));
        lp = w0;
        rp = w1;
        if (!(h0 && h1)) {
            if (threadIdx.x < 2 && need) {
                const int q = lock_max_eager(c);
                bc[BC_E + threadIdx.x] = is_inshold(q) ? DELMOD : q;
            }
            __syncthreads();
            if (aborted()) return false;
            if (!h0) {
                lp = bc[BC_E];
                f0.issue(node(l));
            }
            if (!h1) {
                rp = bc[BC_F];
                f1.issue(node(r));
            }
        }
        if (lin) f0.commit(X1);
        else fill_lds<K, B, NT>(X1, h.init_limits);
        if (rin) f1.commit(X2);
        else fill_lds<K, B, NT>(X2, h.init_limits);
        __syncthreads();
        return true;
    }

    // Folding the partial buffer (X1) into a batch (X0) for B >= kSplitFrom: the
    // partial buffer's new contents, the larger half, are computed and stored
    // first (fold_large), so status[0] can be released before the batch takes
    // the smaller half (fold_small, from the same LDS inputs).
    static constexpr bool kSplitFold = B >= kSplitFrom;
    // The fold merges spread keys with spread keys (mixed windows) and keeps
    // the 128-byte groups: with 64-lane ones `buitdd` measured the same at
    // B = 1024 and 2048.
    __device__ __forceinline__ void fold_large(const K* X0, const K* X1) const {
        HalfMerge<K, B, NT, true, false, merge_group<K>()> hi;
        hi.run(X0, X1);
        hi.put(ToGlobal<K>{node(0)});
    }
    __device__ __forceinline__ void fold_small(K* X0, const K* X1) const {
        HalfMerge<K, B, NT, false, false, merge_group<K>()> lo;
        lo.run(X0, X1);
        __syncthreads();  // every read of X0 is done
        lo.put(ToLds<K>{X0});
        __syncthreads();
    }

    // ---- sift the candidate X0 down from the root, which the caller holds.
    // Each level merges the two children: the larger half goes back to the
    // child with the larger maximum (tgt), the smaller half competes with the
    // candidate (buitdd:422-443), and the sift continues at the other child
    // (oth).  The next deleter waits for the node this level releases, so for
    // B >= kSplitFrom the node's hold is kept short:
    //  - the smaller half is computed first; when the candidate is no smaller
    //    than all of it (the usual case) it is the node's new batch, and the
    //    node is published and released before the larger half is computed;
    //  - the larger child's new batch is then stored without a publish of its
    //    own and released with the next level's publish ("pending"): no other
    //    workgroup can want it before the node below is released, which
    //    happens at that same publish;
    //  - bu: the locks of oth's children are requested without waiting, just
    //    before the node's store (and given back if the sift stops here), and
    //    when both are granted their batches are loaded while the larger half
    //    is computed, so the next level starts with its children in LDS.
    // Every wait is for a node deeper than every node held, as before.
    __device__ void sift_down(K* X0, K* X1, K* X2) const {
        constexpr bool kSplit = B >= kSplitFrom;
        constexpr bool kAhead = BU && kSplit;
        int* bc = bcast_slots();
        int cur = 1;
        int cur_prev = AVAIL;
        int pend = 0, pend_word = AVAIL;  // stored, released with the next publish
        int pre = kNoPre;                 // lanes 0, 1: result of lock_ahead
        bool ready = false;               // X1, X2 hold cur's children, already locked
        int ready_lp = AVAIL, ready_rp = AVAIL;
        while (true) {
            const int l = left_child(cur), r = right_child(cur);
            const bool lin = l <= h.batch_num, rin = r <= h.batch_num;
            BGPQ_EVENT(10);
            int lp, rp, type;
            if (ready) {
                ready = false;
                lp = ready_lp;
                rp = ready_rp;
                type = 2;
            } else if constexpr (BU && kSplit) {
                if (!lock_load_split(l, r, lin, rin, pre, X1, X2, lp, rp)) return;
                pre = kNoPre;
                type = 2;
            } else {
                if constexpr (BU) {
                    // lanes 0 and 1 lock the two children at the same time
                    if (threadIdx.x < 2) {
                        const int c = threadIdx.x == 0 ? l : r;
                        const bool in = threadIdx.x == 0 ? lin : rin;
                        int p = AVAIL;
                        if (in) p = held_ahead(pre) ? pre : lock_max_eager(c);
                        if (is_inshold(p)) p = DELMOD;
                        bc[BC_C + threadIdx.x] = p;
                    }
                    pre = kNoPre;
                    if (threadIdx.x == 0) bc[BC_E] = 2;
                } else {
                    // a child that is an insert target in flight is treated as
                    // absent, and so is the right one after an absent left one
                    // (tditdd:425-446).  Lanes 0 and 1 take the two locks at
                    // the same time (one round trip instead of two); a right
                    // lock taken beside an absent left child is given back
                    // untouched.
                    int s = AVAIL;
                    if (threadIdx.x < 2 && (threadIdx.x == 0 ? lin : rin))
                        s = lock_child_td(threadIdx.x == 0 ? l : r);
                    if (threadIdx.x < 64) {
                        const int sr = __shfl(s, 1, 64);
                        if (threadIdx.x == 0) {
                            const int t2 = s != AVAIL ? 0 : sr != AVAIL ? 1 : 2;
                            if (t2 == 0 && rin && sr == AVAIL) a_store(st(r), AVAIL);
                            bc[BC_C] = AVAIL;
                            bc[BC_D] = AVAIL;
                            bc[BC_E] = t2;
                        }
                    }
                }
                __syncthreads();
                BGPQ_EVENT(11);
                if (aborted()) return;
                lp = bc[BC_C];
                rp = bc[BC_D];
                type = bc[BC_E];
                if (type != 0) {
                    if (lin) g2l_batch<K, B, NT>(X1, node(l));
                    else fill_lds<K, B, NT>(X1, h.init_limits);
                }
                if (type == 2) {
                    if (rin) g2l_batch<K, B, NT>(X2, node(r));
                    else fill_lds<K, B, NT>(X2, h.init_limits);
                }
                __syncthreads();
            }
            BGPQ_EVENT(12);

            if constexpr (!BU) {
                if (type == 0) {  // no child: the candidate stays here
                    l2g_batch<K, B, NT>(node(cur), X0);
                    publish();
                    if (threadIdx.x == 0) {
                        release_st(cur, AVAIL);
                        if (pend) release_st(pend, pend_word);
                    }
                    return;
                }
                if (type == 1) {  // only the left child, which is a leaf
                    wg_merge<K, B, NT>(X0, X1, ToGlobal<K>{node(cur)}, ToGlobal<K>{node(l)});
                    publish();
                    if (threadIdx.x == 0) {
                        release_st(cur, AVAIL);
                        release_st(l, AVAIL);
                        if (pend) release_st(pend, pend_word);
                    }
                    return;
                }
            }

            const bool tr = X1[B - 1] < X2[B - 1];
            const int tgt = tr ? r : l, oth = tr ? l : r;
            const bool tgt_in = tr ? rin : lin, oth_in = tr ? lin : rin;
            const int tgt_prev = tr ? rp : lp, oth_prev = tr ? lp : rp;
            const K cmin = X0[0], cmax = X0[B - 1];
            const K kmin_ = kmin(X1[0], X2[0]);  // the children's smallest key
            const bool stop = (VAR == VAR_BUITDD) ? (cmax < kmin_) : !(kmin_ < cmax);
            if constexpr (kSplit) {
                HalfMerge<K, B, NT, false> lo;
                lo.run(X1, X2);
                BGPQ_EVENT(30);
                if (threadIdx.x == NT - 1) *key_slot<K>() = lo.last();
                // the larger child stored one level up is released at the
                // barrier below: every wave waits for its own stores to it
                // first, before it issues the node's (issued a whole half
                // merge ago, they have nearly always landed by now), so the
                // next deleter can load it while it waits for this node
                if (pend) wait_stores();
                BGPQ_EVENT(31);
                // the lock requests ahead go out after that wait (it would wait
                // for their atomics too); their answers are read at the publish
                if constexpr (kAhead) {
                    if (oth_in) pre = lock_ahead(oth);
                }
                // the node's new batch in the usual case (below), stored
                // before the barrier; the rare paths overwrite it
                if (oth_in) lo.put(ToGlobal<K>{node(cur)});
                __syncthreads();
                if (pend) {
                    if (threadIdx.x == 0) a_store(st(pend), pend_word);
                    pend = 0;
                }
                BGPQ_EVENT(13);
                if (!(cmin < *key_slot<K>()) && oth_in) {
                    // candidate >= all of the small half: the node takes it
                    // (already stored above)
                    if constexpr (kAhead) {
                        if (threadIdx.x < 2) bc[BC_C + threadIdx.x] = pre;
                    }
                    publish();
                    if (threadIdx.x == 0) {
                        release_st(cur, BU ? cur_prev : AVAIL);
                        if (pend) release_st(pend, pend_word);
                    }
                    BGPQ_EVENT(17);
                    bool fetch = false;
                    int plp = AVAIL, prp = AVAIL;
                    BatchLoad<K, B, NT> fl, fr;
                    if constexpr (kAhead) {
                        const int nl = left_child(oth), nr = right_child(oth);
                        plp = bc[BC_C];
                        prp = bc[BC_D];
                        fetch = nl <= h.batch_num && nr <= h.batch_num && held_ahead(plp) && held_ahead(prp);
                        if (fetch) {
                            fl.issue(node(nl));
                            fr.issue(node(nr));
                        }
                    }
                    HalfMerge<K, B, NT, true> hi;
                    hi.run(X1, X2);
                    BGPQ_EVENT(32);
                    if (tgt_in) hi.put(ToGlobal<K>{node(tgt)});
                    pend = tgt_in ? tgt : 0;
                    pend_word = BU ? tgt_prev : AVAIL;
                    __syncthreads();  // X1 and X2 are read
                    BGPQ_EVENT(33);
                    if (fetch) {
                        fl.commit(X1);
                        fr.commit(X2);
                        __syncthreads();
                        ready = true;
                        ready_lp = is_inshold(plp) ? DELMOD : plp;
                        ready_rp = is_inshold(prp) ? DELMOD : prp;
                        pre = kNoPre;
                    }
                    BGPQ_EVENT(15);
                    cur = oth;
                    cur_prev = oth_prev;
                    continue;
                }
                // the candidate takes part: both halves, the small one into X1.
                // The node is written again below: the batch stored above
                // lands first.
                wait_stores();
                HalfMerge<K, B, NT, true> hi;
                hi.run(X1, X2);
                __syncthreads();  // every read of X1 and X2 is done
                lo.put(ToLds<K>{X1});
                if (tgt_in) hi.put(ToGlobal<K>{node(tgt)});
                __syncthreads();
            } else {
                if (tgt_in) wg_merge<K, B, NT>(X1, X2, ToLds<K>{X1}, ToGlobal<K>{node(tgt)});
                else wg_merge<K, B, NT>(X1, X2, ToLds<K>{X1}, ToNone{});
                BGPQ_EVENT(13);
            }

            if (!(cmin < X1[B - 1]) && oth_in) {
                // candidate >= all of the children's small half: the node takes it
                l2g_batch<K, B, NT>(node(cur), X1);
            } else if (stop || !oth_in) {
                // the candidate stays here and the small half goes to oth
                // (without oth, keys equal to init_limits were inserted,
                // which is outside the contract)
                l2g_batch<K, B, NT>(node(cur), X0);
                if (oth_in) l2g_batch<K, B, NT>(node(oth), X1);
                publish();
                if (threadIdx.x == 0) {
                    if (tgt_in) release_st(tgt, BU ? tgt_prev : AVAIL);
                    release_st(cur, BU ? cur_prev : AVAIL);
                    if (oth_in) release_st(oth, BU ? oth_prev : AVAIL);
                    if (pend) release_st(pend, pend_word);
                }
                if constexpr (kAhead) {
                    // give back the locks requested ahead, untouched
                    if (threadIdx.x < 2 && held_ahead(pre))
                        a_store(st(threadIdx.x == 0 ? left_child(oth) : right_child(oth)), pre);
                }
                if (!stop) flag_error(ERR_LOST_KEYS);
                BGPQ_EVENT(16);
                return;
            } else if constexpr (kSplit) {
                // the node takes the B smallest of candidate and small half,
                // the candidate keeps the rest (two halves: fewer registers)
                HalfMerge<K, B, NT, false> lo2;
                lo2.run(X0, X1);
                lo2.put(ToGlobal<K>{node(cur)});
                HalfMerge<K, B, NT, true> hi2;
                hi2.run(X0, X1);
                __syncthreads();  // every read of X0 is done
                hi2.put(ToLds<K>{X0});
                __syncthreads();
            } else {
                wg_merge<K, B, NT>(X0, X1, ToGlobal<K>{node(cur)}, ToLds<K>{X0});
            }
            BGPQ_EVENT(14);
            publish();
            if (threadIdx.x == 0) {
                if (tgt_in) release_st(tgt, BU ? tgt_prev : AVAIL);
                release_st(cur, BU ? cur_prev : AVAIL);
                if (pend) release_st(pend, pend_word);
            }
            pend = 0;
            BGPQ_EVENT(15);
            cur = oth;
            cur_prev = oth_prev;
        }
    }

    // -------------------------------------------------------------- insertion
    __device__ void insertion(const K* items, int size, int smemOffset) const {
        K* X0 = lds_ws(smemOffset);  // the new batch, sorted
        K* X1 = X0 + B;
        int* bc = bcast_slots();
        if (size <= 0) return;
        if (size > B) {
            flag_error(ERR_BAD_SIZE);
            return;
        }
        if (!op_begin()) return;
        BGPQ_EVENT(20);
        wg_sort_batch<K, B, NT>(X0, items, size, h.init_limits);
        BGPQ_EVENT(21);

        // A batch that goes into the tree (case 2.2, or any insertion
        // without a partial buffer) needs slot nb + 1.  On overflow nothing
        // is touched, the partial buffer included: status[0] is released
        // and the error flagged.
        auto to_tree_of = [&](int nb_, int pbs_) {
            return !(PB && (size + pbs_ < B || (nb_ == 0 && size < B)));
        };
        if (threadIdx.x == 0) {
            bc[BC_E] = 0;
            if (lock0()) {
                BGPQ_EVENT(22);
                const int nb0 = a_load(ctr(CTR_BATCH_COUNT));
                const int pbs0 = PB ? a_load(ctr(CTR_PARTIAL)) : 0;
                bc[BC_A] = nb0;
                bc[BC_B] = pbs0;
                // bu, a batch that goes straight into the tree (no partial
                // buffer to fold): lane 0 claims its slot and releases
                // status[0] at once, with no barrier in between
                if constexpr (BU) {
                    if (to_tree_of(nb0, pbs0) && !(PB && !(pbs0 == 0 && size == B))) {
                        bc[BC_C] = claim_slot(nb0);
                        unlock0();
                        bc[BC_E] = 1;
                    }
                }
            }
            // A workgroup runs one insertion at a time and an insertion leaves
            // no tagged word behind when it ends, so the workgroup index is a
            // tag no other live insertion carries.
            if constexpr (BU) bc[BC_TAG] = (((int)blockIdx.x + 1) & kTagMask) << 2 | INSHOLD;
        }
        __syncthreads();
        if (aborted()) return;
        const int nb = bc[BC_A], pbs = bc[BC_B];
        const bool direct = bc[BC_E] != 0;  // slot claimed, status[0] released
        const int mine = BU ? bc[BC_TAG] : 0;  // this insertion's INSHOLD word
        __syncthreads();
        const bool to_tree = to_tree_of(nb, pbs);
        if (!direct && to_tree && nb + 1 > h.batch_num) {
            if (threadIdx.x == 0) unlock0();
            flag_error(ERR_OVERFLOW);
            return;
        }

        // ---- partial-buffer cases 1.1-2.2 (buitdd:522-613, tditdd:571-663).
        // The partial buffer and every counter only change under status[0],
        // so they are read and written with plain coherent accesses here.
        if constexpr (PB) {
            if (nb == 0 && size < B) {
                if (pbs == 0) {  // 1.1: the batch becomes the partial buffer
                    l2g_batch<K, B, NT>(node(0), X0);
                    publish();
                    if (threadIdx.x == 0) {
                        a_store(ctr(CTR_PARTIAL), size);
                        unlock0();
                    }
                    return;
                }
                g2l_batch<K, B, NT>(X1, node(0));
                __syncthreads();
                if (size + pbs < B) {  // 1.2: still no full batch
                    wg_merge<K, B, NT>(X0, X1, ToGlobal<K>{node(0)}, ToNone{});
                    publish();
                    if (threadIdx.x == 0) {
                        a_store(ctr(CTR_PARTIAL), pbs + size);
                        unlock0();
                    }
                    return;
                }
                // 1.3: the smaller B keys become the root
                if (threadIdx.x == 0) {
                    a_store(ctr(CTR_BATCH_COUNT), 1);
                    lock_cas(1, AVAIL, INUSE);
                }
                __syncthreads();
                if (aborted()) return;
                wg_merge<K, B, NT>(X0, X1, ToGlobal<K>{node(1)}, ToGlobal<K>{node(0)});
                publish();
                if (threadIdx.x == 0) {
                    a_store(ctr(CTR_PARTIAL), pbs + size - B);
                    release_st(1, AVAIL);
                    unlock0();
                }
                return;
            }
        }
        // 2.2 with an empty partial buffer: the fold below is the identity
        const bool fold = PB && !(pbs == 0 && size == B);
        if constexpr (PB) {
            if (fold && !direct) {
                g2l_batch<K, B, NT>(X1, node(0));
                __syncthreads();
                if (size + pbs < B) {  // 2.1: fold into the partial buffer via the root
                    wg_merge<K, B, NT>(X0, X1, ToLds<K>{X0}, ToNone{});
                    if (threadIdx.x == 0) lock_cas(1, AVAIL, INUSE);
                    __syncthreads();
                    if (aborted()) return;
                    g2l_batch<K, B, NT>(X1, node(1));
                    __syncthreads();
                    wg_merge<K, B, NT>(X0, X1, ToGlobal<K>{node(1)}, ToGlobal<K>{node(0)});
                    publish();
                    if (threadIdx.x == 0) {
                        a_store(ctr(CTR_PARTIAL), pbs + size);
                        release_st(1, AVAIL);
                        unlock0();
                    }
                    return;
                }
                // 2.2: a full batch goes into the tree, the rest stays partial
                // (bu, B >= kSplitFrom: the batch takes its half after
                // status[0] is released, below)
                if constexpr (BU && kSplitFold) fold_large(X0, X1);
                else wg_merge<K, B, NT>(X0, X1, ToLds<K>{X0}, ToGlobal<K>{node(0)});
                if (threadIdx.x == 0) a_store(ctr(CTR_PARTIAL), pbs + size - B);
                publish();
            }
        }

        if constexpr (BU) {
            // ---- claim the next slot (unless done above) and sift up (buitdd:616-689)
            if (!direct) {
                if (threadIdx.x == 0) {
                    bc[BC_C] = claim_slot(nb);
                    unlock0();
                }
                __syncthreads();
            }
            if (aborted()) return;
            int cur = bc[BC_C];
            __syncthreads();
            if (cur < 0) {
                flag_error(ERR_OVERFLOW);
                return;
            }
            if constexpr (PB && kSplitFold) {
                if (fold) fold_small(X0, X1);
            }
            BGPQ_EVENT(23);
            l2g_batch<K, B, NT>(node(cur), X0);
            publish();
            if (threadIdx.x == 0) release_st(cur, cur == 1 ? AVAIL : mine);
            // A concurrent deleter may pull keys of this insertion above the
            // node it sits in (it marks the node DELMOD, or takes it as its
            // last batch).  Such keys are only repaired by walking every level
            // up to the root, as the reference always does (buitdd:636-683).
            // While the node still carries this insertion's own tag nothing
            // was pulled, and the walk may end as soon as parent.max <=
            // node.min: the merges above would be identities.
            bool dirty = false;
            // kSplitWalk: C holds the node's batch as this insertion last
            // wrote it (reloaded only when a deleter touched the node), P the
            // parent's, S the spare buffer (LDS past 2*B keys) that keeps the
            // parent's new batch.  Otherwise both are reloaded every step and
            // only the first 2*B keys of LDS are used.
            K* C = X0;
            K* P = X1;
            K* S = X1 + B;
            while (cur != 1) {
                const int par = parent(cur);
                int prev;
                if constexpr (kSplitWalk) {
                    // the parent is locked first (top-down order, as deleters
                    // lock); its batch loads while lane 0 takes the node's lock
                    if (threadIdx.x == 0) {
                        BGPQ_EVENT(24);
                        lock_cas_polled(par, AVAIL, INUSE);
                    }
                    __syncthreads();
                    if (aborted()) return;
                    BatchLoad<K, B, NT> pl;
                    pl.issue(node(par));
                    if (threadIdx.x == 0) bc[BC_D] = lock_max(cur);
                    pl.commit(P);
                    __syncthreads();
                    BGPQ_EVENT(25);
                    if (aborted()) return;
                    prev = bc[BC_D];
                    if (prev != mine) {  // a deleter touched the node: reload it
                        g2l_batch<K, B, NT>(C, node(cur));
                        __syncthreads();
                    }
                } else {
                    if (threadIdx.x == 0) {
                        int p = -1;
                        BGPQ_EVENT(24);
                        if (lock_cas(par, AVAIL, INUSE)) p = lock_max(cur);
                        bc[BC_D] = p;
                    }
                    __syncthreads();
                    BGPQ_EVENT(25);
                    if (aborted()) return;
                    prev = bc[BC_D];
                    g2l_batch<K, B, NT>(C, node(cur));
                    g2l_batch<K, B, NT>(P, node(par));
                    __syncthreads();
                }
                BGPQ_EVENT(26);
                dirty = dirty || (prev != mine);
                const bool identity = !(C[0] < P[B - 1]);  // parent.max <= node.min
                const bool done = identity && !dirty;
                const int par_word = (done || par == 1) ? AVAIL : mine;
                if (identity) {
                    publish();
                    if (threadIdx.x == 0) {
                        release_st(par, par_word);
                        release_st(cur, AVAIL);
                    }
                    K* t = C;  // the parent's batch is unchanged
                    C = P;
                    P = t;
                } else {
                    // The parent's new batch is stored and the parent released
                    // first (every insertion that reaches the root waits for
                    // it); the node stays locked until it holds the larger
                    // half, so no one reads it in between.
                    if (!(P[0] < C[B - 1])) {
                        // parent.min >= node.max: swap (buitdd:647-661); the
                        // parent now holds C
                        l2g_batch<K, B, NT>(node(par), C);
                        publish();
                        if (threadIdx.x == 0) release_st(par, par_word);
                        l2g_batch<K, B, NT>(node(cur), P);
                    } else if constexpr (!kSplitWalk) {
                        // B < 1024: one full merge is cheaper than two halves here
                        wg_merge<K, B, NT>(C, P, ToGlobal<K>{node(par)}, ToGlobal<K>{node(cur)});
                        publish();
                        if (threadIdx.x == 0) release_st(par, par_word);
                    } else {
                        HalfMerge<K, B, NT, false, true, kWalkGroup> lo;
                        lo.run(C, P);
                        lo.put(ToGlobal<K>{node(par)});
                        lo.put(ToLds<K>{S});
                        publish();
                        if (threadIdx.x == 0) release_st(par, par_word);
                        HalfMerge<K, B, NT, true, true, kWalkGroup> hi;
                        hi.run(C, P);
                        hi.put(ToGlobal<K>{node(cur)});
                        K* t = C;  // the parent's new batch is S
                        C = S;
                        S = t;
                    }
                    publish();
                    if (threadIdx.x == 0) release_st(cur, AVAIL);
                }
                BGPQ_EVENT(done ? 28 : 27);
                if (done) return;
                cur = par;
            }
        } else {
            // ---- claim a target slot and walk root -> target (tditdd:666-752)
            if (threadIdx.x == 0) {
                const int tgt = nb + 1;
                if (tgt > h.batch_num) {
                    bc[BC_C] = -1;
                    unlock0();
                } else {
                    a_store(ctr(CTR_BATCH_COUNT), tgt);
                    bc[BC_C] = tgt;
                    if (lock_cas(tgt, AVAIL, TARGET) && tgt != 1) lock_cas(1, AVAIL, INUSE);
                }
            }
            __syncthreads();
            if (aborted()) return;
            const int tgt = bc[BC_C];
            __syncthreads();
            if (tgt < 0) {
                flag_error(ERR_OVERFLOW);
                return;
            }
            int cur = 1;
            bool stolen = false;
            while (cur != tgt) {
                if (threadIdx.x == 0) {
                    // release the node above (for the root this is status[0]);
                    // its data was published at the end of the previous step
                    const int m = (a_load(st(tgt)) == MARKED);
                    if (!m) release_above(cur);
                    bc[BC_D] = m;
                }
                __syncthreads();
                stolen = bc[BC_D] != 0;
                __syncthreads();
                if (stolen) break;
                g2l_batch<K, B, NT>(X1, node(cur));
                __syncthreads();
                if (!(X1[0] < X0[B - 1])) {
                    // carried.max <= node.min: node takes the carried batch
                    l2g_batch<K, B, NT>(node(cur), X0);
                    l2l_batch<K, B, NT>(X0, X1);
                } else if (X0[0] < X1[B - 1]) {
                    wg_merge<K, B, NT, (B >= 4096)>(X0, X1, ToGlobal<K>{node(cur)}, ToLds<K>{X0});
                }
                publish();
                const int nxt = next_toward(cur, tgt);
                if (threadIdx.x == 0 && nxt != tgt) lock_cas(nxt, AVAIL, INUSE);
                __syncthreads();
                if (aborted()) return;
                cur = nxt;
            }
            if (!stolen) {
                if (threadIdx.x == 0) {
                    a_cas(st(tgt), TARGET, INUSE);
                    bc[BC_D] = (a_load(st(tgt)) == MARKED);
                }
                __syncthreads();
                stolen = bc[BC_D] != 0;
                __syncthreads();
            }
            if (stolen) {
                // a deleter took our target as its last batch: hand the carried
                // batch over through the root slot, which that deleter holds
                l2g_batch<K, B, NT>(node(1), X0);
                publish();
                if (threadIdx.x == 0) {
                    release_above(cur);
                    if (tgt != cur) release_st(cur, AVAIL);
                    lock_cas(tgt, MARKED, TARGET);
                }
                return;
            }
            l2g_batch<K, B, NT>(node(tgt), X0);
            publish();
            if (threadIdx.x == 0) {
                release_above(cur);
                release_st(cur, AVAIL);
            }
        }
    }

    // Item count under the partial-buffer lock; thread 0 takes the lock
    // (the reference's getItemCount lets every thread spin, buitdd:206-211).
    __device__ int getItemCount() const {
        int* bc = bcast_slots();
        if (!op_begin()) return 0;
        if (threadIdx.x == 0) {
            bc[BC_F] = 0;
            if (lock0()) {
                bc[BC_F] = a_load(ctr(CTR_PARTIAL)) + a_load(ctr(CTR_BATCH_COUNT)) * B;
                unlock0();
            }
        }
        __syncthreads();
        const int n = bc[BC_F];
        __syncthreads();
        return n;
    }
};

}  // namespace bgpq
