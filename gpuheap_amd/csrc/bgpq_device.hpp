// Device API of the batched GPU priority queue (BGPQ) for gfx950.
//
// This is the workgroup-collective interface that the reference exposes as
// the Heap<K> class template (heap/{buitdd,buitdd_np,tditdd,tditdd_np}/heap.cuh):
//   bool deleteRoot(K* items, int& size)               (buitdd:283-343, tditdd:267-341)
//   void deleteUpdate(int smemOffset)                   (buitdd:348-482, tditdd:345-534)
//   void insertion(const K* items, int size, int smemOffset)   (buitdd:484-690, tditdd:536-753)
//   int  getItemCount()                                 (buitdd:206-211)
// Every thread of the workgroup must call them.  `smemOffset` counts ints
// into the kernel's dynamic LDS (`extern __shared__ int[]`), exactly as in
// the reference; it must be a multiple of 4.  LDS needed past the offset:
// insertion 2*B*sizeof(K), deleteUpdate 3*B*sizeof(K) (lds_bytes()).
//
// Storage (HBM): items[(batch_num+1)*B]: slot 0 is the partial buffer,
// slots 1..batchCount the full sorted batches, every other slot holds
// init_limits.  status[batch_num+1] holds one state word per slot.
// counters: [0] batchCount, [1] partialBufferSize, [2] deleteCount (output
// cursor in heap-sort mode), [3] error flags.
//
// Variants (template parameter VAR):
//   VAR_BUITDD    bottom-up insertion, top-down deletion, partial buffer
//   VAR_BUITDD_NP same without the partial buffer (inserts must be full batches)
//   VAR_TDITDD    top-down insertion toward a target slot, partial buffer
//   VAR_TDITDD_NP same without the partial buffer
// bu variants address slot s as logical tree node rev(s) (bit reversal below
// the leading one), so consecutive new slots take disjoint insert paths.
#pragma once
#include "bgpq_prims.hpp"

namespace bgpq {

enum { VAR_BUITDD = 0, VAR_BUITDD_NP = 1, VAR_TDITDD = 2, VAR_TDITDD_NP = 3 };
enum { CTR_BATCH_COUNT = 0, CTR_PARTIAL = 1, CTR_DELETE_COUNT = 2, CTR_ERROR = 3, CTR_ABORT = 4, CTR_TICKET = 5,
       CTR_WORDS = 8 };
enum { ERR_OVERFLOW = 1, ERR_BAD_SIZE = 2, ERR_LOST_KEYS = 4, ERR_SPIN_LIMIT = 8 };

// A lock wait longer than this many polls (each at least one HBM atomic
// round trip, i.e. seconds in total) trips the heap's abort word: every
// workgroup then leaves its current operation at the next hand-off point, so
// a protocol fault ends the kernel with ERR_SPIN_LIMIT instead of hanging.
constexpr int kSpinLimit = 1 << 22;

template <typename K>
struct HeapView {
    K* items;
    int* status;
    int* ctr;
    K init_limits;
    int batch_num;
    int heap_sort;  // nonzero: deleteRoot writes at items + deleteCount (HEAP_SORT mode)
};

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

__device__ __forceinline__ int* bcast_slots() {
    __shared__ int slots[8];
    return slots;
}

template <typename K, int B>
constexpr int insert_lds_bytes() { return 2 * B * (int)sizeof(K); }
template <typename K, int B>
constexpr int delete_lds_bytes() { return 3 * B * (int)sizeof(K); }

template <typename K, int VAR, int B, int NT>
struct DevHeap {
    static_assert((B & (B - 1)) == 0 && B >= 64, "batch size must be a power of two >= 64");
    static_assert(NT % 64 == 0 && NT <= B && NT <= 1024, "threads per workgroup");
    static constexpr bool BU = (VAR == VAR_BUITDD || VAR == VAR_BUITDD_NP);
    static constexpr bool PB = (VAR == VAR_BUITDD || VAR == VAR_TDITDD);
    // state words (buitdd:6-9, tditdd:6-9).  bu words carry a tag: an
    // inserter marks the node it is sifting up as (tag << 2) | INSHOLD with a
    // tag unique to that insertion, and INUSE is INT_MAX so the atomicMax
    // lock idiom still works.  Finding its own tag again proves that no
    // deleter touched the node in between (see the early stop in insertion).
    static constexpr int AVAIL = 0;
    static constexpr int INUSE = BU ? 0x7FFFFFFF : 1;
    static constexpr int INSHOLD = 1, DELMOD = 2;  // bu, low two bits
    static constexpr int TARGET = 2, MARKED = 3;   // td
    static constexpr int kTagMask = (1 << 29) - 1;
    __device__ __forceinline__ static bool is_inshold(int w) { return w != INUSE && (w & 3) == INSHOLD; }

    HeapView<K> h;

    __device__ __forceinline__ K* node(int i) const { return h.items + (size_t)i * B; }
    __device__ __forceinline__ int* st(int i) const { return h.status + i; }
    __device__ __forceinline__ int* ctr(int i) const { return h.ctr + i; }
    __device__ __forceinline__ K* lds_ws(int smemOffset) const {
        extern __shared__ __align__(16) int bgpq_dyn_lds[];
        return reinterpret_cast<K*>(bgpq_dyn_lds + smemOffset);
    }

    // ---- lane-0 lock primitives (only thread 0 calls these).  Each returns
    // false when the heap's abort word is set or its own wait exceeds
    // kSpinLimit; the caller then leaves the operation (see aborted()).
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
        bcast_slots()[7] = 1;
    }
    __device__ __forceinline__ bool lock_cas(int i, int from, int to) const {
        int n = 0;
        while (!a_cas(st(i), from, to)) {
            if (poll_abort(n)) return false;
            backoff();
        }
        return true;
    }
    // bu "lock whatever state, remember it" idiom (buitdd:372-377,409-418);
    // returns the previous state, or -1 on abort
    __device__ __forceinline__ int lock_max(int i) const {
        int p, n = 0;
        while ((p = a_max(st(i), INUSE)) == INUSE) {
            if (poll_abort(n)) return -1;
            backoff();
        }
        return p;
    }
    // td child lock: CAS AVAIL->INUSE, waiting while INUSE; returns the state
    // found (AVAIL means locked), or -1 on abort
    __device__ __forceinline__ int lock_child_td(int i) const {
        int s, n = 0;
        while ((s = a_cas_val(st(i), AVAIL, INUSE)) == INUSE) {
            if (poll_abort(n)) return -1;
            backoff();
        }
        return s;
    }
    // Uniform abort test after a hand-off barrier.
    __device__ __forceinline__ bool aborted() const { return bcast_slots()[7] != 0; }
    // Start of every operation: thread 0 samples the abort word, all threads
    // agree on it through LDS.
    __device__ __forceinline__ bool op_begin() const {
        if (threadIdx.x == 0) bcast_slots()[7] = a_load(ctr(CTR_ABORT)) != 0;
        __syncthreads();
        return !aborted();
    }

    // ---- workgroup hand-off points
    __device__ __forceinline__ void sync_acq() const {
        __syncthreads();
        fence_acq();
    }
    __device__ __forceinline__ void sync_rel() const {
        fence_rel();
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

    // Copy `n` keys (any alignment) between global buffers, optionally
    // resetting the source to init_limits.
    __device__ __forceinline__ void copy_keys(K* dst, K* src, int n, bool reset) const {
        for (int i = threadIdx.x; i < n; i += NT) {
            dst[i] = src[i];
            if (reset) src[i] = h.init_limits;
        }
    }

    // ------------------------------------------------------------- deleteRoot
    // Lock the partial buffer (status[0]); if the heap holds full batches,
    // also lock the root and hand the root batch to the caller.  Returns true
    // when deleteUpdate must follow (both locks stay held until then).
    __device__ bool deleteRoot(K* items, int& size) const {
        int* bc = bcast_slots();
        size = 0;
        if (!op_begin()) return false;
        if (threadIdx.x == 0) {
            do {
                if (!lock_cas(0, AVAIL, INUSE)) break;
                fence_acq();
                const int nb = a_load(ctr(CTR_BATCH_COUNT));
                const int pbs = a_load(ctr(CTR_PARTIAL));
                const int off = h.heap_sort ? a_load(ctr(CTR_DELETE_COUNT)) : 0;
                int mode;
                if (nb == 0 && pbs == 0) mode = 0;
                else if (nb == 0) mode = PB ? 1 : 0;
                else mode = 2;
                if (mode == 2) {
                    if (!lock_cas(1, AVAIL, INUSE)) break;
                    if (h.heap_sort) a_store(ctr(CTR_DELETE_COUNT), off + B);
                }
                bc[0] = mode;
                bc[1] = pbs;
                bc[2] = off;
            } while (0);
        }
        sync_acq();
        if (aborted()) return false;
        const int mode = bc[0], pbs = bc[1], off = bc[2];
        __syncthreads();
        if (mode == 0) {
            if (threadIdx.x == 0) a_store_rel(st(0), AVAIL);
            return false;
        }
        if (mode == 1) {
            // only the partial buffer holds keys: hand all of it out (buitdd:305-320)
            copy_keys(items + off, node(0), pbs, true);
            sync_rel();
            if (threadIdx.x == 0) {
                if (h.heap_sort) a_store(ctr(CTR_DELETE_COUNT), off + pbs);
                a_store(ctr(CTR_PARTIAL), 0);
                a_store_rel(st(0), AVAIL);
            }
            size = pbs;
            return false;
        }
        K* dst = items + off;
        if (((uintptr_t)dst % 16) == 0) copy_batch<K, B, NT>(dst, node(1));
        else copy_keys(dst, node(1), B, false);
        size = B;
        __syncthreads();
        return true;
    }

    // ----------------------------------------------------------- deleteUpdate
    __device__ void deleteUpdate(int smemOffset) const {
        K* X0 = lds_ws(smemOffset);  // candidate carried down the tree
        K* X1 = X0 + B;              // smaller half of the children
        K* X2 = X1 + B;
        int* bc = bcast_slots();
        if (!op_begin()) return;

        if (threadIdx.x == 0) {
            const int last = a_add(ctr(CTR_BATCH_COUNT), -1);
            if (last == 1) {
                a_store_rel(st(1), AVAIL);
                a_store_rel(st(0), AVAIL);
            }
            bc[0] = last;
        }
        __syncthreads();
        const int last = bc[0];
        __syncthreads();
        if (last == 1) return;

        // ---- take the last batch as the candidate
        if constexpr (BU) {
            if (threadIdx.x == 0) lock_max(last);
            sync_acq();
            if (aborted()) return;
            take_batch<K, B, NT>(X0, node(last), h.init_limits);
            sync_rel();
            if (threadIdx.x == 0) a_store_rel(st(last), AVAIL);
        } else {
            // td: the last slot may be an insert target still in flight; then
            // mark it and let the inserter hand its batch over through the
            // root slot (tditdd:371-393, 727-744).
            if (threadIdx.x == 0) {
                int type = -1, n = 0;
                while (true) {
                    if (a_cas(st(last), AVAIL, INUSE)) { type = 0; break; }
                    if (a_cas(st(last), TARGET, MARKED)) { type = 1; break; }
                    if (poll_abort(n)) break;
                    backoff();
                }
                if (type == 1) {
                    if (!lock_cas(last, TARGET, AVAIL)) type = -1;
                }
                bc[1] = type;
            }
            sync_acq();
            if (aborted()) return;
            const int type = bc[1];
            if (type == 1) {
                copy_batch<K, B, NT>(X0, node(1));
                __syncthreads();
            } else {
                take_batch<K, B, NT>(X0, node(last), h.init_limits);
                sync_rel();
                if (threadIdx.x == 0) a_store_rel(st(last), AVAIL);
            }
        }
        __syncthreads();

        // ---- fold the partial buffer in: the larger half stays there
        if constexpr (PB) {
            copy_batch<K, B, NT>(X1, node(0));
            __syncthreads();
            wg_merge<K, B, NT>(X0, X1, ToPtr<K>{X0}, ToPtr<K>{node(0)});
        }
        sync_rel();
        if (threadIdx.x == 0) a_store_rel(st(0), AVAIL);

        // ---- sift the candidate down from the root
        int cur = 1;
        int cur_prev = AVAIL;
        while (true) {
            const int l = left_child(cur), r = right_child(cur);
            const bool lin = l <= h.batch_num, rin = r <= h.batch_num;
            if (threadIdx.x == 0) {
                int lp = AVAIL, rp = AVAIL, type = 2;
                if constexpr (BU) {
                    if (lin) lp = lock_max(l);
                    if (lp >= 0 && rin) rp = lock_max(r);
                    if (is_inshold(lp)) lp = DELMOD;
                    if (is_inshold(rp)) rp = DELMOD;
                } else {
                    // a child that is an insert target in flight is treated as
                    // absent (tditdd:425-446)
                    if (lin) {
                        const int s = lock_child_td(l);
                        if (s != AVAIL) type = 0;
                    }
                    if (type == 2 && rin) {
                        const int s = lock_child_td(r);
                        if (s != AVAIL) type = 1;
                    }
                }
                bc[2] = lp;
                bc[3] = rp;
                bc[4] = type;
            }
            sync_acq();
            if (aborted()) return;
            const int lp = bc[2], rp = bc[3], type = bc[4];
            if (type != 0) {
                if (lin) copy_batch<K, B, NT>(X1, node(l));
                else fill_batch<K, B, NT>(X1, h.init_limits);
            }
            if (type == 2) {
                if (rin) copy_batch<K, B, NT>(X2, node(r));
                else fill_batch<K, B, NT>(X2, h.init_limits);
            }
            __syncthreads();

            if constexpr (!BU) {
                if (type == 0) {  // no child: the candidate stays here
                    copy_batch<K, B, NT>(node(cur), X0);
                    sync_rel();
                    if (threadIdx.x == 0) a_store_rel(st(cur), AVAIL);
                    return;
                }
                if (type == 1) {  // only the left child, which is a leaf
                    wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(cur)}, ToPtr<K>{node(l)});
                    sync_rel();
                    if (threadIdx.x == 0) {
                        a_store_rel(st(cur), AVAIL);
                        a_store_rel(st(l), AVAIL);
                    }
                    return;
                }
            }

            // merge the two children: the larger half goes back to the child
            // with the larger maximum, the smaller half (X1) competes with the
            // candidate (buitdd:422-443)
            const bool tr = X1[B - 1] < X2[B - 1];
            const int tgt = tr ? r : l, oth = tr ? l : r;
            const bool tgt_in = tr ? rin : lin, oth_in = tr ? lin : rin;
            const int tgt_prev = tr ? rp : lp, oth_prev = tr ? lp : rp;
            if (tgt_in) wg_merge<K, B, NT>(X1, X2, ToPtr<K>{X1}, ToPtr<K>{node(tgt)});
            else wg_merge<K, B, NT>(X1, X2, ToPtr<K>{X1}, ToNone{});
            sync_rel();
            if (threadIdx.x == 0 && tgt_in) a_store_rel(st(tgt), BU ? tgt_prev : AVAIL);

            const K cmin = X0[0], cmax = X0[B - 1], kmin_ = X1[0], kmax_ = X1[B - 1];
            const bool stop = (VAR == VAR_BUITDD) ? (cmax < kmin_) : !(kmin_ < cmax);
            if (!(cmin < kmax_)) {
                // candidate >= all of the children's small half: the node takes it
                copy_batch<K, B, NT>(node(cur), X1);
            } else if (stop) {
                copy_batch<K, B, NT>(node(cur), X0);
                if (oth_in) copy_batch<K, B, NT>(node(oth), X1);
                sync_rel();
                if (threadIdx.x == 0) {
                    a_store_rel(st(cur), BU ? cur_prev : AVAIL);
                    if (oth_in) a_store_rel(st(oth), BU ? oth_prev : AVAIL);
                }
                return;
            } else {
                wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(cur)}, ToPtr<K>{X0});
            }
            sync_rel();
            if (threadIdx.x == 0) a_store_rel(st(cur), BU ? cur_prev : AVAIL);
            if (!oth_in) {
                // only reachable when keys equal init_limits (outside the contract)
                flag_error(ERR_LOST_KEYS);
                return;
            }
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
        wg_sort_batch<K, B, NT>(X0, items, size, h.init_limits);

        if (threadIdx.x == 0) {
            if (lock_cas(0, AVAIL, INUSE)) {
                fence_acq();
                bc[0] = a_load(ctr(CTR_BATCH_COUNT));
                bc[1] = a_load(ctr(CTR_PARTIAL));
                if constexpr (BU) {
                    const int ticket = a_load(ctr(CTR_TICKET));
                    a_store(ctr(CTR_TICKET), ticket + 1);
                    bc[6] = ((ticket & kTagMask) << 2) | INSHOLD;
                }
            }
        }
        sync_acq();
        if (aborted()) return;
        const int nb = bc[0], pbs = bc[1];
        const int mine = BU ? bc[6] : 0;  // this insertion's INSHOLD word
        __syncthreads();

        // ---- partial-buffer cases 1.1-2.2 (buitdd:522-613, tditdd:571-663)
        if constexpr (PB) {
            if (nb == 0 && size < B) {
                if (pbs == 0) {  // 1.1: the batch becomes the partial buffer
                    copy_batch<K, B, NT>(node(0), X0);
                    sync_rel();
                    if (threadIdx.x == 0) {
                        a_store(ctr(CTR_PARTIAL), size);
                        a_store_rel(st(0), AVAIL);
                    }
                    return;
                }
                copy_batch<K, B, NT>(X1, node(0));
                __syncthreads();
                if (size + pbs < B) {  // 1.2: still no full batch
                    wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(0)}, ToNone{});
                    sync_rel();
                    if (threadIdx.x == 0) {
                        a_store(ctr(CTR_PARTIAL), pbs + size);
                        a_store_rel(st(0), AVAIL);
                    }
                    return;
                }
                // 1.3: the smaller B keys become the root
                if (threadIdx.x == 0) {
                    a_add(ctr(CTR_BATCH_COUNT), 1);
                    lock_cas(1, AVAIL, INUSE);
                }
                sync_acq();
                if (aborted()) return;
                wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(1)}, ToPtr<K>{node(0)});
                sync_rel();
                if (threadIdx.x == 0) {
                    a_store(ctr(CTR_PARTIAL), pbs + size - B);
                    a_store_rel(st(0), AVAIL);
                    a_store_rel(st(1), AVAIL);
                }
                return;
            }
            copy_batch<K, B, NT>(X1, node(0));
            __syncthreads();
            if (size + pbs < B) {  // 2.1: fold into the partial buffer via the root
                wg_merge<K, B, NT>(X0, X1, ToPtr<K>{X0}, ToNone{});
                if (threadIdx.x == 0) lock_cas(1, AVAIL, INUSE);
                sync_acq();
                if (aborted()) return;
                copy_batch<K, B, NT>(X1, node(1));
                __syncthreads();
                wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(1)}, ToPtr<K>{node(0)});
                sync_rel();
                if (threadIdx.x == 0) {
                    a_store(ctr(CTR_PARTIAL), pbs + size);
                    a_store_rel(st(0), AVAIL);
                    a_store_rel(st(1), AVAIL);
                }
                return;
            }
            // 2.2: a full batch goes into the tree, the rest stays partial
            wg_merge<K, B, NT>(X0, X1, ToPtr<K>{X0}, ToPtr<K>{node(0)});
            if (threadIdx.x == 0) a_store(ctr(CTR_PARTIAL), pbs + size - B);
        }
        sync_rel();

        if constexpr (BU) {
            // ---- claim the next slot and sift up (buitdd:616-689)
            if (threadIdx.x == 0) {
                const int idx = a_add(ctr(CTR_BATCH_COUNT), 1) + 1;
                if (idx > h.batch_num) {
                    a_add(ctr(CTR_BATCH_COUNT), -1);  // no slot: roll the claim back
                    bc[2] = -1;
                } else {
                    bc[2] = lock_cas(idx, AVAIL, INUSE) ? idx : -2;
                }
                a_store_rel(st(0), AVAIL);
            }
            sync_acq();
            if (aborted()) return;
            int cur = bc[2];
            __syncthreads();
            if (cur < 0) {
                flag_error(ERR_OVERFLOW);
                return;
            }
            copy_batch<K, B, NT>(node(cur), X0);
            sync_rel();
            if (threadIdx.x == 0) a_store_rel(st(cur), cur == 1 ? AVAIL : mine);
            // A concurrent deleter may pull keys of this insertion above the
            // node it sits in (it marks the node DELMOD, or takes it as its
            // last batch).  Such keys are only repaired by walking every level
            // up to the root, as the reference always does (buitdd:636-683).
            // While the node still carries this insertion's own tag nothing
            // was pulled, and the walk may end as soon as parent.max <=
            // node.min: the merges above would be identities.
            bool dirty = false;
            while (cur != 1) {
                const int par = parent(cur);
                if (threadIdx.x == 0) {
                    int prev = -1;
                    if (lock_cas(par, AVAIL, INUSE)) prev = lock_max(cur);
                    bc[3] = prev;
                }
                sync_acq();
                if (aborted()) return;
                dirty = dirty || (bc[3] != mine);
                copy_batch<K, B, NT>(X0, node(cur));
                copy_batch<K, B, NT>(X1, node(par));
                __syncthreads();
                const bool identity = !(X0[0] < X1[B - 1]);  // parent.max <= node.min
                const bool done = identity && !dirty;
                if (!identity) {
                    if (!(X1[0] < X0[B - 1])) {
                        // parent.min >= node.max: swap (buitdd:647-661)
                        copy_batch<K, B, NT>(node(par), X0);
                        copy_batch<K, B, NT>(node(cur), X1);
                    } else {
                        wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(par)}, ToPtr<K>{node(cur)});
                    }
                }
                sync_rel();
                if (threadIdx.x == 0) {
                    a_store_rel(st(par), (done || par == 1) ? AVAIL : mine);
                    a_store_rel(st(cur), AVAIL);
                }
                if (done) return;
                cur = par;
            }
        } else {
            // ---- claim a target slot and walk root -> target (tditdd:666-752)
            if (threadIdx.x == 0) {
                const int tgt = a_add(ctr(CTR_BATCH_COUNT), 1) + 1;
                if (tgt > h.batch_num) {
                    a_add(ctr(CTR_BATCH_COUNT), -1);  // no slot: roll the claim back
                    bc[2] = -1;
                    a_store_rel(st(0), AVAIL);
                } else {
                    bc[2] = tgt;
                    if (lock_cas(tgt, AVAIL, TARGET) && tgt != 1) lock_cas(1, AVAIL, INUSE);
                }
            }
            sync_acq();
            if (aborted()) return;
            const int tgt = bc[2];
            __syncthreads();
            if (tgt < 0) {
                flag_error(ERR_OVERFLOW);
                return;
            }
            int cur = 1;
            bool stolen = false;
            while (cur != tgt) {
                if (threadIdx.x == 0) bc[3] = (a_load(st(tgt)) == MARKED);
                __syncthreads();
                stolen = bc[3] != 0;
                __syncthreads();
                if (stolen) break;
                // release the node above (for the root this is status[0])
                sync_rel();
                if (threadIdx.x == 0) a_store_rel(st(cur >> 1), AVAIL);
                copy_batch<K, B, NT>(X1, node(cur));
                __syncthreads();
                if (!(X1[0] < X0[B - 1])) {
                    // carried.max <= node.min: node takes the carried batch
                    copy_batch<K, B, NT>(node(cur), X0);
                    copy_batch<K, B, NT>(X0, X1);
                } else if (X0[0] < X1[B - 1]) {
                    wg_merge<K, B, NT>(X0, X1, ToPtr<K>{node(cur)}, ToPtr<K>{X0});
                }
                __syncthreads();
                const int nxt = next_toward(cur, tgt);
                if (threadIdx.x == 0 && nxt != tgt) lock_cas(nxt, AVAIL, INUSE);
                sync_acq();
                if (aborted()) return;
                cur = nxt;
            }
            if (!stolen) {
                if (threadIdx.x == 0) {
                    a_cas(st(tgt), TARGET, INUSE);
                    bc[3] = (a_load(st(tgt)) == MARKED);
                }
                sync_acq();
                stolen = bc[3] != 0;
                __syncthreads();
            }
            if (stolen) {
                // a deleter took our target as its last batch: hand the carried
                // batch over through the root slot, which that deleter holds
                copy_batch<K, B, NT>(node(1), X0);
                sync_rel();
                if (threadIdx.x == 0) {
                    a_store_rel(st(cur >> 1), AVAIL);
                    if (tgt != cur) a_store_rel(st(cur), AVAIL);
                    lock_cas(tgt, MARKED, TARGET);
                }
                return;
            }
            copy_batch<K, B, NT>(node(tgt), X0);
            sync_rel();
            if (threadIdx.x == 0) {
                a_store_rel(st(cur >> 1), AVAIL);
                a_store_rel(st(cur), AVAIL);
            }
        }
    }

    // Item count under the partial-buffer lock; thread 0 takes the lock
    // (the reference's getItemCount lets every thread spin, buitdd:206-211).
    __device__ int getItemCount() const {
        int* bc = bcast_slots();
        if (!op_begin()) return 0;
        if (threadIdx.x == 0) {
            bc[5] = 0;
            if (lock_cas(0, AVAIL, INUSE)) {
                fence_acq();
                bc[5] = a_load(ctr(CTR_PARTIAL)) + a_load(ctr(CTR_BATCH_COUNT)) * B;
                a_store_rel(st(0), AVAIL);
            }
        }
        __syncthreads();
        const int n = bc[5];
        __syncthreads();
        return n;
    }
};

}  // namespace bgpq
