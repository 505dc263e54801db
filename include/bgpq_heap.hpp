// C++ host facade of the heap over the C ABI (include/bgpq.h): the
// reference's Heap<K> class template (heap/buitdd/heap.cuh:13-231,
// heap/tditdd/heap.cuh:11-231) with the same names and argument meaning,
// plus its harness kernels (test/heap_test.cu:13-74) as bulk operations.
//
// Differences a caller sees: a C ABI status other than BGPQ_OK becomes a
// bgpq::Error exception (the reference prints and exits in its harness);
// checkInsertHeap prints the first violation, as the reference does, and
// returns false; the variant is a constructor argument instead of an
// include directory; device buffers are plain pointers from
// bgpq_dev_malloc or hipMalloc.
#pragma once
#include <stdint.h>

#include <cstdio>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bgpq.h"

namespace bgpq {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

inline int check(int rc) {
    if (rc < 0) throw Error(rc, std::string("bgpq error ") + std::to_string(rc) + ": " + bgpq_last_error());
    return rc;
}

enum class Variant { buitdd = BGPQ_BUITDD, buitdd_np = BGPQ_BUITDD_NP, tditdd = BGPQ_TDITDD, tditdd_np = BGPQ_TDITDD_NP };

template <class K>
struct KeyType;
template <>
struct KeyType<uint32_t> {
    static constexpr int value = BGPQ_KEY_U32;
};
template <>
struct KeyType<uint64_t> {
    static constexpr int value = BGPQ_KEY_U64;
};

template <class K>
class Heap {
   public:
    // Heap(int batchNum, int batchSize, K init_limits) -- heap/buitdd/heap.cuh:34-68
    Heap(int batchNum, int batchSize, K init_limits = std::numeric_limits<K>::max(), Variant v = Variant::buitdd)
        : batchNum_(batchNum), batchSize_(batchSize) {
        check(bgpq_create(&h_, (int)v, KeyType<K>::value, batchNum, batchSize, (uint64_t)init_limits));
    }
    ~Heap() { bgpq_destroy(h_); }
    Heap(const Heap&) = delete;
    Heap& operator=(const Heap&) = delete;

    // reset() -- heap/buitdd/heap.cuh:118 (fill every slot with init_limits)
    void reset() {
        check(bgpq_reset(h_, nullptr));
        check(bgpq_sync());
    }
    // checkInsertHeap() -- heap/buitdd/heap.cuh:118-172
    bool checkInsertHeap() {
        const int rc = check(bgpq_check(h_));
        if (rc != 1) std::printf("%s\n", bgpq_last_error());
        return rc == 1;
    }
    int itemCount() {
        int64_t v = 0;
        check(bgpq_item_count(h_, &v));
        return (int)v;
    }
    int nodeCount() {
        int64_t v = 0;
        check(bgpq_node_count(h_, &v));
        return (int)v;
    }
    bool isEmpty() { return check(bgpq_is_empty(h_)) == 1; }
    // printHeap() -- heap/buitdd/heap.cuh:175-194: the partial buffer, then
    // every batch in use, each line tagged with its slot's state word
    void printHeap() {
        const size_t keys = (size_t)(batchNum_ + 1) * batchSize_;
        std::vector<K> items(keys);
        std::vector<int> status(batchNum_ + 1);
        int ctr[8] = {};
        check(bgpq_snapshot(h_, items.data(), status.data(), ctr));
        // "batch partial <size>_<state>:" and "batch <slot>_<state>:", keys after
        auto line = [&](const char* what, int label, int slot, int n) {
            std::printf("%s %d_%d:", what, label, status[slot]);
            for (int i = 0; i < n; ++i) std::printf(" %llu", (unsigned long long)items[(size_t)slot * batchSize_ + i]);
            std::printf("\n");
        };
        line("batch partial", ctr[1], 0, ctr[1]);
        for (int i = 1; i <= ctr[0]; ++i) line("batch", i, i, batchSize_);
    }

    // The harness kernels.  insertKernel: n keys from d_items in calls of
    // `chunk` keys; deleteKernel: `calls` deleteRoot/deleteUpdate pairs,
    // popped batches written at the heap-sort cursor of d_out (out_keys
    // long); concurrentKernel: half the workgroups insert, half delete.
    void insertKernel(const K* d_items, int64_t n, int chunk, int grid = 32) {
        check(bgpq_insert(h_, d_items, n, chunk, grid, nullptr));
    }
    void deleteKernel(K* d_out, int64_t out_keys, int64_t calls, int grid = 32) {
        check(bgpq_delete(h_, d_out, out_keys, calls, grid, nullptr));
    }
    void concurrentKernel(const K* d_items, int64_t n_insert, int chunk, K* d_out, int64_t out_keys, int64_t calls,
                          int grid = 32) {
        check(bgpq_concurrent(h_, d_items, n_insert, chunk, d_out, out_keys, calls, grid, nullptr));
    }
    // insert all n keys, then delete until empty; returns (insert ms, delete ms)
    std::pair<float, float> heapSort(const K* d_in, K* d_out, int64_t n, int chunk = 0, int grid_insert = 0,
                                     int grid_delete = 0) {
        float a = 0.f, b = 0.f;
        check(bgpq_heapsort(h_, d_in, d_out, n, chunk, grid_insert, grid_delete, &a, &b, nullptr));
        return {a, b};
    }
    void setCursor(int64_t v) { check(bgpq_set_cursor(h_, v)); }
    int errorFlags() {
        int f = 0;
        check(bgpq_error_flags(h_, &f));
        return f;
    }

    int batchNum() const { return batchNum_; }
    int batchSize() const { return batchSize_; }
    bgpq_heap* handle() { return h_; }

   private:
    bgpq_heap* h_ = nullptr;
    int batchNum_, batchSize_;
};

// Device buffer owned by the host (bgpq_dev_malloc / bgpq_dev_free).
template <class K>
class DeviceArray {
   public:
    explicit DeviceArray(size_t n) : n_(n) {
        void* p = nullptr;
        check(bgpq_dev_malloc(&p, n * sizeof(K) + 1));
        p_ = (K*)p;
    }
    ~DeviceArray() { bgpq_dev_free(p_); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void upload(const K* h, size_t n) { check(bgpq_h2d(p_, h, n * sizeof(K))); }
    void download(K* h, size_t n) const { check(bgpq_d2h(h, p_, n * sizeof(K))); }
    K* get() const { return p_; }
    size_t size() const { return n_; }

   private:
    K* p_ = nullptr;
    size_t n_;
};

}  // namespace bgpq
