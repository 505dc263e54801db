// The reference's heap-sort harness (test/heap_test.cu:84-173 and
// test/partial_heap_test.cu) over the C++ facade: insert half the keys,
// insert the other half while deleting as many, delete the rest; the
// output, written in place by the heap-sort cursor, must equal std::sort.
//
//   heap_test <batch size> <log2 keys> [variant] [chunk] [seed]
//
// Keys: the first half below INT_MAX / 2, the second half at or above it,
// as in the reference (there from rand(), here from a seeded mt19937).
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "bgpq_heap.hpp"

using bgpq::Heap;
using bgpq::Variant;

static Variant parse_variant(const char* s) {
    if (!std::strcmp(s, "buitdd")) return Variant::buitdd;
    if (!std::strcmp(s, "buitdd_np")) return Variant::buitdd_np;
    if (!std::strcmp(s, "tditdd")) return Variant::tditdd;
    if (!std::strcmp(s, "tditdd_np")) return Variant::tditdd_np;
    std::fprintf(stderr, "unknown variant %s\n", s);
    std::exit(2);
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// `heap_test print`: printHeap of a small buitdd heap (B = 64) holding
// keys 0..199 (three full batches and a partial buffer of 8).
static int print_small() {
    try {
        Heap<uint32_t> h(8, 64, UINT32_MAX, Variant::buitdd);
        std::vector<uint32_t> keys(200);
        for (uint32_t i = 0; i < 200; ++i) keys[i] = (i * 7919u) % 200u;
        bgpq::DeviceArray<uint32_t> d(200);
        d.upload(keys.data(), 200);
        h.insertKernel(d.get(), 200, 50, 1);
        bgpq::check(bgpq_sync());
        h.printHeap();
    } catch (const bgpq::Error& e) {
        std::printf("%s\n", e.what());
        return -1;
    }
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc == 2 && std::string(argv[1]) == "print") return print_small();
    if (argc < 3) {
        std::printf("%s [batch size] [array num] [variant] [chunk] [seed]\n", argv[0]);
        return -1;
    }
    const uint32_t batchSize = (uint32_t)std::atoi(argv[1]);
    uint32_t arrayNum = 1u << std::atoi(argv[2]);
    const Variant v = argc > 3 ? parse_variant(argv[3]) : Variant::buitdd;
    const bool np = v == Variant::buitdd_np || v == Variant::tditdd_np;
    const int chunk = argc > 4 ? std::atoi(argv[4]) : (int)batchSize;
    const unsigned seed = argc > 5 ? (unsigned)std::atoi(argv[5]) : 1u;
    arrayNum = (arrayNum + 2 * batchSize - 1) / (2 * batchSize) * (2 * batchSize);
    uint32_t batchNum = 1;
    while (batchNum * batchSize < arrayNum * 2) batchNum <<= 1;
    const int blockNum = 32;
    std::printf("%s test size: %u heap[%u/%u] grid %d chunk %d\n", argv[0], arrayNum, batchSize, batchNum, blockNum,
                chunk);
    if (np && chunk != (int)batchSize) {
        std::printf("_np variants insert full batches only\n");
        return -1;
    }

    std::mt19937 rng(seed);
    std::vector<uint32_t> oriItems(arrayNum), oriItems_(arrayNum);
    for (uint32_t i = 0; i < arrayNum / 2; ++i) oriItems_[i] = oriItems[i] = rng() % (INT_MAX / 2);
    for (uint32_t i = arrayNum / 2; i < arrayNum; ++i) oriItems_[i] = oriItems[i] = INT_MAX / 2 + rng() % (INT_MAX / 2);
    std::sort(oriItems_.begin(), oriItems_.end());

    try {
        Heap<uint32_t> h_heap((int)batchNum, (int)batchSize, UINT32_MAX, v);
        bgpq::DeviceArray<uint32_t> heapItems(arrayNum);
        heapItems.upload(oriItems.data(), arrayNum);
        const uint32_t half = arrayNum / 2;
        const int64_t calls = half / batchSize;

        auto t0 = std::chrono::steady_clock::now();
        h_heap.insertKernel(heapItems.get(), half, chunk, blockNum);
        bgpq::check(bgpq_sync());
        const double insertTime = ms_since(t0);
        if (!h_heap.checkInsertHeap()) {
            std::printf("Wrong heap after the insertion phase\n");
            return -1;
        }

        t0 = std::chrono::steady_clock::now();
        h_heap.concurrentKernel(heapItems.get() + half, half, chunk, heapItems.get(), arrayNum, calls, blockNum);
        bgpq::check(bgpq_sync());
        const double concurrentTime = ms_since(t0);

        t0 = std::chrono::steady_clock::now();
        h_heap.deleteKernel(heapItems.get(), arrayNum, calls + 1, blockNum);
        bgpq::check(bgpq_sync());
        const double deleteTime = ms_since(t0);

        heapItems.download(oriItems.data(), arrayNum);
        if (!h_heap.isEmpty() || h_heap.errorFlags() != 0) {
            std::printf("Heap not empty or error flags 0x%x\n", h_heap.errorFlags());
            return -1;
        }
        for (uint32_t i = 0; i < arrayNum; ++i) {
            if (oriItems_[i] != oriItems[i]) {
                std::printf("Wrong Answer! id: %u correct %u heap %u\n", i, oriItems_[i], oriItems[i]);
                return -1;
            }
        }
        std::printf("Success %.3f %.3f %.3f %.3f\n", insertTime, concurrentTime, deleteTime,
                    insertTime + concurrentTime + deleteTime);
    } catch (const bgpq::Error& e) {
        std::printf("%s\n", e.what());
        return -1;
    }
    return 0;
}
