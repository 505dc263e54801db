// ORACLE -- test infrastructure only: the CPU path that bench.py's
// cpu_baseline leg times beside the GPU heap.
//
// The reference's CPU path for heap sort is std::sort of the same keys
// (test/heap_test.cu:118); the CPU priority queue beside it is a
// std::priority_queue push-all / pop-all heap sort.  Both single-threaded.
//
// usage: cpu_baseline <u32|u64> <n> <seed> <sort_reps> [pq_n] [pq_reps]
// sort_reps timed std::sort runs of all n keys after one warm-up (0: none);
// pq_reps push-all / pop-all runs of the first pq_n keys.  Reports the
// median run of each.  Prints one JSON object.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <queue>
#include <vector>

#include "oracle.h"

template <typename K>
static double time_sort(const std::vector<K>& keys, int reps, bool* ok) {
    std::vector<double> t;
    std::vector<K> v;
    for (int r = 0; r < reps + 1; ++r) {
        v = keys;
        auto a = std::chrono::steady_clock::now();
        std::sort(v.begin(), v.end());
        auto b = std::chrono::steady_clock::now();
        if (r > 0) t.push_back(std::chrono::duration<double>(b - a).count());
    }
    *ok = std::is_sorted(v.begin(), v.end());
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

template <typename K>
static double time_pq_once(const std::vector<K>& keys, int64_t n, bool* ok) {
    auto a = std::chrono::steady_clock::now();
    std::priority_queue<K, std::vector<K>, std::greater<K>> pq;
    for (int64_t i = 0; i < n; ++i) pq.push(keys[i]);
    K prev = 0;
    bool sorted = true;
    for (int64_t i = 0; i < n; ++i) {
        K k = pq.top();
        pq.pop();
        if (i && k < prev) sorted = false;
        prev = k;
    }
    auto b = std::chrono::steady_clock::now();
    *ok = sorted;
    return std::chrono::duration<double>(b - a).count();
}

template <typename K>
static double time_pq(const std::vector<K>& keys, int64_t n, int reps, bool* ok) {
    std::vector<double> t;
    *ok = true;
    for (int r = 0; r < reps; ++r) {
        bool o = false;
        t.push_back(time_pq_once(keys, n, &o));
        *ok = *ok && o;
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

template <typename K>
static int run(int64_t n, uint64_t seed, int reps, int64_t pq_n, int pq_reps, const char* name) {
    std::vector<K> keys(n);
    if (sizeof(K) == 4) orc_keys_u32(seed, n, (uint32_t*)keys.data());
    else orc_keys_u64(seed, n, (uint64_t*)keys.data());
    bool ok1 = true, ok2 = true;
    const double ts = reps > 0 ? time_sort(keys, reps, &ok1) : 0.0;
    pq_n = std::min(pq_n, n);
    const double tp = (pq_n > 0 && pq_reps > 0) ? time_pq(keys, pq_n, pq_reps, &ok2) : 0.0;
    printf("{\"key\": \"%s\", \"n\": %lld, \"seed\": %llu, \"reps\": %d, \"std_sort_s\": %.6f, "
           "\"std_sort_keys_per_s\": %.6e, \"pq_n\": %lld, \"pq_reps\": %d, \"pq_s\": %.6f, "
           "\"pq_keys_per_s\": %.6e, \"threads\": 1, \"ok\": %s}\n",
           name, (long long)n, (unsigned long long)seed, reps, ts, ts > 0 ? n / ts : 0.0, (long long)pq_n, pq_reps,
           tp, tp > 0 ? pq_n / tp : 0.0, (ok1 && ok2) ? "true" : "false");
    return (ok1 && ok2) ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s <u32|u64> <n> <seed> <sort_reps> [pq_n] [pq_reps]\n", argv[0]);
        return 2;
    }
    const int64_t n = atoll(argv[2]);
    const uint64_t seed = strtoull(argv[3], nullptr, 10);
    const int reps = atoi(argv[4]);
    const int64_t pq_n = argc > 5 ? atoll(argv[5]) : 0;
    const int pq_reps = argc > 6 ? atoi(argv[6]) : 1;
    if (!strcmp(argv[1], "u64")) return run<uint64_t>(n, seed, reps, pq_n, pq_reps, "u64");
    return run<uint32_t>(n, seed, reps, pq_n, pq_reps, "u32");
}
