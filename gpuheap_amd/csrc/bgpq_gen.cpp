// Host generators of the synthetic driver inputs (include/bgpq_apps.h).
#include <stdint.h>

#include <vector>

#include "bgpq_apps.h"
#include "bgpq_handle.hpp"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += kGolden);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

uint64_t threshold(double p) {
    if (p <= 0.0) return 0;
    if (p >= 1.0) return 1ull << 32;
    return (uint64_t)(p * 4294967296.0);
}

}  // namespace

extern "C" int bgpq_gen_rmat(int scale, int degree, double a, double b, double c, uint64_t seed, uint32_t wmin,
                             uint32_t wmax, int32_t* rowptr, int32_t* col, uint32_t* w) {
    if (scale < 1 || scale > 30 || degree < 1 || !rowptr || !col || !w || wmin > wmax)
        return bgpq_fail(BGPQ_EINVAL, "rmat: bad arguments");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return bgpq_fail(BGPQ_EINVAL, "rmat: bad probabilities");
    const int64_t nv = 1ll << scale;
    const int64_t ne = nv * degree;
    if (ne > INT32_MAX) return bgpq_fail(BGPQ_EINVAL, "rmat: %lld edges exceed int32 indices", (long long)ne);
    const uint64_t ta = threshold(a), tab = threshold(a + b), tabc = threshold(a + b + c);
    const uint64_t span = (uint64_t)wmax - wmin + 1;
    std::vector<uint32_t> src(ne), dst(ne), wt(ne);
    for (int64_t e = 0; e < ne; ++e) {
        SplitMix64 r{seed + (uint64_t)(e + 1) * kGolden};
        uint32_t s = 0, d = 0;
        uint64_t bits = 0;
        for (int l = 0; l < scale; ++l) {
            uint32_t x;
            if ((l & 1) == 0) {
                bits = r.next();
                x = (uint32_t)(bits >> 32);
            } else {
                x = (uint32_t)bits;
            }
            const uint32_t sb = x >= tab ? 1u : 0u;
            const uint32_t db = (x >= ta && x < tab) || x >= tabc ? 1u : 0u;
            s = (s << 1) | sb;
            d = (d << 1) | db;
        }
        src[e] = s;
        dst[e] = d;
        wt[e] = wmin + (uint32_t)(r.next() % span);
    }
    std::vector<int64_t> cnt(nv + 1, 0);
    for (int64_t e = 0; e < ne; ++e) ++cnt[src[e] + 1];
    for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    for (int64_t v = 0; v <= nv; ++v) rowptr[v] = (int32_t)cnt[v];
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t at = cnt[src[e]]++;
        col[at] = (int32_t)dst[e];
        w[at] = wt[e];
    }
    return BGPQ_OK;
}

extern "C" int bgpq_gen_grid(int width, int height, double p, uint64_t seed, uint8_t* free_cells) {
    if (width < 1 || height < 1 || !free_cells || (int64_t)width * height > INT32_MAX)
        return bgpq_fail(BGPQ_EINVAL, "grid: bad arguments");
    const int64_t n = (int64_t)width * height;
    const uint64_t t = threshold(p);
    for (int64_t i = 0; i < n; ++i) {
        SplitMix64 r{seed + (uint64_t)(i + 1) * kGolden};
        free_cells[i] = (r.next() >> 32) < t ? 0 : 1;
    }
    free_cells[0] = 1;
    free_cells[n - 1] = 1;
    return BGPQ_OK;
}

extern "C" int bgpq_gen_knapsack(int n, uint64_t seed, uint32_t wmin, uint32_t wmax, uint32_t value_add,
                                 uint32_t* weight, uint32_t* value, uint64_t* capacity) {
    if (n < 1 || !weight || !value || !capacity || wmin > wmax || wmin == 0)
        return bgpq_fail(BGPQ_EINVAL, "knapsack: bad arguments");
    const uint64_t span = (uint64_t)wmax - wmin + 1;
    uint64_t total = 0;
    for (int i = 0; i < n; ++i) {
        SplitMix64 r{seed + (uint64_t)(i + 1) * kGolden};
        weight[i] = wmin + (uint32_t)(r.next() % span);
        value[i] = weight[i] + value_add;
        total += weight[i];
    }
    *capacity = total / 2;
    return BGPQ_OK;
}
