// One instantiation unit per (key type, variant), selected at compile time
// with -DBGPQ_KEY_BITS=32|64 and -DBGPQ_VAR=0..3, so the build compiles the
// 8 units in parallel.
#include "bgpq_kernels.hpp"

#if BGPQ_KEY_BITS == 32
using KeyT = uint32_t;
#define BGPQ_ROWS_FN(v) rows_u32_v##v
#else
using KeyT = uint64_t;
#define BGPQ_ROWS_FN(v) rows_u64_v##v
#endif
#define BGPQ_CAT(a, b) BGPQ_CAT2(a, b)
#define BGPQ_CAT2(a, b) a##b

namespace bgpq {

#if BGPQ_KEY_BITS == 32
const LaunchRow* BGPQ_CAT(rows_u32_v, BGPQ_VAR)() {
#else
const LaunchRow* BGPQ_CAT(rows_u64_v, BGPQ_VAR)() {
#endif
    static const LaunchRow rows[kNumBatchSizes] = {
        launch_row<KeyT, BGPQ_VAR, 64>(),   launch_row<KeyT, BGPQ_VAR, 128>(),
        launch_row<KeyT, BGPQ_VAR, 256>(),  launch_row<KeyT, BGPQ_VAR, 512>(),
        launch_row<KeyT, BGPQ_VAR, 1024>(), launch_row<KeyT, BGPQ_VAR, 2048>(),
        launch_row<KeyT, BGPQ_VAR, 4096>(), launch_row<KeyT, BGPQ_VAR, 8192>(),
    };
    return rows;
}

#if BGPQ_VAR == 0
#if BGPQ_KEY_BITS == 32
hipError_t fill_u32(void* p, long long n, uint32_t v, hipStream_t s) {
#else
hipError_t fill_u64(void* p, long long n, uint64_t v, hipStream_t s) {
#endif
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    (void)hipGetLastError();  // report this launch only
    hipLaunchKernelGGL(k_fill<KeyT>, dim3((unsigned)blocks), dim3(256), 0, s, (KeyT*)p, n, (KeyT)v);
    return hipGetLastError();
}

#if BGPQ_KEY_BITS == 32
hipError_t clear_u32(void* items, int* status, int* ctr, int batch_num, int batch_size, uint32_t v, hipStream_t s) {
#else
hipError_t clear_u64(void* items, int* status, int* ctr, int batch_num, int batch_size, uint64_t v, hipStream_t s) {
#endif
    (void)hipGetLastError();  // report these launches only
    hipLaunchKernelGGL(k_clear<KeyT>, dim3(1024), dim3(256), 0, s, (KeyT*)items, status, (const int*)ctr, batch_num,
                       batch_size, (KeyT)v);
    hipLaunchKernelGGL(k_clear_ctr<KeyT>, dim3(1), dim3(256), 0, s, ctr);
    return hipGetLastError();
}
#endif

}  // namespace bgpq
