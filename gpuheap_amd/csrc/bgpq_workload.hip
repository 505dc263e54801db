// Synthetic workload generators that run on the device.
//
// Heap-sort keys: SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) evaluated
// counter-style: key i is the mix of state seed + (i+1) * gamma, exactly the
// (i+1)-th output of the sequential generator, so any range of keys can be
// produced by any thread.  u32 key = ((x >> 32) * 0xFFFFFFFF) >> 32, which
// lies in [0, 2^32 - 2] (strictly below the UINT32_MAX init_limits);
// u64 key = (u32 key << 32) | i, a 32-bit key with its input position as
// payload, so every u64 key is distinct.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bgpq.h"

namespace {

__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t key32(uint64_t x) { return (uint32_t)(((x >> 32) * 0xFFFFFFFFull) >> 32); }

__global__ void k_gen_u32(uint64_t seed, long long n, uint32_t* out) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = key32(splitmix64_at(seed, (uint64_t)i));
}

__global__ void k_gen_u64(uint64_t seed, long long n, uint64_t* out) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = ((uint64_t)key32(splitmix64_at(seed, (uint64_t)i)) << 32) | (uint32_t)i;
}

}  // namespace

extern "C" int bgpq_gen_keys(int key_type, uint64_t seed, int64_t n, void* d_out, void* stream) {
    if (n < 0 || (n > 0 && !d_out) || (key_type != BGPQ_KEY_U32 && key_type != BGPQ_KEY_U64)) return BGPQ_EINVAL;
    if (n == 0) return BGPQ_OK;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    (void)hipGetLastError();  // report this launch only
    if (key_type == BGPQ_KEY_U32)
        hipLaunchKernelGGL(k_gen_u32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, seed, (long long)n,
                           (uint32_t*)d_out);
    else
        hipLaunchKernelGGL(k_gen_u64, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, seed, (long long)n,
                           (uint64_t*)d_out);
    return hipGetLastError() == hipSuccess ? BGPQ_OK : BGPQ_EHIP;
}
