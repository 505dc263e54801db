// The heap handle behind the C ABI, shared by the heap entry points
// (bgpq_capi.cpp) and the application drivers (bgpq_apps.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bgpq.h"
#include "bgpq_kernels.hpp"

struct bgpq_heap {
    int variant;
    int key_type;
    int batch_num;
    int batch_size;
    uint64_t init_limits;
    void* items;
    int* status;
    int* ctr;
    const bgpq::LaunchRow* row;
    long long* app_ctl;  // driver control words, allocated on first use
};

// Records the message returned by bgpq_last_error and returns `code`.
int bgpq_fail(int code, const char* fmt, ...);
// Counter words (CTR_*) of the heap, copied to the host.
int bgpq_handle_counters(bgpq_heap* h, int* c);
// Kernel arguments of the heap; heap_sort = 0 hands popped batches to the
// caller's buffer instead of the heap-sort output cursor.
bgpq::HeapArgs bgpq_handle_args(const bgpq_heap* h, int heap_sort);

#define BGPQ_HIPCHK(expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return bgpq_fail(BGPQ_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
