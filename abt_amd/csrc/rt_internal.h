// rt_internal.h -- pieces of the host runtime shared by runtime.cpp and
// batch_plan.cpp (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/abt_diff.h"
#include "records.h"

// Set the thread-local error message (abt_last_error) and return `code`.
int abt_fail(int code, const char* fmt, ...);
// Select the library's device for the calling thread; *cus = its CU count.
int abt_ensure_device(int* cus);
// The calling thread's stream and its grow-only device buffer `slot` (at least `bytes`
// long; slot 8 belongs to the visibility entry points).
int abt_thread_staging(int slot, size_t bytes, hipStream_t* st, void** p);

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return abt_fail(ABT_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

// far-field bins resident on the device
struct abt_dev_grid {
    abt::DevGrid d;
    void* mem = nullptr;
    int64_t nbins = 0;
};
