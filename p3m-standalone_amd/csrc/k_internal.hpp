// Declarations shared by the HIP translation units (not part of any ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "p3m_kernels.h"

int p3mk_fail(int code, const char *what, int hip_code);
int p3mk_launch_check(const char *what);
hipStream_t p3mk_stream(p3mk_queue *q);

#define P3MK_TRY(expr)                \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != P3MK_OK)           \
            return rc_;               \
    } while (0)

// Exclusive prefix sum: start[0..m] (start[m] = total) of count[0..m);
// scratch holds p3mk_scan_scratch_ints(m) ints.
size_t p3mk_scan_scratch_ints(long m);
int p3mk_exclusive_scan(const int *count, long m, int *start, int *scratch, p3mk_queue *q);
