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
