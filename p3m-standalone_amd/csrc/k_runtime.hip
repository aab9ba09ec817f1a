// HIP runtime wrappers of the internal C ABI: streams, events, memory,
// small utility kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "k_internal.hpp"

static thread_local char g_err[512];

extern "C" const char *p3mk_error(void) { return g_err; }

int p3mk_fail(int code, const char *what, int hip_code) {
    if (hip_code != 0)
        std::snprintf(g_err, sizeof(g_err), "%s: %s", what,
                      hipGetErrorString((hipError_t)hip_code));
    else
        std::snprintf(g_err, sizeof(g_err), "%s", what);
    return code;
}

int p3mk_launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, what, (int)e);
    return P3MK_OK;
}

struct p3mk_queue {
    hipStream_t stream;
    hipEvent_t ev[P3MK_NUM_EVENTS];
    // optional ring of event sets: set k receives the events of the k-th
    // step after p3mk_ring_start, so the stage times of several steps can be
    // read after they all ran
    hipEvent_t *ring = nullptr;
    int ring_sets = 0;
    int ring_pos = 0;
};

static void ring_free(p3mk_queue *q) {
    if (q->ring) {
        for (int i = 0; i < q->ring_sets * P3MK_NUM_EVENTS; i++)
            (void)hipEventDestroy(q->ring[i]);
        delete[] q->ring;
    }
    q->ring = nullptr;
    q->ring_sets = q->ring_pos = 0;
}

hipStream_t p3mk_stream(p3mk_queue *q) { return q ? q->stream : (hipStream_t)0; }

extern "C" int p3mk_device_check(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1)
        return p3mk_fail(P3MK_ENODEVICE, "no HIP device available", e == hipSuccess ? 0 : (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_queue_create(p3mk_queue **out) {
    int rc = p3mk_device_check();
    if (rc != P3MK_OK)
        return rc;
    p3mk_queue *q = new (std::nothrow) p3mk_queue;
    if (!q)
        return p3mk_fail(P3MK_ENOMEM, "queue allocation failed", 0);
    hipError_t e = hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete q;
        return p3mk_fail(P3MK_EDEVICE, "hipStreamCreate", (int)e);
    }
    for (int i = 0; i < P3MK_NUM_EVENTS; i++) {
        e = hipEventCreate(&q->ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; j++)
                (void)hipEventDestroy(q->ev[j]);
            (void)hipStreamDestroy(q->stream);
            delete q;
            return p3mk_fail(P3MK_EDEVICE, "hipEventCreate", (int)e);
        }
    }
    *out = q;
    return P3MK_OK;
}

extern "C" void p3mk_queue_destroy(p3mk_queue *q) {
    if (!q)
        return;
    (void)hipStreamSynchronize(q->stream);
    ring_free(q);
    for (int i = 0; i < P3MK_NUM_EVENTS; i++)
        (void)hipEventDestroy(q->ev[i]);
    (void)hipStreamDestroy(q->stream);
    delete q;
}

extern "C" int p3mk_queue_sync(p3mk_queue *q) {
    hipError_t e = hipStreamSynchronize(q->stream);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipStreamSynchronize", (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_record(p3mk_queue *q, int slot) {
    if (slot < 0 || slot >= P3MK_NUM_EVENTS)
        return p3mk_fail(P3MK_EINVAL, "event slot out of range", 0);
    hipError_t e = hipEventRecord(q->ev[slot], q->stream);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipEventRecord", (int)e);
    if (q->ring && q->ring_pos < q->ring_sets) {
        e = hipEventRecord(q->ring[q->ring_pos * P3MK_NUM_EVENTS + slot], q->stream);
        if (e != hipSuccess)
            return p3mk_fail(P3MK_EDEVICE, "hipEventRecord", (int)e);
    }
    return P3MK_OK;
}

extern "C" int p3mk_ring_start(p3mk_queue *q, int nsets) {
    if (nsets < 0)
        return p3mk_fail(P3MK_EINVAL, "negative ring size", 0);
    hipError_t e = hipStreamSynchronize(q->stream);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipStreamSynchronize", (int)e);
    ring_free(q);
    if (nsets == 0)
        return P3MK_OK;
    q->ring = new (std::nothrow) hipEvent_t[(size_t)nsets * P3MK_NUM_EVENTS];
    if (!q->ring)
        return p3mk_fail(P3MK_ENOMEM, "event ring allocation failed", 0);
    for (int i = 0; i < nsets * P3MK_NUM_EVENTS; i++) {
        e = hipEventCreate(&q->ring[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; j++)
                (void)hipEventDestroy(q->ring[j]);
            delete[] q->ring;
            q->ring = nullptr;
            return p3mk_fail(P3MK_EDEVICE, "hipEventCreate", (int)e);
        }
    }
    q->ring_sets = nsets;
    q->ring_pos = 0;
    return P3MK_OK;
}

extern "C" void p3mk_ring_advance(p3mk_queue *q) {
    if (q->ring && q->ring_pos < q->ring_sets)
        q->ring_pos++;
}

extern "C" int p3mk_ring_filled(p3mk_queue *q) { return q->ring ? q->ring_pos : 0; }

extern "C" int p3mk_ring_elapsed_ms(p3mk_queue *q, int set, int a, int b, float *ms) {
    if (!q->ring || set < 0 || set >= q->ring_pos || a < 0 || b < 0 || a >= P3MK_NUM_EVENTS ||
        b >= P3MK_NUM_EVENTS)
        return p3mk_fail(P3MK_EINVAL, "event ring slot out of range", 0);
    hipError_t e = hipEventElapsedTime(ms, q->ring[set * P3MK_NUM_EVENTS + a],
                                       q->ring[set * P3MK_NUM_EVENTS + b]);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipEventElapsedTime", (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_elapsed_ms(p3mk_queue *q, int a, int b, float *ms) {
    if (a < 0 || b < 0 || a >= P3MK_NUM_EVENTS || b >= P3MK_NUM_EVENTS)
        return p3mk_fail(P3MK_EINVAL, "event slot out of range", 0);
    hipError_t e = hipEventElapsedTime(ms, q->ev[a], q->ev[b]);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipEventElapsedTime", (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_alloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0)
        return P3MK_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return p3mk_fail(P3MK_ENOMEM, "hipMalloc", (int)e);
    }
    return P3MK_OK;
}

extern "C" void p3mk_release(void *p) {
    if (p)
        (void)hipFree(p);
}

extern "C" int p3mk_alloc_host(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0)
        return P3MK_OK;
    hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *p = nullptr;
        return p3mk_fail(P3MK_ENOMEM, "hipHostMalloc", (int)e);
    }
    return P3MK_OK;
}

extern "C" void p3mk_release_host(void *p) {
    if (p)
        (void)hipHostFree(p);
}

static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind k, p3mk_queue *q) {
    if (bytes == 0)
        return P3MK_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, k, q->stream);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipMemcpyAsync", (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_copy_h2d(void *dst, const void *src, size_t bytes, p3mk_queue *q) {
    return copy(dst, src, bytes, hipMemcpyHostToDevice, q);
}
extern "C" int p3mk_copy_d2h(void *dst, const void *src, size_t bytes, p3mk_queue *q) {
    return copy(dst, src, bytes, hipMemcpyDeviceToHost, q);
}
extern "C" int p3mk_copy_d2d(void *dst, const void *src, size_t bytes, p3mk_queue *q) {
    return copy(dst, src, bytes, hipMemcpyDeviceToDevice, q);
}

__global__ void k_fill_zero16(uint4 *p, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (; i < n16; i += stride)
        p[i] = z;
}

__global__ void k_fill_zero1(unsigned char *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

extern "C" int p3mk_fill_zero(void *p, size_t bytes, p3mk_queue *q) {
    if (bytes == 0)
        return P3MK_OK;
    // hipMalloc returns 256-byte aligned memory; all callers pass such
    // pointers, so the 16-byte body below covers bytes - bytes % 16.
    const size_t n16 = bytes / 16, rem = bytes % 16;
    if (n16) {
        size_t blocks = (n16 + 255) / 256;
        if (blocks > 8192)
            blocks = 8192;
        hipLaunchKernelGGL(k_fill_zero16, dim3((unsigned)blocks), dim3(256), 0, q->stream,
                           (uint4 *)p, n16);
    }
    if (rem)
        hipLaunchKernelGGL(k_fill_zero1, dim3(1), dim3(64), 0, q->stream,
                           (unsigned char *)p + 16 * n16, rem);
    return p3mk_launch_check("fill_zero");
}

__global__ void k_axpy(size_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        y[i] += a * x[i];
}

extern "C" int p3mk_axpy(size_t n, double a, const double *x, double *y, p3mk_queue *q) {
    if (n == 0)
        return P3MK_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)blocks), dim3(256), 0, q->stream, n, a, x, y);
    return p3mk_launch_check("axpy");
}

// f[i] += fk[i] + fr[i]: the sum of Calculate_forces in its order
// (common.c:315-319), so the total force rounds as the reference's does.
__global__ void k_sum_parts(size_t n, const double *__restrict__ fk, const double *__restrict__ fr,
                            double *__restrict__ f) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        f[i] += fk[i] + fr[i];
}

extern "C" int p3mk_sum_parts(size_t n, const double *fk, const double *fr, double *f, p3mk_queue *q) {
    if (n == 0)
        return P3MK_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)blocks), dim3(256), 0, q->stream, n, fk, fr, f);
    return p3mk_launch_check("sum_parts");
}

__global__ void k_check_positions(int n, double lo, double hi, const double *__restrict__ x,
                                  const double *__restrict__ y, const double *__restrict__ z,
                                  int *bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int b = 0;
    if (i < n) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; d++)
            b |= !(v[d] >= lo && v[d] < hi);
    }
    // one atomic per wavefront
    unsigned long long m = __ballot(b);
    if ((threadIdx.x & 63) == 0 && m)
        atomicAdd(bad, (int)__popcll(m));
}

extern "C" int p3mk_check_positions(int n, double lo, double hi, const double *x, const double *y,
                                    const double *z, int *bad, int *scratch, p3mk_queue *q) {
    *bad = 0;
    if (n <= 0)
        return P3MK_OK;
    int rc = p3mk_fill_zero(scratch, sizeof(int), q);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_check_positions, dim3((n + 255) / 256), dim3(256), 0, q->stream, n, lo,
                       hi, x, y, z, scratch);
    rc = p3mk_launch_check("check_positions");
    if (rc)
        return rc;
    rc = p3mk_copy_d2h(bad, scratch, sizeof(int), q);
    if (rc)
        return rc;
    return p3mk_queue_sync(q);
}

struct p3mk_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" int p3mk_capture_begin(p3mk_queue *q) {
    hipError_t e = hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipStreamBeginCapture", (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_capture_end(p3mk_queue *q, p3mk_graph **out) {
    *out = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(q->stream, &g);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipStreamEndCapture", (int)e);
    p3mk_graph *pg = new (std::nothrow) p3mk_graph;
    if (!pg) {
        (void)hipGraphDestroy(g);
        return p3mk_fail(P3MK_ENOMEM, "graph allocation", 0);
    }
    pg->graph = g;
    e = hipGraphInstantiate(&pg->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete pg;
        return p3mk_fail(P3MK_EDEVICE, "hipGraphInstantiate", (int)e);
    }
    *out = pg;
    return P3MK_OK;
}

extern "C" int p3mk_graph_launch(p3mk_graph *g, p3mk_queue *q) {
    hipError_t e = hipGraphLaunch(g->exec, q->stream);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipGraphLaunch", (int)e);
    return P3MK_OK;
}

extern "C" void p3mk_graph_destroy(p3mk_graph *g) {
    if (!g)
        return;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
}
