// Exclusive prefix sum of int counts, used by the counting sorts of the
// particle <-> mesh and real-space kernels.  Up to 4096 counts: one launch
// of one workgroup (runs of up to four counts per thread; longer serial runs
// lose to the three launches below).  Beyond: three launches, per-block
// scans of 1024 elements, a scan of the block sums, and the block offsets
// added.
#include <hip/hip_runtime.h>

#include "k_internal.hpp"

namespace {

constexpr int kBlock = 1024;

__device__ __forceinline__ int block_exclusive_scan(int v, int *sh, int &total) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        sh[wave] = x;
    __syncthreads();
    if (t < 64) {
        int w = (t < kBlock / 64) ? sh[t] : 0;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(w, off, 64);
            if (t >= off)
                w += y;
        }
        sh[64 + t] = w;
    }
    __syncthreads();
    total = sh[64 + kBlock / 64 - 1];
    const int before = (wave == 0) ? 0 : sh[64 + wave - 1];
    return before + x - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_blocks(const int *__restrict__ in, int m,
                                                        int *__restrict__ out,
                                                        int *__restrict__ sums) {
    __shared__ int sh[128];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int v = (i < m) ? in[i] : 0;
    int total;
    const int e = block_exclusive_scan(v, sh, total);
    if (i < m)
        out[i] = e;
    if (threadIdx.x == 0)
        sums[blockIdx.x] = total;
}

// single block: exclusive scan of nb block sums (nb <= 1024 * 1024 handled by
// a serial carry over 1024-chunks), total written to sums[nb]
__global__ void __launch_bounds__(kBlock) k_scan_sums(int *__restrict__ sums, int nb) {
    __shared__ int sh[128];
    int carry = 0;
    for (int base = 0; base < nb; base += kBlock) {
        const int i = base + threadIdx.x;
        const int v = (i < nb) ? sums[i] : 0;
        int total;
        const int e = block_exclusive_scan(v, sh, total);
        __syncthreads();
        if (i < nb)
            sums[i] = carry + e;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        sums[nb] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(int *__restrict__ out, int m,
                                                     const int *__restrict__ sums, int nb) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m)
        out[i] += sums[blockIdx.x];
    if (i == 0)
        out[m] = sums[nb];
}

// one workgroup: thread t sums its run of ceil(m / 1024) consecutive counts,
// the run totals are scanned across the workgroup, and each thread writes its
// run's prefixes; total to out[m].  Each thread reads its run before writing
// it, so in and out may be the same array.
constexpr long kSmallScan = 4 * kBlock;

__global__ void __launch_bounds__(kBlock) k_scan_small(const int *in, int m, int *out) {
    __shared__ int sh[128];
    const int k = (m + kBlock - 1) / kBlock;
    const int b = min(m, (int)threadIdx.x * k), e = min(m, b + k);
    int sum = 0;
    for (int i = b; i < e; i++)
        sum += in[i];
    int total;
    int run = block_exclusive_scan(sum, sh, total);
    for (int i = b; i < e; i++) {
        const int v = in[i];
        out[i] = run;
        run += v;
    }
    if (threadIdx.x == 0)
        out[m] = total;
}

}  // namespace

size_t p3mk_scan_scratch_ints(long m) { return (size_t)((m + kBlock - 1) / kBlock) + 1; }

int p3mk_exclusive_scan(const int *count, long m, int *start, int *scratch, p3mk_queue *q) {
    hipStream_t s = p3mk_stream(q);
    const int nb = (int)((m + kBlock - 1) / kBlock);
    if (m <= 0)
        return p3mk_fill_zero(start, sizeof(int), q);
    if (m <= kSmallScan) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kBlock), 0, s, count, (int)m, start);
        return p3mk_launch_check("scan");
    }
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kBlock), 0, s, count, (int)m, start, scratch);
    P3MK_TRY(p3mk_launch_check("scan blocks"));
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, scratch, nb);
    P3MK_TRY(p3mk_launch_check("scan sums"));
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlock), 0, s, start, (int)m, scratch, nb);
    return p3mk_launch_check("scan add");
}
