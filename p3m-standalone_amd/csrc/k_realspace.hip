// Real-space (short-range) Ewald force on a linked-cell list.
//
// Pair set and arithmetic of Realteil (realpart.c:109-152): every ordered
// pair i != j whose minimum-image distance r (dx -= round(dx/L) L per axis)
// satisfies r <= rcut contributes
//   f_i += q_i q_j (erfc(alpha r)/r + 2 alpha/sqrt(pi) exp(-alpha^2 r^2)) / r^2 * d
// and 0.5 q_i q_j erfc(alpha r)/r to the energy.  The cell side is >= rcut,
// so with rcut <= L/2 every such pair lies in one of the 27 neighbour cells.
//
// Particles are counting-sorted by cell; one thread owns one particle (in
// cell order, so neighbouring threads read the same neighbour cells) and
// writes its force once, without atomics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "device_common.hpp"
#include "k_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxCellsPerSide = 256;

__device__ __forceinline__ int cell_coord(double r, double L, int nc) {
    int c = (int)floor(r / L * nc);
    c %= nc;
    if (c < 0)
        c += nc;
    return c;
}

__global__ void k_cell_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double L, int nc,
                             int *__restrict__ cell_of, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int c = (cell_coord(x[i], L, nc) * nc + cell_coord(y[i], L, nc)) * nc +
                  cell_coord(z[i], L, nc);
    cell_of[i] = c;
    atomicAdd(&count[c], 1);
}

// Exclusive scan of count[0..m) into start[0..m], one workgroup of 1024.
__global__ void __launch_bounds__(1024) k_scan(const int *__restrict__ count, int m,
                                               int *__restrict__ start) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (m + 1023) / 1024;
    const int b = t * chunk, e = min(m, b + chunk);
    int s = 0;
    for (int k = b; k < e; k++)
        s += count[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = (t == 0) ? 0 : part[t - 1];
    for (int k = b; k < e; k++) {
        start[k] = run;
        run += count[k];
    }
    if (t == 1023)
        start[m] = part[1023];
}

__global__ void k_cell_scatter(int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const double *__restrict__ q,
                               const int *__restrict__ cell_of, const int *__restrict__ start,
                               int *__restrict__ cursor, int *__restrict__ perm,
                               double4 *__restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int c = cell_of[i];
    const int pos = start[c] + atomicAdd(&cursor[c], 1);
    perm[pos] = i;
    sorted[pos] = make_double4(x[i], y[i], z[i], q[i]);
}

// 2 alpha / sqrt(pi) with the reference's sqrt(pi) literal (realpart.c:120)
__device__ __forceinline__ double wupi() { return 1.77245385090551602729816748334; }

template <bool FP32>
__global__ void __launch_bounds__(kThreads)
k_realspace(int n, const double4 *__restrict__ sorted, const int *__restrict__ perm,
            const int *__restrict__ start, int nc, double L, double rcut, double alpha,
            double *__restrict__ fr, double *__restrict__ energy) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, en = 0.0;
    if (t < n) {
        const double4 pi = sorted[t];
        const double Li = 1.0 / L;
        const double two_a_spi = 2.0 * alpha / wupi();
        const int cx = cell_coord(pi.x, L, nc), cy = cell_coord(pi.y, L, nc),
                  cz = cell_coord(pi.z, L, nc);
        const int span = (nc >= 3) ? 1 : 0; // nc < 3: single cell (nc == 1) or 2 cells per side
        const int lo = (nc >= 3) ? -1 : 0, hi = (nc >= 3) ? 1 : nc - 1;
        for (int a = lo; a <= hi; a++) {
            const int ca = span ? (cx + a + nc) % nc : a;
            for (int b = lo; b <= hi; b++) {
                const int cb = span ? (cy + b + nc) % nc : b;
                for (int c = lo; c <= hi; c++) {
                    const int cc = span ? (cz + c + nc) % nc : c;
                    const int cell = (ca * nc + cb) * nc + cc;
                    const int jb = start[cell], je = start[cell + 1];
                    for (int j = jb; j < je; j++) {
                        if (j == t)
                            continue;
                        const double4 pj = sorted[j];
                        double dx, dy, dz, r2;
                        {
#pragma clang fp contract(off)
                            dx = pi.x - pj.x;
                            dx -= round(dx * Li) * L;
                            dy = pi.y - pj.y;
                            dy -= round(dy * Li) * L;
                            dz = pi.z - pj.z;
                            dz -= round(dz * Li) * L;
                            r2 = dx * dx + dy * dy + dz * dz;
                        }
                        if (FP32) {
                            if (r2 <= rcut * rcut) {
                                const float rf = sqrtf((float)r2);
                                const float ar = (float)alpha * rf;
                                const float ec = erfcf(ar);
                                const float fak = (float)(pi.w * pj.w) *
                                                  (ec / rf + (float)two_a_spi * expf(-ar * ar)) /
                                                  (float)r2;
                                f0 += (double)(fak * (float)dx);
                                f1 += (double)(fak * (float)dy);
                                f2 += (double)(fak * (float)dz);
                                en += 0.5 * pi.w * pj.w * (double)(ec / rf);
                            }
                        } else {
                            const double r = sqrt(r2);
                            if (r <= rcut) {
                                const double ar = alpha * r;
                                const double ec = erfc(ar);
                                const double fak =
                                    pi.w * pj.w * (ec / r + two_a_spi * exp(-ar * ar)) / (r * r);
                                f0 += fak * dx;
                                f1 += fak * dy;
                                f2 += fak * dz;
                                en += 0.5 * pi.w * pj.w * ec / r;
                            }
                        }
                    }
                }
            }
        }
        const int i = perm[t];
        fr[i] += f0;
        fr[(long)n + i] += f1;
        fr[2L * n + i] += f2;
    }
    if (energy != nullptr) {
        for (int off = 32; off > 0; off >>= 1)
            en += __shfl_down(en, off, 64);
        if ((threadIdx.x & 63) == 0 && en != 0.0)
            atomicAdd(energy, en);
    }
}

}  // namespace

struct p3mk_cells {
    int nmax;
    int nc;
    long ncell;
    double L, rcut;
    int *cell_of, *count, *start, *cursor, *perm;
    double4 *sorted;
};

extern "C" void p3mk_cells_destroy(p3mk_cells *c) {
    if (!c)
        return;
    p3mk_release(c->cell_of);
    p3mk_release(c->count);
    p3mk_release(c->start);
    p3mk_release(c->cursor);
    p3mk_release(c->perm);
    p3mk_release(c->sorted);
    delete c;
}

extern "C" int p3mk_cells_create(p3mk_cells **out, int n, double L, double rcut) {
    *out = nullptr;
    if (!(rcut > 0.0) || !(L > 0.0) || rcut > 0.5 * L)
        return p3mk_fail(P3MK_EINVAL, "real space needs 0 < rcut <= L/2", 0);
    p3mk_cells *c = new (std::nothrow) p3mk_cells();
    if (!c)
        return p3mk_fail(P3MK_ENOMEM, "cell list", 0);
    c->nmax = n;
    int nc = (int)std::floor(L / rcut);
    if (nc < 1)
        nc = 1;
    if (nc > kMaxCellsPerSide)
        nc = kMaxCellsPerSide;
    c->nc = nc;
    c->ncell = (long)nc * nc * nc;
    c->L = L;
    c->rcut = rcut;
    int rc = P3MK_OK;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->cell_of, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->count, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->start, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->cursor, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->perm, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->sorted, sizeof(double4) * nn);
    if (rc != P3MK_OK) {
        p3mk_cells_destroy(c);
        return rc;
    }
    *out = c;
    return P3MK_OK;
}

extern "C" int p3mk_realspace(p3mk_cells *c, int n, double L, const double *x, const double *y,
                              const double *z, const double *q, double rcut, double alpha,
                              int fp32, double *fr, double *energy, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (n > c->nmax || L != c->L || rcut != c->rcut)
        return p3mk_fail(P3MK_EINVAL, "cell list was built for another system size, box or cutoff", 0);
    hipStream_t s = p3mk_stream(qu);
    const int nb = (n + kThreads - 1) / kThreads;
    P3MK_TRY(p3mk_fill_zero(c->count, sizeof(int) * (c->ncell + 1), qu));
    P3MK_TRY(p3mk_fill_zero(c->cursor, sizeof(int) * (c->ncell + 1), qu));
    hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(kThreads), 0, s, n, x, y, z, L, c->nc,
                       c->cell_of, c->count);
    P3MK_TRY(p3mk_launch_check("cell count"));
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, c->count, (int)c->ncell, c->start);
    P3MK_TRY(p3mk_launch_check("cell scan"));
    hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(kThreads), 0, s, n, x, y, z, q, c->cell_of,
                       c->start, c->cursor, c->perm, c->sorted);
    P3MK_TRY(p3mk_launch_check("cell scatter"));
    if (fp32)
        hipLaunchKernelGGL(k_realspace<true>, dim3(nb), dim3(kThreads), 0, s, n, c->sorted, c->perm,
                           c->start, c->nc, L, rcut, alpha, fr, energy);
    else
        hipLaunchKernelGGL(k_realspace<false>, dim3(nb), dim3(kThreads), 0, s, n, c->sorted,
                           c->perm, c->start, c->nc, L, rcut, alpha, fr, energy);
    return p3mk_launch_check("realspace");
}
