// Real-space (short-range) Ewald force on a linked-cell list.
//
// Pair set and arithmetic of Realteil (realpart.c:109-152): every ordered
// pair i != j whose minimum-image distance r satisfies r <= rcut contributes
//   f_i += q_i q_j (erfc(alpha r)/r + 2 alpha/sqrt(pi) exp(-alpha^2 r^2)) / r^2 * d
// and 0.5 q_i q_j erfc(alpha r)/r to the energy.
//
// Layout: particles are counting-sorted into cells of side >= rcut/R (R = 2
// when the box holds >= 5 cells per side), folded into [0, L).  A work item
// is (home cell, chunk of 64 home particles) with its (2R+1)^3 neighbour
// cells; a persistent grid of single-wavefront workgroups claims
// items from an atomic counter, so the dense cells of a clustered system are
// spread over the whole GPU.  Inside an item the 64 lanes own 64 home
// particles and walk the slab's neighbour cells in the same order:
// neighbours arrive in chunks of 64 with one coalesced load and are
// broadcast lane by lane (readlane), so the candidate loop touches no
// memory.  Periodic images are handled by a per-neighbour-cell shift of 0 or
// +-L, which equals the reference's round(dx/L)*L for every pair inside the
// cutoff (|dx| <= rcut <= L/2), so dx = (x_i - x_j) - shift is the
// reference's dx bit for bit.
//
// A cheap distance test accepts ~30% of the candidates; accepted neighbour
// indices go to a per-lane queue in LDS, and the expensive erfc/exp part
// runs when the queue is drained, so lanes stay busy instead of idling
// through other lanes' accepted pairs (all lanes of an item see the whole
// neighbourhood, so their acceptance rates and queue lengths are alike).
// Forces of an item are added to fr with FP64 atomics.

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "device_common.hpp"
#include "erfcx_coeffs.hpp"
#include "k_internal.hpp"

namespace {

constexpr int kWave = 64;
constexpr int kQueue = 2 * kWave; // flat queue: at most one round pending plus one append
constexpr int kMaxCellsPerSide = 512;

__device__ __forceinline__ double fold(double r, double L) {
    // into [0, L); identity for positions already inside the box
    if (r >= 0.0 && r < L)
        return r;
    double f = r - floor(r / L) * L;
    return (f >= L || f < 0.0) ? 0.0 : f;
}

__device__ __forceinline__ int cell_of(double rf, double inv_side, int nc) {
    int c = (int)(rf * inv_side);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void k_cell_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double L, double inv_side, int nc,
                             int *__restrict__ cell_id, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int c = (cell_of(fold(x[i], L), inv_side, nc) * nc + cell_of(fold(y[i], L), inv_side, nc)) * nc +
                  cell_of(fold(z[i], L), inv_side, nc);
    cell_id[i] = c;
    atomicAdd(&count[c], 1);
}

__global__ void k_cell_scatter(int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const double *__restrict__ q, double L,
                               const int *__restrict__ cell_id, const int *__restrict__ start,
                               int *__restrict__ cursor, int *__restrict__ perm,
                               double4 *__restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int c = cell_id[i];
    const int pos = start[c] + atomicAdd(&cursor[c], 1);
    perm[pos] = i;
    sorted[pos] = make_double4(fold(x[i], L), fold(y[i], L), fold(z[i], L), q[i]);
}

// work items of a cell: ceil(count / 64) chunks times nslab slabs
__global__ void k_cell_items(const int *__restrict__ count, long ncell, int nslab,
                             int *__restrict__ items) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell)
        items[c] = ((count[c] + 63) / 64) * nslab;
}

// value of lane k (uniform) in every lane
__device__ __forceinline__ double bcast(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float bcastf(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1)
        v = max(v, __shfl_xor(v, off, kWave));
    return v;
}

// 2 alpha / sqrt(pi) with the reference's sqrt(pi) literal (realpart.c:120)
__device__ __forceinline__ double sqrt_pi() { return 1.77245385090551602729816748334; }

struct PairAcc {
    double f0, f1, f2, en;
};

// erfcx(x) = exp(x^2) erfc(x), x >= 0, from the generated piecewise tables
__device__ __forceinline__ double erfcx_d(double x) {
    const double y = 1.0 / (1.0 + 0.5 * x);
    const int K = p3m::kErfcxDoublePieces;
    int k = (int)(y * K);
    k = k > K - 1 ? K - 1 : k;
    const double t = (y * K - (double)k) - 0.5;
    const double *c = p3m::kErfcxDoubleCoef[k];
    double acc = c[p3m::kErfcxDoubleDegree];
#pragma unroll
    for (int d = p3m::kErfcxDoubleDegree - 1; d >= 0; d--)
        acc = fma(acc, t, c[d]);
    return acc;
}

__device__ __forceinline__ float erfcx_f(float x) {
    const float y = 1.0f / (1.0f + 0.5f * x);
    const int K = p3m::kErfcxFloatPieces;
    int k = (int)(y * K);
    k = k > K - 1 ? K - 1 : k;
    const float t = (y * K - (float)k) - 0.5f;
    const float *c = p3m::kErfcxFloatCoef[k];
    float acc = c[p3m::kErfcxFloatDegree];
#pragma unroll
    for (int d = p3m::kErfcxFloatDegree - 1; d >= 0; d--)
        acc = fmaf(acc, t, c[d]);
    return acc;
}

// Full pair evaluation with the reference's inclusion test r <= rcut
// (realpart.c:134-148).  erfc(x) = exp(-x^2) erfcx(x) shares the one
// exponential with the exp(-x^2) term of the force; x^2 is split exactly
// (hi + lo by FMA) so exp(-x^2) keeps its accuracy for large x.
template <bool FP32>
__device__ __forceinline__ void pair(const double4 &pi, const double4 &pj, double sx, double sy,
                                     double sz, double rcut, double alpha, double two_a_spi,
                                     PairAcc &acc) {
    double dx, dy, dz, r2;
    {
#pragma clang fp contract(off)
        dx = (pi.x - pj.x) - sx;
        dy = (pi.y - pj.y) - sy;
        dz = (pi.z - pj.z) - sz;
        r2 = dx * dx + dy * dy + dz * dz;
    }
    // r <= rcut decided exactly (correctly rounded sqrt) only near the edge
    const double rc2 = rcut * rcut;
    bool inside = r2 < rc2 * (1.0 - 1e-12);
    if (!inside && r2 <= rc2 * (1.0 + 1e-12))
        inside = sqrt(r2) <= rcut;
    if (!inside)
        return;
    const double qq = pi.w * pj.w;
    if (FP32) {
        const float rinv = rsqrtf((float)r2);
        const float x = (float)alpha * ((float)r2 * rinv);
        const float e = expf(-x * x);
        const float ec = e * erfcx_f(x);
        const float fak = (float)qq * (ec * rinv + (float)two_a_spi * e) * (rinv * rinv);
        acc.f0 += (double)(fak * (float)dx);
        acc.f1 += (double)(fak * (float)dy);
        acc.f2 += (double)(fak * (float)dz);
        acc.en += 0.5 * qq * (double)(ec * rinv);
    } else {
        const double rinv = rsqrt(r2);
        const double x = alpha * (r2 * rinv);
        const double x2 = x * x;
        const double x2lo = fma(x, x, -x2);
        const double e0 = exp(-x2);
        const double e = fma(-e0, x2lo, e0); // exp(-(x2 + x2lo))
        const double ec = e * erfcx_d(x);
        const double fak = qq * (ec * rinv + two_a_spi * e) * (rinv * rinv);
        acc.f0 += fak * dx;
        acc.f1 += fak * dy;
        acc.f2 += fak * dz;
        acc.en += 0.5 * qq * ec * rinv;
    }
}

// Persistent kernel over work items (home cell, chunk of 64 home
// particles).  Accepted candidates go to one flat queue per wavefront:
// lanes that accept the same neighbour append at count + (their rank among
// the accepting lanes), so entries are dense; every drain round hands 64
// entries to the 64 lanes, whoever owns them, and the pair forces are added
// to the owners' accumulators in LDS.  An entry packs the neighbour's sorted
// index (bits 5..31) and its periodic-shift index (bits 0..4: 0..26 for
// -1/0/+1 per axis); the owner lane is stored beside it.
template <bool FP32, bool SMALL_BOX>
__global__ void __launch_bounds__(kWave)
k_realspace(int n, const double4 *__restrict__ sorted, const int *__restrict__ perm,
            const int *__restrict__ start, const int *__restrict__ istart, long ncell, int nc,
            int reach, double L, double rcut, double alpha, int *__restrict__ counter,
            double *__restrict__ fr, double *__restrict__ energy) {
    __shared__ unsigned int qe[kQueue];
    __shared__ unsigned char qo[kQueue];
    __shared__ double4 home_pos[kWave];
    __shared__ double facc[3][kWave];
    const int lane = threadIdx.x;
    const int total = istart[ncell];
    const double two_a_spi = 2.0 * alpha / sqrt_pi();
    const double Li = 1.0 / L, side = L / nc;
    // prefilter thresholds; the exact test runs in pair()
    const double rc2 = rcut * rcut * (1.0 + 1e-12);
    const float rc2f = (float)(rcut * rcut) * (1.0f + 1e-5f);
    const int lo = SMALL_BOX ? 0 : -reach, hi = SMALL_BOX ? nc - 1 : reach;
    double en_tot = 0.0;

    for (;;) {
        int item = 0;
        if (lane == 0)
            item = atomicAdd(counter, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= total)
            break;
        // item -> home cell (binary search in the item prefix sums)
        long clo = 0, chi = ncell;
        while (chi - clo > 1) {
            const long mid = (clo + chi) / 2;
            if (istart[mid] <= item)
                clo = mid;
            else
                chi = mid;
        }
        const int home = (int)clo;
        const int chunk = item - istart[home];
        const int hb = start[home], he = start[home + 1];
        const int hz = home % nc, hy = (home / nc) % nc, hx = home / (nc * nc);
        const int i = hb + chunk * kWave + lane;
        const bool active = i < he;
        const double4 pi = active ? sorted[i] : make_double4(0.0, 0.0, 0.0, 0.0);
        // home-cell-relative single-precision coordinates for the prefilter
        const double ox = hx * side, oy = hy * side, oz = hz * side;
        const float pxf = (float)(pi.x - ox), pyf = (float)(pi.y - oy), pzf = (float)(pi.z - oz);
        home_pos[lane] = pi;
        facc[0][lane] = 0.0;
        facc[1][lane] = 0.0;
        facc[2][lane] = 0.0;
        int count = 0; // wave-uniform queue length
        __syncthreads();

        // one round: lane l takes entry l (if l < m), evaluates it and adds
        // the force to the owner's accumulator
        auto drain_round = [&](int m) {
            if (lane < m) {
                const unsigned int e = qe[lane];
                const int o = qo[lane];
                const int j = (int)(e >> 5), sh = (int)(e & 31u);
                const double sx = (double)(sh / 9 - 1) * L, sy = (double)((sh / 3) % 3 - 1) * L,
                             sz = (double)(sh % 3 - 1) * L;
                PairAcc pa = {0.0, 0.0, 0.0, 0.0};
                pair<FP32>(home_pos[o], sorted[j], sx, sy, sz, rcut, alpha, two_a_spi, pa);
                if (pa.f0 != 0.0 || pa.f1 != 0.0 || pa.f2 != 0.0) {
                    atomicAdd(&facc[0][o], pa.f0);
                    atomicAdd(&facc[1][o], pa.f1);
                    atomicAdd(&facc[2][o], pa.f2);
                }
                en_tot += pa.en;
            }
            __syncthreads();
            // move the rest (< 64 entries) to the front
            const int rest = count - m;
            unsigned int e2 = 0;
            unsigned char o2 = 0;
            if (lane < rest) {
                e2 = qe[m + lane];
                o2 = qo[m + lane];
            }
            __syncthreads();
            if (lane < rest) {
                qe[lane] = e2;
                qo[lane] = o2;
            }
            count = rest;
            __syncthreads();
        };

        auto append = [&](bool ok, unsigned int entry) {
            const unsigned long long mask = __ballot(ok);
            if (mask) {
                const unsigned int rank = __builtin_amdgcn_mbcnt_hi(
                    (unsigned int)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
                if (ok) {
                    qe[count + rank] = entry;
                    qo[count + rank] = (unsigned char)lane;
                }
                count += __popcll(mask);
                if (count >= kWave) {
                    __syncthreads();
                    drain_round(kWave);
                }
            }
        };

        for (int a = lo; a <= hi; a++) {
            int cx = SMALL_BOX ? a : hx + a, wx = 0;
            if (cx < 0) { cx += nc; wx = -1; } else if (cx >= nc) { cx -= nc; wx = 1; }
            for (int b = lo; b <= hi; b++) {
                int cy = SMALL_BOX ? b : hy + b, wy = 0;
                if (cy < 0) { cy += nc; wy = -1; } else if (cy >= nc) { cy -= nc; wy = 1; }
                for (int c = lo; c <= hi; c++) {
                    int cz = SMALL_BOX ? c : hz + c, wz = 0;
                    if (cz < 0) { cz += nc; wz = -1; } else if (cz >= nc) { cz -= nc; wz = 1; }
                    const int cell = (cx * nc + cy) * nc + cz;
                    const int jb = __builtin_amdgcn_readfirstlane(start[cell]);
                    const int je = __builtin_amdgcn_readfirstlane(start[cell + 1]);
                    const double sx = wx * L, sy = wy * L, sz = wz * L;
                    const unsigned int sidx = (unsigned)((wx + 1) * 9 + (wy + 1) * 3 + (wz + 1));
                    // neighbours in chunks of 64: one coalesced load, then
                    // lane k's particle is broadcast to all lanes
                    for (int cb = jb; cb < je; cb += kWave) {
                        const int jj = cb + lane;
                        const double4 mine = (jj < je) ? sorted[jj] : make_double4(0.0, 0.0, 0.0, 0.0);
                        const int cnt = min(kWave, je - cb);
                        if (SMALL_BOX) {
                            for (int k = 0; k < cnt; k++) {
                                const int j = cb + k;
                                double dx = pi.x - bcast(mine.x, k), dy = pi.y - bcast(mine.y, k),
                                       dz = pi.z - bcast(mine.z, k);
                                // nearest image per axis (the reference's round(dx/L))
                                const double rx = round(dx * Li), ry = round(dy * Li), rz = round(dz * Li);
                                dx -= rx * L;
                                dy -= ry * L;
                                dz -= rz * L;
                                const unsigned int sh =
                                    (unsigned)(((int)rx + 1) * 9 + ((int)ry + 1) * 3 + ((int)rz + 1));
                                const double r2 = dx * dx + dy * dy + dz * dz;
                                append(active && j != i && r2 <= rc2, ((unsigned)j << 5) | sh);
                            }
                        } else {
                            // neighbour in the home cell's frame, image shift included
                            const float mxf = (float)((mine.x + sx) - ox), myf = (float)((mine.y + sy) - oy),
                                        mzf = (float)((mine.z + sz) - oz);
                            for (int k = 0; k < cnt; k++) {
                                const int j = cb + k;
                                const float dx = pxf - bcastf(mxf, k);
                                const float dy = pyf - bcastf(myf, k);
                                const float dz = pzf - bcastf(mzf, k);
                                const float r2 = dx * dx + dy * dy + dz * dz;
                                append(active && j != i && r2 <= rc2f, ((unsigned)j << 5) | sidx);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (count > 0)
            drain_round(count);
        if (active) {
            const int p = perm[i];
            atomicAdd(&fr[p], facc[0][lane]);
            atomicAdd(&fr[(long)n + p], facc[1][lane]);
            atomicAdd(&fr[2L * n + p], facc[2][lane]);
        }
        __syncthreads();
    }
    if (energy != nullptr) {
        for (int off = 32; off > 0; off >>= 1)
            en_tot += __shfl_down(en_tot, off, kWave);
        if (lane == 0 && en_tot != 0.0)
            atomicAdd(energy, en_tot);
    }
}

}  // namespace

struct p3mk_cells {
    int nmax;
    int nc, reach;
    long ncell;
    double L, rcut, inv_side;
    int *cell_id, *count, *start, *cursor, *perm, *scan_scratch, *items, *istart, *counter;
    double4 *sorted;
    int grid;
};

extern "C" void p3mk_cells_destroy(p3mk_cells *c) {
    if (!c)
        return;
    p3mk_release(c->cell_id);
    p3mk_release(c->count);
    p3mk_release(c->start);
    p3mk_release(c->cursor);
    p3mk_release(c->perm);
    p3mk_release(c->scan_scratch);
    p3mk_release(c->items);
    p3mk_release(c->istart);
    p3mk_release(c->counter);
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
    // cells of side >= rcut/reach (with a margin, so that rounding of
    // x*nc/L never puts a pair inside the cutoff more than `reach` cells apart)
    int reach = 2;
    int nc = (int)std::floor(L * reach / rcut * (1.0 - 1e-12));
    if (nc < 2 * reach + 1) {
        reach = 1;
        nc = (int)std::floor(L / rcut * (1.0 - 1e-12));
    }
    if (nc < 1)
        nc = 1;
    if (nc > kMaxCellsPerSide)
        nc = kMaxCellsPerSide;
    c->nc = nc;
    c->reach = reach;
    c->ncell = (long)nc * nc * nc;
    c->L = L;
    c->rcut = rcut;
    c->inv_side = (double)nc / L;
    int rc = P3MK_OK;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->cell_id, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->count, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->start, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->cursor, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->perm, sizeof(int) * nn);
    if (rc == P3MK_OK)
        rc = p3mk_alloc((void **)&c->scan_scratch, sizeof(int) * p3mk_scan_scratch_ints(c->ncell));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->items, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->istart, sizeof(int) * (c->ncell + 1));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&c->counter, sizeof(int) * 4);
    if (rc == P3MK_OK) {
        // persistent grid: enough single-wave workgroups to fill every CU
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        c->grid = (cus > 0 ? cus : 256) * 32;
    }
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
    const int nb = (n + 255) / 256;
    P3MK_TRY(p3mk_fill_zero(c->count, sizeof(int) * (c->ncell + 1), qu));
    P3MK_TRY(p3mk_fill_zero(c->cursor, sizeof(int) * (c->ncell + 1), qu));
    hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, s, n, x, y, z, L, c->inv_side, c->nc,
                       c->cell_id, c->count);
    P3MK_TRY(p3mk_launch_check("cell count"));
    P3MK_TRY(p3mk_exclusive_scan(c->count, c->ncell, c->start, c->scan_scratch, qu));
    hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(256), 0, s, n, x, y, z, q, L, c->cell_id,
                       c->start, c->cursor, c->perm, c->sorted);
    P3MK_TRY(p3mk_launch_check("cell scatter"));
    const bool all_cells = (2 * c->reach + 1 > c->nc);
    const int nslab = 1; // one item walks the whole neighbourhood
    hipLaunchKernelGGL(k_cell_items, dim3((unsigned)((c->ncell + 255) / 256)), dim3(256), 0, s,
                       c->count, c->ncell, nslab, c->items);
    P3MK_TRY(p3mk_launch_check("cell items"));
    P3MK_TRY(p3mk_exclusive_scan(c->items, c->ncell, c->istart, c->scan_scratch, qu));
    P3MK_TRY(p3mk_fill_zero(c->counter, sizeof(int) * 4, qu));
    const dim3 grid((unsigned)c->grid);
#define LAUNCH(F32, SB)                                                                            \
    hipLaunchKernelGGL((k_realspace<F32, SB>), grid, dim3(kWave), 0, s, n, c->sorted, c->perm,     \
                       c->start, c->istart, c->ncell, c->nc, c->reach, L, rcut, alpha, c->counter, \
                       fr, energy)
    if (fp32) {
        if (all_cells) LAUNCH(true, true); else LAUNCH(true, false);
    } else {
        if (all_cells) LAUNCH(false, true); else LAUNCH(false, false);
    }
#undef LAUNCH
    return p3mk_launch_check("realspace");
}
