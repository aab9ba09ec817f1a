// Real-space (short-range) Ewald force on a linked-cell list.
//
// Pair set and arithmetic of Realteil (realpart.c:124-167): every ordered
// pair i != j whose minimum-image distance r satisfies r <= rcut contributes
//   f_i += q_i q_j (erfc(alpha r)/r + 2 alpha/sqrt(pi) exp(-alpha^2 r^2)) / r^2 * d
// and 0.5 q_i q_j erfc(alpha r)/r to the energy.
//
// Layout: particles are folded into [0, L) and counting-sorted into cells of
// side >= rcut/R (R = 2 when the box holds >= 5 cells per side), and inside a
// cell by octant, so that 64 consecutive particles of a cell are spatially
// compact.  A work item is (home cell, chunk of 64 home particles); a
// persistent grid of single-wavefront workgroups claims items from an
// atomic counter, so the dense cells of a clustered system are spread over
// the whole GPU.  Every item has a bounding box (computed after the sort),
// and neighbour cells and neighbour chunks whose box is beyond rcut from the
// item's box are skipped.  Periodic images are a per-neighbour-cell shift
// of 0 or +-L, which equals the reference's round(dx/L)*L for every pair
// inside the cutoff (rcut <= L/2), so dx = (x_i - x_j) - shift is the
// reference's dx bit for bit.  (Newton's third law was tried and measured
// slower here: the partner forces need global atomics per neighbour chunk.)

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "batch_desc.hpp"
#include "device_common.hpp"
#include "erfcx_coeffs.hpp"
#include "k_internal.hpp"

namespace {

using p3m::CellArrays;

constexpr int kWave = 64;
constexpr int kQueue = 2 * kWave; // ring: at most one round pending plus one append
constexpr int kMaxCellsPerSide = 512;

__device__ __forceinline__ double fold(double r, double L) {
    // into [0, L); identity for positions already inside the box
    if (r >= 0.0 && r < L)
        return r;
    double f = r - floor(r / L) * L;
    return (f >= L || f < 0.0) ? 0.0 : f;
}

// cell * 8 + octant of the folded position inside its cell
__device__ __forceinline__ int fine_key(double xf, double yf, double zf, double inv_side, int nc) {
    const double ux = xf * inv_side, uy = yf * inv_side, uz = zf * inv_side;
    int cx = (int)ux, cy = (int)uy, cz = (int)uz;
    cx = cx < 0 ? 0 : (cx >= nc ? nc - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= nc ? nc - 1 : cy);
    cz = cz < 0 ? 0 : (cz >= nc ? nc - 1 : cz);
    const int oct = ((ux - cx >= 0.5) ? 4 : 0) | ((uy - cy >= 0.5) ? 2 : 0) | ((uz - cz >= 0.5) ? 1 : 0);
    return ((cx * nc + cy) * nc + cz) * 8 + oct;
}

__global__ void k_cell_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double L, double inv_side, int nc,
                             int *__restrict__ key, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int k = fine_key(fold(x[i], L), fold(y[i], L), fold(z[i], L), inv_side, nc);
    key[i] = k;
    atomicAdd(&count[k], 1);
}

__global__ void k_cell_scatter(int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const double *__restrict__ q, double L,
                               const int *__restrict__ key, const int *__restrict__ start,
                               int *__restrict__ cursor, int *__restrict__ perm,
                               double4 *__restrict__ sorted, float4 *__restrict__ sorted32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int k = key[i];
    const int pos = start[k] + atomicAdd(&cursor[k], 1);
    perm[pos] = i;
    const double4 r = make_double4(fold(x[i], L), fold(y[i], L), fold(z[i], L), q[i]);
    sorted[pos] = r;
    sorted32[pos] = make_float4((float)r.x, (float)r.y, (float)r.z, 0.0f);
}

// cell starts from the octant starts; work items per cell: ceil(count / 64)
__global__ void k_cell_start_items(const int *__restrict__ fstart, long ncell, int *__restrict__ start,
                                   int *__restrict__ items) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell)
        return;
    start[c] = fstart[8 * c];
    if (c < ncell)
        items[c] = (fstart[8 * (c + 1)] - fstart[8 * c] + 63) / 64;
}

// last index k with prefix[k] <= v
__device__ __forceinline__ long prefix_find(const int *__restrict__ prefix, long count, int v) {
    long lo = 0, hi = count;
    while (hi - lo > 1) {
        const long mid = (lo + hi) / 2;
        if (prefix[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float down(double v) {
    float f = (float)v;
    return ((double)f > v) ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float up(double v) {
    float f = (float)v;
    return ((double)f < v) ? nextafterf(f, INFINITY) : f;
}

// bounding box of every work item's particles (outward-rounded FP32):
// bbox[2 t] = (xlo, ylo, zlo, 0), bbox[2 t + 1] = (xhi, yhi, zhi, 0)
__global__ void k_item_bbox(const int *__restrict__ start, const int *__restrict__ istart, long ncell,
                            long max_items, const double4 *__restrict__ sorted,
                            float4 *__restrict__ bbox) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= max_items || t >= istart[ncell])
        return;
    const long c = prefix_find(istart, ncell, (int)t);
    const int p0 = start[c] + (int)(t - istart[c]) * kWave;
    const int p1 = min(start[c + 1], p0 + kWave);
    double xl = 1e300, yl = 1e300, zl = 1e300, xh = -1e300, yh = -1e300, zh = -1e300;
    for (int p = p0; p < p1; p++) {
        const double4 r = sorted[p];
        xl = fmin(xl, r.x); yl = fmin(yl, r.y); zl = fmin(zl, r.z);
        xh = fmax(xh, r.x); yh = fmax(yh, r.y); zh = fmax(zh, r.z);
    }
    bbox[2 * t] = make_float4(down(xl), down(yl), down(zl), 0.0f);
    bbox[2 * t + 1] = make_float4(up(xh), up(yh), up(zh), 0.0f);
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

constexpr int kErfcxTab = p3m::kErfcxDoublePieces * (p3m::kErfcxDoubleDegree + 1);

// erfcx(x) = exp(x^2) erfc(x), x >= 0, from the generated piecewise FP64
// tables (a copy in LDS: the piece index differs from lane to lane)
__device__ __forceinline__ double erfcx_d(double x, const double *tab) {
    const double y = 1.0 / (1.0 + 0.5 * x);
    const int K = p3m::kErfcxDoublePieces;
    int k = (int)(y * K);
    k = k > K - 1 ? K - 1 : k;
    const double t = (y * K - (double)k) - 0.5;
    const double *c = tab + k * (p3m::kErfcxDoubleDegree + 1);
    double acc = c[p3m::kErfcxDoubleDegree];
#pragma unroll
    for (int d = p3m::kErfcxDoubleDegree - 1; d >= 0; d--)
        acc = fma(acc, t, c[d]);
    return acc;
}

// FP32 erfcx: one generated polynomial on [0, 6] (x clamped beyond, where
// erfc < 3e-17 and the pair's contribution vanishes)
__device__ __forceinline__ float erfcx_f(float x) {
    x = fminf(x, p3m::kErfcxSingleX1);
    const float y = __builtin_amdgcn_rcpf(1.0f + 0.5f * x);
    const float t = fmaf(y, p3m::kErfcxSingleA, p3m::kErfcxSingleB);
    float acc = p3m::kErfcxSingleCoef[p3m::kErfcxSingleDegree];
#pragma unroll
    for (int d = p3m::kErfcxSingleDegree - 1; d >= 0; d--)
        acc = fmaf(acc, t, p3m::kErfcxSingleCoef[d]);
    return acc;
}

// Pair evaluation with the reference's inclusion test r <= rcut
// (realpart.c:134-148); dx is the reference's minimum-image difference
// bit for bit.  erfc(x) = exp(-x^2) erfcx(x) shares the one exponential with
// the exp(-x^2) term of the force; in FP64 x^2 is split exactly (hi + lo by
// FMA) so exp(-x^2) keeps its accuracy for large x.
template <bool FP32>
__device__ __forceinline__ void pair(const double4 &pi, const double4 &pj, double sx, double sy,
                                     double sz, double rcut, double alpha, double two_a_spi,
                                     const double *tab, PairAcc &acc) {
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
    const bool edge = !inside && r2 <= rc2 * (1.0 + 1e-12);
    if (__builtin_expect(edge, 0))
        inside = sqrt(r2) <= rcut;
    if (!inside)
        return;
    const double qq = pi.w * pj.w;
    if (FP32) {
        const float r2f = (float)r2;
        const float rinv = __builtin_amdgcn_rsqf(r2f);
        const float x = (float)alpha * (r2f * rinv);
        const float e = __expf(-x * x);
        const float ec = e * erfcx_f(x);
        const float fak = (float)qq * (ec * rinv + (float)two_a_spi * e) * (rinv * rinv);
        acc.f0 = (double)(fak * (float)dx);
        acc.f1 = (double)(fak * (float)dy);
        acc.f2 = (double)(fak * (float)dz);
        acc.en = 0.5 * qq * (double)(ec * rinv);
    } else {
        const double rinv = rsqrt(r2);
        const double x = alpha * (r2 * rinv);
        const double x2 = x * x;
        const double x2lo = fma(x, x, -x2);
        const double e0 = exp(-x2);
        const double e = fma(-e0, x2lo, e0); // exp(-(x2 + x2lo))
        const double ec = e * erfcx_d(x, tab);
        const double fak = qq * (ec * rinv + two_a_spi * e) * (rinv * rinv);
        acc.f0 = fak * dx;
        acc.f1 = fak * dy;
        acc.f2 = fak * dz;
        acc.en = 0.5 * qq * ec * rinv;
    }
}

// LDS of one wavefront's work item: the ring of queued pairs, the home
// particles, their force accumulators and, for FP64, the erfcx table.
template <bool FP32> struct RsShared {
    uint2 q[kQueue]; // neighbour index, owner | shift index << 6
    double4 shift[27]; // image shift (sx, sy, sz) of shift index (wx+1)*9 + (wy+1)*3 + wz+1
    double4 home_pos[kWave];
    double facc[3][kWave];
    double tab[FP32 ? 1 : kErfcxTab];
};

template <bool FP32> __device__ __forceinline__ void load_erfcx_table(RsShared<FP32> &sh) {
    if (!FP32) {
        for (int k = threadIdx.x; k < kErfcxTab; k += blockDim.x)
            sh.tab[k] = (&p3m::kErfcxDoubleCoef[0][0])[k];
    }
    __syncthreads();
}

// queue entry: x = neighbour's sorted index; y = owner lane (bits 0..5) and
// the index of the periodic image shift (wx+1)*9 + (wy+1)*3 + (wz+1) above
__device__ __forceinline__ unsigned int pack_owner_shift(int lane, int wx, int wy, int wz) {
    return (unsigned)lane | ((unsigned)((wx + 1) * 9 + (wy + 1) * 3 + (wz + 1)) << 6);
}

// the 27 image shifts of a box of side L
template <bool FP32> __device__ __forceinline__ void fill_shifts(RsShared<FP32> &sh, double L) {
    if (threadIdx.x < 27) {
        const int t = threadIdx.x;
        sh.shift[t] = make_double4((t / 9 - 1) * L, ((t / 3) % 3 - 1) * L, (t % 3 - 1) * L, 0.0);
    }
}

// squared distance between two boxes (0 if they overlap)
__device__ __forceinline__ float box_dist2(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    const float dx = fmaxf(0.0f, fmaxf(blo.x - ahi.x, alo.x - bhi.x));
    const float dy = fmaxf(0.0f, fmaxf(blo.y - ahi.y, alo.y - bhi.y));
    const float dz = fmaxf(0.0f, fmaxf(blo.z - ahi.z, alo.z - bhi.z));
    return dx * dx + dy * dy + dz * dz;
}

// One work item (home cell, chunk of 64 home particles); every ordered pair
// is evaluated for its home particle.  Lanes own the 64 home particles and
// walk the neighbour cells together: (2 reach + 1)^3 cells around the home
// cell, or, in a box of fewer cells per side (SMALL_BOX), every cell once
// with the nearest image taken per pair.  Neighbour cells, and neighbour
// chunks of 64 particles, whose bounding box lies beyond rcut from the
// item's box are skipped.  The others arrive in chunks of 64 with one
// coalesced load and are broadcast lane by lane (readlane); a cheap FP32
// distance test in the home cell's frame accepts the candidates, and the
// accepted pairs of all lanes go to one ring queue per wavefront (rank by
// mbcnt over the ballot).  Every 64 queued pairs are evaluated by the 64
// lanes at once, whoever owns them, with the exact FP64 difference and
// inclusion test, and the forces are added to the owners' LDS accumulators.
// Cells of the item's system are cell0 + (cx nc + cy) nc + cz with items
// from istart[cell]; forces go to fr[d stride + perm[i]].
template <bool FP32, bool SMALL_BOX>
__device__ __forceinline__ void
realspace_item(int home, int chunk, int cell0, const double4 *__restrict__ sorted,
               const float4 *__restrict__ sorted32, const int *__restrict__ perm,
               const int *__restrict__ start, const int *__restrict__ istart,
               const float4 *__restrict__ bbox, int nc, int reach,
               double L, double rcut, double alpha, long stride, double *__restrict__ fr,
               RsShared<FP32> &sh, double &en_tot) {
    constexpr int kMask = kQueue - 1;
    const int lane = threadIdx.x;
    const double two_a_spi = 2.0 * alpha / sqrt_pi();
    const double Li = 1.0 / L, side = L / nc;
    // prefilter thresholds; the exact test runs in pair().  The FP32 test
    // works on absolute FP32 coordinates (error ~ 6e-8 L per coordinate)
    const double rc2 = rcut * rcut * (1.0 + 1e-12);
    const float rc2f = (float)(rcut * rcut * (1.0 + 1e-5 + 1e-6 * L / rcut));
    // boxes are compared in FP32: margin for their rounding
    const float rskip = (float)(rcut * (1.0 + 1e-5) + 1e-5 * L);
    const float rskip2 = rskip * rskip;
    const int lo = SMALL_BOX ? 0 : -reach, hi = SMALL_BOX ? nc - 1 : reach;
    const int hcell = cell0 + home;
    const int hb = start[hcell], he = start[hcell + 1];
    const int hz = home % nc, hy = (home / nc) % nc, hx = home / (nc * nc);
    const int i0 = hb + chunk * kWave; // sorted index of lane 0's home particle
    const int i = i0 + lane;
    const bool active = i < he;
    const double4 pi = active ? sorted[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int hitem = istart[hcell] + chunk;
    const float4 hlo = bbox[2 * hitem], hhi = bbox[2 * hitem + 1];
    // absolute single-precision coordinates for the prefilter; inactive
    // lanes are placed far away so that they never pass it
    const float hxf = active ? (float)pi.x : 1e30f, hyf = active ? (float)pi.y : 1e30f,
                hzf = active ? (float)pi.z : 1e30f;
    sh.home_pos[lane] = pi;
    sh.facc[0][lane] = 0.0;
    sh.facc[1][lane] = 0.0;
    sh.facc[2][lane] = 0.0;
    fill_shifts(sh, L);
    int head = 0, count = 0; // wave-uniform ring state
    __syncthreads();

    // lane l evaluates ring entry head + l (l < m) for its owner
    auto drain = [&](int m) {
        if (lane < m) {
            const uint2 e = sh.q[(head + lane) & kMask];
            const int o = (int)(e.y & 63u);
            const double4 s3 = sh.shift[e.y >> 6];
            PairAcc pa = {0.0, 0.0, 0.0, 0.0};
            pair<FP32>(sh.home_pos[o], sorted[e.x], s3.x, s3.y, s3.z, rcut, alpha, two_a_spi,
                       sh.tab, pa);
            atomicAdd(&sh.facc[0][o], pa.f0);
            atomicAdd(&sh.facc[1][o], pa.f1);
            atomicAdd(&sh.facc[2][o], pa.f2);
            en_tot += pa.en;
        }
        head = (head + m) & kMask;
        count -= m;
        __syncthreads();
    };

    // the accepting lanes (ok) queue the pair (own home particle, j)
    auto append = [&](bool ok, unsigned int j, unsigned int os) {
        const unsigned long long mask = __ballot(ok);
        if (mask) {
            if (ok) {
                const unsigned int rank = __builtin_amdgcn_mbcnt_hi(
                    (unsigned int)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
                sh.q[(head + count + (int)rank) & kMask] = make_uint2(j, os);
            }
            count += __popcll(mask);
            if (count >= kWave) {
                __syncthreads();
                drain(kWave);
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
                const double sx = wx * L, sy = wy * L, sz = wz * L;
                if (!SMALL_BOX) {
                    // the whole neighbour cell beyond the cutoff from the item's box?
                    const float4 clo = make_float4((float)(cx * side + sx), (float)(cy * side + sy),
                                                   (float)(cz * side + sz), 0.0f);
                    const float4 chi = make_float4((float)((cx + 1) * side + sx),
                                                   (float)((cy + 1) * side + sy),
                                                   (float)((cz + 1) * side + sz), 0.0f);
                    if (box_dist2(hlo, hhi, clo, chi) > rskip2)
                        continue;
                }
                const int cell = cell0 + (cx * nc + cy) * nc + cz;
                const int jb = __builtin_amdgcn_readfirstlane(start[cell]);
                const int je = __builtin_amdgcn_readfirstlane(start[cell + 1]);
                const int item0 = __builtin_amdgcn_readfirstlane(istart[cell]);
                const unsigned int os_cell = pack_owner_shift(lane, wx, wy, wz);
                const float fsx = (float)sx, fsy = (float)sy, fsz = (float)sz;
                // home particles relative to the neighbour cell's image
                const float pxs = hxf - fsx, pys = hyf - fsy, pzs = hzf - fsz;
                // neighbours in chunks of 64
                for (int cb = jb; cb < je; cb += kWave) {
                    const int cnt = min(kWave, je - cb);
                    if (!SMALL_BOX) {
                        const int nitem = item0 + (cb - jb) / kWave;
                        float4 nlo = bbox[2 * nitem], nhi = bbox[2 * nitem + 1];
                        nlo.x += fsx; nlo.y += fsy; nlo.z += fsz;
                        nhi.x += fsx; nhi.y += fsy; nhi.z += fsz;
                        if (box_dist2(hlo, hhi, nlo, nhi) > rskip2)
                            continue;
                    }
                    if (SMALL_BOX) {
                        // one coalesced load, then lane k's particle is broadcast
                        const int jj = cb + lane;
                        const double4 mine = (lane < cnt) ? sorted[jj] : make_double4(0.0, 0.0, 0.0, 0.0);
                        for (int k = 0; k < cnt; k++) {
                            const int j = cb + k;
                            double dx = pi.x - bcast(mine.x, k), dy = pi.y - bcast(mine.y, k),
                                   dz = pi.z - bcast(mine.z, k);
                            // nearest image per axis (the reference's round(dx/L))
                            const double rx = round(dx * Li), ry = round(dy * Li), rz = round(dz * Li);
                            dx -= rx * L;
                            dy -= ry * L;
                            dz -= rz * L;
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            append(active && j != i && r2 <= rc2, (unsigned)j,
                                   pack_owner_shift(lane, (int)rx, (int)ry, (int)rz));
                        }
                    } else if (cb == i0) {
                        // the item's own chunk: exclude the self pair
                        for (int k = 0; k < cnt; k++) {
                            const float4 v = sorted32[cb + k];
                            const float dx = pxs - v.x, dy = pys - v.y, dz = pzs - v.z;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            append(r2 <= rc2f && k != lane, (unsigned)(cb + k), os_cell);
                        }
                    } else {
                        // neighbours read with scalar (wave-uniform) loads, four
                        // at a time so that their latencies overlap
                        auto test = [&](const float4 &v, int k) {
                            const float dx = pxs - v.x, dy = pys - v.y, dz = pzs - v.z;
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            append(r2 <= rc2f, (unsigned)(cb + k), os_cell);
                        };
                        int k = 0;
                        for (; k + 4 <= cnt; k += 4) {
                            const float4 v0 = sorted32[cb + k], v1 = sorted32[cb + k + 1],
                                         v2 = sorted32[cb + k + 2], v3 = sorted32[cb + k + 3];
                            test(v0, k);
                            test(v1, k + 1);
                            test(v2, k + 2);
                            test(v3, k + 3);
                        }
                        for (; k < cnt; k++)
                            test(sorted32[cb + k], k);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (count > 0)
        drain(count);
    if (active) {
        const int p = perm[i];
        atomicAdd(&fr[p], sh.facc[0][lane]);
        atomicAdd(&fr[stride + p], sh.facc[1][lane]);
        atomicAdd(&fr[2 * stride + p], sh.facc[2][lane]);
    }
    __syncthreads();
}

// Persistent grid: each single-wavefront workgroup claims work items from an
// atomic counter until none are left.
template <bool FP32, bool SMALL_BOX>
__global__ void __launch_bounds__(kWave)
k_realspace(int n, const double4 *__restrict__ sorted, const float4 *__restrict__ sorted32,
            const int *__restrict__ perm,
            const int *__restrict__ start, const int *__restrict__ istart,
            const float4 *__restrict__ bbox, long ncell, int nc, int reach, double L, double rcut,
            double alpha, int *__restrict__ counter, double *__restrict__ fr,
            double *__restrict__ energy) {
    __shared__ RsShared<FP32> sh;
    const int total = istart[ncell];
    double en_tot = 0.0;
    load_erfcx_table(sh);
    for (;;) {
        int item = 0;
        if (threadIdx.x == 0)
            item = atomicAdd(counter, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= total)
            break;
        const int home = (int)prefix_find(istart, ncell, item);
        realspace_item<FP32, SMALL_BOX>(home, item - istart[home], 0, sorted, sorted32, perm, start,
                                        istart, bbox, nc, reach, L, rcut, alpha, n, fr, sh, en_tot);
    }
    if (energy != nullptr) {
        for (int off = 32; off > 0; off >>= 1)
            en_tot += __shfl_down(en_tot, off, kWave);
        if (threadIdx.x == 0 && en_tot != 0.0)
            atomicAdd(energy, en_tot);
    }
}

}  // namespace

// ---- cell-list arrays and launchers ---------------------------------------

namespace p3m {

void cells_release(CellArrays &a) {
    void *ptrs[] = {a.key, a.fcount, a.fstart, a.fcursor, a.start, a.items, a.istart, a.perm,
                    a.scan, a.counter, a.sorted, a.sorted32, a.bbox};
    for (void *p : ptrs)
        p3mk_release(p);
    a = CellArrays{};
}

int cells_alloc(CellArrays &a, long n, long ncell) {
    a = CellArrays{};
    const size_t nn = (size_t)(n > 0 ? n : 1), nf = (size_t)(8 * ncell + 1), nc1 = (size_t)ncell + 1;
    a.max_items = ncell + n / kWave + 1;
    int rc = P3MK_OK;
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.key, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.fcount, sizeof(int) * nf);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.fstart, sizeof(int) * nf);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.fcursor, sizeof(int) * nf);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.start, sizeof(int) * nc1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.items, sizeof(int) * nc1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.istart, sizeof(int) * nc1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.perm, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.scan, sizeof(int) * p3mk_scan_scratch_ints(8 * ncell));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.counter, sizeof(int) * 4);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.sorted, sizeof(double4) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.sorted32, sizeof(float4) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.bbox, sizeof(float4) * 2 * (size_t)a.max_items);
    if (rc != P3MK_OK)
        cells_release(a);
    return rc;
}

// zero the octant counters before the keys are counted
static int cells_begin(CellArrays &a, long ncell, p3mk_queue *qu) {
    P3MK_TRY(p3mk_fill_zero(a.fcount, sizeof(int) * (8 * ncell + 1), qu));
    return p3mk_fill_zero(a.fcursor, sizeof(int) * (8 * ncell + 1), qu);
}

// octant starts after the count (before the scatter)
static int cells_scan(CellArrays &a, long ncell, p3mk_queue *qu) {
    return p3mk_exclusive_scan(a.fcount, 8 * ncell, a.fstart, a.scan, qu);
}

// after the scatter: cell starts, work items, item boxes, counter reset
static int cells_end(CellArrays &a, long ncell, p3mk_queue *qu) {
    hipStream_t s = p3mk_stream(qu);
    hipLaunchKernelGGL(k_cell_start_items, dim3((unsigned)((ncell + 1 + 255) / 256)), dim3(256), 0, s,
                       a.fstart, ncell, a.start, a.items);
    P3MK_TRY(p3mk_launch_check("cell starts"));
    P3MK_TRY(p3mk_exclusive_scan(a.items, ncell, a.istart, a.scan, qu));
    hipLaunchKernelGGL(k_item_bbox, dim3((unsigned)((a.max_items + 255) / 256)), dim3(256), 0, s,
                       a.start, a.istart, ncell, a.max_items, a.sorted, a.bbox);
    P3MK_TRY(p3mk_launch_check("item boxes"));
    return p3mk_fill_zero(a.counter, sizeof(int) * 4, qu);
}

}  // namespace p3m

struct p3mk_cells {
    int nmax;
    int nc, reach;
    long ncell;
    double L, rcut, inv_side;
    CellArrays a;
    int grid;
};

extern "C" void p3mk_cells_destroy(p3mk_cells *c) {
    if (!c)
        return;
    p3m::cells_release(c->a);
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
    int rc = p3m::cells_alloc(c->a, n, c->ncell);
    if (rc == P3MK_OK) {
        // persistent grid: enough single-wave workgroups to fill every CU
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        c->grid = (cus > 0 ? cus : 256) * 32;
    }
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
    CellArrays &a = c->a;
    hipStream_t s = p3mk_stream(qu);
    const int nb = (n + 255) / 256;
    P3MK_TRY(p3m::cells_begin(a, c->ncell, qu));
    hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, s, n, x, y, z, L, c->inv_side, c->nc,
                       a.key, a.fcount);
    P3MK_TRY(p3mk_launch_check("cell count"));
    P3MK_TRY(p3m::cells_scan(a, c->ncell, qu));
    hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(256), 0, s, n, x, y, z, q, L, a.key, a.fstart,
                       a.fcursor, a.perm, a.sorted, a.sorted32);
    P3MK_TRY(p3mk_launch_check("cell scatter"));
    P3MK_TRY(p3m::cells_end(a, c->ncell, qu));
    const bool all_cells = (2 * c->reach + 1 > c->nc);
    const dim3 grid((unsigned)c->grid);
#define LAUNCH(F32, SB)                                                                            \
    hipLaunchKernelGGL((k_realspace<F32, SB>), grid, dim3(kWave), 0, s, n, a.sorted, a.sorted32,   \
                       a.perm,                                                                     \
                       a.start, a.istart, a.bbox, c->ncell, c->nc, c->reach, L, rcut, alpha,        \
                       a.counter, fr, energy)
    if (fp32) {
        if (all_cells) LAUNCH(true, true); else LAUNCH(true, false);
    } else {
        if (all_cells) LAUNCH(false, true); else LAUNCH(false, false);
    }
#undef LAUNCH
    return p3mk_launch_check("realspace");
}

// ---- batched real space over many systems -----------------------------------
//
// Cells of system s are global cells [cell0_s, cell0_s + nc_s^3); one
// persistent launch claims work items of all systems from one counter.

namespace {

using p3m::BSys;

__global__ void k_bcell_count(int ntot, const BSys *__restrict__ sys, const int *__restrict__ psys,
                              const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, int *__restrict__ key,
                              int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot)
        return;
    const BSys &S = sys[psys[i]];
    const double L = S.L;
    const int k = 8 * S.cell0 + fine_key(fold(x[i], L), fold(y[i], L), fold(z[i], L), (double)S.nc / L, S.nc);
    key[i] = k;
    atomicAdd(&count[k], 1);
}

__global__ void k_bcell_scatter(int ntot, const BSys *__restrict__ sys, const int *__restrict__ psys,
                                const double *__restrict__ x, const double *__restrict__ y,
                                const double *__restrict__ z, const double *__restrict__ q,
                                const int *__restrict__ key, const int *__restrict__ start,
                                int *__restrict__ cursor, int *__restrict__ perm,
                                double4 *__restrict__ sorted, float4 *__restrict__ sorted32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot)
        return;
    const double L = sys[psys[i]].L;
    const int k = key[i];
    const int pos = start[k] + atomicAdd(&cursor[k], 1);
    perm[pos] = i;
    const double4 r = make_double4(fold(x[i], L), fold(y[i], L), fold(z[i], L), q[i]);
    sorted[pos] = r;
    sorted32[pos] = make_float4((float)r.x, (float)r.y, (float)r.z, 0.0f);
}

template <bool FP32>
__global__ void __launch_bounds__(kWave)
k_brealspace(const BSys *__restrict__ sys, const int *__restrict__ cell0s, int nsys, int ntot,
             const double4 *__restrict__ sorted, const float4 *__restrict__ sorted32,
             const int *__restrict__ perm,
             const int *__restrict__ start, const int *__restrict__ istart,
             const float4 *__restrict__ bbox, long ncell, int *__restrict__ counter,
             double *__restrict__ fr, double *__restrict__ energy) {
    __shared__ RsShared<FP32> sh;
    const int total = istart[ncell];
    load_erfcx_table(sh);
    for (;;) {
        int item = 0;
        if (threadIdx.x == 0)
            item = atomicAdd(counter, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= total)
            break;
        const int home = (int)prefix_find(istart, ncell, item);
        const int s = p3m::fft::prefix_slot(cell0s, nsys, (long)home);
        const BSys &S = sys[s];
        if (!(S.rcut > 0.0))
            continue; // no real-space part for this system
        double en = 0.0;
        if (S.small)
            realspace_item<FP32, true>(home - S.cell0, item - istart[home], S.cell0, sorted, sorted32,
                                       perm, start, istart, bbox, S.nc, S.reach, S.L, S.rcut, S.alpha,
                                       ntot, fr, sh, en);
        else
            realspace_item<FP32, false>(home - S.cell0, item - istart[home], S.cell0, sorted, sorted32,
                                        perm, start, istart, bbox, S.nc, S.reach, S.L, S.rcut, S.alpha,
                                        ntot, fr, sh, en);
        if (energy != nullptr) {
            for (int off = 32; off > 0; off >>= 1)
                en += __shfl_down(en, off, kWave);
            if (threadIdx.x == 0 && en != 0.0)
                atomicAdd(&energy[s], en);
        }
    }
}

}  // namespace

namespace p3m {

int brs_forces(const BForm &f, const BSys *sys, const int *psys, const int *cell0s, int nsys,
               long ncell, int ntot, const double *x, const double *y, const double *z,
               const double *q, CellArrays &a, int grid, double *fr, double *energy,
               p3mk_queue *qu) {
    if (ntot <= 0)
        return P3MK_OK;
    hipStream_t s = p3mk_stream(qu);
    const int nb = (ntot + 255) / 256;
    P3MK_TRY(cells_begin(a, ncell, qu));
    hipLaunchKernelGGL(k_bcell_count, dim3(nb), dim3(256), 0, s, ntot, sys, psys, x, y, z, a.key,
                       a.fcount);
    P3MK_TRY(p3mk_launch_check("batched cell count"));
    P3MK_TRY(cells_scan(a, ncell, qu));
    hipLaunchKernelGGL(k_bcell_scatter, dim3(nb), dim3(256), 0, s, ntot, sys, psys, x, y, z, q, a.key,
                       a.fstart, a.fcursor, a.perm, a.sorted, a.sorted32);
    P3MK_TRY(p3mk_launch_check("batched cell scatter"));
    P3MK_TRY(cells_end(a, ncell, qu));
    if (f.fp32)
        hipLaunchKernelGGL(k_brealspace<true>, dim3((unsigned)grid), dim3(kWave), 0, s, sys, cell0s,
                           nsys, ntot, a.sorted, a.sorted32, a.perm, a.start, a.istart, a.bbox, ncell, a.counter,
                           fr, energy);
    else
        hipLaunchKernelGGL(k_brealspace<false>, dim3((unsigned)grid), dim3(kWave), 0, s, sys, cell0s,
                           nsys, ntot, a.sorted, a.sorted32, a.perm, a.start, a.istart, a.bbox, ncell, a.counter,
                           fr, energy);
    return p3mk_launch_check("batched real space");
}

}  // namespace p3m
