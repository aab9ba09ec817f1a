// Real-space (short-range) Ewald force on a linked-cell list.
//
// Pair set and arithmetic of Realteil (realpart.c:124-167): every ordered
// pair i != j whose minimum-image distance r satisfies r <= rcut contributes
//   f_i += q_i q_j (erfc(alpha r)/r + 2 alpha/sqrt(pi) exp(-alpha^2 r^2)) / r^2 * d
// and 0.5 q_i q_j erfc(alpha r)/r to the energy.
//
// Layout: particles are folded into [0, L) and counting-sorted into cells of
// side >= rcut/R (R = 2, or R = 1 when cells of side rcut/2 would hold fewer
// than 6 particles on average or the box has fewer than 5 of them per side;
// cell_geometry), and inside a cell by octant, so that 64 consecutive particles of a cell are spatially
// compact.  A work item is (home cell, chunk of 64 home particles); a
// persistent grid of single-wavefront workgroups claims items from an
// atomic counter, so the dense cells of a clustered system are spread over
// the whole GPU.  Every item has a bounding box (computed after the sort),
// and neighbour cells and neighbour chunks whose box is beyond rcut from the
// item's box are skipped.
//
// Two kernels walk this layout.
//  - Cell grids (at least 2 R + 1 cells per side): the half shell
//    (k_realspace_half, further down), each unordered pair once with its
//    reaction added to the partner; neighbour particles are staged into a
//    64-slot LDS window and paired with the item's home particles on the
//    matrix cores, K = 4 products of
//      [-2 n', 1] . [h, |h|^2 - rho^2 - m] + |n'|^2 = |h - n'|^2 - rho^2 - m
//    with n' and h relative to the home cell centre in units of the cell
//    side, rho = rcut / side, FP16 operands and FP32 sums; the margin
//    m = 0.02 covers the FP16 rounding, so no pair within rcut is lost.
//    Home cells of at most four particles take a vector search instead.
//  - Boxes of fewer cells per side (nearest image per pair) and the adjoint:
//    the full shell (k_realspace, k_realspace_vjp), every ordered pair for
//    its home particle, candidates found by the vector ALUs.
// Both queue the accepted pairs in LDS (rank by mbcnt over the ballot) and
// evaluate every 64 queued pairs with the 64 lanes at once, whoever owns
// them, with the exact FP64 difference and the reference's inclusion test.
//
// Periodic images are a per-neighbour-cell shift of 0 or +-L, which equals
// the reference's round(dx/L)*L for every pair inside the cutoff
// (rcut <= L/2), so for positions inside [0, L) dx = (x_i - x_j) - shift is
// the reference's dx bit for bit.  Positions outside the box are folded into
// it first; their dx then differs from the reference's raw difference by
// rounding only.

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <type_traits>

#include "batch_desc.hpp"
#include "device_common.hpp"
#include "erfcx_coeffs.hpp"
#include "k_internal.hpp"

namespace {

using p3m::CellArrays;

constexpr int kWave = 64;
// full-shell kernels (nearest images, adjoint): at most 63 pending pairs plus
// one append of the vector search
constexpr int kQueue = 2 * kWave;
// work-item counters of a cell list (one per kernel launched on it)
constexpr int kCounters = 4;
// sorted arrays carry kPad zero-charge entries past the last particle
constexpr int kPad = 64;
// home cells of more particles than this use the matrix-core candidate search
// of the half shell; sparser ones the vector search, whose per-cell cost is
// lower
constexpr int kMfmaMinCell = 4;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));

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

// Fixed-point coordinates relative to a cell centre: fix_bits(reach)
// fractional bits per cell side, so differences of two such coordinates in
// one frame are exact integers.  A frame spans the (2 reach + 1)^3 cells
// around its centre: reach 2, |v| <= 2.5 sides < 2^31 / 2^29; reach 1,
// |v| <= 1.5 sides < 2^31 / 2^30.  A difference of two coordinates of one
// frame stays below 2^31 in magnitude except at exactly 2 sides apart
// (reach 1), where it wraps to -2^31 of the same magnitude; such a pair is
// beyond rcut, which is at most one side.
__host__ __device__ constexpr int fix_bits(int reach) { return reach == 1 ? 30 : 29; }

__device__ __forceinline__ int to_fixed(double v, double inv_quantum) {
    return (int)rint(v * inv_quantum);
}

// Sorted copies of one particle: FP64 (x, y, z, q); relative to the centre
// of its cell (cx, cy, cz) in FP32 the row (-2 x, -2 y, -2 z, q) the
// candidate search builds its matrix rows from; and in fixed point
// (x, y, z, q as FP32 bits) for the FP32 pair difference.
__device__ __forceinline__ void store_sorted(int pos, double4 r, int cell, int nc, double side, int fixbits,
                                             double4 *__restrict__ sorted,
                                             float4 *__restrict__ sorted32,
                                             int4 *__restrict__ sortedq) {
    const int cz = cell % nc, cy = (cell / nc) % nc, cx = cell / (nc * nc);
    const double rx = r.x - (cx + 0.5) * side, ry = r.y - (cy + 0.5) * side, rz = r.z - (cz + 0.5) * side;
    sorted[pos] = r;
    sorted32[pos] = make_float4(-2.0f * (float)rx, -2.0f * (float)ry, -2.0f * (float)rz, (float)r.w);
    const double iq = ldexp(1.0, fixbits) / side;
    sortedq[pos] = make_int4(to_fixed(rx, iq), to_fixed(ry, iq), to_fixed(rz, iq), __float_as_int((float)r.w));
}

// the kPad entries past the last particle: finite rows of zero charge
__device__ __forceinline__ void store_pad(int pos, double4 *__restrict__ sorted,
                                          float4 *__restrict__ sorted32, int4 *__restrict__ sortedq) {
    sorted[pos] = make_double4(0.0, 0.0, 0.0, 0.0);
    sorted32[pos] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    sortedq[pos] = make_int4(0, 0, 0, 0);
}

__global__ void k_cell_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double L, double inv_side, int nc,
                             int *__restrict__ key, int *__restrict__ rank, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int k = fine_key(fold(x[i], L), fold(y[i], L), fold(z[i], L), inv_side, nc);
    key[i] = k;
    rank[i] = atomicAdd(&count[k], 1); // the scatter then needs no atomics
}

// n + kPad threads; key[i] is replaced by the particle's sorted position (the
// inverse of perm); with g (3n, x/y/z blocks) also gsorted[pos] = (g_x, g_y,
// g_z, 0) of the particle at pos (the adjoint's cotangent in sorted order)
__global__ void k_cell_scatter(int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const double *__restrict__ q, double L,
                               int nc, int fixbits, int *__restrict__ key, const int *__restrict__ start,
                               const int *__restrict__ rank, int *__restrict__ perm,
                               double4 *__restrict__ sorted, float4 *__restrict__ sorted32,
                               int4 *__restrict__ sortedq, const double *__restrict__ g,
                               double4 *__restrict__ gsorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n + kPad)
        return;
    if (i >= n) {
        store_pad(i, sorted, sorted32, sortedq);
        if (g != nullptr)
            gsorted[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const int k = key[i];
    const int pos = start[k] + rank[i];
    perm[pos] = i;
    key[i] = pos;
    const double4 r = make_double4(fold(x[i], L), fold(y[i], L), fold(z[i], L), q[i]);
    store_sorted(pos, r, k >> 3, nc, L / nc, fixbits, sorted, sorted32, sortedq);
    if (g != nullptr)
        gsorted[pos] = make_double4(g[i], g[n + i], g[2 * (long)n + i], 0.0);
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
// bbox[2 t] = (xlo, ylo, zlo, cell), bbox[2 t + 1] = (xhi, yhi, zhi, chunk),
// the item's cell and chunk stored as int bits (read by the real-space
// kernels instead of a search of istart; nothing reads them as floats).  Eight
// lanes per item, each over every eighth of its (up to 64) particles, then
// min / max across the eight: a thread per item serialises 64 loads, a
// wavefront per item multiplies the cell searches by 64.
constexpr int kBoxLanes = 8;

__global__ void k_item_bbox(const int *__restrict__ start, const int *__restrict__ istart, long ncell,
                            long max_items, const double4 *__restrict__ sorted,
                            float4 *__restrict__ bbox) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long t = g / kBoxLanes;
    const int sub = (int)(g % kBoxLanes);
    if (t >= max_items || t >= istart[ncell])
        return; // all eight lanes of the item
    const long c = prefix_find(istart, ncell, (int)t);
    const int p0 = start[c] + (int)(t - istart[c]) * kWave;
    const int p1 = min(start[c + 1], p0 + kWave);
    double xl = 1e300, yl = 1e300, zl = 1e300, xh = -1e300, yh = -1e300, zh = -1e300;
    for (int p = p0 + sub; p < p1; p += kBoxLanes) {
        const double4 r = sorted[p];
        xl = fmin(xl, r.x); yl = fmin(yl, r.y); zl = fmin(zl, r.z);
        xh = fmax(xh, r.x); yh = fmax(yh, r.y); zh = fmax(zh, r.z);
    }
    for (int off = kBoxLanes / 2; off > 0; off >>= 1) {
        xl = fmin(xl, __shfl_xor(xl, off, kBoxLanes));
        yl = fmin(yl, __shfl_xor(yl, off, kBoxLanes));
        zl = fmin(zl, __shfl_xor(zl, off, kBoxLanes));
        xh = fmax(xh, __shfl_xor(xh, off, kBoxLanes));
        yh = fmax(yh, __shfl_xor(yh, off, kBoxLanes));
        zh = fmax(zh, __shfl_xor(zh, off, kBoxLanes));
    }
    if (sub == 0) {
        bbox[2 * t] = make_float4(down(xl), down(yl), down(zl), __int_as_float((int)c));
        bbox[2 * t + 1] = make_float4(up(xh), up(yh), up(zh), __int_as_float((int)(t - istart[c])));
    }
}

// value of lane k (uniform) in every lane
__device__ __forceinline__ double bcast(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 2 alpha / sqrt(pi) with the reference's sqrt(pi) literal (realpart.c:135)
__device__ __forceinline__ double sqrt_pi() { return 1.77245385090551602729816748334; }

struct PairAcc {
    double f0, f1, f2;
    double en;  // FP64 form
    float enf;  // FP32 form
};

constexpr int kErfcxTab = p3m::kErfcxDoublePieces * (p3m::kErfcxDoubleDegree + 1);

// erfcx(x) = exp(x^2) erfc(x), x >= 0, from the generated piecewise FP64
// tables (a copy in LDS: the piece index differs from lane to lane)
__device__ __forceinline__ double erfcx_d(double x, const double *tab) {
    // y = 1 / (1 + x / 2): hardware reciprocal refined by two Newton steps
    // (the piece index and argument need y to a few ulp, not correctly rounded)
    const double den = fma(0.5, x, 1.0);
    double y = __builtin_amdgcn_rcp(den);
    y = fma(fma(-den, y, 1.0), y, y);
    y = fma(fma(-den, y, 1.0), y, y);
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

// The reference's inclusion test r <= rcut (realpart.c:149-163) on its dx,
// decided exactly: r2 against rcut^2, with the correctly rounded sqrt only
// within 1e-12 of the edge.
__device__ __forceinline__ bool within_cutoff(double dx, double dy, double dz, double rcut) {
    double r2;
    {
#pragma clang fp contract(off)
        r2 = dx * dx + dy * dy + dz * dz;
    }
    const double rc2 = rcut * rcut;
    bool inside = r2 < rc2 * (1.0 - 1e-12);
    if (__builtin_expect(!inside && r2 <= rc2 * (1.0 + 1e-12), 0))
        inside = sqrt(r2) <= rcut;
    return inside;
}

// Pair evaluation; dx is the reference's minimum-image difference bit for
// bit and the pair set is the reference's.  erfc(x) = exp(-x^2) erfcx(x)
// shares the one exponential with the exp(-x^2) term of the force.
//  - FP64: x^2 is split exactly (hi + lo by FMA) so exp(-x^2) keeps its
//    accuracy for large x.
//  - FP32: d = dx rounded to FP32 gives r2f within 5 2^-24 of r2, so the
//    inclusion test is decided on r2f except within 1e-6 of rcut^2, where
//    the FP64 test runs.
// FP32 pair force from the FP32 difference (fx, fy, fz), r2f = |f|^2, the
// exact inclusion decision and the charge product
__device__ __forceinline__ void pair_force_f32(float fx, float fy, float fz, float r2f, bool inside,
                                               float qq, double alpha, double two_a_spi,
                                               PairAcc &acc) {
    const float rinv = __builtin_amdgcn_rsqf(r2f);
    const float x = (float)alpha * (r2f * rinv);
    // e^{-alpha^2 r^2} as 2^{-alpha^2 log2(e) r^2}, the charge product folded
    // into it: qe = q_i q_j e^{-x^2}, enf = q_i q_j erfc(x) / r
    const float qe = (inside ? qq : 0.0f) *
                     __builtin_amdgcn_exp2f((float)(-alpha * alpha * 1.4426950408889634) * r2f);
    const float enf = qe * erfcx_f(x) * rinv;
    const float fak = fmaf((float)two_a_spi, qe, enf) * (rinv * rinv);
    acc.f0 = (double)(fak * fx);
    acc.f1 = (double)(fak * fy);
    acc.f2 = (double)(fak * fz);
    acc.enf = enf; // twice the energy
}

template <bool FP32>
__device__ __forceinline__ void pair(const double4 &pi, const double4 &pj, double sx, double sy,
                                     double sz, double rcut, double alpha, double two_a_spi,
                                     const double *tab, PairAcc &acc) {
    double dx, dy, dz;
    {
#pragma clang fp contract(off)
        dx = (pi.x - pj.x) - sx;
        dy = (pi.y - pj.y) - sy;
        dz = (pi.z - pj.z) - sz;
    }
    const double qq = pi.w * pj.w;
    if (FP32) {
        const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
        const float r2f = fx * fx + fy * fy + fz * fz;
        const double rc2 = rcut * rcut;
        bool inside = r2f <= (float)(rc2 * (1.0 - 1e-6));
        if (__builtin_expect(!inside && r2f <= (float)(rc2 * (1.0 + 1e-6)), 0))
            inside = within_cutoff(dx, dy, dz, rcut);
        pair_force_f32(fx, fy, fz, r2f, inside, (float)qq, alpha, two_a_spi, acc);
    } else {
        if (!within_cutoff(dx, dy, dz, rcut))
            return;
        double r2;
        {
#pragma clang fp contract(off)
            r2 = dx * dx + dy * dy + dz * dz;
        }
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

// Adjoint of the pair force for the ordered pair (i, j) of the reference's
// pair set.  With f_ij = q_i q_j s(r) d, s = (erfc(alpha r)/r + c e^{-alpha^2
// r^2}) / r^2 (c = 2 alpha / sqrt(pi)), the Jacobian d f_ij / d d is
// q_i q_j (s I + (s'/r) d d^T), symmetric and even in d, so for a cotangent g
// of the forces the owner i collects, over its partners j,
//   d pos_i += q_i q_j (s dg + (s'/r) (d . dg) d),  d q_i += q_j s (d . dg),
// dg = g_i - g_j (each ordered pair is evaluated once for its owner).
// acc = (d pos_i, d q_i).  FP64.
__device__ __forceinline__ void pair_vjp(const double4 &pi, const double4 &pj, double sx, double sy,
                                         double sz, double rcut, double alpha, double two_a_spi,
                                         const double *tab, const double4 &gi, const double4 &gj,
                                         double acc[4]) {
    double dx, dy, dz;
    {
#pragma clang fp contract(off)
        dx = (pi.x - pj.x) - sx;
        dy = (pi.y - pj.y) - sy;
        dz = (pi.z - pj.z) - sz;
    }
    if (!within_cutoff(dx, dy, dz, rcut))
        return;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rinv = rsqrt(r2);
    const double r = r2 * rinv;
    const double x = alpha * r;
    const double x2 = x * x;
    const double e0 = exp(-x2);
    const double e = fma(-e0, fma(x, x, -x2), e0); // exp(-alpha^2 r^2)
    const double ec = e * erfcx_d(x, tab);           // erfc(alpha r)
    const double phi = ec * rinv + two_a_spi * e;
    const double dphi = -ec * (rinv * rinv) - two_a_spi * e * (rinv + 2.0 * alpha * alpha * r);
    const double s = phi * (rinv * rinv);
    const double ds_r = (dphi * (rinv * rinv) - 2.0 * phi * (rinv * rinv * rinv)) * rinv;
    const double gx = gi.x - gj.x, gy = gi.y - gj.y, gz = gi.z - gj.z;
    const double ddg = dx * gx + dy * gy + dz * gz;
    const double qq = pi.w * pj.w;
    acc[0] += qq * (s * gx + ds_r * ddg * dx);
    acc[1] += qq * (s * gy + ds_r * ddg * dy);
    acc[2] += qq * (s * gz + ds_r * ddg * dz);
    acc[3] += pj.w * s * ddg;
}

// LDS of one wavefront's work item: the queue of accepted pairs, the home
// particles, their accumulators (force, or for the adjoint d pos and d q and
// the home cotangents) and, for FP64, the erfcx table.
template <bool FP32, bool VJP = false> struct RsShared {
    unsigned int qj[kQueue]; // neighbour's sorted index
    unsigned int qo[kQueue]; // pack_pair(owner lane, shift index)
    double4 shift[27]; // image shift (sx, sy, sz) of shift index (wx+1)*9 + (wy+1)*3 + wz+1
    double4 home_pos[kWave];
    double facc[VJP ? 4 : 3][kWave];
    double4 home_g[VJP ? kWave : 1];
    double tab[FP32 ? 1 : kErfcxTab];
};

template <bool FP32, bool VJP>
__device__ __forceinline__ void load_erfcx_table(RsShared<FP32, VJP> &sh) {
    if (!FP32) {
        for (int k = threadIdx.x; k < kErfcxTab; k += blockDim.x)
            sh.tab[k] = (&p3m::kErfcxDoubleCoef[0][0])[k];
    }
    __syncthreads();
}

// Work-item claims of the persistent grids: a claim takes `per` consecutive
// items with one device atomic on the shared counter, which is a serial point
// (190k single-item claims of sparse cells took most of a 2.2 ms batched
// launch).  per = total / (8 grid), at least 1, leaves a wavefront about
// eight claims, for the balance of clustered systems.
__device__ __forceinline__ int claim_size(int total) { return max(1, total / ((int)gridDim.x * 8)); }

// this wavefront's next item (next, end: its current claim), total when the
// items are exhausted
__device__ __forceinline__ int next_item(int *counter, int total, int per, int &next, int &end) {
    if (next >= end) {
        int first = 0;
        if (threadIdx.x == 0)
            first = atomicAdd(counter, per);
        first = __builtin_amdgcn_readfirstlane(first);
        if (first >= total)
            return total;
        next = first;
        end = min(total, first + per);
    }
    return next++;
}

// index of the periodic image shift (wx, wy, wz) in {-1, 0, 1}^3
__device__ __forceinline__ unsigned int shift_index(int wx, int wy, int wz) {
    return (unsigned)((wx + 1) * 9 + (wy + 1) * 3 + (wz + 1));
}

// queue entry of a pair: the LDS byte offsets of the owner's home_pos entry
// (bits 0..15) and of the image shift (bits 16..31), both double4
__device__ __forceinline__ unsigned int pack_pair(int owner, unsigned int shift) {
    return (unsigned)owner * 32u | (shift * 32u) << 16;
}

// the 27 image shifts of a box of side L
template <bool FP32, bool VJP>
__device__ __forceinline__ void fill_shifts(RsShared<FP32, VJP> &sh, double L) {
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

// distance along one axis between [alo, ahi] and [blo, bhi]
__device__ __forceinline__ float gap(float alo, float ahi, float blo, float bhi) {
    return fmaxf(0.0f, fmaxf(blo - ahi, alo - bhi));
}

// One neighbour chunk [cb, cb + cnt) against the item's home particles on
// the vector ALUs (the FP64 form): neighbours are read with scalar
// (wave-uniform) loads, four at a time so that their latencies overlap, and
// each lane tests its own home particle hl (sorted32 row of the home
// particle) against them in the frame of the home cell centre; (ox, oy, oz)
// is the neighbour cell centre relative to it.  os: owner lane | shift.
template <typename Append, typename Flush>
__device__ __forceinline__ void scan_chunk_scalar(Append &append, Flush &flush,
                                                  const float4 *__restrict__ sorted32, int cb,
                                                  int cnt, int i0, int lane, float4 hl, double ox,
                                                  double oy, double oz, float rc2f,
                                                  unsigned int os) {
    // h - o; the neighbour is n = -v / 2 of its row v, so h - o - n = p + v / 2
    const float px = -0.5f * hl.x - (float)ox, py = -0.5f * hl.y - (float)oy,
                pz = -0.5f * hl.z - (float)oz;
    auto test = [&](const float4 &v, int k) {
        const float dx = fmaf(0.5f, v.x, px), dy = fmaf(0.5f, v.y, py), dz = fmaf(0.5f, v.z, pz);
        const float r2 = dx * dx + dy * dy + dz * dz;
        append(r2 <= rc2f, (unsigned)(cb + k), os);
        flush();
    };
    if (cb == i0) {
        // the item's own chunk: exclude the self pair
        for (int k = 0; k < cnt; k++) {
            const float4 v = sorted32[cb + k];
            const float dx = fmaf(0.5f, v.x, px), dy = fmaf(0.5f, v.y, py), dz = fmaf(0.5f, v.z, pz);
            append(dx * dx + dy * dy + dz * dz <= rc2f && k != lane, (unsigned)(cb + k), os);
            flush();
        }
        return;
    }
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
        const float4 v0 = sorted32[cb + k], v1 = sorted32[cb + k + 1], v2 = sorted32[cb + k + 2],
                     v3 = sorted32[cb + k + 3];
        test(v0, k);
        test(v1, k + 1);
        test(v2, k + 2);
        test(v3, k + 3);
    }
    for (; k < cnt; k++)
        test(sorted32[cb + k], k);
}

// One work item of the full shell (home cell, chunk of up to 64 home
// particles; boxes of fewer than 2 reach + 1 cells per side and the
// adjoint); every ordered pair is evaluated for its home particle.
// Neighbour cells are the (2 reach + 1)^3 cells around the home cell, or,
// in a box of fewer cells per side (SMALL_BOX), every cell once with the
// nearest image taken per pair.  Neighbour cells and neighbour chunks whose
// bounding box lies beyond rcut from the item's box are skipped.  Cells of
// the item's system are cell0 + (cx nc + cy) nc + cz with items from
// istart[cell]; forces go to fr[d stride + perm[i]].
// VJP: the adjoint instead of the force (FP64 only): gsorted holds the
// cotangent in sorted order, fr receives d pos and dq d q.
template <bool FP32, bool SMALL_BOX, bool VJP = false>
__device__ __forceinline__ void
realspace_item(int home, int chunk, int cell0, const double4 *__restrict__ sorted,
               const float4 *__restrict__ sorted32, const int *__restrict__ perm,
               const int *__restrict__ start, const int *__restrict__ istart,
               const float4 *__restrict__ bbox, int nc, int reach, double L, double rcut,
               double alpha, long stride, double *__restrict__ fr,
               RsShared<FP32, VJP> &sh, double &en_tot, const double4 *__restrict__ gsorted = nullptr,
               double *__restrict__ dq = nullptr) {
    static_assert(!(VJP && FP32), "the adjoint is FP64");
    const int lane = threadIdx.x;
    const double two_a_spi = 2.0 * alpha / sqrt_pi();
    const double side = L / nc;
    // boxes are compared in FP32: margin for their rounding
    const float rskip = (float)(rcut * (1.0 + 1e-5) + 1e-5 * L);
    const float rskip2 = rskip * rskip;
    const int hcell = cell0 + home;
    const int hb = start[hcell], he = start[hcell + 1];
    const int hz = home % nc, hy = (home / nc) % nc, hx = home / (nc * nc);
    const int i0 = hb + chunk * kWave; // sorted index of lane 0's home particle
    const int nh = min(kWave, he - i0);
    const int i = i0 + lane;
    const bool active = lane < nh;
    const double4 pi = active ? sorted[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    sh.home_pos[lane] = pi;
    sh.facc[0][lane] = 0.0;
    sh.facc[1][lane] = 0.0;
    sh.facc[2][lane] = 0.0;
    if constexpr (VJP) {
        sh.facc[3][lane] = 0.0;
        sh.home_g[lane] = active ? gsorted[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    fill_shifts(sh, L);
    int count = 0; // wave-uniform queue length
    __syncthreads();

    float en32 = 0.0f; // FP32 form: twice this item's energy, per lane

    // lane l < m evaluates queue entry count - m + l for its owner
    auto drain = [&](int m) {
        const int base = count - m;
        if (lane < m) {
            const unsigned int j = sh.qj[base + lane], os = sh.qo[base + lane];
            const unsigned int ho = os & 0xffffu; // owner * 32
            const int o = (int)(ho >> 5);
            const char *lds_home = reinterpret_cast<const char *>(sh.home_pos);
            const char *lds_shift = reinterpret_cast<const char *>(sh.shift);
            const double4 s3 = *reinterpret_cast<const double4 *>(lds_shift + (os >> 16));
            // 32-bit byte offset (the particle count is bounded on the host)
            const double4 pj = *reinterpret_cast<const double4 *>(
                reinterpret_cast<const char *>(sorted) + (j << 5));
            if constexpr (VJP) {
                double acc[4] = {0.0, 0.0, 0.0, 0.0};
                pair_vjp(*reinterpret_cast<const double4 *>(lds_home + ho), pj, s3.x, s3.y, s3.z,
                         rcut, alpha, two_a_spi, sh.tab, sh.home_g[o], gsorted[j], acc);
                for (int d = 0; d < 4; d++)
                    atomicAdd(&sh.facc[d][o], acc[d]);
            } else {
                PairAcc pa = {0.0, 0.0, 0.0, 0.0, 0.0f};
                pair<FP32>(*reinterpret_cast<const double4 *>(lds_home + ho), pj, s3.x, s3.y, s3.z,
                           rcut, alpha, two_a_spi, sh.tab, pa);
                atomicAdd(&sh.facc[0][o], pa.f0);
                atomicAdd(&sh.facc[1][o], pa.f1);
                atomicAdd(&sh.facc[2][o], pa.f2);
                if (FP32)
                    en32 += pa.enf;
                else
                    en_tot += pa.en;
            }
        }
        count = base;
        __syncthreads();
    };

    // the accepting lanes (ok) queue the pair (neighbour j, owner | shift)
    auto append = [&](bool ok, unsigned int j, unsigned int os) {
        const unsigned long long mask = __ballot(ok);
        if (mask) {
            if (ok) {
                const unsigned int slot = __builtin_amdgcn_mbcnt_hi(
                    (unsigned int)(mask >> 32),
                    __builtin_amdgcn_mbcnt_lo((unsigned int)mask, (unsigned int)count));
                sh.qj[slot] = j;
                sh.qo[slot] = os;
            }
            count += __popcll(mask);
        }
    };
    // evaluate full rounds of 64 queued pairs
    auto flush = [&]() {
        if (count >= kWave) {
            __syncthreads();
            do
                drain(kWave);
            while (count >= kWave);
        }
    };

    if (SMALL_BOX) {
        // every cell once, the nearest image per pair (the reference's
        // round(dx/L)); FP64 test with a margin, the exact one in pair()
        const double Li = 1.0 / L;
        const double rc2 = rcut * rcut * (1.0 + 1e-12);
        const int ncell = nc * nc * nc;
        for (int cl = 0; cl < ncell; cl++) {
            const int cell = cell0 + cl;
            const int jb = __builtin_amdgcn_readfirstlane(start[cell]);
            const int je = __builtin_amdgcn_readfirstlane(start[cell + 1]);
            for (int cb = jb; cb < je; cb += kWave) {
                const int cnt = min(kWave, je - cb);
                // one coalesced load, then lane k's particle is broadcast
                const double4 mine = (lane < cnt) ? sorted[cb + lane] : make_double4(0.0, 0.0, 0.0, 0.0);
                for (int k = 0; k < cnt; k++) {
                    const int j = cb + k;
                    double dx = pi.x - bcast(mine.x, k), dy = pi.y - bcast(mine.y, k),
                           dz = pi.z - bcast(mine.z, k);
                    const double rx = round(dx * Li), ry = round(dy * Li), rz = round(dz * Li);
                    dx -= rx * L;
                    dy -= ry * L;
                    dz -= rz * L;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    append(active && j != i && r2 <= rc2, (unsigned)j,
                           pack_pair(lane, shift_index((int)rx, (int)ry, (int)rz)));
                    flush();
                }
            }
        }
    } else {
        const int hitem = istart[hcell] + chunk;
        const float4 hlo = bbox[2 * hitem], hhi = bbox[2 * hitem + 1];
        const float rc2f = (float)(rcut * rcut * (1.0 + 2e-5));
        // this lane's home particle relative to the home cell centre (inactive
        // lanes far away)
        const float4 hl = active ? sorted32[i] : make_float4(-2e30f, -2e30f, -2e30f, 0.0f);
        for (int a = -reach; a <= reach; a++) {
            int cx = hx + a, wx = 0;
            if (cx < 0) { cx += nc; wx = -1; } else if (cx >= nc) { cx -= nc; wx = 1; }
            // the neighbour cell's image spans [(hx + a) side, (hx + a + 1) side]
            const float gx = gap(hlo.x, hhi.x, (float)((hx + a) * side), (float)((hx + a + 1) * side));
            for (int b = -reach; b <= reach; b++) {
                int cy = hy + b, wy = 0;
                if (cy < 0) { cy += nc; wy = -1; } else if (cy >= nc) { cy -= nc; wy = 1; }
                const float gy = gap(hlo.y, hhi.y, (float)((hy + b) * side), (float)((hy + b + 1) * side));
                if (gx * gx + gy * gy > rskip2)
                    continue;
                for (int c = -reach; c <= reach; c++) {
                    int cz = hz + c, wz = 0;
                    if (cz < 0) { cz += nc; wz = -1; } else if (cz >= nc) { cz -= nc; wz = 1; }
                    const float gz = gap(hlo.z, hhi.z, (float)((hz + c) * side), (float)((hz + c + 1) * side));
                    // the whole neighbour cell beyond the cutoff from the item's box?
                    if (gx * gx + gy * gy + gz * gz > rskip2)
                        continue;
                    const int cell = cell0 + (cx * nc + cy) * nc + cz;
                    const int jb = __builtin_amdgcn_readfirstlane(start[cell]);
                    const int je = __builtin_amdgcn_readfirstlane(start[cell + 1]);
                    if (jb == je)
                        continue;
                    const int item0 = __builtin_amdgcn_readfirstlane(istart[cell]);
                    const unsigned int shp = pack_pair(0, shift_index(wx, wy, wz));
                    const float fsx = (float)(wx * L), fsy = (float)(wy * L), fsz = (float)(wz * L);
                    for (int cb = jb; cb < je; cb += kWave) {
                        const int cnt = min(kWave, je - cb);
                        const int nitem = item0 + (cb - jb) / kWave;
                        float4 nlo = bbox[2 * nitem], nhi = bbox[2 * nitem + 1];
                        nlo.x += fsx; nlo.y += fsy; nlo.z += fsz;
                        nhi.x += fsx; nhi.y += fsy; nhi.z += fsz;
                        if (box_dist2(hlo, hhi, nlo, nhi) > rskip2)
                            continue;
                        scan_chunk_scalar(append, flush, sorted32, cb, cnt, i0, lane, hl, a * side,
                                          b * side, c * side, rc2f, pack_pair(lane, 0) | shp);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (count > 0)
        drain(count);
    if (FP32)
        en_tot += 0.5 * (double)en32;
    if (active) {
        const int p = perm[i];
        atomicAdd(&fr[p], sh.facc[0][lane]);
        atomicAdd(&fr[stride + p], sh.facc[1][lane]);
        atomicAdd(&fr[2 * stride + p], sh.facc[2][lane]);
        if constexpr (VJP)
            atomicAdd(&dq[p], sh.facc[3][lane]);
    }
    __syncthreads();
}

// Persistent grid of the full shell (launched for boxes of fewer than
// 2 reach + 1 cells per side): each single-wavefront workgroup claims work
// items from an atomic counter until none are left.
template <bool FP32, bool SMALL_BOX>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(FP32 ? 5 : 1)))
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
    const int per = claim_size(total);
    int next = 0, end = 0;
    for (;;) {
        const int item = next_item(counter, total, per, next, end);
        if (item >= total)
            break;
        const int home = __float_as_int(bbox[2 * item].w), chunk = __float_as_int(bbox[2 * item + 1].w);
        realspace_item<FP32, SMALL_BOX>(home, chunk, 0, sorted, sorted32, perm,
                                        start, istart, bbox, nc, reach, L, rcut, alpha, n, fr, sh,
                                        en_tot);
    }
    if (energy != nullptr) {
        for (int off = 32; off > 0; off >>= 1)
            en_tot += __shfl_down(en_tot, off, kWave);
        if (threadIdx.x == 0 && en_tot != 0.0)
            atomicAdd(energy, en_tot);
    }
}

// Adjoint of the real-space force over the same pair set (FP64), persistent
// grid as k_realspace: dpos (3 n) += d pos, dq (n) += d q.
template <bool SMALL_BOX>
__global__ void __launch_bounds__(kWave)
k_realspace_vjp(int n, const double4 *__restrict__ sorted, const float4 *__restrict__ sorted32,
                const int *__restrict__ perm,
                const int *__restrict__ start, const int *__restrict__ istart,
                const float4 *__restrict__ bbox, long ncell, int nc, int reach, double L,
                double rcut, double alpha, int *__restrict__ counter,
                const double4 *__restrict__ gsorted, double *__restrict__ dpos,
                double *__restrict__ dq) {
    __shared__ RsShared<false, true> sh;
    const int total = istart[ncell];
    double en_unused = 0.0;
    load_erfcx_table(sh);
    const int per = claim_size(total);
    int next = 0, end = 0;
    for (;;) {
        const int item = next_item(counter, total, per, next, end);
        if (item >= total)
            break;
        const int home = __float_as_int(bbox[2 * item].w), chunk = __float_as_int(bbox[2 * item + 1].w);
        realspace_item<false, SMALL_BOX, true>(home, chunk, 0, sorted, sorted32,
                                               perm, start, istart, bbox, nc, reach, L, rcut, alpha,
                                               n, dpos, sh, en_unused, gsorted, dq);
    }
}


// ---- half shell (Newton's third law) ---------------------------------------
//
// Each unordered pair is evaluated once: the item of a home cell visits the
// neighbour cells of lexicographically positive offset, its own chunk and the
// later chunks of its cell (in its own chunk only the neighbours after the
// home particle).  The owner receives f and the neighbour -f.  The reverse
// pair's difference is -dx exactly, so its r, its inclusion and its force
// factor are the same, and -f is the force the reference computes for it bit
// for bit (realpart.c:149-163).  Forces are summed in sorted order
// (fs[d stride + i]) and added to the caller's array by perm afterwards.
//
// Neighbour particles within rcut of the item's box are staged in a window
// of 64 slots (several chunks, of several cells), relative to the HOME cell
// centre in units of the cell side; the reactions on them are summed per
// slot in LDS.  The candidate search of a dense item pairs the window with
// its home particles on the matrix cores in 4 x 4 blocks
// (v_mfma_f32_4x4x4f16: lane l = 4 b + col of block b holds row l % 4 of A
// and column col of B, K = 4 each, and result register r is D[row r][col]),
// one product of A = (-2 n, 1) and B = (h, |h|^2 - rho^2 - margin) in FP16
// with C = |n|^2 in FP32.  Issue s pairs block b with home group b mod G (G the
// item's group count rounded up to a power of two) and window group (b + s)
// mod 16, and each lane queues its results in the rotated order (col + q)
// mod 4: the entries of one append then have distinct window slots and, for
// a full item, distinct owners.  LDS atomics of one drained round that hit
// one address from several lanes serialise: a chunk tiled 16 x 16 puts 16
// lanes on each neighbour, a plain 4 x 4 order four, and both were measured
// several times slower in the reaction sums.

constexpr unsigned int kIdxMask = (1u << 27) - 1u; // sorted index bits of a queue / window entry

// The half shell's neighbour cells, per lane, for reach 1 and 2: offsets
// (a, b, c) with a > 0, or a = 0 and b > 0, or a = b = 0 and c >= 0 (the own
// cell first), packed (a + 2) | (b + 2) << 3 | (c + 2) << 6; 0 past the
// last one (14 for reach 1, 63 for reach 2).  A table rather than lane
// arithmetic, so that nothing of it stays live across a kernel's items.
struct HalfShellTable {
    unsigned short v[2][kWave];
};
constexpr HalfShellTable make_half_shell() {
    HalfShellTable t{};
    for (int reach = 1; reach <= 2; reach++) {
        int lane = 0;
        for (int a = 0; a <= reach; a++)
            for (int b = (a == 0 ? 0 : -reach); b <= reach; b++)
                for (int c = (a == 0 && b == 0 ? 0 : -reach); c <= reach; c++)
                    t.v[reach - 1][lane++] = (unsigned short)((a + 2) | (b + 2) << 3 | (c + 2) << 6);
    }
    return t;
}
__constant__ HalfShellTable kHalfShell = make_half_shell();
// 127 pending (paired rounds) + the eight appends of one search step
constexpr int kHalfQueue = 10 * kWave;
constexpr unsigned int kNoShift = 13;              // shift_index(0, 0, 0)
// queue entry: owner o as o * 8 in bits 3..8, window slot in bits 19..24, so
// that the entry masked or shifted is the byte offset of an 8- or 16-byte
// element of either (e >> 16 = slot * 8, e >> 15 = slot * 16)
constexpr int kSlotShift = 19;
__device__ __forceinline__ unsigned int queue_entry(unsigned int o, unsigned int slot) {
    return o << 3 | slot << kSlotShift;
}

template <bool FP32> struct RsHalfShared {
    unsigned int qs[kHalfQueue]; // queue_entry(owner, window slot)
    double4 shift[FP32 ? 1 : 27]; // image shifts by shift index (the FP32 form computes them)
    double facc[3][kWave];       // the home particles' forces, each in kWave / owner-slots replicas
    double facc_nb[3][kWave];    // the window particles' share, f summed (the reaction is -f)
    halfx4 win_a[kWave];         // (-2 n, 1) in FP16, n = n' / side, n' relative to the home cell centre
    float win_n2[kWave];         // |n'|^2; 1e30 in empty slots
    unsigned int win_js[kWave];  // sorted index | shift index << 27
    int4 cells[kWave];           // the half shell's neighbour cells: range, first item, offsets
    // FP32 form: (n', q as FP32 bits), n' the fixed-point position of the
    // window particle's image relative to the home cell centre
    int4 win_q[FP32 ? kWave : 1];
    double tab[FP32 ? 1 : kErfcxTab];
};

template <bool FP32, bool PAIRED_ROUNDS>
__device__ __forceinline__ void
realspace_item_half(int home, int chunk, int cell0, const double4 *__restrict__ sorted,
                    const float4 *__restrict__ sorted32, const int4 *__restrict__ sortedq,
                    const int *__restrict__ start, const int *__restrict__ istart,
                    const float4 *__restrict__ bbox, int nc, int reach, double L, double rcut,
                    double alpha, long stride,
                    double *__restrict__ fs, RsHalfShared<FP32> &sh, double &en_tot) {
    const int lane = threadIdx.x;
    const double two_a_spi = 2.0 * alpha / sqrt_pi();
    const double side = L / nc;
    const float rskip = (float)(rcut * (1.0 + 1e-5) + 1e-5 * L);
    const float rskip2 = rskip * rskip;
    const int hb = start[cell0 + home], he = start[cell0 + home + 1];
    const int hz = home % nc, hy = (home / nc) % nc, hx = home / (nc * nc);
    const int i0 = hb + chunk * kWave;
    const int nh = min(kWave, he - i0);
    const int i = i0 + lane;
    const bool active = lane < nh;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        sh.facc[d][lane] = 0.0;
        sh.facc_nb[d][lane] = 0.0;
    }
    sh.win_a[lane] = halfx4{(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)1.0f};
    sh.win_n2[lane] = 1e30f;
    sh.win_js[lane] = 0u;
    if (!FP32 && lane < 27)
        sh.shift[lane] = make_double4((lane / 9 - 1) * L, ((lane / 3) % 3 - 1) * L, (lane % 3 - 1) * L, 0.0);
    const bool tiles = he - hb > kMfmaMinCell;
    // home groups of four rounded up to a power of two (2 ... 16)
    const int gh = (nh + 3) >> 2;
    const int G = gh <= 2 ? 2 : (gh <= 4 ? 4 : (gh <= 8 ? 8 : 16));
    // the home particle of this lane's matrix column (tiles) or of the lane
    const int hidx = tiles ? 4 * ((lane >> 2) & (G - 1)) + (lane & 3) : lane;
    const bool hval = hidx < nh;
    const unsigned int i_own = (unsigned)(i0 + hidx);
    // owner o of the lane evaluating a pair adds to facc[.][o + oslots r],
    // r the lane's low bits: a round holds up to 64 / oslots pairs per owner
    // (16 with the four home particles of the vector search), which the
    // replicas spread over distinct LDS words
    const int oslots = (tiles || nh > 4) ? 4 * G : 4;
    const int orep = kWave / oslots;
    const int oofs = oslots * (lane & (orep - 1));
    // candidate test in units of the cell side, FP16 operands: coordinates
    // within 2.5 sides carry at most 1.3e-3, so |h - n|^2 is within 1.3e-2
    // side^2 of its value; the test passes pairs up to rcut^2 + 0.02 side^2
    const float inv_side = (float)(1.0 / side);
    const float rho2m = (float)(rcut * rcut / (side * side)) + 0.02f;
    halfx4 bop = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)30000.0f};
    float hxf = 1e4f, hyf = 1e4f, hzf = 1e4f; // vector search: far away unless valid
    if (hval) {
        const float4 v = sorted32[i_own];
        const _Float16 hx16 = (_Float16)(-0.5f * v.x * inv_side), hy16 = (_Float16)(-0.5f * v.y * inv_side),
                       hz16 = (_Float16)(-0.5f * v.z * inv_side);
        hxf = (float)hx16;
        hyf = (float)hy16;
        hzf = (float)hz16;
        bop = halfx4{hx16, hy16, hz16, (_Float16)(hxf * hxf + hyf * hyf + hzf * hzf - rho2m)};
    }
    int count = 0;        // wave-uniform queue length
    int w = 0;            // filled window slots
    bool selfwin = false; // the window holds the item's own chunk
    float en32 = 0.0f;
    __syncthreads();

    // FP32 form: the pair difference d = h - n' is the exact integer
    // difference of the home particle h and the window particle's image n',
    // both in fixed point relative to the home cell centre (quantum
    // s = side 2^-fix_bits(reach); each rounded once from FP64 relative to its own
    // cell centre, whose offset from the home one is a whole number of
    // sides), scaled by s in FP32.  Per component |d~ - d| <= s + 2 u |d|, u = 2^-24, and r2f, the
    // FP32 sum of squares, is within (2 sqrt(3) s / rcut + 8 u) rcut^2 of r^2
    // near the cutoff.  The inclusion test is decided on r2f outside a band
    // of four times that width around rcut^2; inside it the reference's FP64
    // test runs on the FP64 positions.  The relative error of d is at most
    // 2 u + sqrt(3) s / r: 3.2e-7 for a pair side / 100 apart.  The pair
    // arithmetic runs in units of s: D = d / s, the window charge staged as
    // q / s^2, alpha and 2 alpha / sqrt(pi) taken times s, so that
    // (q_i q_j / s^2)(erfc(alpha s R) / R + 2 alpha s e^{-alpha^2 s^2 R^2} / sqrt(pi)) D / R^2
    // is the force, and the energy sum, in units of 1 / s, is scaled once.
    const int fb = fix_bits(reach);
    const double quantum = ldexp(side, -fb);
    const double band = 4.0 * (3.4642 * quantum / rcut + 8.0 * 0x1p-24);
    const double rq2 = (rcut / quantum) * (rcut / quantum);
    const float r2_in = (float)(rq2 * (1.0 - band)), r2_out = (float)(rq2 * (1.0 + band));
    const double alpha_q = alpha * quantum, two_a_spi_q = two_a_spi * quantum;
    // one queued pair: what a round reads before its arithmetic (FP32: the
    // window particle from LDS, the home particle from sortedq through L1;
    // FP64: both positions from global memory, through L1)
    struct Fetched {
        unsigned int o8, s8; // owner * 8, window slot * 8
        unsigned int js;
        double4 pj, ph;
        int4 nq, hq;
    };
    // the item's home particles (wave-uniform bases, byte offsets per lane)
    const int4 *hq_base = FP32 ? sortedq + i0 : sortedq;
    const double4 *ph_base = sorted + i0;
    auto win_js_at = [&](unsigned int s8) {
        return *reinterpret_cast<const unsigned int *>(reinterpret_cast<const char *>(sh.win_js) + (s8 >> 1));
    };
    auto fetch = [&](int base, Fetched &f) {
        const unsigned int e = sh.qs[base + lane];
        f.o8 = e & 0x1F8u;
        f.s8 = e >> (kSlotShift - 3);
        if constexpr (FP32) {
            f.nq = *reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(sh.win_q) +
                                                   (e >> (kSlotShift - 4)));
            f.hq = *reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(hq_base) + 2u * f.o8);
        } else {
            f.js = win_js_at(f.s8);
            f.pj = sorted[f.js & kIdxMask];
            f.ph = *reinterpret_cast<const double4 *>(reinterpret_cast<const char *>(ph_base) + 4u * f.o8);
        }
    };
    // the image shift of window entry js (0 unless the pair wraps)
    auto image_shift = [&](unsigned int js, double &sx, double &sy, double &sz) {
        sx = 0.0;
        sy = 0.0;
        sz = 0.0;
        const unsigned int sidx = js >> 27;
        if (sidx != kNoShift) {
            if constexpr (FP32) {
                sx = (double)((int)(sidx / 9u) - 1) * L;
                sy = (double)((int)((sidx / 3u) % 3u) - 1) * L;
                sz = (double)((int)(sidx % 3u) - 1) * L;
            } else {
                const double4 s3 = sh.shift[sidx];
                sx = s3.x;
                sy = s3.y;
                sz = s3.z;
            }
        }
    };
    auto evaluate = [&](const Fetched &f) {
        PairAcc pa = {0.0, 0.0, 0.0, 0.0, 0.0f};
        if constexpr (FP32) {
            const float dx = (float)(f.hq.x - f.nq.x), dy = (float)(f.hq.y - f.nq.y),
                        dz = (float)(f.hq.z - f.nq.z);
            const float r2f = dx * dx + dy * dy + dz * dz;
            bool inside = r2f <= r2_in;
            if (__builtin_expect(!inside && r2f <= r2_out, 0)) {
                const unsigned int js = win_js_at(f.s8);
                const double4 pj = sorted[js & kIdxMask],
                              ph = *reinterpret_cast<const double4 *>(reinterpret_cast<const char *>(ph_base) +
                                                                      4u * f.o8);
                double sx, sy, sz;
                image_shift(js, sx, sy, sz);
                double ex, ey, ez;
                {
#pragma clang fp contract(off)
                    ex = (ph.x - pj.x) - sx;
                    ey = (ph.y - pj.y) - sy;
                    ez = (ph.z - pj.z) - sz;
                }
                inside = within_cutoff(ex, ey, ez, rcut);
            }
            pair_force_f32(dx, dy, dz, r2f, inside, __int_as_float(f.hq.w) * __int_as_float(f.nq.w), alpha_q,
                           two_a_spi_q, pa);
        } else {
            double sx, sy, sz;
            image_shift(f.js, sx, sy, sz);
            pair<false>(f.ph, f.pj, sx, sy, sz, rcut, alpha, two_a_spi, sh.tab, pa);
        }
        auto at = [](double *row, unsigned int ofs) {
            return reinterpret_cast<double *>(reinterpret_cast<char *>(row) + ofs);
        };
        const unsigned int fo = f.o8 + 8u * (unsigned)oofs;
        atomicAdd(at(sh.facc[0], fo), pa.f0);
        atomicAdd(at(sh.facc[1], fo), pa.f1);
        atomicAdd(at(sh.facc[2], fo), pa.f2);
        atomicAdd(at(sh.facc_nb[0], f.s8), pa.f0);
        atomicAdd(at(sh.facc_nb[1], f.s8), pa.f1);
        atomicAdd(at(sh.facc_nb[2], f.s8), pa.f2);
        if (FP32)
            en32 += pa.enf;
        else
            en_tot += 2.0 * pa.en;
    };
    // evaluate the last m <= 64 queued pairs
    // (no LDS wait after a round: the workgroup is this one wavefront, whose
    // LDS accesses are performed in program order, and the queue entries read
    // were consumed; the accumulators are read only after the barrier that
    // ends the window)
    auto drain = [&](int m) {
        const int base = count - m;
        if (lane < m) {
            Fetched f;
            fetch(base, f);
            evaluate(f);
        }
        count = base;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    // two full rounds, the second one's reads issued before the first one's
    // arithmetic so that their latencies overlap
    auto drain2 = [&]() {
        Fetched fa, fb;
        fetch(count - kWave, fa);
        fetch(count - 2 * kWave, fb);
        evaluate(fa);
        evaluate(fb);
        count -= 2 * kWave;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    auto append = [&](bool ok, unsigned int entry) {
        const unsigned long long mask = __ballot(ok);
        if (mask) {
            if (ok) {
                const unsigned int slot = __builtin_amdgcn_mbcnt_hi(
                    (unsigned int)(mask >> 32),
                    __builtin_amdgcn_mbcnt_lo((unsigned int)mask, (unsigned int)count));
                sh.qs[slot] = entry;
            }
            count += __popcll(mask);
        }
    };
    // PAIRED_ROUNDS: two rounds at a time (single-system kernel: FP64
    // 11.4 ms against 12.0 at the Gaussian 1M workload, FP32 at four waves
    // per SIMD 7.07 ms against 7.35 at five with single rounds); the batched
    // FP64 kernel, held to 128 registers, has no room for the second round's
    // data
    auto flush = [&]() {
        if constexpr (!PAIRED_ROUNDS) {
            if (count >= kWave) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                do
                    drain(kWave);
                while (count >= kWave);
            }
        } else {
            if (count >= 2 * kWave) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                do
                    drain2();
                while (count >= 2 * kWave);
            }
        }
    };
    // the reference's exclusion in the own chunk: keep j > i only
    auto keep = [&](unsigned int js, unsigned int iown) {
        const unsigned int j = js & kIdxMask;
        return !(j - (unsigned)i0 < (unsigned)nh && j <= iown);
    };
    // candidate search of the window against the home particles, then all
    // queued pairs, then the reactions to fs; the window is emptied
    auto process_window = [&]() {
        __syncthreads();
        if (tiles) {
            // two issues at a time (G is even), so that their product chains overlap
            auto issues = [&](auto self_tag) {
                constexpr bool SELFW = decltype(self_tag)::value;
                // append q takes result register (col + q) mod 4 of each lane:
                // the 64 entries of one append then have 64 distinct window
                // slots and (G = 16) 64 distinct owners
                const int col = lane & 3;
                const bool c1 = col & 1, c2 = col & 2;
                auto emit = [&](const floatx4 &d, int base) {
                    // v[q] = d[(col + q) mod 4] by two selects per value
                    const float u0 = c1 ? d[1] : d[0], u1 = c1 ? d[2] : d[1], u2 = c1 ? d[3] : d[2],
                                u3 = c1 ? d[0] : d[3];
                    const float v[4] = {c2 ? u2 : u0, c2 ? u3 : u1, c2 ? u0 : u2, c2 ? u1 : u3};
                    const unsigned int e0 = queue_entry((unsigned)hidx, (unsigned)base);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const unsigned int rq = (unsigned)((col + q) & 3);
                        bool ok = v[q] <= 0.0f;
                        if (SELFW)
                            ok = ok && keep(sh.win_js[base + rq], i_own);
                        append(ok, e0 + (rq << kSlotShift));
                    }
                };
                for (int s = 0; s < G; s += 2) {
                    const int base0 = ((lane & ~3) + 4 * s) & 63, base1 = (base0 + 4) & 63;
                    const halfx4 a0 = sh.win_a[(lane + 4 * s) & 63];
                    const halfx4 a1 = sh.win_a[(lane + 4 * s + 4) & 63];
                    const float4 c0 = *reinterpret_cast<const float4 *>(&sh.win_n2[base0]);
                    const float4 c1 = *reinterpret_cast<const float4 *>(&sh.win_n2[base1]);
                    floatx4 d0 = {c0.x, c0.y, c0.z, c0.w}, d1 = {c1.x, c1.y, c1.z, c1.w};
                    d0 = __builtin_amdgcn_mfma_f32_4x4x4f16(a0, bop, d0, 0, 0, 0);
                    d1 = __builtin_amdgcn_mfma_f32_4x4x4f16(a1, bop, d1, 0, 0, 0);
                    emit(d0, base0);
                    emit(d1, base1);
                    // one drain site per step keeps the kernel's code small
                    flush();
                }
            };
            if (selfwin)
                issues(std::true_type{});
            else
                issues(std::false_type{});
        } else {
            // vector search (at most four home particles): the lanes take
            // the window slots, pass r tests slot l against home particle
            // (l + r) mod nh, so that an append has distinct slots and spreads
            // evenly over the home particles
            const halfx4 a = sh.win_a[lane];
            const unsigned int js = sh.win_js[lane];
            const bool filled = lane < w;
            int o = lane % nh;
            for (int r = 0; r < nh; r++) {
                const float dx = fmaf(0.5f, (float)a.x, __shfl(hxf, o, kWave)),
                            dy = fmaf(0.5f, (float)a.y, __shfl(hyf, o, kWave)),
                            dz = fmaf(0.5f, (float)a.z, __shfl(hzf, o, kWave));
                bool ok = filled && dx * dx + dy * dy + dz * dz <= rho2m;
                if (selfwin)
                    ok = ok && keep(js, (unsigned)(i0 + o));
                append(ok, queue_entry((unsigned)o, (unsigned)lane));
                o = (o + 1 == nh) ? 0 : o + 1;
            }
        }
        __syncthreads();
        flush();
        if (count >= kWave)
            drain(kWave);
        if (count > 0)
            drain(count);
        if (lane < w) {
            const unsigned int j = sh.win_js[lane] & kIdxMask;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = sh.facc_nb[d][lane];
                if (v != 0.0)
                    atomicAdd(&fs[d * stride + j], -v);
                sh.facc_nb[d][lane] = 0.0;
            }
        }
        sh.win_n2[lane] = 1e30f;
        w = 0;
        selfwin = false;
        __syncthreads();
    };

    const int hitem = istart[cell0 + home] + chunk;
    const float4 hlo = bbox[2 * hitem], hhi = bbox[2 * hitem + 1];
    // the item's box relative to the home cell centre (the window's frame)
    const float hcx = (float)((hx + 0.5) * side), hcy = (float)((hy + 0.5) * side),
                hcz = (float)((hz + 0.5) * side);
    const float4 hlo_r = make_float4(hlo.x - hcx, hlo.y - hcy, hlo.z - hcz, 0.0f);
    const float4 hhi_r = make_float4(hhi.x - hcx, hhi.y - hcy, hhi.z - hcz, 0.0f);
    // the half shell's neighbour cells, one per lane (63 for reach 2, 14 for
    // reach 1; the own cell first): offsets, image flags, particle range and
    // first item read by one vector load each, cells that are empty or beyond
    // rcut from the item's box dropped; the loop then walks the live ones
    bool live = false;
    int cjb = 0, cje = 0, citem = 0;
    unsigned int cinfo = 0;
    const unsigned int abc = kHalfShell.v[reach - 1][lane];
    if (abc != 0u) {
        const int a = (int)(abc & 7u) - 2, b = (int)((abc >> 3) & 7u) - 2, c = (int)((abc >> 6) & 7u) - 2;
        int cx = hx + a, cy = hy + b, cz = hz + c, wx = 0, wy = 0, wz = 0;
        if (cx >= nc) { cx -= nc; wx = 1; }
        if (cy < 0) { cy += nc; wy = -1; } else if (cy >= nc) { cy -= nc; wy = 1; }
        if (cz < 0) { cz += nc; wz = -1; } else if (cz >= nc) { cz -= nc; wz = 1; }
        // the neighbour cell's image spans [(h + a) side, (h + a + 1) side] per axis
        const float gx = gap(hlo.x, hhi.x, (float)((hx + a) * side), (float)((hx + a + 1) * side));
        const float gy = gap(hlo.y, hhi.y, (float)((hy + b) * side), (float)((hy + b + 1) * side));
        const float gz = gap(hlo.z, hhi.z, (float)((hz + c) * side), (float)((hz + c + 1) * side));
        if (gx * gx + gy * gy + gz * gz <= rskip2) {
            const int cell = cell0 + (cx * nc + cy) * nc + cz;
            cjb = start[cell];
            cje = start[cell + 1];
            citem = istart[cell];
            live = cjb < cje;
        }
        cinfo = (unsigned)((a + 2) | (b + 2) << 3 | (c + 2) << 6 | (wx + 1) << 9 | (wy + 1) << 11 |
                           (wz + 1) << 13);
    }
    // kept in LDS for the walk (read back uniformly), not in registers
    sh.cells[lane] = make_int4(cjb, cje, citem, (int)cinfo);
    unsigned long long cells = __ballot(live);
    while (cells) {
        const int k = __builtin_ctzll(cells);
        cells &= cells - 1;
        {
            {
                const int4 ck = sh.cells[k];
                const int jb = __builtin_amdgcn_readfirstlane(ck.x);
                const int je = __builtin_amdgcn_readfirstlane(ck.y);
                const int item0 = __builtin_amdgcn_readfirstlane(ck.z);
                const unsigned int info = (unsigned)__builtin_amdgcn_readfirstlane(ck.w);
                const int a = (int)(info & 7u) - 2, b = (int)((info >> 3) & 7u) - 2,
                          c = (int)((info >> 6) & 7u) - 2;
                const int wx = (int)((info >> 9) & 3u) - 1, wy = (int)((info >> 11) & 3u) - 1,
                          wz = (int)((info >> 13) & 3u) - 1;
                const unsigned int sidx = shift_index(wx, wy, wz) << 27;
                const float fsx = (float)(wx * L), fsy = (float)(wy * L), fsz = (float)(wz * L);
                // FP32 form: the neighbour cell's centre relative to the home
                // cell's is (a, b, c) sides, exactly 2^fb quanta each
                // (the image shift w L is the whole-mesh part of it), so a
                // window particle's fixed-point position in the home frame is
                // its own-cell one plus that integer offset
                const int fox = a * (1 << fb), foy = b * (1 << fb), foz = c * (1 << fb);
                const double inv_quantum = ldexp(1.0, fb) / side;
                const float q_scale = (float)(inv_quantum * inv_quantum);
                // window rows: A' = (-2 (n + o), 1), o the cell centre relative to the home's
                const float ox2 = (float)(2.0 * a * side), oy2 = (float)(2.0 * b * side),
                            oz2 = (float)(2.0 * c * side);
                const bool own = a == 0 && b == 0 && c == 0;
                for (int cb = own ? i0 : jb; cb < je; cb += kWave) {
                    const int cnt = min(kWave, je - cb);
                    const int nitem = item0 + (cb - jb) / kWave;
                    float4 nlo = bbox[2 * nitem], nhi = bbox[2 * nitem + 1];
                    nlo.x += fsx; nlo.y += fsy; nlo.z += fsz;
                    nhi.x += fsx; nhi.y += fsy; nhi.z += fsz;
                    if (box_dist2(hlo, hhi, nlo, nhi) > rskip2)
                        continue;
                    if (own && cb == i0)
                        selfwin = true; // the first chunk staged: w == 0 here
                    // the chunk's particles within rcut of the item's box (all of
                    // the own chunk), compacted into the window
                    bool keep_l = false;
                    halfx4 arow = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)1.0f};
                    float n2 = 0.0f;
                    int4 nq = make_int4(0, 0, 0, 0);
                    if (lane < cnt) {
                        const float4 v = sorted32[cb + lane];
                        if constexpr (FP32) {
                            const int4 q0 = sortedq[cb + lane];
                            nq = make_int4(q0.x + fox, q0.y + foy, q0.z + foz,
                                           __float_as_int(__int_as_float(q0.w) * q_scale));
                        }
                        const float ax = v.x - ox2, ay = v.y - oy2, az = v.z - oz2;
                        const float nx = -0.5f * ax, ny = -0.5f * ay, nz = -0.5f * az;
                        const float gxx = gap(hlo_r.x, hhi_r.x, nx, nx), gyy = gap(hlo_r.y, hhi_r.y, ny, ny),
                                    gzz = gap(hlo_r.z, hhi_r.z, nz, nz);
                        keep_l = gxx * gxx + gyy * gyy + gzz * gzz <= rskip2;
                        // the matrix row in side units, rounded to FP16, and |n|^2
                        // of the rounded row
                        arow = halfx4{(_Float16)(ax * inv_side), (_Float16)(ay * inv_side),
                                      (_Float16)(az * inv_side), (_Float16)1.0f};
                        const float rx = -0.5f * (float)arow.x, ry = -0.5f * (float)arow.y,
                                    rz = -0.5f * (float)arow.z;
                        n2 = rx * rx + ry * ry + rz * rz;
                    }
                    unsigned long long mask = __ballot(keep_l);
                    while (mask) {
                        const int take = min(kWave - w, __popcll(mask));
                        const int rank = (int)__builtin_amdgcn_mbcnt_hi(
                            (unsigned int)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
                        if (keep_l && rank < take) {
                            sh.win_a[w + rank] = arow;
                            sh.win_n2[w + rank] = n2;
                            sh.win_js[w + rank] = (unsigned)(cb + lane) | sidx;
                            if constexpr (FP32)
                                sh.win_q[w + rank] = nq;
                        }
                        w += take;
                        keep_l = keep_l && rank >= take;
                        mask = __ballot(keep_l);
                        if (w == kWave)
                            process_window();
                    }
                }
            }
        }
    }
    if (w > 0)
        process_window();
    if (FP32)
        en_tot += quantum * (double)en32;
    if (active) {
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        for (int r = 0; r < orep; r++) {
            f0 += sh.facc[0][lane + oslots * r];
            f1 += sh.facc[1][lane + oslots * r];
            f2 += sh.facc[2][lane + oslots * r];
        }
        atomicAdd(&fs[i], f0);
        atomicAdd(&fs[stride + i], f1);
        atomicAdd(&fs[2 * stride + i], f2);
    }
    __syncthreads();
}

// Persistent grid as k_realspace over the half shell; fs (3 n, sorted order)
// is zeroed by the caller and added to the forces by k_unsort_add.  Queue
// rounds are evaluated two at a time in both forms; FP32: four waves per
// SIMD (7.07 ms at the Gaussian 1M workload; five with single rounds 7.35,
// five with paired rounds 8.06 from spills); FP64: three (12.0 ms against
// 13.3 unconstrained and 12.0 at four).
template <bool FP32>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(FP32 ? 4 : 3)))
k_realspace_half(const double4 *__restrict__ sorted, const float4 *__restrict__ sorted32,
                 const int4 *__restrict__ sortedq, const int *__restrict__ start, const int *__restrict__ istart,
                 const float4 *__restrict__ bbox, long ncell, int nc, int reach, double L,
                 double rcut, double alpha, int *__restrict__ counter, long stride,
                 double *__restrict__ fs, double *__restrict__ energy) {
    __shared__ RsHalfShared<FP32> sh;
    const int total = istart[ncell];
    double en_tot = 0.0;
    if (!FP32) {
        for (int k = threadIdx.x; k < kErfcxTab; k += blockDim.x)
            sh.tab[k] = (&p3m::kErfcxDoubleCoef[0][0])[k];
    }
    __syncthreads();
    const int per = claim_size(total);
    int next = 0, end = 0;
    for (;;) {
        const int item = next_item(counter, total, per, next, end);
        if (item >= total)
            break;
        // the item's cell and chunk, stored with its box
        const int home = __float_as_int(bbox[2 * item].w), chunk = __float_as_int(bbox[2 * item + 1].w);
        realspace_item_half<FP32, true>(home, chunk, 0, sorted, sorted32, sortedq, start, istart, bbox, nc, reach, L,
                                        rcut, alpha, stride, fs, sh, en_tot);
    }
    if (energy != nullptr) {
        for (int off = 32; off > 0; off >>= 1)
            en_tot += __shfl_down(en_tot, off, kWave);
        if (threadIdx.x == 0 && en_tot != 0.0)
            atomicAdd(energy, en_tot);
    }
}

// fr[d stride + i] += fs[d n + pos[i]]: sorted-order forces to the caller's
// order, read through the inverse permutation so that the caller's array is
// updated in order
__global__ void k_unsort_add(int n, const int *__restrict__ pos, const double *__restrict__ fs,
                             long stride, double *__restrict__ fr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int p = pos[i];
    fr[i] += fs[p];
    fr[stride + i] += fs[(long)n + p];
    fr[2 * stride + i] += fs[2 * (long)n + p];
}

}  // namespace


namespace p3m {

void cells_release(CellArrays &a) {
    // counter lives in the fcount allocation
    void *ptrs[] = {a.key, a.fcount, a.fstart, a.start, a.items, a.istart, a.perm,
                    a.scan, a.sorted, a.sorted32, a.sortedq, a.bbox, a.gsorted, a.fsorted, a.rank};
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
    // fcount and the work-item counters in one allocation, cleared by one
    // fill: [fcount | counters], fcount a multiple of 16 bytes
    const size_t nf4 = (nf + 3) & ~(size_t)3;
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.fcount, sizeof(int) * (nf4 + kCounters));
    if (rc == P3MK_OK)
        a.counter = a.fcount + nf4;
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.fstart, sizeof(int) * nf);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.start, sizeof(int) * nc1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.items, sizeof(int) * nc1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.istart, sizeof(int) * nc1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.perm, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.rank, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.scan, sizeof(int) * p3mk_scan_scratch_ints(8 * ncell));
    const size_t np = nn + kPad;
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.sorted, sizeof(double4) * np);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.sorted32, sizeof(float4) * np);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.sortedq, sizeof(int4) * np);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.bbox, sizeof(float4) * 2 * (size_t)a.max_items);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&a.fsorted, sizeof(double) * 3 * nn);
    if (rc != P3MK_OK)
        cells_release(a);
    return rc;
}

// zero the octant counts and the work-item counters (one allocation) before
// the keys are counted
static int cells_begin(CellArrays &a, long ncell, p3mk_queue *qu) {
    (void)ncell;
    return p3mk_fill_zero(a.fcount, sizeof(int) * (size_t)(a.counter + kCounters - a.fcount), qu);
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
    hipLaunchKernelGGL(k_item_bbox, dim3((unsigned)((a.max_items * kBoxLanes + 255) / 256)), dim3(256), 0, s,
                       a.start, a.istart, ncell, a.max_items, a.sorted, a.bbox);
    return p3mk_launch_check("item boxes");
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
    if ((long)n > p3m::kCellMaxParticles)
        return p3mk_fail(P3MK_EINVAL, "real space supports at most 2^27 - 64 particles", 0);
    p3mk_cells *c = new (std::nothrow) p3mk_cells();
    if (!c)
        return p3mk_fail(P3MK_ENOMEM, "cell list", 0);
    c->nmax = n;
    int nc, reach;
    p3m::cell_geometry(n, L, rcut, nc, reach);
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

// sort the particles into the cell list (with g: also the cotangent)
static int build_cells(p3mk_cells *c, int n, double L, const double *x, const double *y,
                       const double *z, const double *q, double rcut, const double *g,
                       p3mk_queue *qu) {
    if (n > c->nmax || L != c->L || rcut != c->rcut)
        return p3mk_fail(P3MK_EINVAL, "cell list was built for another system size, box or cutoff", 0);
    CellArrays &a = c->a;
    hipStream_t s = p3mk_stream(qu);
    const int nb = (n + 255) / 256, nbp = (n + kPad + 255) / 256;
    P3MK_TRY(p3m::cells_begin(a, c->ncell, qu));
    hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, s, n, x, y, z, L, c->inv_side, c->nc,
                       a.key, a.rank, a.fcount);
    P3MK_TRY(p3mk_launch_check("cell count"));
    P3MK_TRY(p3m::cells_scan(a, c->ncell, qu));
    hipLaunchKernelGGL(k_cell_scatter, dim3(nbp), dim3(256), 0, s, n, x, y, z, q, L, c->nc,
                       fix_bits(c->reach), a.key, a.fstart, a.rank, a.perm, a.sorted, a.sorted32,
                       a.sortedq, g, a.gsorted);
    P3MK_TRY(p3mk_launch_check("cell scatter"));
    return p3m::cells_end(a, c->ncell, qu);
}

extern "C" int p3mk_realspace_vjp(p3mk_cells *c, int n, double L, const double *x, const double *y,
                                  const double *z, const double *q, double rcut, double alpha,
                                  const double *g, double *dpos, double *dq, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    CellArrays &a = c->a;
    if (a.gsorted == nullptr)
        P3MK_TRY(p3mk_alloc((void **)&a.gsorted, sizeof(double4) * ((size_t)c->nmax + kPad)));
    P3MK_TRY(build_cells(c, n, L, x, y, z, q, rcut, g, qu));
    const bool all_cells = (2 * c->reach + 1 > c->nc);
    const dim3 grid((unsigned)c->grid);
    hipStream_t s = p3mk_stream(qu);
    if (all_cells)
        hipLaunchKernelGGL(k_realspace_vjp<true>, grid, dim3(kWave), 0, s, n, a.sorted, a.sorted32,
                           a.perm, a.start, a.istart, a.bbox, c->ncell, c->nc, c->reach, L,
                           rcut, alpha, a.counter, a.gsorted, dpos, dq);
    else
        hipLaunchKernelGGL(k_realspace_vjp<false>, grid, dim3(kWave), 0, s, n, a.sorted, a.sorted32,
                           a.perm, a.start, a.istart, a.bbox, c->ncell, c->nc, c->reach, L,
                           rcut, alpha, a.counter, a.gsorted, dpos, dq);
    return p3mk_launch_check("real-space adjoint");
}

extern "C" int p3mk_realspace(p3mk_cells *c, int n, double L, const double *x, const double *y,
                              const double *z, const double *q, double rcut, double alpha,
                              int fp32, double *fr, double *energy, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    CellArrays &a = c->a;
    hipStream_t s = p3mk_stream(qu);
    const bool all_cells = (2 * c->reach + 1 > c->nc);
    const dim3 grid((unsigned)c->grid);
    if (!all_cells) {
        // half shell: forces summed in sorted order, then added by perm
        P3MK_TRY(build_cells(c, n, L, x, y, z, q, rcut, nullptr, qu));
        P3MK_TRY(p3mk_fill_zero(a.fsorted, sizeof(double) * 3 * (size_t)n, qu));
        // the pair kernel alone between two events
        P3MK_TRY(p3mk_record(qu, P3MK_EV_PAIRS));
        if (fp32)
            hipLaunchKernelGGL(k_realspace_half<true>, grid, dim3(kWave), 0, s, a.sorted, a.sorted32,
                               a.sortedq, a.start, a.istart, a.bbox, c->ncell, c->nc, c->reach, L,
                               rcut, alpha, a.counter, (long)n, a.fsorted, energy);
        else
            hipLaunchKernelGGL(k_realspace_half<false>, grid, dim3(kWave), 0, s, a.sorted, a.sorted32,
                               a.sortedq, a.start, a.istart, a.bbox, c->ncell, c->nc, c->reach, L,
                               rcut, alpha, a.counter, (long)n, a.fsorted, energy);
        P3MK_TRY(p3mk_launch_check("realspace (half shell)"));
        P3MK_TRY(p3mk_record(qu, P3MK_EV_PAIRS + 1));
        hipLaunchKernelGGL(k_unsort_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a.key,
                           a.fsorted, (long)n, fr);
        return p3mk_launch_check("realspace forces to caller order");
    }
    P3MK_TRY(build_cells(c, n, L, x, y, z, q, rcut, nullptr, qu));
#define LAUNCH(F32, SB)                                                                            \
    hipLaunchKernelGGL((k_realspace<F32, SB>), grid, dim3(kWave), 0, s, n, a.sorted, a.sorted32,   \
                       a.perm,                                                                     \
                       a.start, a.istart, a.bbox, c->ncell, c->nc, c->reach, L, rcut, alpha,        \
                       a.counter, fr, energy)
    P3MK_TRY(p3mk_record(qu, P3MK_EV_PAIRS));
    if (fp32)
        LAUNCH(true, true);
    else
        LAUNCH(false, true);
#undef LAUNCH
    P3MK_TRY(p3mk_launch_check("realspace"));
    return p3mk_record(qu, P3MK_EV_PAIRS + 1);
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
                              int *__restrict__ rank, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot)
        return;
    const BSys &S = sys[psys[i]];
    const double L = S.L;
    const int k = 8 * S.cell0 + fine_key(fold(x[i], L), fold(y[i], L), fold(z[i], L), (double)S.nc / L, S.nc);
    key[i] = k;
    rank[i] = atomicAdd(&count[k], 1);
}

// as k_cell_scatter: key[i] becomes the particle's sorted position
__global__ void k_bcell_scatter(int ntot, const BSys *__restrict__ sys, const int *__restrict__ psys,
                                const double *__restrict__ x, const double *__restrict__ y,
                                const double *__restrict__ z, const double *__restrict__ q,
                                int *__restrict__ key, const int *__restrict__ start,
                                const int *__restrict__ rank, int *__restrict__ perm,
                                double4 *__restrict__ sorted, float4 *__restrict__ sorted32,
                                int4 *__restrict__ sortedq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot + kPad)
        return;
    if (i >= ntot) {
        store_pad(i, sorted, sorted32, sortedq);
        return;
    }
    const BSys &S = sys[psys[i]];
    const double L = S.L;
    const int k = key[i];
    const int pos = start[k] + rank[i];
    perm[pos] = i;
    key[i] = pos;
    const double4 r = make_double4(fold(x[i], L), fold(y[i], L), fold(z[i], L), q[i]);
    store_sorted(pos, r, (k >> 3) - S.cell0, S.nc, L / S.nc, fix_bits(S.reach), sorted, sorted32, sortedq);
}

template <bool FP32>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(FP32 ? 5 : 1)))
k_brealspace(const BSys *__restrict__ sys, const int *__restrict__ cell0s, int nsys, int ntot,
             const double4 *__restrict__ sorted, const float4 *__restrict__ sorted32,
             const int *__restrict__ perm,
             const int *__restrict__ start, const int *__restrict__ istart,
             const float4 *__restrict__ bbox, long ncell, int *__restrict__ counter,
             double *__restrict__ fr, double *__restrict__ energy) {
    __shared__ RsShared<FP32> sh;
    const int total = istart[ncell];
    load_erfcx_table(sh);
    const int per = claim_size(total);
    int next = 0, end = 0;
    for (;;) {
        const int item = next_item(counter, total, per, next, end);
        if (item >= total)
            break;
        const int home = __float_as_int(bbox[2 * item].w), chunk = __float_as_int(bbox[2 * item + 1].w);
        const int s = p3m::fft::prefix_slot(cell0s, nsys, (long)home);
        const BSys &S = sys[s];
        if (!(S.rcut > 0.0) || !S.small)
            continue; // no real-space part, or a cell grid (k_brealspace_half)
        double en = 0.0;
        realspace_item<FP32, true>(home - S.cell0, chunk, S.cell0, sorted, sorted32,
                                   perm, start, istart, bbox, S.nc, S.reach, S.L, S.rcut, S.alpha, ntot,
                                   fr, sh, en);
        if (energy != nullptr) {
            for (int off = 32; off > 0; off >>= 1)
                en += __shfl_down(en, off, kWave);
            if (threadIdx.x == 0 && en != 0.0)
                atomicAdd(&energy[s], en);
        }
    }
}

// Cell-grid systems of a batch over the half shell: forces to fs (3 ntot,
// sorted order, zeroed by the caller), energy[s] += system s's energy.
// Four waves per SIMD: the batch's sparse cells are latency bound (FP64:
// 21.1 ms against 29.1 ms unconstrained at the 256-system batch); the FP32
// form evaluates its queue rounds two at a time, the FP64 form one.
template <bool FP32>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4)))
k_brealspace_half(const BSys *__restrict__ sys, const int *__restrict__ cell0s, int nsys, int ntot,
                  const double4 *__restrict__ sorted, const float4 *__restrict__ sorted32,
                  const int4 *__restrict__ sortedq, const int *__restrict__ start,
                  const int *__restrict__ istart, const float4 *__restrict__ bbox, long ncell,
                  int *__restrict__ counter,
                  double *__restrict__ fs, double *__restrict__ energy) {
    __shared__ RsHalfShared<FP32> sh;
    const int total = istart[ncell];
    if (!FP32) {
        for (int k = threadIdx.x; k < kErfcxTab; k += blockDim.x)
            sh.tab[k] = (&p3m::kErfcxDoubleCoef[0][0])[k];
    }
    __syncthreads();
    const int per = claim_size(total);
    int next = 0, end = 0;
    for (;;) {
        const int item = next_item(counter, total, per, next, end);
        if (item >= total)
            break;
        const int home = __float_as_int(bbox[2 * item].w), chunk = __float_as_int(bbox[2 * item + 1].w);
        const int s = p3m::fft::prefix_slot(cell0s, nsys, (long)home);
        const BSys &S = sys[s];
        if (!(S.rcut > 0.0) || S.small)
            continue; // no real-space part, or nearest images (k_brealspace)
        double en = 0.0;
        realspace_item_half<FP32, FP32>(home - S.cell0, chunk, S.cell0, sorted, sorted32, sortedq, start,
                                  istart, bbox, S.nc, S.reach, S.L, S.rcut, S.alpha, ntot, fs, sh, en);
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
               const double *q, CellArrays &a, int grid, bool any_small, bool any_cells,
               bool time_pairs, double *fr, double *energy, p3mk_queue *qu) {
    if (ntot <= 0)
        return P3MK_OK;
    hipStream_t s = p3mk_stream(qu);
    const int nb = (ntot + 255) / 256;
    P3MK_TRY(cells_begin(a, ncell, qu));
    hipLaunchKernelGGL(k_bcell_count, dim3(nb), dim3(256), 0, s, ntot, sys, psys, x, y, z, a.key,
                       a.rank, a.fcount);
    P3MK_TRY(p3mk_launch_check("batched cell count"));
    P3MK_TRY(cells_scan(a, ncell, qu));
    hipLaunchKernelGGL(k_bcell_scatter, dim3((ntot + kPad + 255) / 256), dim3(256), 0, s, ntot, sys,
                       psys, x, y, z, q, a.key, a.fstart, a.rank, a.perm, a.sorted, a.sorted32,
                       a.sortedq);
    P3MK_TRY(p3mk_launch_check("batched cell scatter"));
    P3MK_TRY(cells_end(a, ncell, qu));
    if (any_small) {
        // nearest-image systems: forces straight to fr
        if (f.fp32)
            hipLaunchKernelGGL(k_brealspace<true>, dim3((unsigned)grid), dim3(kWave), 0, s, sys, cell0s,
                               nsys, ntot, a.sorted, a.sorted32, a.perm, a.start, a.istart,
                               a.bbox, ncell, a.counter, fr, energy);
        else
            hipLaunchKernelGGL(k_brealspace<false>, dim3((unsigned)grid), dim3(kWave), 0, s, sys, cell0s,
                               nsys, ntot, a.sorted, a.sorted32, a.perm, a.start, a.istart,
                               a.bbox, ncell, a.counter, fr, energy);
        P3MK_TRY(p3mk_launch_check("batched real space (nearest images)"));
    }
    if (any_cells) {
        // cell-grid systems: half shell in sorted order, then added by perm
        P3MK_TRY(p3mk_fill_zero(a.fsorted, sizeof(double) * 3 * (size_t)ntot, qu));
        // the half-shell kernel alone between two events, on request (not
        // in captured graphs)
        if (time_pairs)
            P3MK_TRY(p3mk_record(qu, P3MK_EV_PAIRS));
        if (f.fp32)
            hipLaunchKernelGGL(k_brealspace_half<true>, dim3((unsigned)grid), dim3(kWave), 0, s, sys,
                               cell0s, nsys, ntot, a.sorted, a.sorted32, a.sortedq, a.start, a.istart, a.bbox, ncell,
                               a.counter + 1, a.fsorted, energy);
        else
            hipLaunchKernelGGL(k_brealspace_half<false>, dim3((unsigned)grid), dim3(kWave), 0, s, sys,
                               cell0s, nsys, ntot, a.sorted, a.sorted32, a.sortedq, a.start, a.istart, a.bbox, ncell,
                               a.counter + 1, a.fsorted, energy);
        P3MK_TRY(p3mk_launch_check("batched real space (half shell)"));
        if (time_pairs)
            P3MK_TRY(p3mk_record(qu, P3MK_EV_PAIRS + 1));
        hipLaunchKernelGGL(k_unsort_add, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, s, ntot,
                           a.key, a.fsorted, (long)ntot, fr);
        P3MK_TRY(p3mk_launch_check("batched real-space forces to caller order"));
    }
    return P3MK_OK;
}

}  // namespace p3m
