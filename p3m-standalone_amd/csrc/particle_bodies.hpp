// Per-item / per-particle bodies of charge assignment and back-interpolation,
// shared by the single-system kernels and the batched kernels.  Particles
// are in tile-sorted order (double4 x, y, z, q) with `perm` mapping a sorted
// slot to the caller's index.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace p3m {
namespace part {

constexpr int kThreads = 256;
// particles per assignment / interpolation work item: 512 balances the dense
// tiles of a clustered system over the GPU (Gaussian 1M, assignment +
// interpolation mixed: 1024 0.76 ms, 512 0.67, 256 0.67, 128 0.74)
constexpr int kChunk = 512;
// threads of a row-deposit workgroup: its waves share one FP64 region in LDS,
// so larger workgroups keep more waves per CU within the LDS (assignment,
// mixed: Gaussian 1M 0.496 ms at 256, 0.487 at 512, 0.615 at 1024;
// uniform 100k 0.078, 0.068, 0.068)
constexpr int kAssignThreads = 512;

// work item -> tile: the last tile whose chunk prefix sum is <= item
__device__ __forceinline__ long item_tile(const int *__restrict__ cstart, long ntiles, int item) {
    long lo = 0, hi = ntiles;
    while (hi - lo > 1) {
        const long mid = (lo + hi) / 2;
        if (cstart[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// What a particle deposits (MODE): kCharge q W; kWeighted q g_i W with a
// per-particle factor g (the adjoint's q g_d); kDipole
// q sum_d g_{i,d} W^(d), W^(d) = W with the d-axis weight replaced by its
// derivative (the adjoint of ad back-interpolation).
enum { kCharge = 0, kWeighted = 1, kDipole = 2 };

// ---- row-parallel halo-tile assignment ----------------------------------------
//
// One mesh at a time, a work item (<= kChunk particles of one tile) deposits
// into an LDS copy of the tile's halo region.  Per round each wavefront
// stages kRowBatch particles (a lane each: quantised weights of the three
// axes, q folded into the x weights, local base offsets) in LDS; then the
// wavefront deposits them one after another with its lanes spread over the
// stencil: lane l takes stencil row (a0, a1) = ((l mod P^2) / P, l mod P) of
// particle l / P^2 of the current group (64 / P^2 particles per step) and
// adds its P points along z.  The lanes of one particle hit distinct
// addresses, and the region's strides (z 1, y Sy odd, x Sx = P Sy mod 64)
// spread the P^2 rows of a step over the LDS banks.  Consecutive particles
// with the same base (a tile's particles are ordered by base) are summed in
// the lanes' registers, in the accumulator's precision, and added to LDS
// once per run.  The products are the per-particle form's ((q w0) w1) w2.
// Charge (MODE kCharge) and
// charge-times-factor (kWeighted) deposits; the dipole deposit of the ad
// adjoint keeps assign_item.

constexpr int kRowBatch = 32; // particles staged per wavefront and round

struct RowRegion {
    int R, Sy, Sx, size; // side, y and x strides (z stride 1), points incl. padding
};

__host__ __device__ inline RowRegion row_region(int Tt, int P, bool il) {
    RowRegion g;
    g.R = Tt + P - 1 + (il ? 1 : 0);
    g.Sy = g.R + ((g.R & 1) ? 0 : 1);
    const int want = (P * g.Sy) & 63, base = g.R * g.Sy;
    g.Sx = base + ((want - (base & 63)) & 63);
    g.size = (g.R * g.Sx + 1) & ~1; // even: the staging area after it stays 16-byte aligned
    return g;
}

// staged record of one particle: z weights w2 in a block of row_w2<P>()
// values (read by 16-byte loads), then q w0[P] and w1[P]; records are
// 16-byte aligned
__host__ __device__ constexpr int row_w2_n(int P) { return (P + 3) & ~3; }
__host__ __device__ constexpr int row_record_n(int P) { return (row_w2_n(P) + 2 * P + 3) & ~3; }
template <int P> __host__ __device__ constexpr int row_w2() { return row_w2_n(P); }
template <int P> __host__ __device__ constexpr int row_record() { return row_record_n(P); }

// LDS bytes of one wavefront's staging area: kRowBatch records of T and the
// particles' region offsets
__host__ __device__ constexpr int row_stage_bytes_n(int P, int tsize) {
    return kRowBatch * (row_record_n(P) * tsize + (int)sizeof(int));
}
template <int P, typename T> __host__ __device__ constexpr int row_stage_bytes() {
    return row_stage_bytes_n(P, (int)sizeof(T));
}

// n values of a 16-byte-aligned LDS row by 16-byte loads
template <typename T, int n> __device__ __forceinline__ void load_row(const T *src, T *dst) {
    if constexpr (sizeof(T) == 4) {
        const float4 *s = reinterpret_cast<const float4 *>(src);
#pragma unroll
        for (int k = 0; k < n / 4; k++) {
            const float4 v = s[k];
            dst[4 * k] = v.x;
            dst[4 * k + 1] = v.y;
            dst[4 * k + 2] = v.z;
            dst[4 * k + 3] = v.w;
        }
    } else {
        const double2 *s = reinterpret_cast<const double2 *>(src);
#pragma unroll
        for (int k = 0; k < n / 2; k++) {
            const double2 v = s[k];
            dst[2 * k] = v.x;
            dst[2 * k + 1] = v.y;
        }
    }
}

template <int P, typename T, bool R2C, bool IL, int MODE, typename A>
__device__ __forceinline__ void
assign_rows_item(const double4 *__restrict__ sorted, int pb, int pe, long tl, int nt, int Tt, int M,
                 double Hi, T *__restrict__ mesh, const int *__restrict__ perm,
                 const double *__restrict__ gv, A *acc, T *stage) {
    static_assert(MODE == kCharge || MODE == kWeighted, "row assignment deposits charges");
    constexpr int NM = IL ? 2 : 1;
    constexpr int REC = row_record<P>();
    constexpr int W2 = row_w2<P>();
    constexpr int P2 = P * P;
    constexpr int G = 64 / P2; // particles per deposit step
    const RowRegion rg = row_region(Tt, P, IL);
    const int tz = (int)(tl % nt), ty = (int)((tl / nt) % nt), tx = (int)(tl / ((long)nt * nt));
    const int t0x = tx * Tt, t0y = ty * Tt, t0z = tz * Tt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    T *wst = reinterpret_cast<T *>(reinterpret_cast<unsigned char *>(stage) +
                                   wave * row_stage_bytes<P, T>());
    int *woff = reinterpret_cast<int *>(wst + kRowBatch * REC);
    const int grp = lane / P2, row = lane - grp * P2, ra0 = row / P, ra1 = row - ra0 * P;
    const bool row_lane = grp < G;
    const int cnt = pe - pb, per_round = nwaves * kRowBatch;
    const int rounds = (cnt + per_round - 1) / per_round;

    // the particle (and factor) of this lane in round rd, read one round ahead
    auto fetch = [&](int rd, double4 &r, double &gq) {
        const int p = pb + rd * per_round + wave * kRowBatch + lane;
        if (rd < rounds && lane < kRowBatch && p < pe) {
            r = sorted[p];
            gq = (MODE == kWeighted) ? gv[perm[p]] : 1.0;
        }
    };
    for (int ii = 0; ii < NM; ii++) {
        const double half = 0.5 * ii;
        double4 rn = make_double4(0.0, 0.0, 0.0, 0.0);
        double gn = 1.0;
        fetch(0, rn, gn);
        for (int k = threadIdx.x; k < rg.size; k += blockDim.x)
            acc[k] = A(0);
        __syncthreads();
        for (int rd = 0; rd < rounds; rd++) {
            const int b0 = pb + rd * per_round + wave * kRowBatch;
            const int nb = max(0, min(kRowBatch, pe - b0));
            const double4 r = rn;
            const double gq = gn;
            fetch(rd + 1, rn, gn);
            if (lane < nb) {
                T qv = (T)r.w;
                if (MODE == kWeighted)
                    qv = (T)(r.w * gq);
                int c0, c1, c2;
                T w0[P], w1[P], w2[P];
                axis_weights<P, T>(r.x, Hi, half, M, c0, w0);
                axis_weights<P, T>(r.y, Hi, half, M, c1, w1);
                axis_weights<P, T>(r.z, Hi, half, M, c2, w2);
                int o0 = c0 - t0x, o1 = c1 - t0y, o2 = c2 - t0z;
                o0 += (o0 < 0) ? M : 0;
                o1 += (o1 < 0) ? M : 0;
                o2 += (o2 < 0) ? M : 0;
                T *rec = wst + lane * REC;
#pragma unroll
                for (int a = 0; a < P; a++) {
                    rec[a] = w2[a];
                    rec[W2 + a] = qv * w0[a];
                    rec[W2 + P + a] = w1[a];
                }
                woff[lane] = o0 * rg.Sx + o1 * rg.Sy + o2;
            }
            __syncthreads();
            // a lane's row sums over consecutive particles of one base (the
            // tile order puts them together) stay in registers until the base
            // changes, in the form's precision (a run is a few particles; the
            // FP32 form's sums are FMAs, widened to FP64 at the LDS add)
            T run[P];
#pragma unroll
            for (int a2 = 0; a2 < P; a2++)
                run[a2] = T(0);
            for (int j = 0; j < nb; j += G) {
                const int jj = j + grp;
                if (row_lane && jj < nb) {
                    const T *rec = wst + jj * REC;
                    T wz[W2];
                    load_row<T, W2>(rec, wz);
                    const T tr = rec[W2 + ra0] * rec[W2 + P + ra1];
#pragma unroll
                    for (int a2 = 0; a2 < P; a2++)
                        run[a2] += tr * wz[a2];
                    const int off = woff[jj];
                    if (jj + G >= nb || woff[jj + G] != off) {
                        A *a = acc + off + ra0 * rg.Sx + ra1 * rg.Sy;
#pragma unroll
                        for (int a2 = 0; a2 < P; a2++) {
                            atomicAdd(&a[a2], (A)run[a2]);
                            run[a2] = T(0);
                        }
                    }
                }
            }
            __syncthreads();
        }
        // add the region to mesh ii (periodic wrap of the halo)
        const int R = rg.R;
        for (int k = threadIdx.x; k < R * R * R; k += blockDim.x) {
            const int lz = k % R, ly = (k / R) % R, lx = k / (R * R);
            const T v = (T)acc[lx * rg.Sx + ly * rg.Sy + lz];
            if (v != T(0)) {
                const int gx = (t0x + lx) % M, gy = (t0y + ly) % M, gz = (t0z + lz) % M;
                const long g = ((long)gx * M + gy) * M + gz;
                if (R2C)
                    atomicAdd(&mesh[g], v);
                else
                    atomicAdd(&mesh[2 * g + ii], v);
            }
        }
        __syncthreads();
    }
}

// ---- halo-tile assignment ------------------------------------------------------

// LDS accumulation is FP64 in both forms: fire-and-forget ds_add_f64 (no
// returned value to wait for), and the FP32 form's products are summed
// without rounding error growth; the mesh itself is stored in the form's
// precision.

template <int P, typename T, bool R2C, bool IL, int MODE>
__device__ __forceinline__ void
assign_item(const double4 *__restrict__ sorted, int pb, int pe, long tl, int nt, int Tt, int M,
            double Hi, T *__restrict__ mesh, const int *__restrict__ perm,
            const double *__restrict__ gv, int n, double *acc) {
    constexpr int NM = IL ? 2 : 1;
    const int R = Tt + P - 1 + (IL ? 1 : 0); // halo region side
    const int npts = R * R * R;
    const int nall = NM * npts;
    const int tz = (int)(tl % nt), ty = (int)((tl / nt) % nt), tx = (int)(tl / ((long)nt * nt));
    const int t0x = tx * Tt, t0y = ty * Tt, t0z = tz * Tt;

    for (int k = threadIdx.x; k < nall; k += blockDim.x)
        acc[k] = 0.0;
    __syncthreads();
    for (int p = pb + threadIdx.x; p < pe; p += blockDim.x) {
        const double4 r = sorted[p];
        T qv = (T)r.w;
        T g0 = T(0), g1 = T(0), g2 = T(0);
        if (MODE == kWeighted) {
            qv = (T)(r.w * gv[perm[p]]);
        } else if (MODE == kDipole) {
            const int i = perm[p];
            g0 = (T)gv[i];
            g1 = (T)gv[(long)n + i];
            g2 = (T)gv[2L * n + i];
        }
#pragma unroll
        for (int ii = 0; ii < NM; ii++) {
            int b0, b1, b2;
            T w0[P], w1[P], w2[P];
            const double half = 0.5 * ii;
            if (MODE == kDipole) {
                // W^(d) summed with g_d: fold g_d into the derivative weights
                // of axis d, then deposit (d0 w1 w2 + w0 d1 w2 + w0 w1 d2)
                T d0[P], d1[P], d2[P];
                axis_weights_d<P, T>(r.x, Hi, half, M, b0, w0, d0);
                axis_weights_d<P, T>(r.y, Hi, half, M, b1, w1, d1);
                axis_weights_d<P, T>(r.z, Hi, half, M, b2, w2, d2);
                int o0 = b0 - t0x, o1 = b1 - t0y, o2 = b2 - t0z;
                o0 += (o0 < 0) ? M : 0;
                o1 += (o1 < 0) ? M : 0;
                o2 += (o2 < 0) ? M : 0;
                double *a = acc + ii * npts;
                for (int a0 = 0; a0 < P; a0++) {
#pragma unroll
                    for (int a1 = 0; a1 < P; a1++) {
                        const T u = qv * (g0 * d0[a0] * w1[a1] + g1 * w0[a0] * d1[a1]);
                        const T v = qv * (g2 * w0[a0] * w1[a1]);
                        const int row = ((o0 + a0) * R + (o1 + a1)) * R + o2;
#pragma unroll
                        for (int a2 = 0; a2 < P; a2++)
                            atomicAdd(&a[row + a2], (double)(u * w2[a2] + v * d2[a2]));
                    }
                }
                continue;
            }
            axis_weights<P, T>(r.x, Hi, half, M, b0, w0);
            axis_weights<P, T>(r.y, Hi, half, M, b1, w1);
            axis_weights<P, T>(r.z, Hi, half, M, b2, w2);
            // local offsets inside the halo region (base is within T + 1 of
            // the tile start, modulo the periodic wrap)
            int o0 = b0 - t0x, o1 = b1 - t0y, o2 = b2 - t0z;
            o0 += (o0 < 0) ? M : 0;
            o1 += (o1 < 0) ? M : 0;
            o2 += (o2 < 0) ? M : 0;
            double *a = acc + ii * npts;
            for (int a0 = 0; a0 < P; a0++) {
                const T t0 = qv * w0[a0];
#pragma unroll
                for (int a1 = 0; a1 < P; a1++) {
                    const T t1 = t0 * w1[a1];
                    const int row = ((o0 + a0) * R + (o1 + a1)) * R + o2;
#pragma unroll
                    for (int a2 = 0; a2 < P; a2++)
                        atomicAdd(&a[row + a2], (double)(t1 * w2[a2]));
                }
            }
        }
    }
    __syncthreads();
    // add the region to the mesh (periodic wrap of the halo)
    for (int k = threadIdx.x; k < npts; k += blockDim.x) {
        const int lz = k % R, ly = (k / R) % R, lx = k / (R * R);
        const int gx = (t0x + lx) % M, gy = (t0y + ly) % M, gz = (t0z + lz) % M;
        const long g = ((long)gx * M + gy) * M + gz;
#pragma unroll
        for (int ii = 0; ii < NM; ii++) {
            const T v = (T)acc[ii * npts + k];
            if (v != T(0)) {
                if (R2C)
                    atomicAdd(&mesh[g], v);
                else
                    atomicAdd(&mesh[2 * g + ii], v);
            }
        }
    }
}

// ---- back-interpolation -----------------------------------------------------

template <typename T, bool R2C>
__device__ __forceinline__ T mesh_val(const T *__restrict__ m, long idx, int ii) {
    return R2C ? m[idx] : m[2 * idx + ii];
}

template <typename T> struct Acc;
template <> struct Acc<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double v) { s += v; }
    __device__ __forceinline__ double value() const { return s; }
};
template <> struct Acc<float> {
    CompSum<float> s;
    __device__ __forceinline__ void add(float v) { s.add(v); }
    __device__ __forceinline__ double value() const { return s.value(); }
};

// ---- tile-staged ik back-interpolation ------------------------------------------
//
// As interp_ad_item, one field component at a time: F_d's halo region (one
// plane per mesh) is staged in LDS and every particle of the item adds
// f_k,d -= s q sum_a w0[a] sum_b w1[b] sum_c w2[c] F_d(a,b,c) (s = prefac,
// x 0.5 per mesh when interlaced), the sums over a in FP64.
template <int P, typename T, bool R2C, bool IL>
__device__ __forceinline__ void
interp_ik_item(const double4 *__restrict__ sorted, const int *__restrict__ perm, int pb, int pe,
               long tl, int nt, int Tt, int M, double Hi, const T *__restrict__ F0,
               const T *__restrict__ F1, const T *__restrict__ F2, double prefac,
               double *__restrict__ fk, long stride, T *__restrict__ lds) {
    constexpr int NM = IL ? 2 : 1;
    const int R = Tt + P - 1 + (IL ? 1 : 0);
    const int npts = R * R * R;
    const int tz = (int)(tl % nt), ty = (int)((tl / nt) % nt), tx = (int)(tl / ((long)nt * nt));
    const int t0x = tx * Tt, t0y = ty * Tt, t0z = tz * Tt;
    const double sc = IL ? 0.5 * prefac : prefac;
    for (int d = 0; d < 3; d++) {
        const T *F = d == 0 ? F0 : (d == 1 ? F1 : F2);
        for (int k = threadIdx.x; k < npts; k += blockDim.x) {
            const int lz = k % R, ly = (k / R) % R, lx = k / (R * R);
            const int gx = (t0x + lx) % M, gy = (t0y + ly) % M, gz = (t0z + lz) % M;
            const long g = ((long)gx * M + gy) * M + gz;
            if (R2C) {
                lds[k] = F[g];
            } else {
                lds[k] = F[2 * g];
                if (IL)
                    lds[npts + k] = F[2 * g + 1];
            }
        }
        __syncthreads();
        for (int p = pb + threadIdx.x; p < pe; p += blockDim.x) {
            const double4 r = sorted[p];
            double acc = 0.0;
#pragma unroll
            for (int ii = 0; ii < NM; ii++) {
                int b0, b1, b2;
                T w0[P], w1[P], w2[P];
                const double half = 0.5 * ii;
                axis_weights<P, T>(r.x, Hi, half, M, b0, w0);
                axis_weights<P, T>(r.y, Hi, half, M, b1, w1);
                axis_weights<P, T>(r.z, Hi, half, M, b2, w2);
                int o0 = b0 - t0x, o1 = b1 - t0y, o2 = b2 - t0z;
                o0 += (o0 < 0) ? M : 0;
                o1 += (o1 < 0) ? M : 0;
                o2 += (o2 < 0) ? M : 0;
                const T *pl = lds + ii * npts;
                for (int a = 0; a < P; a++) {
                    T sa = T(0);
#pragma unroll
                    for (int b = 0; b < P; b++) {
                        const T *row = pl + ((o0 + a) * R + (o1 + b)) * R + o2;
                        T sb = T(0);
#pragma unroll
                        for (int c = 0; c < P; c++)
                            sb = fma(w2[c], row[c], sb);
                        sa = fma(w1[b], sb, sa);
                    }
                    acc += (double)w0[a] * (double)sa;
                }
            }
            fk[d * stride + perm[p]] -= sc * r.w * acc;
        }
        __syncthreads();
    }
}

// ---- tile-staged ad back-interpolation ------------------------------------------
//
// A work item is a chunk of the particles of one mesh tile (as in assignment).
// The item's halo region of phi -- (Tt + P - 1 + IL)^3 points per mesh from
// the tile origin, periodic wrap -- is copied to LDS once (plane ii holds mesh
// ii), and every particle of the item gathers its P^3 stencil from LDS.  The
// sums factorise over the axes:
//   F_x = sum_a d0[a] sum_b w1[b] S1(a,b),  F_y = sum_a w0[a] sum_b d1[b] S1(a,b),
//   F_z = sum_a w0[a] sum_b w1[b] S2(a,b),
//   S1 = sum_c w2[c] phi(a,b,c),  S2 = sum_c d2[c] phi(a,b,c),
// two multiply-adds per stencil point; the sums over a run in FP64.  The force
// is f_k -= prefac q / L sum phi W^(d) (x 0.5 per mesh when interlaced;
// charge-assign.c:1352-1428, 727-806), and for ad / ad-r the self force is
// subtracted from f.
template <int P, typename T, bool R2C, bool IL>
__device__ __forceinline__ void
interp_ad_item(const double4 *__restrict__ sorted, const int *__restrict__ perm, int pb, int pe,
               long tl, int nt, int Tt, int M, double Hi, double Leni, const T *__restrict__ phi,
               double prefac, const double *__restrict__ selfb, double h, double *__restrict__ fk,
               double *__restrict__ f, long stride, T *__restrict__ lds) {
    constexpr int NM = IL ? 2 : 1;
    const int R = Tt + P - 1 + (IL ? 1 : 0);
    const int npts = R * R * R;
    const int tz = (int)(tl % nt), ty = (int)((tl / nt) % nt), tx = (int)(tl / ((long)nt * nt));
    const int t0x = tx * Tt, t0y = ty * Tt, t0z = tz * Tt;
    for (int k = threadIdx.x; k < npts; k += blockDim.x) {
        const int lz = k % R, ly = (k / R) % R, lx = k / (R * R);
        const int gx = (t0x + lx) % M, gy = (t0y + ly) % M, gz = (t0z + lz) % M;
        const long g = ((long)gx * M + gy) * M + gz;
        if (R2C) {
            lds[k] = phi[g];
        } else {
            lds[k] = phi[2 * g];
            if (IL)
                lds[npts + k] = phi[2 * g + 1];
        }
    }
    __syncthreads();
    for (int p = pb + threadIdx.x; p < pe; p += blockDim.x) {
        const double4 r = sorted[p];
        double F[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int ii = 0; ii < NM; ii++) {
            int b0, b1, b2;
            T w0[P], w1[P], w2[P], d0[P], d1[P], d2[P];
            const double half = 0.5 * ii;
            axis_weights_d<P, T>(r.x, Hi, half, M, b0, w0, d0);
            axis_weights_d<P, T>(r.y, Hi, half, M, b1, w1, d1);
            axis_weights_d<P, T>(r.z, Hi, half, M, b2, w2, d2);
            int o0 = b0 - t0x, o1 = b1 - t0y, o2 = b2 - t0z;
            o0 += (o0 < 0) ? M : 0;
            o1 += (o1 < 0) ? M : 0;
            o2 += (o2 < 0) ? M : 0;
            const T *pl = lds + ii * npts;
            for (int a = 0; a < P; a++) {
                T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
                for (int b = 0; b < P; b++) {
                    const T *row = pl + ((o0 + a) * R + (o1 + b)) * R + o2;
                    T s1 = T(0), s2 = T(0);
#pragma unroll
                    for (int c = 0; c < P; c++) {
                        const T v = row[c];
                        s1 = fma(w2[c], v, s1);
                        s2 = fma(d2[c], v, s2);
                    }
                    sx = fma(w1[b], s1, sx);
                    sy = fma(d1[b], s1, sy);
                    sz = fma(w1[b], s2, sz);
                }
                F[0] += (double)d0[a] * (double)sx;
                F[1] += (double)w0[a] * (double)sy;
                F[2] += (double)w0[a] * (double)sz;
            }
        }
        const double scale = (IL ? 0.5 * prefac : prefac) * (r.w * Leni);
        const int i = perm[p];
        fk[i] -= scale * F[0];
        fk[stride + i] -= scale * F[1];
        fk[2 * stride + i] -= scale * F[2];
        if (selfb != nullptr) {
            // f -= sum_m' b(m') q^2 sin(2 pi m'.r / h)  (p3m-ad-self-forces.c:122-135)
            const double q2 = r.w * r.w;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            int ind = 0;
            for (int m0 = -1; m0 <= 1; m0++)
                for (int m1 = -1; m1 <= 1; m1++)
                    for (int m2 = -1; m2 <= 1; m2++) {
                        const double st = q2 * sin(2 * 3.14159265358979323846 *
                                                   (m0 * r.x / h + m1 * r.y / h + m2 * r.z / h));
                        s0 += selfb[ind] * st;
                        s1 += selfb[ind + 1] * st;
                        s2 += selfb[ind + 2] * st;
                        ind += 3;
                    }
            f[i] -= s0;
            f[stride + i] -= s1;
            f[2 * stride + i] -= s2;
        }
    }
}

}  // namespace part
}  // namespace p3m
