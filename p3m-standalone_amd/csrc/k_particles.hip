// Particle <-> mesh kernels: charge assignment and back-interpolation.
//
// Charge assignment (charge-assign.c:119-176, 215-267, 1067-1133,
// 1243-1311): particle i adds q W_a(x) W_b(y) W_c(z) to the P^3 mesh points
// of its stencil, on mesh 0 and, interlaced, on mesh 1 (h/2-shifted, held in
// the imaginary part).  Weights are evaluated at the reference's quantised
// offsets (device_common.hpp) instead of read from 7 MB tables.
//
// Assignment: particles are counting-sorted by the mesh tile (T^3 points) of
// their mesh-0 stencil base.  A work item is a chunk of at most kChunk
// particles of one tile; its workgroup accumulates the full P^3 stencils of
// its particles in an LDS copy of the tile plus halo ((T+P)^3 points per
// mesh, every stencil fits, so the P^3 loop has no clipping and no
// divergence) and then adds the non-zero LDS points to the mesh with global
// atomics.  Dense (clustered) tiles are split over many workgroups; the
// global atomics are one per LDS point, not one per particle and stencil
// point.

// Back-interpolation (charge-assign.c:601-643, 646-688, 727-769, 849-885,
// 1352-1391, 1468-1502) walks the same sorted order (neighbouring threads
// gather from neighbouring mesh points) and recomputes the weights instead
// of caching cf / dQ / ca_ind per particle.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <new>

#include "batch_desc.hpp"
#include "device_common.hpp"
#include "k_internal.hpp"
#include "particle_bodies.hpp"

using namespace p3m;

namespace {

using namespace p3m::part;

// ---- sort by mesh tile ------------------------------------------------------

__device__ __forceinline__ int axis_base(double r, double Hi, int P, int M) {
    int b;
    double xq;
    locate(r, Hi, (double)((P - 1) / 2), 0.0, M, b, xq);
    return b;
}

__global__ void k_tile_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double Hi, int P, int M, int T, int nt,
                             int *__restrict__ key, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int tx = axis_base(x[i], Hi, P, M) / T, ty = axis_base(y[i], Hi, P, M) / T,
              tz = axis_base(z[i], Hi, P, M) / T;
    const int k = (tx * nt + ty) * nt + tz;
    key[i] = k;
    atomicAdd(&count[k], 1);
}

__global__ void k_tile_scatter(int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const double *__restrict__ q,
                               const int *__restrict__ key, const int *__restrict__ start,
                               int *__restrict__ cursor, int *__restrict__ perm,
                               double4 *__restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int k = key[i];
    const int p = start[k] + atomicAdd(&cursor[k], 1);
    perm[p] = i;
    sorted[p] = make_double4(x[i], y[i], z[i], q[i]);
}

// The same sort with per-block aggregation, for tile counts that fit in
// LDS: a block takes a contiguous range of particles, counts them per tile in
// LDS and adds each nonzero count to the global counter once; the scatter
// reserves the block's range of every tile with one global atomic and ranks
// its particles inside it with LDS atomics.  (Device-scope atomics are the
// cost of the plain sort: one per particle.)
constexpr int kSortThreads = 1024;
constexpr long kSortMaxTiles = 8192; // two int arrays of this size in LDS

__global__ void __launch_bounds__(kSortThreads)
k_tile_count_agg(int n, int per_block, const double *__restrict__ x, const double *__restrict__ y,
                 const double *__restrict__ z, double Hi, int P, int M, int T, int nt, int ntiles,
                 int *__restrict__ key, int *__restrict__ count) {
    extern __shared__ int hist[];
    for (int k = threadIdx.x; k < ntiles; k += blockDim.x)
        hist[k] = 0;
    __syncthreads();
    const int b = blockIdx.x * per_block, e = min(n, b + per_block);
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const int tx = axis_base(x[i], Hi, P, M) / T, ty = axis_base(y[i], Hi, P, M) / T,
                  tz = axis_base(z[i], Hi, P, M) / T;
        const int k = (tx * nt + ty) * nt + tz;
        key[i] = k;
        atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ntiles; k += blockDim.x)
        if (hist[k] != 0)
            atomicAdd(&count[k], hist[k]);
}

__global__ void __launch_bounds__(kSortThreads)
k_tile_scatter_agg(int n, int per_block, const double *__restrict__ x, const double *__restrict__ y,
                   const double *__restrict__ z, const double *__restrict__ q, int ntiles,
                   const int *__restrict__ key, const int *__restrict__ start,
                   int *__restrict__ cursor, int *__restrict__ perm, double4 *__restrict__ sorted) {
    extern __shared__ int lds_i[];
    int *hist = lds_i, *base = lds_i + ntiles;
    for (int k = threadIdx.x; k < ntiles; k += blockDim.x)
        hist[k] = 0;
    __syncthreads();
    const int b = blockIdx.x * per_block, e = min(n, b + per_block);
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x)
        atomicAdd(&hist[key[i]], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < ntiles; k += blockDim.x) {
        if (hist[k] != 0)
            base[k] = start[k] + atomicAdd(&cursor[k], hist[k]);
        hist[k] = 0;
    }
    __syncthreads();
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const int k = key[i];
        const int p = base[k] + atomicAdd(&hist[k], 1);
        perm[p] = i;
        sorted[p] = make_double4(x[i], y[i], z[i], q[i]);
    }
}

// Sort by base mesh point: the key of a particle is its tile (tile-major)
// and then the base's position inside the tile (the lowest stencil point on
// mesh 0; T^3 <= 512 per tile), so a tile's particles are contiguous and
// ordered by base.  Particles sharing a base become consecutive: the row
// deposit sums their rows in registers before its LDS atomics, and
// neighbouring lanes of the gathers read the same LDS words (a broadcast) or
// neighbouring ones.  One device atomic per particle, spread over M^3 keys
// (at most a few dozen particles share one).  Meshes of more than
// kFineMaxKeys keys keep the tile sort above.
constexpr long kFineMaxKeys = 1L << 24;

// the particle's key and its rank among the particles of that key (the
// value the count's atomic returns), so that the scatter needs no atomics
__global__ void k_fine_count(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double Hi, int P, int M, int T, int nt,
                             int *__restrict__ key, int *__restrict__ rank, int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int bx = axis_base(x[i], Hi, P, M), by = axis_base(y[i], Hi, P, M), bz = axis_base(z[i], Hi, P, M);
    const int tile = ((bx / T) * nt + by / T) * nt + bz / T;
    const int k = tile * T * T * T + ((bx % T) * T + by % T) * T + bz % T;
    key[i] = k;
    rank[i] = atomicAdd(&count[k], 1);
}

__global__ void k_fine_scatter(int n, const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const double *__restrict__ q,
                               const int *__restrict__ key, const int *__restrict__ rank,
                               const int *__restrict__ start, int *__restrict__ perm,
                               double4 *__restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int p = start[key[i]] + rank[i];
    perm[p] = i;
    sorted[p] = make_double4(x[i], y[i], z[i], q[i]);
}

// tile starts, counts and chunk counts from the fine keys' prefix sums
// (start[ntiles] = the total)
__global__ void k_tile_ranges(const int *__restrict__ fstart, long ntiles, int T3, int *__restrict__ start,
                              int *__restrict__ count, int *__restrict__ chunks) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles)
        return;
    const int s = fstart[t * T3];
    start[t] = s;
    if (t < ntiles) {
        const int c = fstart[(t + 1) * T3] - s;
        count[t] = c;
        chunks[t] = (c + kChunk - 1) / kChunk;
    }
}

// chunks per tile: ceil(count / kChunk)
__global__ void k_tile_chunks(const int *__restrict__ count, long ntiles, int *__restrict__ chunks) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles)
        chunks[t] = (count[t] + kChunk - 1) / kChunk;
}

// work item -> (tile, chunk) for every item of the chunk prefix sums, so
// that a workgroup finds its tile with one load (a binary search over the
// tiles is a chain of a dozen dependent loads per workgroup)
__global__ void k_item_map(const int *__restrict__ cstart, long ntiles, int2 *__restrict__ imap) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
        return;
    const int b = cstart[t], e = cstart[t + 1];
    for (int c = b; c < e; c++)
        imap[c] = make_int2((int)t, c - b);
}

// the batched form: (global tile, chunk, system)
__global__ void k_bitem_map(const int *__restrict__ cstart, long ntiles, const long *__restrict__ tile0s,
                            int nsys, int4 *__restrict__ imap) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
        return;
    const int b = cstart[t], e = cstart[t + 1];
    if (b == e)
        return;
    const int s = fft::prefix_slot(tile0s, nsys, t);
    for (int c = b; c < e; c++)
        imap[c] = make_int4((int)t, c - b, s, 0);
}

template <int P, typename T, bool R2C, bool IL, int MODE>
__global__ void __launch_bounds__(kThreads)
k_assign_halo(const double4 *__restrict__ sorted, const int *__restrict__ tstart,
              const int *__restrict__ cstart, const int2 *__restrict__ imap, long ntiles, int nt, int Tt, int M, double Hi,
              T *__restrict__ mesh, const int *__restrict__ perm, const double *__restrict__ gv,
              int n) {
    extern __shared__ unsigned char smem_raw[];
    // work item -> (tile, chunk): binary search in the chunk prefix sums
    const int item = blockIdx.x;
    if (item >= cstart[ntiles])
        return;
    const int2 tc = imap[item];
    const long tile = tc.x;
    const int chunk = tc.y;
    const int pb = tstart[tile] + chunk * kChunk;
    const int pe = min(tstart[tile + 1], pb + kChunk);
    assign_item<P, T, R2C, IL, MODE>(sorted, pb, pe, tile, nt, Tt, M, Hi, mesh, perm, gv, n,
                                     reinterpret_cast<double *>(smem_raw));
}

// Row-parallel form of k_assign_halo (charge and charge-times-factor
// deposits): FP64 LDS accumulation, one mesh of an interlaced pair at a time.
template <int P, typename T, bool R2C, bool IL, int MODE>
__global__ void __launch_bounds__(kAssignThreads)
k_assign_rows(const double4 *__restrict__ sorted, const int *__restrict__ tstart,
              const int *__restrict__ cstart, const int2 *__restrict__ imap, long ntiles, int nt, int Tt, int M, double Hi,
              T *__restrict__ mesh, const int *__restrict__ perm, const double *__restrict__ gv,
              int region_doubles) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    if (item >= cstart[ntiles])
        return;
    const int2 tc = imap[item];
    const long tile = tc.x;
    const int chunk = tc.y;
    const int pb = tstart[tile] + chunk * kChunk;
    const int pe = min(tstart[tile + 1], pb + kChunk);
    double *acc = reinterpret_cast<double *>(smem_raw);
    T *stage = reinterpret_cast<T *>(smem_raw + sizeof(double) * (size_t)region_doubles);
    assign_rows_item<P, T, R2C, IL, MODE, double>(sorted, pb, pe, tile, nt, Tt, M, Hi, mesh, perm,
                                                  gv, acc, stage);
}

template <int P, typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_interp_ik_tile(const double4 *__restrict__ sorted, const int *__restrict__ perm,
                 const int *__restrict__ tstart, const int *__restrict__ cstart,
                 const int2 *__restrict__ imap, long ntiles, int nt,
                 int Tt, int M, double Hi, const T *__restrict__ F0, const T *__restrict__ F1,
                 const T *__restrict__ F2, double prefac, double *__restrict__ fk, long stride) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    if (item >= cstart[ntiles])
        return;
    const int2 tc = imap[item];
    const long tile = tc.x;
    const int chunk = tc.y;
    const int pb = tstart[tile] + chunk * kChunk;
    const int pe = min(tstart[tile + 1], pb + kChunk);
    interp_ik_item<P, T, R2C, IL>(sorted, perm, pb, pe, tile, nt, Tt, M, Hi, F0, F1, F2, prefac, fk,
                                  stride, reinterpret_cast<T *>(smem_raw));
}

template <int P, typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_interp_ad_tile(const double4 *__restrict__ sorted, const int *__restrict__ perm,
                 const int *__restrict__ tstart, const int *__restrict__ cstart,
                 const int2 *__restrict__ imap, long ntiles, int nt,
                 int Tt, int M, double Hi, double Leni, const T *__restrict__ phi, double prefac,
                 const double *__restrict__ selfb, double h, double *__restrict__ fk,
                 double *__restrict__ f, long stride) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    if (item >= cstart[ntiles])
        return;
    const int2 tc = imap[item];
    const long tile = tc.x;
    const int chunk = tc.y;
    const int pb = tstart[tile] + chunk * kChunk;
    const int pe = min(tstart[tile + 1], pb + kChunk);
    interp_ad_item<P, T, R2C, IL>(sorted, perm, pb, pe, tile, nt, Tt, M, Hi, Leni, phi, prefac,
                                  selfb, h, fk, f, stride, reinterpret_cast<T *>(smem_raw));
}

// ---- adjoint (vector-Jacobian product) of the k-space force ---------------
//
// ik: f_{i,d} = -s q_i sum_m W_i(m) phi_d(m), phi_d = K_d Q, Q = sum_j q_j W_j,
// s = prefac (x 0.5 per mesh when interlaced).  For a cotangent g_{i,d}:
//   dq_j     = -s      sum_m W_j(m)    H_j(m)
//   dx_{j,e} = -s q_j  sum_m d_e W_j(m) H_j(m),
//   H_j(m)   = sum_d g_{j,d} phi_d(m) + lambda(m),
//   lambda   = sum_d K_d^T psi_d,  psi_d = sum_i q_i g_{i,d} W_i.
// ad: f_{i,d} = -s q_i sum_m W^(d)_i(m) phi(m), phi = K Q (K symmetric):
//   dq_j     = -s     sum_m [sum_d g_{j,d} W^(d)_j phi + W_j lambda]
//   dx_{j,e} = -s q_j sum_m [sum_d g_{j,d} d_e W^(d)_j phi + d_e W_j lambda],
//   lambda   = K psi,  psi = sum_i q_i sum_d g_{i,d} W^(d)_i.
// Derivatives are those of the B-spline weights at the quantised offsets
// (d/dx = Hi d/du); outputs accumulate (+=) in the caller's order.
template <int P, typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_adjoint_ik(int n, const double4 *__restrict__ sorted, const int *__restrict__ perm,
             const double *__restrict__ q, const double *__restrict__ gv, double Hi, int M,
             const T *__restrict__ F0, const T *__restrict__ F1, const T *__restrict__ F2,
             const T *__restrict__ lam, double prefac, double *__restrict__ dpos,
             double *__restrict__ dq) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    constexpr int NM = IL ? 2 : 1;
    const int i = perm[t];
    const double4 r = sorted[t];
    const double qi = q[i];
    const T g0 = (T)gv[i], g1 = (T)gv[(long)n + i], g2 = (T)gv[2L * n + i];
    Acc<T> aq, ax, ay, az;
    for (int ii = 0; ii < NM; ii++) {
        int b0, b1, b2;
        T w0[P], w1[P], w2[P], d0[P], d1[P], d2[P];
        const double half = 0.5 * ii;
        axis_weights_d<P, T>(r.x, Hi, half, M, b0, w0, d0);
        axis_weights_d<P, T>(r.y, Hi, half, M, b1, w1, d1);
        axis_weights_d<P, T>(r.z, Hi, half, M, b2, w2, d2);
        int kk[P];
#pragma unroll
        for (int c = 0; c < P; c++)
            kk[c] = wrap1(b2 + c, M);
        for (int a = 0; a < P; a++) {
            const int ia = wrap1(b0 + a, M);
#pragma unroll
            for (int b = 0; b < P; b++) {
                const int jb = wrap1(b1 + b, M);
                const long row = ((long)ia * M + jb) * M;
                const T wab = w0[a] * w1[b], dab = d0[a] * w1[b], adb = w0[a] * d1[b];
#pragma unroll
                for (int c = 0; c < P; c++) {
                    const long idx = row + kk[c];
                    const T H = g0 * mesh_val<T, R2C>(F0, idx, ii) + g1 * mesh_val<T, R2C>(F1, idx, ii) +
                                g2 * mesh_val<T, R2C>(F2, idx, ii) + mesh_val<T, R2C>(lam, idx, ii);
                    aq.add(wab * w2[c] * H);
                    ax.add(dab * w2[c] * H);
                    ay.add(adb * w2[c] * H);
                    az.add(wab * d2[c] * H);
                }
            }
        }
    }
    const double s = IL ? 0.5 * prefac : prefac;
    dq[i] += -s * aq.value();
    dpos[i] += -s * qi * Hi * ax.value();
    dpos[(long)n + i] += -s * qi * Hi * ay.value();
    dpos[2L * n + i] += -s * qi * Hi * az.value();
}

template <int P, typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_adjoint_ad(int n, const double4 *__restrict__ sorted, const int *__restrict__ perm,
             const double *__restrict__ q, const double *__restrict__ gv, double Hi, int M,
             const T *__restrict__ phi, const T *__restrict__ lam, double prefac,
             const double *__restrict__ selfb, double h, double *__restrict__ dpos,
             double *__restrict__ dq) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    constexpr int NM = IL ? 2 : 1;
    const int i = perm[t];
    const double4 r = sorted[t];
    const double qi = q[i];
    const double gd0 = gv[i], gd1 = gv[(long)n + i], gd2 = gv[2L * n + i];
    const T g0 = (T)gd0, g1 = (T)gd1, g2 = (T)gd2;
    Acc<T> aq, ax, ay, az;
    for (int ii = 0; ii < NM; ii++) {
        int b0, b1, b2;
        T w0[P], w1[P], w2[P], d0[P], d1[P], d2[P], e0[P], e1[P], e2[P];
        double xq;
        const double half = 0.5 * ii;
        axis_weights_d<P, T>(r.x, Hi, half, M, b0, w0, d0);
        axis_weights_d<P, T>(r.y, Hi, half, M, b1, w1, d1);
        axis_weights_d<P, T>(r.z, Hi, half, M, b2, w2, d2);
        // second derivatives at the same quantised offsets
        locate(r.x, Hi, (double)((P - 1) / 2), half, M, b0, xq);
        Bspline<P, T>::d2w((T)xq, e0);
        locate(r.y, Hi, (double)((P - 1) / 2), half, M, b1, xq);
        Bspline<P, T>::d2w((T)xq, e1);
        locate(r.z, Hi, (double)((P - 1) / 2), half, M, b2, xq);
        Bspline<P, T>::d2w((T)xq, e2);
        int kk[P];
#pragma unroll
        for (int c = 0; c < P; c++)
            kk[c] = wrap1(b2 + c, M);
        for (int a = 0; a < P; a++) {
            const int ia = wrap1(b0 + a, M);
#pragma unroll
            for (int b = 0; b < P; b++) {
                const int jb = wrap1(b1 + b, M);
                const long row = ((long)ia * M + jb) * M;
                // x-y factors of the stencil terms (z factor applied per c)
                const T ww = w0[a] * w1[b], dw_ = d0[a] * w1[b], wd = w0[a] * d1[b];
                const T ew = e0[a] * w1[b], dd = d0[a] * d1[b], we = w0[a] * e1[b];
                // sum_d g_d W^(d) = (g0 dw_ + g1 wd) w2 + g2 ww d2
                const T s_w = g0 * dw_ + g1 * wd, s_d = g2 * ww;
                // d_x: (g0 ew + g1 dd) w2 + g2 dw_ d2 ; d_y: (g0 dd + g1 we) w2 + g2 wd d2
                // d_z: (g0 dw_ + g1 wd) d2 + g2 ww e2
                const T xw = g0 * ew + g1 * dd, xd = g2 * dw_;
                const T yw = g0 * dd + g1 * we, yd = g2 * wd;
#pragma unroll
                for (int c = 0; c < P; c++) {
                    const long idx = row + kk[c];
                    const T ph = mesh_val<T, R2C>(phi, idx, ii);
                    const T la = mesh_val<T, R2C>(lam, idx, ii);
                    aq.add((s_w * w2[c] + s_d * d2[c]) * ph + ww * w2[c] * la);
                    ax.add((xw * w2[c] + xd * d2[c]) * ph + dw_ * w2[c] * la);
                    ay.add((yw * w2[c] + yd * d2[c]) * ph + wd * w2[c] * la);
                    az.add((s_w * d2[c] + s_d * e2[c]) * ph + ww * d2[c] * la);
                }
            }
        }
    }
    const double s = IL ? 0.5 * prefac : prefac;
    double dqi = -s * aq.value();
    double dx = -s * qi * Hi * ax.value(), dy = -s * qi * Hi * ay.value(),
           dz = -s * qi * Hi * az.value();
    if (selfb != nullptr) {
        // f_{i,d} -= q_i^2 sum_m' b_d(m') sin(theta), theta = 2 pi m'.r / h
        const double c0 = 2 * 3.14159265358979323846 / h;
        const double q2 = qi * qi;
        double ss = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        int ind = 0;
        for (int m0 = -1; m0 <= 1; m0++)
            for (int m1 = -1; m1 <= 1; m1++)
                for (int m2 = -1; m2 <= 1; m2++) {
                    const double th = 2 * 3.14159265358979323846 *
                                      (m0 * r.x / h + m1 * r.y / h + m2 * r.z / h);
                    const double gb = gd0 * selfb[ind] + gd1 * selfb[ind + 1] + gd2 * selfb[ind + 2];
                    double sn, cs;
                    sincos(th, &sn, &cs);
                    ss += gb * sn;
                    cx += gb * cs * m0;
                    cy += gb * cs * m1;
                    cz += gb * cs * m2;
                    ind += 3;
                }
        dqi -= 2.0 * qi * ss;
        dx -= q2 * c0 * cx;
        dy -= q2 * c0 * cy;
        dz -= q2 * c0 * cz;
    }
    dq[i] += dqi;
    dpos[i] += dx;
    dpos[(long)n + i] += dy;
    dpos[2L * n + i] += dz;
}

// tile side: 8 mesh points (LDS region (8 + P)^3 per mesh)
int tile_side(int M) { return M < 8 ? M : 8; }

// Dynamic LDS beyond 64 KB per workgroup (a gfx950 workgroup may use the
// CU's whole 160 KB) is allowed per kernel before its launch.
constexpr size_t kMaxGroupLds = 160 * 1024;

int allow_lds(const void *kernel, size_t lds) {
    if (lds <= 64 * 1024)
        return P3MK_OK;
    if (lds > kMaxGroupLds)
        return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return e == hipSuccess ? P3MK_OK : p3mk_fail(P3MK_EDEVICE, "hipFuncSetAttribute", (int)e);
}

// launch of a kernel of `threads` threads per workgroup and `lds` bytes of
// dynamic LDS
template <typename... Params, typename... Args>
int launch_lds_t(void (*kernel)(Params...), dim3 grid, int threads, size_t lds, hipStream_t st,
                 Args... args) {
    P3MK_TRY(allow_lds((const void *)kernel, lds));
    hipLaunchKernelGGL(kernel, grid, dim3(threads), lds, st, args...);
    return P3MK_OK;
}

template <typename... Params, typename... Args>
int launch_lds(void (*kernel)(Params...), dim3 grid, size_t lds, hipStream_t st, Args... args) {
    return launch_lds_t(kernel, grid, kThreads, lds, st, args...);
}

}  // namespace

struct p3mk_psort {
    int nmax, M, T, nt;
    long ntiles;
    int *key, *count, *start, *cursor, *perm, *chunks, *cstart, *scan_scratch;
    int2 *imap; // work item -> (tile, chunk)
    double4 *sorted;
    // sort by base: ntiles T^3 keys (0: the tile sort), their counts and
    // prefix sums
    long nf;
    int *fcount, *fstart, *fscan_scratch;
    int *rank; // per particle: its rank within its key
    int n; // particles of the last sort
};

extern "C" void p3mk_psort_destroy(p3mk_psort *s) {
    if (!s)
        return;
    p3mk_release(s->key);
    p3mk_release(s->count); // cursor lives in this allocation
    p3mk_release(s->start);
    p3mk_release(s->perm);
    p3mk_release(s->chunks);
    p3mk_release(s->cstart);
    p3mk_release(s->imap);
    p3mk_release(s->scan_scratch);
    p3mk_release(s->sorted);
    p3mk_release(s->fcount);
    p3mk_release(s->fstart);
    p3mk_release(s->fscan_scratch);
    p3mk_release(s->rank);
    delete s;
}

extern "C" int p3mk_psort_create(p3mk_psort **out, const p3mk_geom *g, int n) {
    *out = nullptr;
    p3mk_psort *s = new (std::nothrow) p3mk_psort();
    if (!s)
        return p3mk_fail(P3MK_ENOMEM, "particle sort", 0);
    s->nmax = n;
    s->M = g->M;
    s->T = tile_side(g->M);
    s->nt = (g->M + s->T - 1) / s->T;
    s->ntiles = (long)s->nt * s->nt * s->nt;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const size_t nt1 = (size_t)s->ntiles + 1;
    int rc = P3MK_OK;
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->key, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = alloc_count_cursor(&s->count, &s->cursor, nt1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->start, sizeof(int) * nt1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->chunks, sizeof(int) * nt1);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->cstart, sizeof(int) * nt1);
    // work items: at most one partial chunk per tile plus the full chunks
    if (rc == P3MK_OK)
        rc = p3mk_alloc((void **)&s->imap, sizeof(int2) * ((size_t)nn / kChunk + nt1));
    if (rc == P3MK_OK)
        rc = p3mk_alloc((void **)&s->scan_scratch, sizeof(int) * p3mk_scan_scratch_ints(s->ntiles));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->perm, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->sorted, sizeof(double4) * nn);
    const long T3 = (long)s->T * s->T * s->T;
    s->nf = s->ntiles * T3 <= kFineMaxKeys ? s->ntiles * T3 : 0;
    if (s->nf > 0) {
        const size_t nf1 = (size_t)s->nf + 1;
        if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->fcount, sizeof(int) * nf1);
        if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->fstart, sizeof(int) * nf1);
        if (rc == P3MK_OK)
            rc = p3mk_alloc((void **)&s->fscan_scratch, sizeof(int) * p3mk_scan_scratch_ints(s->nf));
        if (rc == P3MK_OK) rc = p3mk_alloc((void **)&s->rank, sizeof(int) * nn);
    }
    if (rc != P3MK_OK) {
        p3mk_psort_destroy(s);
        return rc;
    }
    *out = s;
    return P3MK_OK;
}

namespace {

template <typename T, bool R2C, int MODE>
int launch_assign(const p3mk_geom *g, int n, p3mk_psort *s, void *mesh, const double *gv,
                  p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    const int NM = g->interlaced ? 2 : 1;
    const int R = s->T + g->cao - 1 + (g->interlaced ? 1 : 0);
    // upper bound of the number of work items; surplus workgroups exit
    const long items = (long)(n + kChunk - 1) / kChunk + s->ntiles;
    const dim3 grid((unsigned)items);
    T *m = (T *)mesh;
    hipStream_t st = p3mk_stream(qu);
    const bool il = g->interlaced != 0;
    if constexpr (MODE != kDipole) {
        const RowRegion rg = row_region(s->T, g->cao, il);
        size_t stage = 0;
        switch (g->cao) {
#define SB(P)                                                                                      \
    case P:                                                                                        \
        stage = (size_t)(kAssignThreads / 64) * row_stage_bytes<P, T>();                          \
        break;
            SB(1) SB(2) SB(3) SB(4) SB(5) SB(6) SB(7)
#undef SB
        default:
            return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
        }
        const size_t lds = sizeof(double) * (size_t)rg.size + stage;
        if (lds > kMaxGroupLds)
            return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
        switch (g->cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        if (il)                                                                                    \
            P3MK_TRY(launch_lds_t(k_assign_rows<P, T, R2C, true, MODE>, grid, kAssignThreads, lds, \
                               st, s->sorted, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T, g->M,   \
                               Hi, m, s->perm, gv, rg.size));                                       \
        else                                                                                       \
            P3MK_TRY(launch_lds_t(k_assign_rows<P, T, R2C, false, MODE>, grid, kAssignThreads, lds, \
                               st, s->sorted, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T, g->M,   \
                               Hi, m, s->perm, gv, rg.size));                                       \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
        }
        return p3mk_launch_check("assign");
    }
    const size_t lds = sizeof(double) * NM * (size_t)R * R * R;
    if (lds > kMaxGroupLds)
        return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
    switch (g->cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        if (il)                                                                                    \
            P3MK_TRY(launch_lds(k_assign_halo<P, T, R2C, true, MODE>, grid, lds,  \
                               st, s->sorted, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T, g->M,   \
                               Hi, m, s->perm, gv, n));                                             \
        else                                                                                       \
            P3MK_TRY(launch_lds(k_assign_halo<P, T, R2C, false, MODE>, grid, lds, \
                               st, s->sorted, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T, g->M,   \
                               Hi, m, s->perm, gv, n));                                             \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("assign");
}

template <typename T, bool R2C>
int launch_interp_ik(const p3mk_geom *g, int n, const p3mk_psort *s, void *const F[3],
                     double prefac, double *fk, p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    const T *F0 = (const T *)F[0], *F1 = (const T *)F[1], *F2 = (const T *)F[2];
    hipStream_t st = p3mk_stream(qu);
    const bool il = g->interlaced != 0;
    // one field component's halo region in LDS at a time (one plane per mesh)
    const int NM = il ? 2 : 1;
    const int R = s->T + g->cao - 1 + (il ? 1 : 0);
    const size_t lds = sizeof(T) * NM * (size_t)R * R * R;
    if (lds > kMaxGroupLds)
        return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
    const dim3 grid((unsigned)((long)(n + kChunk - 1) / kChunk + s->ntiles));
    switch (g->cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        if (il)                                                                                    \
            P3MK_TRY(launch_lds(k_interp_ik_tile<P, T, R2C, true>, grid, lds, st, \
                               s->sorted, s->perm, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T,    \
                               g->M, Hi, F0, F1, F2, prefac, fk, (long)n));                         \
        else                                                                                       \
            P3MK_TRY(launch_lds(k_interp_ik_tile<P, T, R2C, false>, grid, lds,    \
                               st, s->sorted, s->perm, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T, \
                               g->M, Hi, F0, F1, F2, prefac, fk, (long)n));                         \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("interp_ik");
}

template <typename T, bool R2C>
int launch_interp_ad(const p3mk_geom *g, int n, const p3mk_psort *s, const void *phi,
                     double prefac, const double *selfb, double *fk, double *f, p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    const double h = g->L / g->M;
    const T *ph = (const T *)phi;
    hipStream_t st = p3mk_stream(qu);
    const bool il = g->interlaced != 0;
    // the item's halo region of phi in LDS (one plane per mesh)
    const int NM = il ? 2 : 1;
    const int R = s->T + g->cao - 1 + (il ? 1 : 0);
    const size_t lds = sizeof(T) * NM * (size_t)R * R * R;
    if (lds > kMaxGroupLds)
        return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
    const dim3 grid((unsigned)((long)(n + kChunk - 1) / kChunk + s->ntiles));
    switch (g->cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        if (il)                                                                                    \
            P3MK_TRY(launch_lds(k_interp_ad_tile<P, T, R2C, true>, grid, lds, st, \
                               s->sorted, s->perm, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T,    \
                               g->M, Hi, 1.0 / g->L, ph, prefac, selfb, h, fk, f, (long)n));        \
        else                                                                                       \
            P3MK_TRY(launch_lds(k_interp_ad_tile<P, T, R2C, false>, grid, lds,    \
                               st, s->sorted, s->perm, s->start, s->cstart, s->imap, s->ntiles, s->nt, s->T, \
                               g->M, Hi, 1.0 / g->L, ph, prefac, selfb, h, fk, f, (long)n));        \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("interp_ad");
}

template <typename T, bool R2C>
int assign_mode_t(const p3mk_geom *g, int n, p3mk_psort *s, void *mesh, int mode, const double *gv,
                  p3mk_queue *qu) {
    switch (mode) {
    case kCharge: return launch_assign<T, R2C, kCharge>(g, n, s, mesh, gv, qu);
    case kWeighted: return launch_assign<T, R2C, kWeighted>(g, n, s, mesh, gv, qu);
    default: return launch_assign<T, R2C, kDipole>(g, n, s, mesh, gv, qu);
    }
}

int assign_mode(const p3mk_geom *g, int n, p3mk_psort *s, void *mesh, int mode, const double *gv,
                p3mk_queue *qu) {
    if (g->fp32)
        return g->r2c ? assign_mode_t<float, true>(g, n, s, mesh, mode, gv, qu)
                      : assign_mode_t<float, false>(g, n, s, mesh, mode, gv, qu);
    return g->r2c ? assign_mode_t<double, true>(g, n, s, mesh, mode, gv, qu)
                  : assign_mode_t<double, false>(g, n, s, mesh, mode, gv, qu);
}

template <typename T, bool R2C>
int launch_adjoint(const p3mk_geom *g, int ad, int n, const p3mk_psort *s, const double *q,
                   const double *gv, void *const F[3], const void *lam, double prefac,
                   const double *selfb, double *dpos, double *dq, p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    const double h = g->L / g->M;
    dim3 grid((n + kThreads - 1) / kThreads);
    const T *F0 = (const T *)F[0], *F1 = (const T *)F[1], *F2 = (const T *)F[2];
    const T *la = (const T *)lam;
    hipStream_t st = p3mk_stream(qu);
    const bool il = g->interlaced != 0;
    switch (g->cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        if (ad) {                                                                                  \
            if (il)                                                                                \
                hipLaunchKernelGGL((k_adjoint_ad<P, T, R2C, true>), grid, dim3(kThreads), 0, st, n, \
                                   s->sorted, s->perm, q, gv, Hi, g->M, F0, la, prefac, selfb, h,  \
                                   dpos, dq);                                                      \
            else                                                                                   \
                hipLaunchKernelGGL((k_adjoint_ad<P, T, R2C, false>), grid, dim3(kThreads), 0, st,  \
                                   n, s->sorted, s->perm, q, gv, Hi, g->M, F0, la, prefac, selfb,  \
                                   h, dpos, dq);                                                   \
        } else {                                                                                   \
            if (il)                                                                                \
                hipLaunchKernelGGL((k_adjoint_ik<P, T, R2C, true>), grid, dim3(kThreads), 0, st, n, \
                                   s->sorted, s->perm, q, gv, Hi, g->M, F0, F1, F2, la, prefac,    \
                                   dpos, dq);                                                      \
            else                                                                                   \
                hipLaunchKernelGGL((k_adjoint_ik<P, T, R2C, false>), grid, dim3(kThreads), 0, st,  \
                                   n, s->sorted, s->perm, q, gv, Hi, g->M, F0, F1, F2, la, prefac, \
                                   dpos, dq);                                                      \
        }                                                                                          \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("adjoint");
}

}  // namespace

extern "C" int p3mk_assign(const p3mk_geom *g, int n, const double *x, const double *y,
                           const double *z, const double *q, p3mk_psort *s, void *mesh,
                           p3mk_queue *qu) {
    if (n > s->nmax || g->M != s->M)
        return p3mk_fail(P3MK_EINVAL, "particle sort was made for another size", 0);
    hipStream_t st = p3mk_stream(qu);
    s->n = n;
    P3MK_TRY(p3mk_fill_zero(mesh, p3mk_mesh_bytes(g), qu));
    if (n <= 0)
        return P3MK_OK;
    const double Hi = (double)g->M / g->L;
    const int nb = (n + 255) / 256;
    const size_t nt1 = (size_t)s->ntiles + 1;
    if (s->nf > 0) {
        const int T3 = s->T * s->T * s->T;
        P3MK_TRY(p3mk_fill_zero(s->fcount, sizeof(int) * ((size_t)s->nf + 1), qu));
        hipLaunchKernelGGL(k_fine_count, dim3(nb), dim3(256), 0, st, n, x, y, z, Hi, g->cao, g->M, s->T,
                           s->nt, s->key, s->rank, s->fcount);
        P3MK_TRY(p3mk_launch_check("base count"));
        P3MK_TRY(p3mk_exclusive_scan(s->fcount, s->nf, s->fstart, s->fscan_scratch, qu));
        hipLaunchKernelGGL(k_fine_scatter, dim3(nb), dim3(256), 0, st, n, x, y, z, q, s->key, s->rank,
                           s->fstart, s->perm, s->sorted);
        P3MK_TRY(p3mk_launch_check("base scatter"));
        hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)((s->ntiles + 256) / 256)), dim3(256), 0, st,
                           s->fstart, s->ntiles, T3, s->start, s->count, s->chunks);
        P3MK_TRY(p3mk_launch_check("tile ranges"));
    } else if (s->ntiles <= kSortMaxTiles) {
        P3MK_TRY(zero_count_cursor(s->count, s->cursor, nt1, qu));
        // about 64 blocks: one global atomic per (block, tile) instead of per particle
        const int per_block = std::max(4096, (n + 63) / 64);
        const int blocks = (n + per_block - 1) / per_block;
        const int ntiles = (int)s->ntiles;
        hipLaunchKernelGGL(k_tile_count_agg, dim3(blocks), dim3(kSortThreads),
                           sizeof(int) * (size_t)ntiles, st, n, per_block, x, y, z, Hi, g->cao,
                           g->M, s->T, s->nt, ntiles, s->key, s->count);
        P3MK_TRY(p3mk_launch_check("tile count"));
        P3MK_TRY(p3mk_exclusive_scan(s->count, s->ntiles, s->start, s->scan_scratch, qu));
        hipLaunchKernelGGL(k_tile_scatter_agg, dim3(blocks), dim3(kSortThreads),
                           2 * sizeof(int) * (size_t)ntiles, st, n, per_block, x, y, z, q, ntiles,
                           s->key, s->start, s->cursor, s->perm, s->sorted);
        P3MK_TRY(p3mk_launch_check("tile scatter"));
    } else {
        P3MK_TRY(zero_count_cursor(s->count, s->cursor, nt1, qu));
        hipLaunchKernelGGL(k_tile_count, dim3(nb), dim3(256), 0, st, n, x, y, z, Hi, g->cao, g->M,
                           s->T, s->nt, s->key, s->count);
        P3MK_TRY(p3mk_launch_check("tile count"));
        P3MK_TRY(p3mk_exclusive_scan(s->count, s->ntiles, s->start, s->scan_scratch, qu));
        hipLaunchKernelGGL(k_tile_scatter, dim3(nb), dim3(256), 0, st, n, x, y, z, q, s->key,
                           s->start, s->cursor, s->perm, s->sorted);
        P3MK_TRY(p3mk_launch_check("tile scatter"));
    }
    if (s->nf == 0) {
        hipLaunchKernelGGL(k_tile_chunks, dim3((unsigned)((s->ntiles + 255) / 256)), dim3(256), 0, st,
                           s->count, s->ntiles, s->chunks);
        P3MK_TRY(p3mk_launch_check("tile chunks"));
    }
    P3MK_TRY(p3mk_exclusive_scan(s->chunks, s->ntiles, s->cstart, s->scan_scratch, qu));
    hipLaunchKernelGGL(k_item_map, dim3((unsigned)((s->ntiles + 255) / 256)), dim3(256), 0, st,
                       s->cstart, s->ntiles, s->imap);
    P3MK_TRY(p3mk_launch_check("item map"));
    return assign_mode(g, n, s, mesh, kCharge, nullptr, qu);
}

extern "C" int p3mk_assign_weighted(const p3mk_geom *g, int n, p3mk_psort *s, int mode,
                                    const double *gv, void *mesh, p3mk_queue *qu) {
    if (n != s->n || g->M != s->M)
        return p3mk_fail(P3MK_EINVAL, "weighted assignment needs the sort of the same particles", 0);
    if (mode != 1 && mode != 2)
        return p3mk_fail(P3MK_EINVAL, "weighted assignment mode must be 1 or 2", 0);
    P3MK_TRY(p3mk_fill_zero(mesh, p3mk_mesh_bytes(g), qu));
    if (n <= 0)
        return P3MK_OK;
    return assign_mode(g, n, s, mesh, mode, gv, qu);
}

extern "C" int p3mk_interp_ik(const p3mk_geom *g, int n, const p3mk_psort *s, void *const F[3],
                              double prefac, double *fk, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (n != s->n)
        return p3mk_fail(P3MK_EINVAL, "interpolation needs the sort of the same particles", 0);
    if (g->fp32)
        return g->r2c ? launch_interp_ik<float, true>(g, n, s, F, prefac, fk, qu)
                      : launch_interp_ik<float, false>(g, n, s, F, prefac, fk, qu);
    return g->r2c ? launch_interp_ik<double, true>(g, n, s, F, prefac, fk, qu)
                  : launch_interp_ik<double, false>(g, n, s, F, prefac, fk, qu);
}

extern "C" int p3mk_interp_ad(const p3mk_geom *g, int n, const p3mk_psort *s, const void *phi,
                              double prefac, const double *selfb, double *fk, double *f,
                              p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (n != s->n)
        return p3mk_fail(P3MK_EINVAL, "interpolation needs the sort of the same particles", 0);
    if (selfb != nullptr && f == nullptr)
        return p3mk_fail(P3MK_EINVAL, "self-force correction needs the total-force array", 0);
    if (g->fp32)
        return g->r2c ? launch_interp_ad<float, true>(g, n, s, phi, prefac, selfb, fk, f, qu)
                      : launch_interp_ad<float, false>(g, n, s, phi, prefac, selfb, fk, f, qu);
    return g->r2c ? launch_interp_ad<double, true>(g, n, s, phi, prefac, selfb, fk, f, qu)
                  : launch_interp_ad<double, false>(g, n, s, phi, prefac, selfb, fk, f, qu);
}

extern "C" int p3mk_adjoint_gather(const p3mk_geom *g, int ad, int n, const p3mk_psort *s,
                                   const double *q, const double *gv, void *const F[3],
                                   const void *lam, double prefac, const double *selfb,
                                   double *dpos, double *dq, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (n != s->n)
        return p3mk_fail(P3MK_EINVAL, "adjoint needs the sort of the same particles", 0);
    if (g->fp32)
        return g->r2c ? launch_adjoint<float, true>(g, ad, n, s, q, gv, F, lam, prefac, selfb, dpos, dq, qu)
                      : launch_adjoint<float, false>(g, ad, n, s, q, gv, F, lam, prefac, selfb, dpos, dq, qu);
    return g->r2c ? launch_adjoint<double, true>(g, ad, n, s, q, gv, F, lam, prefac, selfb, dpos, dq, qu)
                  : launch_adjoint<double, false>(g, ad, n, s, q, gv, F, lam, prefac, selfb, dpos, dq, qu);
}

// ---- batched kernels over many systems ----------------------------------------
//
// Particles of system s occupy slots [p0_s, p0_s + n_s) of the concatenated
// arrays; tiles of system s are global tiles [tile0_s, tile0_s + nt_s^3), so
// after the sort system s's particles are again the slots [p0_s, p0_s + n_s).

namespace {

// The batched sort with per-block aggregation (as k_tile_count_agg and
// k_tile_scatter_agg).  A block's particles are a contiguous range, so they
// belong to a run of neighbouring systems whose tiles are the key window
// [lo, hi).  A block whose window fits in LDS counts and ranks there, with
// one global atomic per (block, tile); any other block takes one global
// atomic per particle.
__device__ __forceinline__ void btile_window(const BSys *__restrict__ sys, const int *__restrict__ psys,
                                             int b, int e, long &lo, long &hi) {
    const BSys &S0 = sys[psys[b]], &S1 = sys[psys[e - 1]];
    lo = S0.tile0;
    hi = S1.tile0 + (long)S1.nt * S1.nt * S1.nt;
}

__global__ void __launch_bounds__(kSortThreads)
k_btile_count_agg(int ntot, int per_block, const BSys *__restrict__ sys, const int *__restrict__ psys,
                  const double *__restrict__ x, const double *__restrict__ y,
                  const double *__restrict__ z, int *__restrict__ key, int *__restrict__ count) {
    extern __shared__ int hist[];
    const int b = blockIdx.x * per_block, e = min(ntot, b + per_block);
    if (b >= e)
        return;
    long lo, hi;
    btile_window(sys, psys, b, e, lo, hi);
    const int w = (int)min(hi - lo, kSortMaxTiles + 1);
    const bool agg = w <= kSortMaxTiles;
    if (agg)
        for (int k = threadIdx.x; k < w; k += blockDim.x)
            hist[k] = 0;
    __syncthreads();
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const BSys &S = sys[psys[i]];
        const int tx = axis_base(x[i], S.Hi, S.cao, S.M) / S.T, ty = axis_base(y[i], S.Hi, S.cao, S.M) / S.T,
                  tz = axis_base(z[i], S.Hi, S.cao, S.M) / S.T;
        const int k = (int)(S.tile0 + (tx * S.nt + ty) * S.nt + tz);
        key[i] = k;
        if (agg)
            atomicAdd(&hist[k - lo], 1);
        else
            atomicAdd(&count[k], 1);
    }
    if (agg) {
        __syncthreads();
        for (int k = threadIdx.x; k < w; k += blockDim.x)
            if (hist[k] != 0)
                atomicAdd(&count[lo + k], hist[k]);
    }
}

__global__ void __launch_bounds__(kSortThreads)
k_btile_scatter_agg(int ntot, int per_block, const BSys *__restrict__ sys, const int *__restrict__ psys,
                    const double *__restrict__ x, const double *__restrict__ y,
                    const double *__restrict__ z, const double *__restrict__ q,
                    const int *__restrict__ key, const int *__restrict__ start,
                    int *__restrict__ cursor, int *__restrict__ perm, double4 *__restrict__ sorted) {
    extern __shared__ int lds_i[];
    const int b = blockIdx.x * per_block, e = min(ntot, b + per_block);
    if (b >= e)
        return;
    long lo, hi;
    btile_window(sys, psys, b, e, lo, hi);
    const int w = (int)min(hi - lo, kSortMaxTiles + 1);
    if (w > kSortMaxTiles) {
        for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
            const int k = key[i];
            const int p = start[k] + atomicAdd(&cursor[k], 1);
            perm[p] = i;
            sorted[p] = make_double4(x[i], y[i], z[i], q[i]);
        }
        return;
    }
    int *hist = lds_i, *base = lds_i + kSortMaxTiles;
    for (int k = threadIdx.x; k < w; k += blockDim.x)
        hist[k] = 0;
    __syncthreads();
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x)
        atomicAdd(&hist[key[i] - lo], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < w; k += blockDim.x) {
        if (hist[k] != 0)
            base[k] = start[lo + k] + atomicAdd(&cursor[lo + k], hist[k]);
        hist[k] = 0;
    }
    __syncthreads();
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const int k = key[i] - (int)lo;
        const int p = base[k] + atomicAdd(&hist[k], 1);
        perm[p] = i;
        sorted[p] = make_double4(x[i], y[i], z[i], q[i]);
    }
}

template <typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_bassign(const BSys *__restrict__ sys, long ntiles, const int *__restrict__ start,
          const int *__restrict__ cstart, const int4 *__restrict__ imap,
          const double4 *__restrict__ sorted) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    if (item >= cstart[ntiles])
        return;
    const int4 m = imap[item];
    const long tile = m.x;
    const BSys &S = sys[m.z];
    const int chunk = m.y;
    const int pb = start[tile] + chunk * kChunk;
    const int pe = min(start[tile + 1], pb + kChunk);
    T *mesh = (T *)S.buf[kBufMesh];
    double *acc = reinterpret_cast<double *>(smem_raw);
    const long tl = tile - S.tile0;
    // the row-parallel deposit of the single-system form: the system's
    // region, then the wavefronts' staging areas
    switch (S.cao) {
#define CASE(P)                                                                                    \
    case P: {                                                                                      \
        const RowRegion rg = row_region(S.T, P, IL);                                               \
        T *stage = reinterpret_cast<T *>(smem_raw + sizeof(double) * (size_t)rg.size);             \
        assign_rows_item<P, T, R2C, IL, kCharge, double>(sorted, pb, pe, tl, S.nt, S.T, S.M, S.Hi, \
                                                         mesh, nullptr, nullptr, acc, stage);      \
        break;                                                                                     \
    }
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default: break;
    }
}

// back-interpolation of every system: work items as in assignment (a chunk
// of one tile's particles), the tile's halo region staged in LDS
template <typename T, bool R2C, bool IL, bool AD>
__global__ void __launch_bounds__(kThreads)
k_binterp(const BSys *__restrict__ sys, long ntiles, const int *__restrict__ start,
          const int *__restrict__ cstart, const int4 *__restrict__ imap,
          const double4 *__restrict__ sorted, const int *__restrict__ perm, int ntot,
          double *__restrict__ fk, double *__restrict__ ftot) {
    extern __shared__ unsigned char smem_raw[];
    const int item = blockIdx.x;
    if (item >= cstart[ntiles])
        return;
    const int4 m = imap[item];
    const long tile = m.x;
    const BSys &S = sys[m.z];
    const int chunk = m.y;
    const int pb = start[tile] + chunk * kChunk;
    const int pe = min(start[tile + 1], pb + kChunk);
    const long tl = tile - S.tile0;
    T *lds = reinterpret_cast<T *>(smem_raw);
    if (AD) {
        const T *phi = (const T *)S.buf[kBufMesh];
        switch (S.cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        interp_ad_item<P, T, R2C, IL>(sorted, perm, pb, pe, tl, S.nt, S.T, S.M, S.Hi, 1.0 / S.L,    \
                                      phi, S.prefac, S.selfb, S.h, fk, ftot, ntot, lds);          \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
        default: break;
        }
    } else {
        const T *F0 = (const T *)S.buf[kBufF0], *F1 = (const T *)S.buf[kBufF0 + 1],
                *F2 = (const T *)S.buf[kBufF0 + 2];
        switch (S.cao) {
#define CASE(P)                                                                                    \
    case P:                                                                                        \
        interp_ik_item<P, T, R2C, IL>(sorted, perm, pb, pe, tl, S.nt, S.T, S.M, S.Hi, F0, F1, F2,   \
                                      S.prefac, fk, ntot, lds);                                   \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
        default: break;
        }
    }
}

template <typename T, bool R2C, bool IL>
void launch_bassign(const BSys *sys, long ntiles, long max_items, size_t lds, const int *start,
                    const int *cstart, const int4 *imap, const double4 *sorted, hipStream_t st) {
    hipLaunchKernelGGL((k_bassign<T, R2C, IL>), dim3((unsigned)max_items), dim3(kThreads), lds, st,
                       sys, ntiles, start, cstart, imap, sorted);
}

template <typename T, bool R2C, bool IL>
void launch_binterp(bool ad, const BSys *sys, long ntiles, long max_items, size_t lds,
                    const int *start, const int *cstart, const int4 *imap, const double4 *sorted,
                    const int *perm, int ntot, double *fk, double *ftot, hipStream_t st) {
    const dim3 grid((unsigned)max_items);
    if (ad)
        hipLaunchKernelGGL((k_binterp<T, R2C, IL, true>), grid, dim3(kThreads), lds, st, sys, ntiles,
                           start, cstart, imap, sorted, perm, ntot, fk, ftot);
    else
        hipLaunchKernelGGL((k_binterp<T, R2C, IL, false>), grid, dim3(kThreads), lds, st, sys,
                           ntiles, start, cstart, imap, sorted, perm, ntot, fk, ftot);
}

}  // namespace

namespace p3m {

int bpart_sort(const BForm &f, const BSys *sys, const int *psys, const long *tile0s, int nsys,
               long ntiles, int ntot, const double *x, const double *y, const double *z,
               const double *q, int *key, int *count, int *start, int *cursor, int *perm,
               int *chunks, int *cstart, int4 *imap, int *scan_scratch, double4 *sorted,
               p3mk_queue *qu) {
    (void)f;
    if (ntot <= 0)
        return P3MK_OK;
    hipStream_t st = p3mk_stream(qu);
    P3MK_TRY(zero_count_cursor(count, cursor, (size_t)ntiles + 1, qu));
    // blocks of 4096 particles: one system or a few neighbouring ones, whose
    // tiles fit in LDS unless the tiles are very small
    constexpr int per_block = 4096;
    const int blocks = (ntot + per_block - 1) / per_block;
    hipLaunchKernelGGL(k_btile_count_agg, dim3(blocks), dim3(kSortThreads),
                       sizeof(int) * (size_t)kSortMaxTiles, st, ntot, per_block, sys, psys, x, y, z,
                       key, count);
    P3MK_TRY(p3mk_launch_check("batched tile count"));
    P3MK_TRY(p3mk_exclusive_scan(count, ntiles, start, scan_scratch, qu));
    hipLaunchKernelGGL(k_btile_scatter_agg, dim3(blocks), dim3(kSortThreads),
                       2 * sizeof(int) * (size_t)kSortMaxTiles, st, ntot, per_block, sys, psys, x, y,
                       z, q, key, start, cursor, perm, sorted);
    P3MK_TRY(p3mk_launch_check("batched tile scatter"));
    hipLaunchKernelGGL(k_tile_chunks, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0, st,
                       count, ntiles, chunks);
    P3MK_TRY(p3mk_launch_check("batched tile chunks"));
    P3MK_TRY(p3mk_exclusive_scan(chunks, ntiles, cstart, scan_scratch, qu));
    hipLaunchKernelGGL(k_bitem_map, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0, st, cstart,
                       ntiles, tile0s, nsys, imap);
    return p3mk_launch_check("batched item map");
}

int bpart_assign(const BForm &f, const BSys *sys, long ntiles, long max_items, size_t lds,
                 const int *start, const int *cstart, const int4 *imap, const double4 *sorted,
                 p3mk_queue *qu) {
    if (max_items <= 0)
        return P3MK_OK;
    if (lds > kMaxGroupLds)
        return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
    hipStream_t st = p3mk_stream(qu);
#define GO(T, R, I) launch_bassign<T, R, I>(sys, ntiles, max_items, lds, start, cstart, imap, sorted, st)
    if (f.fp32) {
        if (f.r2c) { if (f.interlaced) GO(float, true, true); else GO(float, true, false); }
        else { if (f.interlaced) GO(float, false, true); else GO(float, false, false); }
    } else {
        if (f.r2c) { if (f.interlaced) GO(double, true, true); else GO(double, true, false); }
        else { if (f.interlaced) GO(double, false, true); else GO(double, false, false); }
    }
#undef GO
    return p3mk_launch_check("batched assignment");
}

int bpart_interp(const BForm &f, const BSys *sys, long ntiles, long max_items, size_t lds,
                 const int *start, const int *cstart, const int4 *imap, const double4 *sorted,
                 const int *perm, int ntot, double *fk, double *ftot, p3mk_queue *qu) {
    if (ntot <= 0 || max_items <= 0)
        return P3MK_OK;
    if (lds > kMaxGroupLds)
        return p3mk_fail(P3MK_EINVAL, "tile does not fit in LDS", 0);
    hipStream_t st = p3mk_stream(qu);
    const bool ad = f.ad != 0;
#define GO(T, R, I) launch_binterp<T, R, I>(ad, sys, ntiles, max_items, lds, start, cstart, imap, sorted, perm, ntot, fk, ftot, st)
    if (f.fp32) {
        if (f.r2c) { if (f.interlaced) GO(float, true, true); else GO(float, true, false); }
        else { if (f.interlaced) GO(float, false, true); else GO(float, false, false); }
    } else {
        if (f.r2c) { if (f.interlaced) GO(double, true, true); else GO(double, true, false); }
        else { if (f.interlaced) GO(double, false, true); else GO(double, false, false); }
    }
#undef GO
    return p3mk_launch_check("batched back-interpolation");
}

}  // namespace p3m
