// Particle <-> mesh kernels: charge assignment and back-interpolation.
//
// Charge assignment (charge-assign.c:119-176, 215-267, 1067-1133,
// 1243-1311): particle i adds q W_a(x) W_b(y) W_c(z) to the P^3 mesh points
// of its stencil, on mesh 0 and, interlaced, on mesh 1 (h/2-shifted, held in
// the imaginary part).  Weights are evaluated at the reference's quantised
// offsets (device_common.hpp) instead of read from 7 MB tables.
//
// Back-interpolation (charge-assign.c:601-643, 646-688, 727-769, 849-885,
// 1352-1391, 1468-1502) recomputes the same weights instead of caching
// cf / dQ / ca_ind per particle, so it streams positions, not N*P^3 caches.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "k_internal.hpp"

using namespace p3m;

namespace {

constexpr int kThreads = 256;

// ---- assignment ---------------------------------------------------------

template <int P, typename T, bool R2C>
__global__ void __launch_bounds__(kThreads)
k_assign_atomic(int n, const double *__restrict__ x, const double *__restrict__ y,
                const double *__restrict__ z, const double *__restrict__ q, double Hi, int M,
                T *__restrict__ mesh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ii = blockIdx.y; // 0: mesh 0, 1: h/2-shifted mesh 1
    if (i >= n)
        return;
    const double half = 0.5 * ii;
    int b0, b1, b2;
    T w0[P], w1[P], w2[P];
    axis_weights<P, T>(x[i], Hi, half, M, b0, w0);
    axis_weights<P, T>(y[i], Hi, half, M, b1, w1);
    axis_weights<P, T>(z[i], Hi, half, M, b2, w2);
    const T qi = (T)q[i];
    int kk[P];
#pragma unroll
    for (int c = 0; c < P; c++)
        kk[c] = wrap1(b2 + c, M);
#pragma unroll
    for (int a = 0; a < P; a++) {
        const int ia = wrap1(b0 + a, M);
        const T t0 = qi * w0[a];
#pragma unroll
        for (int b = 0; b < P; b++) {
            const int jb = wrap1(b1 + b, M);
            const T t1 = t0 * w1[b];
            const long row = ((long)ia * M + jb) * M;
#pragma unroll
            for (int c = 0; c < P; c++) {
                const T v = t1 * w2[c];
                if (R2C)
                    atomicAdd(&mesh[row + kk[c]], v);
                else
                    atomicAdd(&mesh[2 * (row + kk[c]) + ii], v);
            }
        }
    }
}

template <typename T, bool R2C>
int launch_assign(const p3mk_geom *g, int n, const double *x, const double *y, const double *z,
                  const double *q, void *mesh, p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    dim3 grid((n + kThreads - 1) / kThreads, g->interlaced ? 2 : 1);
    T *m = (T *)mesh;
    hipStream_t s = p3mk_stream(qu);
    switch (g->cao) {
#define CASE(P)                                                                                  \
    case P:                                                                                      \
        hipLaunchKernelGGL((k_assign_atomic<P, T, R2C>), grid, dim3(kThreads), 0, s, n, x, y, z, \
                           q, Hi, g->M, m);                                                     \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("assign");
}

// ---- ik back-interpolation ----------------------------------------------

// Reads mesh value of component ii (c2c: re / im of a complex mesh) or the
// real mesh (r2c).
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

template <int P, typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_interp_ik(int n, const double *__restrict__ x, const double *__restrict__ y,
            const double *__restrict__ z, const double *__restrict__ q, double Hi, int M,
            const T *__restrict__ F0, const T *__restrict__ F1, const T *__restrict__ F2,
            double prefac, double *__restrict__ fk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    constexpr int NM = IL ? 2 : 1;
    Acc<T> acc[3];
    const double xi = x[i], yi = y[i], zi = z[i];
    // interlaced: each mesh contributes with 0.5 (charge-assign.c:674-675)
    const T scale = (T)(IL ? 0.5 * prefac : prefac) * (T)q[i];
#pragma unroll
    for (int ii = 0; ii < NM; ii++) {
        int b0, b1, b2;
        T w0[P], w1[P], w2[P];
        const double half = 0.5 * ii;
        axis_weights<P, T>(xi, Hi, half, M, b0, w0);
        axis_weights<P, T>(yi, Hi, half, M, b1, w1);
        axis_weights<P, T>(zi, Hi, half, M, b2, w2);
        int kk[P];
#pragma unroll
        for (int c = 0; c < P; c++)
            kk[c] = wrap1(b2 + c, M);
#pragma unroll
        for (int a = 0; a < P; a++) {
            const int ia = wrap1(b0 + a, M);
#pragma unroll
            for (int b = 0; b < P; b++) {
                const int jb = wrap1(b1 + b, M);
                const T t1 = w0[a] * w1[b];
                const long row = ((long)ia * M + jb) * M;
#pragma unroll
                for (int c = 0; c < P; c++) {
                    const long idx = row + kk[c];
                    const T B = scale * (t1 * w2[c]);
                    acc[0].add(-mesh_val<T, R2C>(F0, idx, ii) * B);
                    acc[1].add(-mesh_val<T, R2C>(F1, idx, ii) * B);
                    acc[2].add(-mesh_val<T, R2C>(F2, idx, ii) * B);
                }
            }
        }
    }
    fk[i] += acc[0].value();
    fk[(long)n + i] += acc[1].value();
    fk[2L * n + i] += acc[2].value();
}

template <typename T, bool R2C>
int launch_interp_ik(const p3mk_geom *g, int n, const double *x, const double *y, const double *z,
                     const double *q, void *const F[3], double prefac, double *fk,
                     p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    dim3 grid((n + kThreads - 1) / kThreads);
    const T *F0 = (const T *)F[0], *F1 = (const T *)F[1], *F2 = (const T *)F[2];
    hipStream_t s = p3mk_stream(qu);
    const bool il = g->interlaced != 0;
    switch (g->cao) {
#define CASE(P)                                                                                   \
    case P:                                                                                       \
        if (il)                                                                                   \
            hipLaunchKernelGGL((k_interp_ik<P, T, R2C, true>), grid, dim3(kThreads), 0, s, n, x, \
                               y, z, q, Hi, g->M, F0, F1, F2, prefac, fk);                       \
        else                                                                                      \
            hipLaunchKernelGGL((k_interp_ik<P, T, R2C, false>), grid, dim3(kThreads), 0, s, n, x, \
                               y, z, q, Hi, g->M, F0, F1, F2, prefac, fk);                       \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("interp_ik");
}

// ---- ad back-interpolation + self force ----------------------------------

template <int P, typename T, bool R2C, bool IL>
__global__ void __launch_bounds__(kThreads)
k_interp_ad(int n, const double *__restrict__ x, const double *__restrict__ y,
            const double *__restrict__ z, const double *__restrict__ q, double Hi, double Leni,
            int M, const T *__restrict__ phi, double prefac, const double *__restrict__ selfb,
            double h, double *__restrict__ fk, double *__restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    constexpr int NM = IL ? 2 : 1;
    Acc<T> acc[3];
    const double xi = x[i], yi = y[i], zi = z[i], qi = q[i];
    // dQ_d = W'.W.W q/L; force_d -= prefac * phi * dQ_d, interlaced x 0.5
    // (charge-assign.c:1123-1125, 1374-1377, 758-760)
    const T scale = (T)((IL ? 0.5 * prefac : prefac) * (qi * Leni));
#pragma unroll
    for (int ii = 0; ii < NM; ii++) {
        int b0, b1, b2;
        T w0[P], w1[P], w2[P], d0[P], d1[P], d2[P];
        const double half = 0.5 * ii;
        axis_weights_d<P, T>(xi, Hi, half, M, b0, w0, d0);
        axis_weights_d<P, T>(yi, Hi, half, M, b1, w1, d1);
        axis_weights_d<P, T>(zi, Hi, half, M, b2, w2, d2);
        int kk[P];
#pragma unroll
        for (int c = 0; c < P; c++)
            kk[c] = wrap1(b2 + c, M);
#pragma unroll
        for (int a = 0; a < P; a++) {
            const int ia = wrap1(b0 + a, M);
#pragma unroll
            for (int b = 0; b < P; b++) {
                const int jb = wrap1(b1 + b, M);
                const long row = ((long)ia * M + jb) * M;
                const T wx = d0[a] * w1[b], wy = w0[a] * d1[b], wz = w0[a] * w1[b];
#pragma unroll
                for (int c = 0; c < P; c++) {
                    const T B = scale * mesh_val<T, R2C>(phi, row + kk[c], ii);
                    acc[0].add(-B * (wx * w2[c]));
                    acc[1].add(-B * (wy * w2[c]));
                    acc[2].add(-B * (wz * d2[c]));
                }
            }
        }
    }
    fk[i] += acc[0].value();
    fk[(long)n + i] += acc[1].value();
    fk[2L * n + i] += acc[2].value();
    if (selfb != nullptr) {
        // f -= sum_m' b(m') q^2 sin(2 pi m'.r / h)  (p3m-ad-self-forces.c:122-135)
        const double q2 = qi * qi;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        int ind = 0;
        for (int m0 = -1; m0 <= 1; m0++)
            for (int m1 = -1; m1 <= 1; m1++)
                for (int m2 = -1; m2 <= 1; m2++) {
                    const double st =
                        q2 * sin(2 * 3.14159265358979323846 * (m0 * xi / h + m1 * yi / h + m2 * zi / h));
                    s0 += selfb[ind] * st;
                    s1 += selfb[ind + 1] * st;
                    s2 += selfb[ind + 2] * st;
                    ind += 3;
                }
        f[i] -= s0;
        f[(long)n + i] -= s1;
        f[2L * n + i] -= s2;
    }
}

template <typename T, bool R2C>
int launch_interp_ad(const p3mk_geom *g, int n, const double *x, const double *y, const double *z,
                     const double *q, const void *phi, double prefac, const double *selfb,
                     double *fk, double *f, p3mk_queue *qu) {
    const double Hi = (double)g->M / g->L;
    const double h = g->L / g->M;
    dim3 grid((n + kThreads - 1) / kThreads);
    const T *ph = (const T *)phi;
    hipStream_t s = p3mk_stream(qu);
    const bool il = g->interlaced != 0;
    switch (g->cao) {
#define CASE(P)                                                                                   \
    case P:                                                                                       \
        if (il)                                                                                   \
            hipLaunchKernelGGL((k_interp_ad<P, T, R2C, true>), grid, dim3(kThreads), 0, s, n, x, \
                               y, z, q, Hi, 1.0 / g->L, g->M, ph, prefac, selfb, h, fk, f);      \
        else                                                                                      \
            hipLaunchKernelGGL((k_interp_ad<P, T, R2C, false>), grid, dim3(kThreads), 0, s, n, x, \
                               y, z, q, Hi, 1.0 / g->L, g->M, ph, prefac, selfb, h, fk, f);      \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
    default:
        return p3mk_fail(P3MK_EINVAL, "charge assignment order must be 1..7", 0);
    }
    return p3mk_launch_check("interp_ad");
}

}  // namespace

extern "C" int p3mk_assign(const p3mk_geom *g, int n, const double *x, const double *y,
                           const double *z, const double *q, void *mesh, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (g->fp32)
        return g->r2c ? launch_assign<float, true>(g, n, x, y, z, q, mesh, qu)
                      : launch_assign<float, false>(g, n, x, y, z, q, mesh, qu);
    return g->r2c ? launch_assign<double, true>(g, n, x, y, z, q, mesh, qu)
                  : launch_assign<double, false>(g, n, x, y, z, q, mesh, qu);
}

extern "C" int p3mk_interp_ik(const p3mk_geom *g, int n, const double *x, const double *y,
                              const double *z, const double *q, void *const F[3], double prefac,
                              double *fk, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (g->fp32)
        return g->r2c ? launch_interp_ik<float, true>(g, n, x, y, z, q, F, prefac, fk, qu)
                      : launch_interp_ik<float, false>(g, n, x, y, z, q, F, prefac, fk, qu);
    return g->r2c ? launch_interp_ik<double, true>(g, n, x, y, z, q, F, prefac, fk, qu)
                  : launch_interp_ik<double, false>(g, n, x, y, z, q, F, prefac, fk, qu);
}

extern "C" int p3mk_interp_ad(const p3mk_geom *g, int n, const double *x, const double *y,
                              const double *z, const double *q, const void *phi, double prefac,
                              const double *selfb, double *fk, double *f, p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    if (selfb != nullptr && f == nullptr)
        return p3mk_fail(P3MK_EINVAL, "self-force correction needs the total-force array", 0);
    if (g->fp32)
        return g->r2c ? launch_interp_ad<float, true>(g, n, x, y, z, q, phi, prefac, selfb, fk, f, qu)
                      : launch_interp_ad<float, false>(g, n, x, y, z, q, phi, prefac, selfb, fk, f, qu);
    return g->r2c ? launch_interp_ad<double, true>(g, n, x, y, z, q, phi, prefac, selfb, fk, f, qu)
                  : launch_interp_ad<double, false>(g, n, x, y, z, q, phi, prefac, selfb, fk, f, qu);
}
