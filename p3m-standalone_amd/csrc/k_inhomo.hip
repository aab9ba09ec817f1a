// Kernels of the inhomogeneous error estimate and the error map
// (Generic_error_estimate_inhomo and Error_map, p3m-common.c:402-719).
//
// All FP64 on complex M^3 meshes (x-major, as the reference's c_ind).  The
// error kernel of aliasing copy m != 0 is, in the reference, a double sum
// over mesh points n and n2 (p3m-common.c:571-612); its terms factorise,
//   K_d(n2) += -(n2_d / L) k2(n2) * sum_n (n_d / L) k1(n),
// so the sum over n is one reduction per copy (two-stage: per-block partials
// in a fixed order, summed on the host), and the fill is one pass over n2.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "k_internal.hpp"

using namespace p3m;

namespace {

constexpr int kThreads = 256;
constexpr int kSumBlocks = 1024;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double sinc(double d) {
    const double pd = kPi * d;
    return (d == 0.0) ? 1.0 : sin(pd) / pd;
}

// mesh index -> signed wave number in [-M/2, M/2) (p3m-common.c:498-504)
__device__ __forceinline__ int signed_n(int i, int M) { return (i >= M / 2) ? i - M : i; }

// U and U^2 of the window (p3m-common.c:370-377): pow of the sinc product
__device__ __forceinline__ double window_u(int nx, int ny, int nz, int M, int p, int power) {
    return pow(sinc((double)nx / M) * sinc((double)ny / M) * sinc((double)nz / M),
               (double)(power * p));
}

// |k|^2 and the continuous influence function 4 pi / k^2 exp(-k^2 / 4 alpha^2)
// (p3m-common.c:313-320)
__device__ __forceinline__ double k_sq(int nx, int ny, int nz, double L) {
#pragma clang fp contract(off)
    const double f = 2.0 * kPi / L;
    return (f * f) * ((double)nx * nx + (double)ny * ny + (double)nz * nz);
}
__device__ __forceinline__ double g_cont(int nx, int ny, int nz, double L, double alpha) {
#pragma clang fp contract(off)
    const double k2 = k_sq(nx, ny, nz, L);
    return (4.0 * kPi / k2) * exp(-k2 / (4.0 * alpha * alpha));
}

// b / a of the reference's B and A without aliasing copies (mc = 0,
// p3m-common.c:356-396): (U^2 G K^2) / (U^2 (K^2 U^2))
__device__ __forceinline__ double g_opt(int nx, int ny, int nz, double L, double alpha, int M, int p) {
#pragma clang fp contract(off)
    const double u2 = window_u(nx, ny, nz, M, p, 2);
    const double k2 = k_sq(nx, ny, nz, L);
    const double b = (u2 * g_cont(nx, ny, nz, L, alpha)) * k2;
    const double a = u2 * (k2 * u2);
    return b / a;
}

__device__ __forceinline__ void point_n(long ind, int M, int &nx, int &ny, int &nz) {
    const int iz = (int)(ind % M), iy = (int)((ind / M) % M), ix = (int)(ind / ((long)M * M));
    nx = signed_n(ix, M);
    ny = signed_n(iy, M);
    nz = signed_n(iz, M);
}

// homogeneous part (p3m-common.c:506-537): K_d(n) = i (-2 pi n_d / L) (U^2 G - G_m)
__global__ void __launch_bounds__(kThreads)
k_homo(int M, double L, double alpha, int p, double2 *K0, double2 *K1, double2 *K2) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= (long)M * M * M)
        return;
    int n[3];
    point_n(ind, M, n[0], n[1], n[2]);
    double v[3] = {0.0, 0.0, 0.0};
    if (n[0] != 0 || n[1] != 0 || n[2] != 0) {
#pragma clang fp contract(off)
        const double G = g_opt(n[0], n[1], n[2], L, alpha, M, p);
        const double u2 = window_u(n[0], n[1], n[2], M, p, 2);
        const double gm = g_cont(n[0], n[1], n[2], L, alpha);
        const double k2 = (u2 * G) - gm;
        for (int d = 0; d < 3; d++)
            v[d] = ((-2.0 * kPi * n[d]) / L) * k2;
    }
    K0[ind] = make_double2(0.0, v[0]);
    K1[ind] = make_double2(0.0, v[1]);
    K2[ind] = make_double2(0.0, v[2]);
}

// per-block partial sums of (n_d / L) k1(n), k1 = 2 pi U(n) U(n + m M) G(n)
__global__ void __launch_bounds__(kThreads)
k_inhomo_sums(int M, double L, double alpha, int p, int mx, int my, int mz, double *__restrict__ part) {
    __shared__ double red[3][kThreads];
    const long M3 = (long)M * M * M;
    double s[3] = {0.0, 0.0, 0.0};
    for (long ind = (long)blockIdx.x * blockDim.x + threadIdx.x; ind < M3;
         ind += (long)gridDim.x * blockDim.x) {
        int n[3];
        point_n(ind, M, n[0], n[1], n[2]);
        if (n[0] == 0 && n[1] == 0 && n[2] == 0)
            continue;
        {
#pragma clang fp contract(off)
            const double G = g_opt(n[0], n[1], n[2], L, alpha, M, p);
            const double u = window_u(n[0], n[1], n[2], M, p, 1);
            const double um = window_u(n[0] + mx * M, n[1] + my * M, n[2] + mz * M, M, p, 1);
            const double k1 = ((2.0 * kPi * u) * um) * G;
            for (int d = 0; d < 3; d++)
                s[d] += ((double)n[d] / L) * k1;
        }
    }
    for (int d = 0; d < 3; d++)
        red[d][threadIdx.x] = s[d];
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int d = 0; d < 3; d++)
                red[d][threadIdx.x] += red[d][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; d++)
            part[3 * blockIdx.x + d] = red[d][0];
}

// aliasing copy m (p3m-common.c:571-612, factorised):
// K_d(n2) = i (-(n2_d / L) k2(n2) S_d), k2 = 2 pi U(n2) U(-n2 - m M) G(n2)
__global__ void __launch_bounds__(kThreads)
k_inhomo_fill(int M, double L, double alpha, int p, int mx, int my, int mz, double S0, double S1,
              double S2, double2 *K0, double2 *K1, double2 *K2) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= (long)M * M * M)
        return;
    int n[3];
    point_n(ind, M, n[0], n[1], n[2]);
    double v[3] = {0.0, 0.0, 0.0};
    if (n[0] != 0 || n[1] != 0 || n[2] != 0) {
#pragma clang fp contract(off)
        const double G = g_opt(n[0], n[1], n[2], L, alpha, M, p);
        const double u = window_u(n[0], n[1], n[2], M, p, 1);
        const double um = window_u(-n[0] - mx * M, -n[1] - my * M, -n[2] - mz * M, M, p, 1);
        const double k2 = ((2.0 * kPi * u) * um) * G;
        const double S[3] = {S0, S1, S2};
        for (int d = 0; d < 3; d++)
            v[d] = (-((double)n[d] / L) * k2) * S[d];
    }
    K0[ind] = make_double2(0.0, v[0]);
    K1[ind] = make_double2(0.0, v[1]);
    K2[ind] = make_double2(0.0, v[2]);
}

// K3 (+)= sum_d Re(K_d)^2 (p3m-common.c:543-560, 619-636)
__global__ void __launch_bounds__(kThreads)
k_square_add(long M3, const double2 *K0, const double2 *K1, const double2 *K2, double2 *K3, int first) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= M3)
        return;
    {
#pragma clang fp contract(off)
        const double kr = (K0[ind].x * K0[ind].x + K1[ind].x * K1[ind].x) + K2[ind].x * K2[ind].x;
        if (first)
            K3[ind] = make_double2(kr, 0.0);
        else
            K3[ind].x += kr;
    }
}

// the reference's product of rho^2(k) and K^2(k) (p3m-common.c:653-661):
// real parts with real parts, imaginary with imaginary
__global__ void __launch_bounds__(kThreads) k_partwise_product(long M3, double2 *Km, const double2 *K3) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= M3)
        return;
    Km[ind].x *= K3[ind].x;
    Km[ind].y *= K3[ind].y;
}

// a constant real part (the uniform density of p3m-common.c:477-479)
__global__ void __launch_bounds__(kThreads) k_fill_real(long M3, double v, double2 *mesh) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind < M3)
        mesh[ind] = make_double2(v, 0.0);
}

// per-particle charges of the two diagnostics: q_i^2, or |a_i - b_i|^2 of
// two force fields (3n, x/y/z blocks)
__global__ void __launch_bounds__(kThreads) k_square_charges(int n, const double *q, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = q[i] * q[i];
}
__global__ void __launch_bounds__(kThreads)
k_diff_charges(int n, const double *a, const double *b, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    {
#pragma clang fp contract(off)
        double s = 0.0;
        for (int d = 0; d < 3; d++) {
            const double t = b[(long)d * n + i] - a[(long)d * n + i];
            s += t * t;
        }
        out[i] = s;
    }
}

// rms_i = sum over the stencil of q_i^2 W Re(mesh) (collect_rms_nocf,
// charge-assign.c:1655-1702), weights of order P at the mesh's own spacing
template <int P>
__global__ void __launch_bounds__(kThreads)
k_collect(int n, int M, double L, const double *x, const double *y, const double *z,
          const double *q, const double2 *mesh, double *rms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double Hi = (double)M / L;
    int b[3];
    double w[3][P];
    axis_weights<P, double>(x[i], Hi, 0.0, M, b[0], w[0]);
    axis_weights<P, double>(y[i], Hi, 0.0, M, b[1], w[1]);
    axis_weights<P, double>(z[i], Hi, 0.0, M, b[2], w[2]);
    const double q2 = q[i] * q[i];
    double acc = 0.0;
    for (int a = 0; a < P; a++) {
        const int ia = wrap1(b[0] + a, M);
        const double t0 = q2 * w[0][a];
        for (int c = 0; c < P; c++) {
            const int ic = wrap1(b[1] + c, M);
            const double t1 = t0 * w[1][c];
            for (int e = 0; e < P; e++) {
                const int ie = wrap1(b[2] + e, M);
                acc += mesh[((long)ia * M + ic) * M + ie].x * (t1 * w[2][e]);
            }
        }
    }
    rms[i] = acc;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

extern "C" int p3mk_inhomo_scratch_doubles(void) { return 3 * kSumBlocks; }

extern "C" int p3mk_inhomo_homo(int M, double L, double alpha, int cao, void *const K[3],
                                p3mk_queue *q) {
    const long M3 = (long)M * M * M;
    hipLaunchKernelGGL(k_homo, dim3(blocks_for(M3)), dim3(kThreads), 0, p3mk_stream(q), M, L, alpha,
                       cao, (double2 *)K[0], (double2 *)K[1], (double2 *)K[2]);
    return p3mk_launch_check("inhomogeneous estimate: homogeneous kernel");
}

extern "C" int p3mk_inhomo_sums(int M, double L, double alpha, int cao, const int m[3], double S[3],
                                double *scratch, p3mk_queue *q) {
    double part[3 * kSumBlocks];
    hipLaunchKernelGGL(k_inhomo_sums, dim3(kSumBlocks), dim3(kThreads), 0, p3mk_stream(q), M, L,
                       alpha, cao, m[0], m[1], m[2], scratch);
    P3MK_TRY(p3mk_launch_check("inhomogeneous estimate: sums"));
    P3MK_TRY(p3mk_copy_d2h(part, scratch, sizeof part, q));
    P3MK_TRY(p3mk_queue_sync(q));
    S[0] = S[1] = S[2] = 0.0;
    for (int b = 0; b < kSumBlocks; b++)
        for (int d = 0; d < 3; d++)
            S[d] += part[3 * b + d];
    return P3MK_OK;
}

extern "C" int p3mk_inhomo_fill(int M, double L, double alpha, int cao, const int m[3],
                                const double S[3], void *const K[3], p3mk_queue *q) {
    const long M3 = (long)M * M * M;
    hipLaunchKernelGGL(k_inhomo_fill, dim3(blocks_for(M3)), dim3(kThreads), 0, p3mk_stream(q), M, L,
                       alpha, cao, m[0], m[1], m[2], S[0], S[1], S[2], (double2 *)K[0],
                       (double2 *)K[1], (double2 *)K[2]);
    return p3mk_launch_check("inhomogeneous estimate: aliasing kernel");
}

extern "C" int p3mk_inhomo_square_add(int M, void *const K[3], void *K3, int first, p3mk_queue *q) {
    const long M3 = (long)M * M * M;
    hipLaunchKernelGGL(k_square_add, dim3(blocks_for(M3)), dim3(kThreads), 0, p3mk_stream(q), M3,
                       (const double2 *)K[0], (const double2 *)K[1], (const double2 *)K[2],
                       (double2 *)K3, first);
    return p3mk_launch_check("inhomogeneous estimate: squared kernel");
}

extern "C" int p3mk_inhomo_product(int M, void *Km, const void *K3, p3mk_queue *q) {
    const long M3 = (long)M * M * M;
    hipLaunchKernelGGL(k_partwise_product, dim3(blocks_for(M3)), dim3(kThreads), 0, p3mk_stream(q),
                       M3, (double2 *)Km, (const double2 *)K3);
    return p3mk_launch_check("inhomogeneous estimate: product");
}

extern "C" int p3mk_fill_real(int M, double v, void *mesh, p3mk_queue *q) {
    const long M3 = (long)M * M * M;
    hipLaunchKernelGGL(k_fill_real, dim3(blocks_for(M3)), dim3(kThreads), 0, p3mk_stream(q), M3, v,
                       (double2 *)mesh);
    return p3mk_launch_check("constant mesh");
}

extern "C" int p3mk_square_charges(int n, const double *q, double *out, p3mk_queue *qu) {
    hipLaunchKernelGGL(k_square_charges, dim3(blocks_for(n)), dim3(kThreads), 0, p3mk_stream(qu), n,
                       q, out);
    return p3mk_launch_check("squared charges");
}

extern "C" int p3mk_diff_charges(int n, const double *a, const double *b, double *out,
                                 p3mk_queue *qu) {
    hipLaunchKernelGGL(k_diff_charges, dim3(blocks_for(n)), dim3(kThreads), 0, p3mk_stream(qu), n, a,
                       b, out);
    return p3mk_launch_check("force differences");
}

extern "C" int p3mk_collect_rms(int M, int cao, double L, int n, const double *x, const double *y,
                                const double *z, const double *q, const void *mesh, double *rms,
                                p3mk_queue *qu) {
    if (n <= 0)
        return P3MK_OK;
    hipStream_t s = p3mk_stream(qu);
    const dim3 g(blocks_for(n)), b(kThreads);
    const double2 *m = (const double2 *)mesh;
    switch (cao) {
    case 1: hipLaunchKernelGGL(k_collect<1>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    case 2: hipLaunchKernelGGL(k_collect<2>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    case 3: hipLaunchKernelGGL(k_collect<3>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    case 4: hipLaunchKernelGGL(k_collect<4>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    case 5: hipLaunchKernelGGL(k_collect<5>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    case 6: hipLaunchKernelGGL(k_collect<6>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    case 7: hipLaunchKernelGGL(k_collect<7>, g, b, 0, s, n, M, L, x, y, z, q, m, rms); break;
    default: return p3mk_fail(P3MK_EINVAL, "collect: cao must be 1..7", 0);
    }
    return p3mk_launch_check("collect");
}
