// Influence functions, self-force coefficients and k-space error sums.
//
// All FP64.  One thread per mesh point; the aliasing sums over
// m in [-mc, mc]^3 run in registers.  Reductions are two-stage (per-block
// partials, then one block in a fixed order), so results do not depend on
// scheduling.
#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.hpp"
#include "k_internal.hpp"

using namespace p3m;

namespace {

constexpr int kThreads = 256;
constexpr double kPi = 3.14159265358979323846;

// p3m-common.c:133-143 / :95-116
__device__ __forceinline__ double nshift_of(int i, int M) {
    return (double)i - round((double)i / (double)M) * (double)M;
}
__device__ __forceinline__ double dn_of(int i, int M) {
    return (i == M / 2) ? 0.0 : nshift_of(i, M);
}

// p3m-common.c:71-75
__device__ __forceinline__ double sinc(double d) {
    const double pd = kPi * d;
    return (d == 0.0) ? 1.0 : sin(pd) / pd;
}

__device__ __forceinline__ double ipow(double b, int e) {
    double r = 1.0;
    for (int k = 0; k < e; k++)
        r *= b;
    return r;
}

// U_hat (window-functions.c:40-44)
__device__ __forceinline__ double bspline_k(int cao, double d) {
    const double pd = kPi * d;
    return (d == 0.0) ? 1.0 : pow(sin(pd) / pd, (double)cao);
}

__global__ void k_dn(int M, double *Dn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M)
        Dn[i] = dn_of(i, M);
}

// kind: 0 ik, 1 ik-i, 2 ad, 3 ad-i
__global__ void __launch_bounds__(kThreads)
k_influence(int kind, int M, int cao, double L, double alpha, int mc, double *__restrict__ G) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long M3 = (long)M * M * M;
    if (ind >= M3)
        return;
    const int NX = (int)(ind / ((long)M * M));
    const int NY = (int)((ind / M) % M);
    const int NZ = (int)(ind % M);
    const int h = M / 2;
    if (NX == 0 && NY == 0 && NZ == 0) {
        G[ind] = 0.0;
        return;
    }
    if (kind <= 1 && (NX % h == 0) && (NY % h == 0) && (NZ % h == 0)) {
        // ik corner zeros (p3m-ik.c:156-157, p3m-ik-i.c:147-148)
        G[ind] = 0.0;
        return;
    }
    const double Leni = 1.0 / L, fak1 = 1.0 / (double)M, fak2 = (kPi / alpha) * (kPi / alpha);
    const double nsx = nshift_of(NX, M), nsy = nshift_of(NY, M), nsz = nshift_of(NZ, M);
    double Z0 = 0.0, Z1 = 0.0, Z2 = 0.0, N1 = 0.0, N2 = 0.0, N3 = 0.0, N4 = 0.0, Zad = 0.0;
    for (int MX = -mc; MX <= mc; MX++) {
        const double NMX = nsx + M * MX;
        const double S1 = (kind == 0) ? ipow(bspline_k(cao, fak1 * NMX), 2)
                                      : ipow(sinc(fak1 * NMX), 2 * cao);
        for (int MY = -mc; MY <= mc; MY++) {
            const double NMY = nsy + M * MY;
            const double S2 = S1 * ((kind == 0) ? ipow(bspline_k(cao, fak1 * NMY), 2)
                                                : ipow(sinc(fak1 * NMY), 2 * cao));
            for (int MZ = -mc; MZ <= mc; MZ++) {
                const double NMZ = nsz + M * MZ;
                const double S3 = S2 * ((kind == 0) ? ipow(bspline_k(cao, fak1 * NMZ), 2)
                                                    : ipow(sinc(fak1 * NMZ), 2 * cao));
                const double NM2 = (NMX * Leni) * (NMX * Leni) + (NMY * Leni) * (NMY * Leni) +
                                   (NMZ * Leni) * (NMZ * Leni);
                const double TE = exp(-fak2 * NM2);
                const bool even = ((MX + MY + MZ) % 2) == 0;
                N1 += S3;
                if (kind <= 1) {
                    // p3m-ik.c:119-128, p3m-ik-i.c:104-118
                    const double zwi = S3 * TE / NM2;
                    Z0 += NMX * zwi * Leni;
                    Z1 += NMY * zwi * Leni;
                    Z2 += NMZ * zwi * Leni;
                    N2 += even ? S3 : -S3;
                } else {
                    // p3m-ad.c:97-105, p3m-ad-i.c:98-114
                    N2 += S3 * NM2;
                    N3 += even ? S3 : -S3;
                    N4 += even ? S3 * NM2 : -S3 * NM2;
                    Zad += S3 * TE;
                }
            }
        }
    }
    double g;
    if (kind <= 1) {
        const double Dnx = dn_of(NX, M), Dny = dn_of(NY, M), Dnz = dn_of(NZ, M);
        double zwi = Dnx * Z0 * Leni + Dny * Z1 * Leni + Dnz * Z2 * Leni;
        const double d2 = (Dnx * Leni) * (Dnx * Leni) + (Dny * Leni) * (Dny * Leni) +
                          (Dnz * Leni) * (Dnz * Leni);
        if (kind == 0) {
            zwi /= d2 * (N1 * N1); // p3m-ik.c:163-165
        } else {
            zwi /= d2;               // p3m-ik-i.c:157-161
            zwi /= 0.5 * (N1 * N1 + N2 * N2);
        }
        g = 2.0 * zwi / kPi;
    } else if (kind == 2) {
        g = Zad / (kPi * N1 * N2); // p3m-ad.c:143
    } else {
        g = Zad / (0.5 * kPi * (N1 * N2 + N3 * N4)); // p3m-ad-i.c:145
    }
    G[ind] = g;
}

// ---- block reduction helpers ----------------------------------------------

template <int K>
__device__ __forceinline__ void block_reduce_store(double (&v)[K], double *__restrict__ out) {
    __shared__ double red[kThreads / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = v[k];
        for (int off = 32; off > 0; off >>= 1)
            s += __shfl_down(s, off, 64);
        if (lane == 0)
            red[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; w++)
            s += red[w][threadIdx.x];
        out[(long)blockIdx.x * K + threadIdx.x] = s;
    }
}

// Sum partials[nblk][K] over blocks (fixed order per k) -> out[K]
template <int K>
__global__ void k_sum_partials(const double *__restrict__ partials, int nblk, double *out) {
    const int k = threadIdx.x;
    if (k >= K)
        return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++)
        s += partials[(long)b * K + k];
    out[k] = s;
}

// ---- self-force coefficients (p3m-ad-self-forces.c:30-112) ----------------
//
// S(m') = sum_{n: G != 0} G(n) sum_{m in [-1,1]^3} U(n + M m) U(n + M (m+m')),
// U(v) = prod_d sinc(v_d / M)^cao.  U factorises over the axes, so
// S(m') = sum_n G(n) Cx(m'_x) Cy(m'_y) Cz(m'_z), C_d(s) = sum_m u_d[m] u_d[m+s].
__global__ void __launch_bounds__(kThreads)
k_self_force_partials(int M, int cao, const double *__restrict__ G, double *__restrict__ partials) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long M3 = (long)M * M * M;
    double S[27];
#pragma unroll
    for (int k = 0; k < 27; k++)
        S[k] = 0.0;
    if (ind < M3) {
        const double g = G[ind];
        if (g != 0.0) {
            const int n[3] = {(int)(ind / ((long)M * M)), (int)((ind / M) % M), (int)(ind % M)};
            const double mesh_i = 1.0 / M;
            double C[3][3];
            for (int d = 0; d < 3; d++) {
                const int t = (int)nshift_of(n[d], M);
                double u[5];
                for (int m = -2; m <= 2; m++)
                    u[m + 2] = pow(sinc(mesh_i * (t + m * M)), (double)cao);
                for (int s = -1; s <= 1; s++)
                    C[d][s + 1] = u[1] * u[1 + s] + u[2] * u[2 + s] + u[3] * u[3 + s];
            }
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        S[(a * 3 + b) * 3 + c] = g * C[0][a] * C[1][b] * C[2][c];
        }
    }
    block_reduce_store<27>(S, partials);
}

__global__ void k_self_force_finish(const double *__restrict__ partials, int nblk, int M, double L,
                                    double *__restrict__ b) {
    __shared__ double S[27];
    const int k = threadIdx.x;
    if (k < 27) {
        double s = 0.0;
        for (int blk = 0; blk < nblk; blk++)
            s += partials[(long)blk * 27 + k];
        S[k] = s;
    }
    __syncthreads();
    if (k < 81) {
        const int a = k / 3, d = k % 3;
        const int m[3] = {a / 9 - 1, (a / 3) % 3 - 1, a % 3 - 1};
        const double L4 = (L * L) * (L * L);
        b[k] = (m[d] == 0) ? 0.0 : (kPi * m[d] * M) / L4 * S[a];
    }
}

// ---- k-space error sums (p3m-ik.c:249-337, p3m-ik-i.c:249-320,
//      p3m-ad.c:436-497, p3m-ad-i.c:205-284) ----------------------------------

__device__ __forceinline__ double cotangent_sum(int n, double mesh_i, int cao) {
    const double cc = cos(kPi * mesh_i * (double)n);
    const double c = cc * cc;
    switch (cao) {
    case 1: return 1.0;
    case 2: return (1.0 + c * 2.0) / 3.0;
    case 3: return (2.0 + c * (11.0 + c * 2.0)) / 15.0;
    case 4: return (17.0 + c * (180.0 + c * (114.0 + c * 4.0))) / 315.0;
    case 5: return (62.0 + c * (1072.0 + c * (1452.0 + c * (247.0 + c * 2.0)))) / 2835.0;
    case 6: return (1382.0 + c * (35396.0 + c * (83021.0 + c * (34096.0 + c * (2026.0 + c * 4.0))))) / 155925.0;
    case 7: return (21844.0 + c * (776661.0 + c * (2801040.0 + c * (2123860.0 + c * (349500.0 + c * (8166.0 + c * 4.0)))))) / 6081075.0;
    }
    return 0.0;
}

__global__ void __launch_bounds__(kThreads)
k_error_partials(int kind, int M, int cao, double L, double alpha, int mc,
                 double *__restrict__ partials) {
    const long ind = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long M3 = (long)M * M * M;
    double v[1] = {0.0};
    if (ind < M3) {
        const int h = M / 2;
        const int nx = (int)(ind / ((long)M * M)) - h, ny = (int)((ind / M) % M) - h,
                  nz = (int)(ind % M) - h;
        const bool any = (nx != 0) || (ny != 0) || (nz != 0);
        const bool all = (nx != 0) && (ny != 0) && (nz != 0);
        if ((kind <= 1 && any) || (kind >= 2 && all)) {
            const double mesh_i = 1.0 / M, f1 = (kPi / (alpha * L)) * (kPi / (alpha * L));
            double a1 = 0, a2ik = 0, a2ad = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0;
            for (int mx = -mc; mx <= mc; mx++) {
                const double nmx = nx + mx * M;
                for (int my = -mc; my <= mc; my++) {
                    const double nmy = ny + my * M;
                    for (int mz = -mc; mz <= mc; mz++) {
                        const double nmz = nz + mz * M;
                        const double nm2 = nmx * nmx + nmy * nmy + nmz * nmz;
                        const double ex = exp(-f1 * nm2), ex2 = ex * ex;
                        const double U2 =
                            ipow(sinc(mesh_i * nmx) * sinc(mesh_i * nmy) * sinc(mesh_i * nmz), 2 * cao);
                        const bool even = ((mx + my + mz) % 2) == 0;
                        a1 += ex2 / nm2;
                        a2ik += U2 * ex * (nx * nmx + ny * nmy + nz * nmz) / nm2;
                        a2ad += U2 * ex;
                        a3 += U2 * nm2;
                        a4 += U2;
                        a5 += even ? U2 : -U2;
                        a6 += even ? U2 * nm2 : -U2 * nm2;
                    }
                }
            }
            const double n2 = (double)nx * nx + (double)ny * ny + (double)nz * nz;
            if (kind == 0) {
                const double cs = cotangent_sum(nz, mesh_i, cao) *
                                  (cotangent_sum(nx, mesh_i, cao) * cotangent_sum(ny, mesh_i, cao));
                v[0] = a1 - (a2ik / cs) * (a2ik / cs) / n2;
            } else if (kind == 1) {
                v[0] = a1 - a2ik * a2ik / (0.5 * n2 * (a4 * a4 + a5 * a5));
            } else if (kind == 2) {
                v[0] = (a3 == 0.0 || a4 == 0.0) ? 0.0 : a1 - (a2ad * a2ad) / (a3 * a4);
            } else {
                v[0] = a1 - a2ad * a2ad / (0.5 * (a3 * a4 + a6 * a5));
            }
        }
    }
    block_reduce_store<1>(v, partials);
}

int kind_of(int method) {
    switch (method) {
    case 0: case 6: return 0;
    case 1: return 1;
    case 2: case 7: return 2;
    case 3: return 3;
    }
    return -1;
}

}  // namespace

extern "C" size_t p3mk_reduce_scratch_doubles(int M) {
    const size_t M3 = (size_t)M * M * M;
    return ((M3 + kThreads - 1) / kThreads) * 27 + 64;
}

extern "C" int p3mk_influence(int method, const p3mk_geom *g, double alpha, int mc, double *G,
                              double *Dn, p3mk_queue *q) {
    const int kind = kind_of(method);
    if (kind < 0)
        return p3mk_fail(P3MK_EINVAL, "unknown method id", 0);
    const int M = g->M;
    hipStream_t s = p3mk_stream(q);
    if (Dn != nullptr) {
        hipLaunchKernelGGL(k_dn, dim3((M + 63) / 64), dim3(64), 0, s, M, Dn);
        P3MK_TRY(p3mk_launch_check("Dn"));
    }
    const long M3 = (long)M * M * M;
    if (kind == 2 && alpha == 0.0) // p3m-ad.c:120-123
        return p3mk_fill_zero(G, sizeof(double) * M3, q);
    hipLaunchKernelGGL(k_influence, dim3((unsigned)((M3 + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, s, kind, M, g->cao, g->L, alpha, mc, G);
    return p3mk_launch_check("influence");
}

extern "C" int p3mk_self_force_coeffs(const p3mk_geom *g, const double *G, double *b,
                                      double *scratch, p3mk_queue *q) {
    const int M = g->M;
    const long M3 = (long)M * M * M;
    const int nblk = (int)((M3 + kThreads - 1) / kThreads);
    hipStream_t s = p3mk_stream(q);
    hipLaunchKernelGGL(k_self_force_partials, dim3(nblk), dim3(kThreads), 0, s, M, g->cao, G, scratch);
    P3MK_TRY(p3mk_launch_check("self force partials"));
    hipLaunchKernelGGL(k_self_force_finish, dim3(1), dim3(128), 0, s, scratch, nblk, M, g->L, b);
    return p3mk_launch_check("self force finish");
}

extern "C" int p3mk_kspace_error_sum(int method, int M, int cao, double L, double alpha, int mc,
                                     double *he_q, double *scratch, p3mk_queue *q) {
    const int kind = kind_of(method);
    if (kind < 0)
        return p3mk_fail(P3MK_EINVAL, "unknown method id", 0);
    const long M3 = (long)M * M * M;
    const int nblk = (int)((M3 + kThreads - 1) / kThreads);
    hipStream_t s = p3mk_stream(q);
    hipLaunchKernelGGL(k_error_partials, dim3(nblk), dim3(kThreads), 0, s, kind, M, cao, L, alpha, mc,
                       scratch);
    P3MK_TRY(p3mk_launch_check("error partials"));
    double *out = scratch + (size_t)nblk;
    hipLaunchKernelGGL(k_sum_partials<1>, dim3(1), dim3(64), 0, s, scratch, nblk, out);
    P3MK_TRY(p3mk_launch_check("error sum"));
    P3MK_TRY(p3mk_copy_d2h(he_q, out, sizeof(double), q));
    return p3mk_queue_sync(q);
}
