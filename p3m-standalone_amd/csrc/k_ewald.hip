// FP64 Ewald sum at a sample of particles: the reference's Ewald method
// (ewald.c:84-229) as an accuracy oracle for large systems, where a full
// Ewald evaluation is out of reach but the forces on a few thousand
// particles are not.
//
// k space: rho(n) = sum_j q_j exp(-i 2 pi n.r_j / L) over the lattice vectors
// of the half sphere n^2 <= kmax^2 (rho(-n) = conj rho(n), so each vector
// counts twice), then for every sampled particle
//   F_i = 2 sum_n n q_i G(n) (Re rho sin(k.r_i) + Im rho cos(k.r_i)),
//   G(n) = 2/L^2 exp(-(pi n / (alpha L))^2) / n^2        (ewald.c:84-110, :127-172)
// real space: the minimum-image erfc sum over every particle within rcut
// (rcut <= L/2), for the Ewald alpha and, optionally, untruncated for a
// second alpha (to split the total force at the P3M alpha).
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "k_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr double kPi = 3.14159265358979323846;

// phase 2 pi (n . r) / L reduced to one period before the sine/cosine
__device__ __forceinline__ double phase(int nx, int ny, int nz, double x, double y, double z,
                                        double Li) {
    double t = ((double)nx * x + (double)ny * y + (double)nz * z) * Li;
    t -= rint(t);
    return 2.0 * kPi * t;
}

__global__ void __launch_bounds__(kThreads)
k_ewald_rho(int n, const double *__restrict__ x, const double *__restrict__ y,
            const double *__restrict__ z, const double *__restrict__ q, int nk,
            const int4 *__restrict__ kv, double Li, double2 *__restrict__ rho) {
    __shared__ double4 tile[kThreads];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int4 v = k < nk ? kv[k] : make_int4(0, 0, 0, 0);
    double re = 0.0, im = 0.0;
    for (int j0 = 0; j0 < n; j0 += kThreads) {
        const int j = j0 + threadIdx.x;
        tile[threadIdx.x] = j < n ? make_double4(x[j], y[j], z[j], q[j]) : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = min(kThreads, n - j0);
        for (int l = 0; l < cnt; l++) {
            const double4 p = tile[l];
            double s, c;
            sincos(phase(v.x, v.y, v.z, p.x, p.y, p.z, Li), &s, &c);
            re += p.w * c;
            im -= p.w * s;
        }
        __syncthreads();
    }
    if (k < nk)
        rho[k] = make_double2(re, im);
}

// k-space force on each sampled particle (one thread per sample)
__global__ void __launch_bounds__(kThreads)
k_ewald_sample_k(int ns, const int *__restrict__ idx, const double *__restrict__ x,
                 const double *__restrict__ y, const double *__restrict__ z,
                 const double *__restrict__ q, int nk, const int4 *__restrict__ kv,
                 const double *__restrict__ G, const double2 *__restrict__ rho, double Li,
                 double *__restrict__ f) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns)
        return;
    const int i = idx[s];
    const double xi = x[i], yi = y[i], zi = z[i], qi = q[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int k = 0; k < nk; k++) {
        const int4 v = kv[k];
        double sn, cs;
        sincos(phase(v.x, v.y, v.z, xi, yi, zi, Li), &sn, &cs);
        const double2 r = rho[k];
        const double ff = 2.0 * qi * G[k] * (r.x * sn + r.y * cs);
        fx += v.x * ff;
        fy += v.y * ff;
        fz += v.z * ff;
    }
    f[s] = fx;
    f[(long)ns + s] = fy;
    f[2L * ns + s] = fz;
}

// real-space sums for each sampled particle (one block per sample) over the
// pairs within rcut: r1 = erfc sum at alpha, r2 = erfc sum at alpha2
__global__ void __launch_bounds__(kThreads)
k_ewald_sample_r(int n, int ns, const int *__restrict__ idx, const double *__restrict__ x,
                 const double *__restrict__ y, const double *__restrict__ z,
                 const double *__restrict__ q, double L, double rcut, double alpha,
                 double alpha2, double *__restrict__ r1, double *__restrict__ r2) {
    __shared__ double red[6][kThreads];
    const int s = blockIdx.x;
    const int i = idx[s];
    const double xi = x[i], yi = y[i], zi = z[i], qi = q[i];
    const double Li = 1.0 / L, rc2 = rcut * rcut;
    const double c1 = 2.0 * alpha / sqrt(kPi), c2 = 2.0 * alpha2 / sqrt(kPi);
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        if (j == i)
            continue;
        double dx = xi - x[j], dy = yi - y[j], dz = zi - z[j];
        dx -= rint(dx * Li) * L;
        dy -= rint(dy * Li) * L;
        dz -= rint(dz * Li) * L;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > rc2)
            continue;
        const double r = sqrt(r2), qq = qi * q[j];
        const double f1 = qq * (erfc(alpha * r) / r + c1 * exp(-alpha * alpha * r2)) / r2;
        a[0] += f1 * dx;
        a[1] += f1 * dy;
        a[2] += f1 * dz;
        if (alpha2 > 0.0) {
            const double f2 = qq * (erfc(alpha2 * r) / r + c2 * exp(-alpha2 * alpha2 * r2)) / r2;
            a[3] += f2 * dx;
            a[4] += f2 * dy;
            a[5] += f2 * dz;
        }
    }
    for (int c = 0; c < 6; c++)
        red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 6; c++)
                red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) {
            r1[(long)c * ns + s] = red[c][0];
            r2[(long)c * ns + s] = red[3 + c][0];
        }
}

// f_total += k + r1;  f_split += k + r1 - r2 (the k-space force at alpha2)
__global__ void k_ewald_combine(long m, const double *__restrict__ k, const double *__restrict__ r1,
                                const double *__restrict__ r2, double *__restrict__ f_total,
                                double *__restrict__ f_split) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m)
        return;
    const double tot = k[t] + r1[t];
    f_total[t] += tot;
    if (f_split != nullptr)
        f_split[t] += tot - r2[t];
}

// ---- allocation-free form ---------------------------------------------------
//
// The sampler's working memory -- the lattice vectors of the half sphere,
// their influence function, the structure factors and the per-sample partial
// forces -- in one file-scope device array instead of per-call allocations,
// for callers that cannot allocate (graph capture, tight serving loops).
// Limits: kmax <= 63, at most kNoAllocMaxK lattice vectors and
// kNoAllocMaxSample sampled particles.

constexpr int kNoAllocMaxKmax = 63;
constexpr int kNoAllocMaxLines = (kNoAllocMaxKmax + 1) * (2 * kNoAllocMaxKmax + 1);
constexpr long kNoAllocMaxK = 1L << 18;
constexpr long kNoAllocMaxSample = 1L << 15;
constexpr size_t kOffKv = 0;
constexpr size_t kOffG = kOffKv + sizeof(int4) * kNoAllocMaxK;
constexpr size_t kOffRho = kOffG + sizeof(double) * kNoAllocMaxK;
constexpr size_t kOffTmp = kOffRho + sizeof(double2) * kNoAllocMaxK;
constexpr size_t kScratchBytes = kOffTmp + sizeof(double) * 9 * kNoAllocMaxSample;

// Lattice vectors n of the half sphere 0 < n^2 <= kmax^2 (n_x > 0, or n_x = 0
// and (n_y, n_z) lexicographically positive) and G(n), in the order nx, ny,
// nz ascending -- the order of the allocating form's host enumeration.  One
// workgroup: each (nx, ny) line counts its vectors, an LDS scan gives the
// line offsets, then each line writes its run.
__device__ __forceinline__ bool lattice_keep(int nx, int ny, int nz, long kmax2) {
    const long n2 = (long)nx * nx + (long)ny * ny + (long)nz * nz;
    if (n2 == 0 || n2 > kmax2)
        return false;
    return !(nx == 0 && (ny < 0 || (ny == 0 && nz < 0)));
}

__global__ void __launch_bounds__(1024)
k_ewald_lattice(int kmax, double fak1, double fak2, int4 *__restrict__ kv, double *__restrict__ G) {
    __shared__ int off[kNoAllocMaxLines];
    __shared__ int part[1024];
    const int t = threadIdx.x, nyn = 2 * kmax + 1, nlines = (kmax + 1) * nyn;
    const long kmax2 = (long)kmax * kmax;
    const int chunk = (nlines + 1023) / 1024, b = min(nlines, t * chunk), e = min(nlines, b + chunk);
    int sum = 0;
    for (int l = b; l < e; l++) {
        const int nx = l / nyn, ny = l % nyn - kmax;
        int c = 0;
        for (int nz = -kmax; nz <= kmax; nz++)
            c += lattice_keep(nx, ny, nz, kmax2);
        off[l] = c;
        sum += c;
    }
    part[t] = sum;
    __syncthreads();
    // inclusive scan of the 1024 chunk sums (Hillis-Steele)
    for (int w = 1; w < 1024; w <<= 1) {
        const int v = t >= w ? part[t - w] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int l = b; l < e; l++) {
        const int c = off[l];
        off[l] = run;
        run += c;
    }
    __syncthreads();
    for (int l = t; l < nlines; l += 1024) {
        const int nx = l / nyn, ny = l % nyn - kmax;
        int k = off[l];
        for (int nz = -kmax; nz <= kmax; nz++) {
            if (!lattice_keep(nx, ny, nz, kmax2))
                continue;
            const double n2 = (double)((long)nx * nx + (long)ny * ny + (long)nz * nz);
            kv[k] = make_int4(nx, ny, nz, 0);
            G[k] = fak1 / n2 * exp(-fak2 * n2);
            k++;
        }
    }
}

__global__ void k_zero_doubles(long m, double *__restrict__ p) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m)
        p[t] = 0.0;
}

// The workspace.  In host code this name is only a stand-in for the symbol:
// kernels receive the device address that hipGetSymbolAddress returns for
// the current device (noalloc_scratch_address).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device, resolved once per
// device on its first use and kept.
int noalloc_scratch_address(unsigned char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "hipGetDevice", (int)e);
    if (dev < 0 || dev >= kMaxDevices)
        return p3mk_fail(P3MK_EDEVICE, "device index beyond the workspace table", dev);
    void *p = __atomic_load_n(&g_scratch_ptr[dev], __ATOMIC_ACQUIRE);
    if (p == nullptr) {
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return p3mk_fail(P3MK_EDEVICE, "hipGetSymbolAddress(scratch)", (int)e);
        __atomic_store_n(&g_scratch_ptr[dev], p, __ATOMIC_RELEASE);
    }
    *out = (unsigned char *)p;
    return P3MK_OK;
}

}  // namespace

extern "C" int p3mk_ewald_sample(int n, double L, const double *x, const double *y, const double *z,
                                 const double *q, int ns, const int *idx, double alpha,
                                 double rcut, int kmax, double alpha2, double *f_total,
                                 double *f_split, p3mk_queue *qu) {
    if (n <= 0 || ns <= 0)
        return P3MK_OK;
    if (!(L > 0.0) || !(alpha > 0.0) || rcut < 0.0 || rcut > 0.5 * L || kmax < 1 || kmax > 512)
        return p3mk_fail(P3MK_EINVAL, "Ewald sample needs alpha > 0, 0 <= rcut <= L/2, 1 <= kmax <= 512", 0);
    // lattice vectors of the half sphere and their influence function
    std::vector<int4> kv;
    std::vector<double> G;
    const double fak1 = 2.0 / (L * L), fak2 = (kPi / (alpha * L)) * (kPi / (alpha * L));
    const long kmax2 = (long)kmax * kmax;
    for (int nx = 0; nx <= kmax; nx++)
        for (int ny = -kmax; ny <= kmax; ny++)
            for (int nz = -kmax; nz <= kmax; nz++) {
                const long n2 = (long)nx * nx + (long)ny * ny + (long)nz * nz;
                if (n2 == 0 || n2 > kmax2)
                    continue;
                if (nx == 0 && (ny < 0 || (ny == 0 && nz < 0)))
                    continue; // the other half
                kv.push_back(make_int4(nx, ny, nz, 0));
                G.push_back(fak1 / (double)n2 * std::exp(-fak2 * (double)n2));
            }
    const int nk = (int)kv.size();
    const long m3 = 3L * ns;
    int4 *dkv = nullptr;
    double *dG = nullptr, *tmp = nullptr;
    double2 *drho = nullptr;
    int rc = p3mk_alloc((void **)&dkv, sizeof(int4) * nk);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&dG, sizeof(double) * nk);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&drho, sizeof(double2) * nk);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&tmp, sizeof(double) * 3 * m3);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(dkv, kv.data(), sizeof(int4) * nk, qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(dG, G.data(), sizeof(double) * nk, qu);
    if (rc == P3MK_OK) rc = p3mk_fill_zero(tmp, sizeof(double) * 3 * m3, qu);
    hipStream_t st = p3mk_stream(qu);
    const double Li = 1.0 / L;
    double *tk = tmp, *tr1 = tmp + m3, *tr2 = tmp + 2 * m3;
    if (rc == P3MK_OK) {
        hipLaunchKernelGGL(k_ewald_rho, dim3((nk + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                           n, x, y, z, q, nk, dkv, Li, drho);
        rc = p3mk_launch_check("Ewald structure factor");
    }
    if (rc == P3MK_OK) {
        hipLaunchKernelGGL(k_ewald_sample_k, dim3((ns + kThreads - 1) / kThreads), dim3(kThreads), 0,
                           st, ns, idx, x, y, z, q, nk, dkv, dG, drho, Li, tk);
        rc = p3mk_launch_check("Ewald k-space sample");
    }
    if (rc == P3MK_OK && rcut > 0.0) {
        hipLaunchKernelGGL(k_ewald_sample_r, dim3(ns), dim3(kThreads), 0, st, n, ns, idx, x, y, z, q,
                           L, rcut, alpha, alpha2, tr1, tr2);
        rc = p3mk_launch_check("Ewald real-space sample");
    }
    if (rc == P3MK_OK) {
        hipLaunchKernelGGL(k_ewald_combine, dim3((unsigned)((m3 + 255) / 256)), dim3(256), 0, st, m3,
                           tk, tr1, tr2, f_total, f_split);
        rc = p3mk_launch_check("Ewald combine");
    }
    // the temporaries are freed after the work that reads them
    const int rs = p3mk_queue_sync(qu);
    if (rc == P3MK_OK)
        rc = rs;
    p3mk_release(dkv);
    p3mk_release(dG);
    p3mk_release(drho);
    p3mk_release(tmp);
    return rc;
}

extern "C" int p3mk_ewald_sample_noalloc(int n, double L, const double *x, const double *y,
                                         const double *z, const double *q, int ns, const int *idx,
                                         double alpha, double rcut, int kmax, double alpha2,
                                         double *f_total, double *f_split, void *stream) {
    if (n <= 0 || ns <= 0)
        return P3MK_OK;
    if (!(L > 0.0) || !(alpha > 0.0) || rcut < 0.0 || rcut > 0.5 * L || kmax < 1)
        return p3mk_fail(P3MK_EINVAL, "Ewald sample needs alpha > 0, 0 <= rcut <= L/2, kmax >= 1", 0);
    if (kmax > kNoAllocMaxKmax)
        return p3mk_fail(P3MK_EINVAL, "allocation-free Ewald sample: kmax beyond the maximum 63",
                         kmax);
    if ((long)ns > kNoAllocMaxSample)
        return p3mk_fail(P3MK_EINVAL,
                         "allocation-free Ewald sample: more sampled particles than the maximum 32768",
                         ns);
    // lattice vectors of the half sphere: the same count the kernel finds
    const long kmax2 = (long)kmax * kmax;
    long nk = 0;
    for (int nx = 0; nx <= kmax; nx++)
        for (int ny = -kmax; ny <= kmax; ny++)
            for (int nz = -kmax; nz <= kmax; nz++) {
                const long n2 = (long)nx * nx + (long)ny * ny + (long)nz * nz;
                if (n2 == 0 || n2 > kmax2 || (nx == 0 && (ny < 0 || (ny == 0 && nz < 0))))
                    continue;
                nk++;
            }
    if (nk > kNoAllocMaxK)
        return p3mk_fail(P3MK_EINVAL, "allocation-free Ewald sample: too many lattice vectors", (int)nk);
    unsigned char *base = nullptr;
    int rc = noalloc_scratch_address(&base);
    if (rc != P3MK_OK)
        return rc;
    int4 *dkv = (int4 *)(base + kOffKv);
    double *dG = (double *)(base + kOffG);
    double2 *drho = (double2 *)(base + kOffRho);
    double *tmp = (double *)(base + kOffTmp);
    const long m3 = 3L * ns;
    hipStream_t st = (hipStream_t)stream;
    const double Li = 1.0 / L;
    const double fak1 = 2.0 / (L * L), fak2 = (kPi / (alpha * L)) * (kPi / (alpha * L));
    double *tk = tmp, *tr1 = tmp + m3, *tr2 = tmp + 2 * m3;
    hipLaunchKernelGGL(k_ewald_lattice, dim3(1), dim3(1024), 0, st, kmax, fak1, fak2, dkv, dG);
    hipLaunchKernelGGL(k_zero_doubles, dim3((unsigned)((3 * m3 + 255) / 256)), dim3(256), 0, st,
                       3 * m3, tmp);
    hipLaunchKernelGGL(k_ewald_rho, dim3((unsigned)((nk + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, st, n, x, y, z, q, (int)nk, dkv, Li, drho);
    hipLaunchKernelGGL(k_ewald_sample_k, dim3((ns + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                       ns, idx, x, y, z, q, (int)nk, dkv, dG, drho, Li, tk);
    if (rcut > 0.0)
        hipLaunchKernelGGL(k_ewald_sample_r, dim3(ns), dim3(kThreads), 0, st, n, ns, idx, x, y, z, q,
                           L, rcut, alpha, alpha2, tr1, tr2);
    hipLaunchKernelGGL(k_ewald_combine, dim3((unsigned)((m3 + 255) / 256)), dim3(256), 0, st, m3, tk,
                       tr1, tr2, f_total, f_split);
    return p3mk_launch_check("allocation-free Ewald sample");
}
