// Mixed-radix Stockham FFT for the P3M meshes (radices 2,3,4,5,7,8,11,13).
//
// A 3D transform is three passes of batched 1D transforms, one per axis.
// Each workgroup loads a tile of `lt` lines into LDS (coalesced: along the
// line for the contiguous z axis, across neighbouring lines for the strided
// y and x axes), runs all radix passes in LDS (autosort Stockham, two
// ping-pong buffers) and writes the tile back.  The x pass of the forward
// transform is fused with the multiplication by the influence function (and,
// for ik, the gradient i 2pi/L Dn) and with the x pass of the backward
// transform, so the spectrum never makes an extra round trip through HBM.
//
// Conventions are FFTW's (the reference's transforms, p3m-ik.c:84-88 etc.):
// forward X[k] = sum x[n] exp(-2 pi i nk/N), backward with +, unnormalised.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>

#include "batch_desc.hpp"
#include "device_common.hpp"
#include "fft_tiles.hpp"
#include "k_internal.hpp"

using namespace p3m;
using namespace p3m::fft;

namespace {

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fft_pass(c2_t<T> *data, LineMap m, Radices rp, const c2_t<T> *__restrict__ tw, int inv, int lt,
           int contig) {
    extern __shared__ unsigned char smem[];
    fft_pass_tile<T>(data, m, rp, tw, inv, lt, contig, (long)blockIdx.x * lt, smem);
}

// Power-of-two lines: register-resident passes (p2_fft), 256 / (N / 8)
// lines per workgroup, one LDS row per line.  Contiguous lines (z axis):
// a line's threads are consecutive, so each load of the eight covers N / 8
// neighbouring elements; strided lines (y, x): consecutive threads take
// neighbouring lines.
template <int N, typename T, bool CONTIG>
__global__ void __launch_bounds__(kThreads)
k_fft_p2(c2_t<T> *data, LineMap m, const c2_t<T> *__restrict__ tw, int inv) {
    using V = c2_t<T>;
    constexpr int TPL = N / 8, LPW = kThreads / TPL, PITCH = N + 1;
    __shared__ V buf[LPW * PITCH];
    const int tid = threadIdx.x;
    const int l = CONTIG ? tid / TPL : tid % LPW;
    const int t = CONTIG ? tid % TPL : tid / LPW;
    const long L = (long)blockIdx.x * LPW + l;
    const bool ok = L < m.nlines;
    const long base = ok ? line_start(m, L) : 0;
    V v[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        if (ok) {
            v[s] = data[base + (long)(t + TPL * s) * m.es];
        } else {
            v[s].x = T(0);
            v[s].y = T(0);
        }
    }
    p2_fft<N>(v, buf + l * PITCH, t, tw, inv);
    if (ok) {
#pragma unroll
        for (int s = 0; s < 8; s++)
            data[base + (long)(t + TPL * s) * m.es] = v[s];
    }
}

// The fused x pass (convolve_x_tile) for power-of-two meshes on registers:
// forward along x, X <- G X (FP64 product rounded to T), then for nout == 1
// backward into out0, for nout == 3 backward of i (2 pi / L) Dn[n_d] X into
// out_d.  Lines c = j W + k (strided by nl = M W), element e = n_x.
template <int N, typename T>
__global__ void __launch_bounds__(kThreads)
k_convolve_p2(const c2_t<T> *in, c2_t<T> *out0, c2_t<T> *out1, c2_t<T> *out2, long nl, int W,
              const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
              const double *__restrict__ Dn, double twopi_L, int nout) {
    using V = c2_t<T>;
    constexpr int TPL = N / 8, LPW = kThreads / TPL, PITCH = N + 1;
    __shared__ V buf[LPW * PITCH];
    const int tid = threadIdx.x;
    const int l = tid % LPW, t = tid / LPW;
    const long L = (long)blockIdx.x * LPW + l;
    const bool ok = L < nl;
    const int j = ok ? (int)(L / W) : 0, k = ok ? (int)(L - (long)j * W) : 0;
    V v[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        if (ok) {
            v[s] = in[L + (long)(t + TPL * s) * nl];
        } else {
            v[s].x = T(0);
            v[s].y = T(0);
        }
    }
    V *line = buf + l * PITCH;
    p2_fft<N>(v, line, t, tw, 0);
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int e = t + TPL * s;
        const double g = ok ? G[((long)e * N + j) * N + k] : 0.0;
        v[s].x = (T)((double)v[s].x * g);
        v[s].y = (T)((double)v[s].y * g);
    }
    if (nout == 1) {
        p2_fft<N>(v, line, t, tw, 1);
        if (ok) {
#pragma unroll
            for (int s = 0; s < 8; s++)
                out0[L + (long)(t + TPL * s) * nl] = v[s];
        }
        return;
    }
    V *outs[3] = {out0, out1, out2};
    for (int d = 0; d < 3; d++) {
        V w[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int e = t + TPL * s;
            const int nd = (d == 0) ? e : ((d == 1) ? j : k);
            const double f = twopi_L * Dn[nd];
            w[s].x = (T)(-f * (double)v[s].y);
            w[s].y = (T)(f * (double)v[s].x);
        }
        p2_fft<N>(w, line, t, tw, 1);
        if (ok) {
#pragma unroll
            for (int s = 0; s < 8; s++)
                outs[d][L + (long)(t + TPL * s) * nl] = w[s];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fft_r2c_z(const T *__restrict__ in, c2_t<T> *__restrict__ out, long nlines, int n, Radices rp,
            const c2_t<T> *__restrict__ tw, int lt) {
    extern __shared__ unsigned char smem[];
    fft_r2c_z_tile<T>(in, out, nlines, n, rp, tw, lt, (long)blockIdx.x * lt, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fft_c2r_z(const c2_t<T> *__restrict__ in, T *__restrict__ out, long nlines, int n, Radices rp,
            const c2_t<T> *__restrict__ tw, int lt) {
    extern __shared__ unsigned char smem[];
    fft_c2r_z_tile<T>(in, out, nlines, n, rp, tw, lt, (long)blockIdx.x * lt, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_convolve_x(const c2_t<T> *in, c2_t<T> *out0, c2_t<T> *out1, c2_t<T> *out2, LineMap m, int W,
             Radices rp, const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
             const double *__restrict__ Dn, double twopi_L, int nout, int lt) {
    extern __shared__ unsigned char smem[];
    convolve_x_tile<T>(in, out0, out1, out2, m, W, rp, tw, G, Dn, twopi_L, nout, lt,
                       (long)blockIdx.x * lt, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_divergence_x(const c2_t<T> *in0, const c2_t<T> *in1, const c2_t<T> *in2, c2_t<T> *out, LineMap m,
               int W, Radices rp, const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
               const double *__restrict__ Dn, double twopi_L, int lt) {
    extern __shared__ unsigned char smem[];
    divergence_x_tile<T>(in0, in1, in2, out, m, W, rp, tw, G, Dn, twopi_L, lt,
                         (long)blockIdx.x * lt, smem);
}

}  // namespace

struct p3mk_fft {
    p3mk_geom g;
    int M, H;
    Radices rp;
    double2 *tw64;
    float2 *tw32;
};

extern "C" size_t p3mk_mesh_bytes(const p3mk_geom *g) {
    const size_t M3 = (size_t)g->M * g->M * g->M;
    const size_t s = g->fp32 ? sizeof(float) : sizeof(double);
    return g->r2c ? M3 * s : 2 * M3 * s;
}

extern "C" size_t p3mk_spec_bytes(const p3mk_geom *g) {
    const size_t s = g->fp32 ? sizeof(float) : sizeof(double);
    const size_t H = g->r2c ? (size_t)(g->M / 2 + 1) : (size_t)g->M;
    return 2 * (size_t)g->M * g->M * H * s;
}

extern "C" int p3mk_fft_create(p3mk_fft **out, const p3mk_geom *g) {
    *out = nullptr;
    Radices rp;
    if (g->M < 2 || !make_radices(g->M, rp))
        return p3mk_fail(P3MK_EINVAL, "mesh size is not a product of 2,3,5,7,11,13", 0);
    p3mk_fft *f = new (std::nothrow) p3mk_fft;
    if (!f)
        return p3mk_fail(P3MK_ENOMEM, "fft plan allocation", 0);
    f->g = *g;
    f->M = g->M;
    f->H = g->M / 2 + 1;
    f->rp = rp;
    f->tw64 = nullptr;
    f->tw32 = nullptr;
    const int n = g->M;
    double2 *h64 = new (std::nothrow) double2[n];
    float2 *h32 = new (std::nothrow) float2[n];
    if (!h64 || !h32) {
        delete[] h64;
        delete[] h32;
        delete f;
        return p3mk_fail(P3MK_ENOMEM, "twiddle table", 0);
    }
    for (int k = 0; k < n; k++) {
        // exp(-2 pi i k/n), from long double for correctly rounded doubles
        const long double a = -2.0L * 3.141592653589793238462643383279502884L * k / n;
        h64[k] = make_double2((double)cosl(a), (double)sinl(a));
        h32[k] = make_float2((float)cosl(a), (float)sinl(a));
    }
    int rc = p3mk_alloc((void **)&f->tw64, sizeof(double2) * n);
    if (rc == P3MK_OK)
        rc = p3mk_alloc((void **)&f->tw32, sizeof(float2) * n);
    if (rc == P3MK_OK) {
        hipError_t e = hipMemcpy(f->tw64, h64, sizeof(double2) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(f->tw32, h32, sizeof(float2) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess)
            rc = p3mk_fail(P3MK_EDEVICE, "twiddle upload", (int)e);
    }
    delete[] h64;
    delete[] h32;
    if (rc != P3MK_OK) {
        p3mk_fft_destroy(f);
        return rc;
    }
    *out = f;
    return P3MK_OK;
}

extern "C" void p3mk_fft_destroy(p3mk_fft *f) {
    if (!f)
        return;
    p3mk_release(f->tw64);
    p3mk_release(f->tw32);
    delete f;
}

namespace {

// Lines per workgroup of a launch over one mesh: pow2_tile_lines, halved
// (down to 8) while the launch would have fewer workgroups than the GPU has
// CUs (256).
template <typename T> int launch_lines(int n, int nbuf, long nlines) {
    int lt = pow2_tile_lines<T>(n, nbuf);
    while (lt > 8 && (nlines + lt - 1) / lt < 256)
        lt /= 2;
    return lt;
}

template <int N, typename T>
int pass_p2(p3mk_fft *f, c2_t<T> *data, const LineMap &m, int inv, bool contig, p3mk_queue *q) {
    constexpr int LPW = kThreads / (N / 8);
    const long blocks = (m.nlines + LPW - 1) / LPW;
    const c2_t<T> *tw = sizeof(T) == 8 ? (const c2_t<T> *)f->tw64 : (const c2_t<T> *)f->tw32;
    if (contig)
        hipLaunchKernelGGL((k_fft_p2<N, T, true>), dim3((unsigned)blocks), dim3(kThreads), 0,
                           p3mk_stream(q), data, m, tw, inv);
    else
        hipLaunchKernelGGL((k_fft_p2<N, T, false>), dim3((unsigned)blocks), dim3(kThreads), 0,
                           p3mk_stream(q), data, m, tw, inv);
    return p3mk_launch_check("fft pass");
}

template <typename T>
int pass(p3mk_fft *f, c2_t<T> *data, const LineMap &m, int inv, bool contig, p3mk_queue *q) {
    switch (m.n) {
    case 16: return pass_p2<16, T>(f, data, m, inv, contig, q);
    case 32: return pass_p2<32, T>(f, data, m, inv, contig, q);
    case 64: return pass_p2<64, T>(f, data, m, inv, contig, q);
    case 128: return pass_p2<128, T>(f, data, m, inv, contig, q);
    case 256: return pass_p2<256, T>(f, data, m, inv, contig, q);
    default: break;
    }
    const int lt = launch_lines<T>(m.n, 2, m.nlines);
    const size_t lds = 2 * (size_t)lt * (m.n + 1) * sizeof(c2_t<T>);
    const long blocks = (m.nlines + lt - 1) / lt;
    const c2_t<T> *tw = sizeof(T) == 8 ? (const c2_t<T> *)f->tw64 : (const c2_t<T> *)f->tw32;
    hipLaunchKernelGGL(k_fft_pass<T>, dim3((unsigned)blocks), dim3(kThreads), lds, p3mk_stream(q),
                       data, m, f->rp, tw, inv, lt, contig ? 1 : 0);
    return p3mk_launch_check("fft pass");
}

template <typename T> int forward_zy(p3mk_fft *f, void *mesh, void *spec, p3mk_queue *q) {
    using V = c2_t<T>;
    const int M = f->M, H = f->H;
    const c2_t<T> *tw = sizeof(T) == 8 ? (const V *)f->tw64 : (const V *)f->tw32;
    if (!f->g.r2c) {
        V *d = (V *)mesh;
        LineMap z = {M, 1, 1, M, (long)M * M};
        P3MK_TRY(pass<T>(f, d, z, 0, true, q));
        LineMap y = {M, M, M, (long)M * M, (long)M * M};
        return pass<T>(f, d, y, 0, false, q);
    }
    const long nl = (long)M * M;
    const int lt = launch_lines<T>(M, 2, nl);
    const size_t lds = 2 * (size_t)lt * (M + 1) * sizeof(V);
    hipLaunchKernelGGL(k_fft_r2c_z<T>, dim3((unsigned)((nl + lt - 1) / lt)), dim3(kThreads), lds,
                       p3mk_stream(q), (const T *)mesh, (V *)spec, nl, M, f->rp, tw, lt);
    P3MK_TRY(p3mk_launch_check("fft r2c"));
    LineMap y = {M, H, H, (long)M * H, (long)M * H};
    return pass<T>(f, (V *)spec, y, 0, false, q);
}

template <typename T> int backward_yz(p3mk_fft *f, void *spec, void *mesh, p3mk_queue *q) {
    using V = c2_t<T>;
    const int M = f->M, H = f->H;
    const c2_t<T> *tw = sizeof(T) == 8 ? (const V *)f->tw64 : (const V *)f->tw32;
    if (!f->g.r2c) {
        V *d = (V *)spec;
        LineMap y = {M, M, M, (long)M * M, (long)M * M};
        P3MK_TRY(pass<T>(f, d, y, 1, false, q));
        LineMap z = {M, 1, 1, M, (long)M * M};
        return pass<T>(f, d, z, 1, true, q);
    }
    LineMap y = {M, H, H, (long)M * H, (long)M * H};
    P3MK_TRY(pass<T>(f, (V *)spec, y, 1, false, q));
    const long nl = (long)M * M;
    const int lt = launch_lines<T>(M, 2, nl);
    const size_t lds = 2 * (size_t)lt * (M + 1) * sizeof(V);
    hipLaunchKernelGGL(k_fft_c2r_z<T>, dim3((unsigned)((nl + lt - 1) / lt)), dim3(kThreads), lds,
                       p3mk_stream(q), (const V *)spec, (T *)mesh, nl, M, f->rp, tw, lt);
    return p3mk_launch_check("fft c2r");
}

template <int N, typename T>
int convolve_p2(p3mk_fft *f, int nout, const double *G, const double *Dn, void *spec,
                void *const out[3], long nl, int W, p3mk_queue *q) {
    using V = c2_t<T>;
    constexpr int LPW = kThreads / (N / 8);
    const c2_t<T> *tw = sizeof(T) == 8 ? (const V *)f->tw64 : (const V *)f->tw32;
    V *o0 = (V *)out[0];
    V *o1 = nout == 3 ? (V *)out[1] : nullptr;
    V *o2 = nout == 3 ? (V *)out[2] : nullptr;
    hipLaunchKernelGGL((k_convolve_p2<N, T>), dim3((unsigned)((nl + LPW - 1) / LPW)), dim3(kThreads), 0,
                       p3mk_stream(q), (const V *)spec, o0, o1, o2, nl, W, tw, G, Dn,
                       2.0 * 3.14159265358979323846 / f->g.L, nout);
    return p3mk_launch_check("convolve");
}

template <typename T>
int convolve(p3mk_fft *f, int nout, const double *G, const double *Dn, void *spec,
             void *const out[3], p3mk_queue *q) {
    using V = c2_t<T>;
    const int M = f->M;
    const int W = f->g.r2c ? f->H : M;
    const long nl = (long)M * W;
    switch (M) {
    case 16: return convolve_p2<16, T>(f, nout, G, Dn, spec, out, nl, W, q);
    case 32: return convolve_p2<32, T>(f, nout, G, Dn, spec, out, nl, W, q);
    case 64: return convolve_p2<64, T>(f, nout, G, Dn, spec, out, nl, W, q);
    case 128: return convolve_p2<128, T>(f, nout, G, Dn, spec, out, nl, W, q);
    case 256: return convolve_p2<256, T>(f, nout, G, Dn, spec, out, nl, W, q);
    default: break;
    }
    LineMap x = {M, nl, nl, 0, nl};
    const int nbuf = (nout == 3) ? 3 : 2;
    const int lt = launch_lines<T>(M, nbuf, nl);
    const size_t lds = (size_t)nbuf * lt * (M + 1) * sizeof(V);
    const c2_t<T> *tw = sizeof(T) == 8 ? (const V *)f->tw64 : (const V *)f->tw32;
    const double twopi_L = 2.0 * 3.14159265358979323846 / f->g.L;
    V *o0 = (V *)out[0];
    V *o1 = nout == 3 ? (V *)out[1] : nullptr;
    V *o2 = nout == 3 ? (V *)out[2] : nullptr;
    hipLaunchKernelGGL(k_convolve_x<T>, dim3((unsigned)((nl + lt - 1) / lt)), dim3(kThreads), lds,
                       p3mk_stream(q), (const V *)spec, o0, o1, o2, x, W, f->rp, tw, G, Dn, twopi_L,
                       nout, lt);
    return p3mk_launch_check("convolve");
}

template <typename T>
int divergence(p3mk_fft *f, const double *G, const double *Dn, void *const in[3], void *out,
               p3mk_queue *q) {
    using V = c2_t<T>;
    const int M = f->M;
    const int W = f->g.r2c ? f->H : M;
    const long nl = (long)M * W;
    LineMap x = {M, nl, nl, 0, nl};
    const int lt = launch_lines<T>(M, 3, nl);
    const size_t lds = (size_t)3 * lt * (M + 1) * sizeof(V);
    const c2_t<T> *tw = sizeof(T) == 8 ? (const V *)f->tw64 : (const V *)f->tw32;
    const double twopi_L = 2.0 * 3.14159265358979323846 / f->g.L;
    hipLaunchKernelGGL(k_divergence_x<T>, dim3((unsigned)((nl + lt - 1) / lt)), dim3(kThreads), lds,
                       p3mk_stream(q), (const V *)in[0], (const V *)in[1], (const V *)in[2],
                       (V *)out, x, W, f->rp, tw, G, Dn, twopi_L, lt);
    return p3mk_launch_check("divergence");
}

}  // namespace

// ---- batched passes over many systems ------------------------------------
//
// Block b belongs to the system s with blk[s] <= b < blk[s+1]; its tile
// starts at line (b - blk[s]) * lt(M_s).  Buffers are the systems' own.

namespace {

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_bfft_pass(const BSys *__restrict__ sys, const long *__restrict__ blk, int nsys, int axis, int slot,
            int inv, int r2c) {
    extern __shared__ unsigned char smem[];
    using V = c2_t<T>;
    const int s = prefix_slot(blk, nsys, (long)blockIdx.x);
    const BSys &S = sys[s];
    const int M = S.M, W = r2c ? M / 2 + 1 : M;
    LineMap m;
    if (axis == kPassZ)
        m = {M, 1, 1, M, (long)M * M};
    else
        m = {M, W, W, (long)M * W, (long)M * W};
    const int lt = batch_tile_lines<T>(M, 2);
    fft_pass_tile<T>((V *)S.buf[slot], m, S.rp, (const V *)S.tw, inv, lt, axis == kPassZ ? 1 : 0,
                     (long)(blockIdx.x - blk[s]) * lt, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_bfft_r2c_z(const BSys *__restrict__ sys, const long *__restrict__ blk, int nsys, int in_slot,
             int out_slot) {
    extern __shared__ unsigned char smem[];
    using V = c2_t<T>;
    const int s = prefix_slot(blk, nsys, (long)blockIdx.x);
    const BSys &S = sys[s];
    const int M = S.M, lt = batch_tile_lines<T>(M, 2);
    fft_r2c_z_tile<T>((const T *)S.buf[in_slot], (V *)S.buf[out_slot], (long)M * M, M, S.rp,
                      (const V *)S.tw, lt, (long)(blockIdx.x - blk[s]) * lt, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_bfft_c2r_z(const BSys *__restrict__ sys, const long *__restrict__ blk, int nsys, int in_slot,
             int out_slot) {
    extern __shared__ unsigned char smem[];
    using V = c2_t<T>;
    const int s = prefix_slot(blk, nsys, (long)blockIdx.x);
    const BSys &S = sys[s];
    const int M = S.M, lt = batch_tile_lines<T>(M, 2);
    fft_c2r_z_tile<T>((const V *)S.buf[in_slot], (T *)S.buf[out_slot], (long)M * M, M, S.rp,
                      (const V *)S.tw, lt, (long)(blockIdx.x - blk[s]) * lt, smem);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_bconvolve_x(const BSys *__restrict__ sys, const long *__restrict__ blk, int nsys, int nout,
              int in_slot, int out_slot, int r2c) {
    extern __shared__ unsigned char smem[];
    using V = c2_t<T>;
    const int s = prefix_slot(blk, nsys, (long)blockIdx.x);
    const BSys &S = sys[s];
    const int M = S.M, W = r2c ? M / 2 + 1 : M;
    const long nl = (long)M * W;
    const LineMap m = {M, nl, nl, 0, nl};
    const int lt = batch_tile_lines<T>(M, nout == 3 ? 3 : 2);
    V *o0 = (V *)S.buf[out_slot];
    V *o1 = nout == 3 ? (V *)S.buf[out_slot + 1] : nullptr;
    V *o2 = nout == 3 ? (V *)S.buf[out_slot + 2] : nullptr;
    convolve_x_tile<T>((const V *)S.buf[in_slot], o0, o1, o2, m, W, S.rp, (const V *)S.tw, S.G, S.Dn,
                       2.0 * 3.14159265358979323846 / S.L, nout, lt,
                       (long)(blockIdx.x - blk[s]) * lt, smem);
}

}  // namespace

namespace p3m {

int bfft_lines_per_tile(const BForm &f, int M, int kind) {
    const int nbuf = (kind == kPassX && !f.ad) ? 3 : 2;
    return f.fp32 ? batch_tile_lines<float>(M, nbuf) : batch_tile_lines<double>(M, nbuf);
}

size_t bfft_lds_bytes(const BForm &f, int M, int kind) {
    const int nbuf = (kind == kPassX && !f.ad) ? 3 : 2;
    const size_t v = f.fp32 ? sizeof(float2) : sizeof(double2);
    return (size_t)nbuf * bfft_lines_per_tile(f, M, kind) * (M + 1) * v;
}

int bfft_pass(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks, size_t lds, int axis,
              int slot, int inv, p3mk_queue *q) {
    if (nblocks <= 0)
        return P3MK_OK;
    if (f.fp32)
        hipLaunchKernelGGL(k_bfft_pass<float>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, axis, slot, inv, f.r2c);
    else
        hipLaunchKernelGGL(k_bfft_pass<double>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, axis, slot, inv, f.r2c);
    return p3mk_launch_check("batched fft pass");
}

int bfft_r2c_z(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks, size_t lds, int in_slot,
               int out_slot, p3mk_queue *q) {
    if (nblocks <= 0)
        return P3MK_OK;
    if (f.fp32)
        hipLaunchKernelGGL(k_bfft_r2c_z<float>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, in_slot, out_slot);
    else
        hipLaunchKernelGGL(k_bfft_r2c_z<double>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, in_slot, out_slot);
    return p3mk_launch_check("batched fft r2c");
}

int bfft_c2r_z(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks, size_t lds, int in_slot,
               int out_slot, p3mk_queue *q) {
    if (nblocks <= 0)
        return P3MK_OK;
    if (f.fp32)
        hipLaunchKernelGGL(k_bfft_c2r_z<float>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, in_slot, out_slot);
    else
        hipLaunchKernelGGL(k_bfft_c2r_z<double>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, in_slot, out_slot);
    return p3mk_launch_check("batched fft c2r");
}

int bfft_convolve_x(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks, size_t lds,
                    int nout, int in_slot, int out_slot, p3mk_queue *q) {
    if (nblocks <= 0)
        return P3MK_OK;
    if (f.fp32)
        hipLaunchKernelGGL(k_bconvolve_x<float>, dim3((unsigned)nblocks), dim3(kThreads), lds,
                           p3mk_stream(q), sys, blk, nsys, nout, in_slot, out_slot, f.r2c);
    else
        hipLaunchKernelGGL(k_bconvolve_x<double>, dim3((unsigned)nblocks), dim3(kThreads),
                           lds, p3mk_stream(q), sys, blk, nsys, nout, in_slot, out_slot,
                           f.r2c);
    return p3mk_launch_check("batched convolve");
}

}  // namespace p3m

extern "C" int p3mk_fft_forward_zy(p3mk_fft *f, void *mesh, void *spec, p3mk_queue *q) {
    return f->g.fp32 ? forward_zy<float>(f, mesh, spec, q) : forward_zy<double>(f, mesh, spec, q);
}

extern "C" int p3mk_fft_backward_yz(p3mk_fft *f, void *spec, void *mesh, p3mk_queue *q) {
    return f->g.fp32 ? backward_yz<float>(f, spec, mesh, q) : backward_yz<double>(f, spec, mesh, q);
}

extern "C" int p3mk_convolve(p3mk_fft *f, int nout, const double *G, const double *Dn, void *spec,
                             void *const out[3], p3mk_queue *q) {
    if (nout != 1 && nout != 3)
        return p3mk_fail(P3MK_EINVAL, "convolve: nout must be 1 or 3", 0);
    return f->g.fp32 ? convolve<float>(f, nout, G, Dn, spec, out, q)
                     : convolve<double>(f, nout, G, Dn, spec, out, q);
}

extern "C" int p3mk_divergence(p3mk_fft *f, const double *G, const double *Dn, void *const in[3],
                               void *out, p3mk_queue *q) {
    return f->g.fp32 ? divergence<float>(f, G, Dn, in, out, q)
                     : divergence<double>(f, G, Dn, in, out, q);
}

template <typename T> static int fft3d(p3mk_fft *f, void *data, int sign, p3mk_queue *q) {
    using V = c2_t<T>;
    const int M = f->M;
    const int inv = sign > 0;
    V *d = (V *)data;
    LineMap z = {M, 1, 1, M, (long)M * M};
    LineMap y = {M, M, M, (long)M * M, (long)M * M};
    LineMap x = {M, (long)M * M, (long)M * M, 0, (long)M * M};
    P3MK_TRY(pass<T>(f, d, z, inv, true, q));
    P3MK_TRY(pass<T>(f, d, y, inv, false, q));
    return pass<T>(f, d, x, inv, false, q);
}

extern "C" int p3mk_fft3d(p3mk_fft *f, void *data, int sign, p3mk_queue *q) {
    return f->g.fp32 ? fft3d<float>(f, data, sign, q) : fft3d<double>(f, data, sign, q);
}
