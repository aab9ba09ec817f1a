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

#include "device_common.hpp"
#include "k_internal.hpp"

using namespace p3m;

namespace {

constexpr int kMaxPasses = 12;
constexpr int kThreads = 256;
constexpr size_t kLdsBudget = 64 * 1024;

struct Radices {
    int n;
    int r[kMaxPasses];
};

// Geometry of a batch of lines: line L = a*C + c starts at a*sa + c,
// element e sits at start + e*es (units: complex elements).
struct LineMap {
    int n;       // transform length
    long es;     // element stride
    long C;      // lines with unit stride between neighbours
    long sa;     // stride between groups of C lines
    long nlines; // number of lines
};

__device__ __forceinline__ long line_start(const LineMap &m, long L) {
    const long a = L / m.C;
    return a * m.sa + (L - a * m.C);
}

template <typename V> __device__ __forceinline__ V twid(const V *tw, int idx, int inv) {
    V w = tw[idx];
    if (inv)
        w.y = -w.y;
    return w;
}

template <typename V> __device__ __forceinline__ V mul_mi(V a, int inv) {
    // multiply by -i (forward) or +i (backward)
    V r;
    if (inv) {
        r.x = -a.y;
        r.y = a.x;
    } else {
        r.x = a.y;
        r.y = -a.x;
    }
    return r;
}

template <int R, typename V> struct Dft;

template <typename V> struct Dft<2, V> {
    __device__ static __forceinline__ void run(V *v, const V *, int, int) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename V> struct Dft<4, V> {
    __device__ static __forceinline__ void run(V *v, const V *, int, int inv) {
        V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        V t2 = cadd(v[1], v[3]), t3 = mul_mi(csub(v[1], v[3]), inv);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <typename V> struct Dft<8, V> {
    __device__ static __forceinline__ void run(V *v, const V *tw, int n, int inv) {
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, V>::run(e, tw, n, inv);
        Dft<4, V>::run(o, tw, n, inv);
        const int s = n / 8;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            V t = (k == 0) ? o[0] : cmul(o[k], twid(tw, k * s, inv));
            v[k] = cadd(e[k], t);
            v[k + 4] = csub(e[k], t);
        }
    }
};

// generic odd radix: X_m = sum_r v_r W^(rm), W = exp(-+2 pi i / R)
template <int R, typename V> struct Dft {
    __device__ static __forceinline__ void run(V *v, const V *tw, int n, int inv) {
        const int s = n / R;
        V w[R];
#pragma unroll
        for (int q = 0; q < R; q++)
            w[q] = twid(tw, q * s, inv);
        V out[R];
#pragma unroll
        for (int m = 0; m < R; m++) {
            V acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                const V p = cmul(v[r], w[(r * m) % R]);
                acc = cadd(acc, p);
            }
            out[m] = acc;
        }
#pragma unroll
        for (int m = 0; m < R; m++)
            v[m] = out[m];
    }
};

template <int R, typename V>
__device__ __forceinline__ void butterfly(const V *__restrict__ src, V *__restrict__ dst, int n,
                                          int Ns, int j, const V *__restrict__ tw, int inv) {
    const int k = j % Ns;
    const int str = n / R;
    V v[R];
#pragma unroll
    for (int r = 0; r < R; r++)
        v[r] = src[j + r * str];
    if (Ns > 1) {
        const int step = k * (n / (Ns * R));
#pragma unroll
        for (int r = 1; r < R; r++)
            v[r] = cmul(v[r], twid(tw, r * step, inv));
    }
    Dft<R, V>::run(v, tw, n, inv);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++)
        dst[base + r * Ns] = v[r];
}

// All radix passes over `nl` lines held in LDS (line l at src + l*pitch).
// Returns the buffer that holds the result.
template <typename V>
__device__ V *fft_tile(V *src, V *dst, int nl, int n, int pitch, const Radices &rp,
                       const V *__restrict__ tw, int inv) {
    int Ns = 1;
    for (int p = 0; p < rp.n; p++) {
        const int R = rp.r[p];
        const int nb = n / R;
        const int total = nl * nb;
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int l = t / nb, j = t - l * nb;
            const V *s = src + l * pitch;
            V *d = dst + l * pitch;
            switch (R) {
            case 2: butterfly<2>(s, d, n, Ns, j, tw, inv); break;
            case 3: butterfly<3>(s, d, n, Ns, j, tw, inv); break;
            case 4: butterfly<4>(s, d, n, Ns, j, tw, inv); break;
            case 5: butterfly<5>(s, d, n, Ns, j, tw, inv); break;
            case 7: butterfly<7>(s, d, n, Ns, j, tw, inv); break;
            case 8: butterfly<8>(s, d, n, Ns, j, tw, inv); break;
            case 11: butterfly<11>(s, d, n, Ns, j, tw, inv); break;
            case 13: butterfly<13>(s, d, n, Ns, j, tw, inv); break;
            default: break;
            }
        }
        __syncthreads();
        V *t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

// Tile load / store.  Strided lines: consecutive threads take consecutive
// lines (adjacent in memory).  Contiguous lines: consecutive threads take
// consecutive elements.
template <typename V>
__device__ __forceinline__ void load_tile(const V *__restrict__ g, V *t, const LineMap &m, long L0,
                                          int lt, int pitch, bool contig) {
    const int n = m.n, total = lt * n;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        int l, e;
        if (contig) {
            l = idx / n;
            e = idx - l * n;
        } else {
            e = idx / lt;
            l = idx - e * lt;
        }
        const long L = L0 + l;
        if (L < m.nlines)
            t[l * pitch + e] = g[line_start(m, L) + (long)e * m.es];
    }
}

template <typename V>
__device__ __forceinline__ void store_tile(V *__restrict__ g, const V *t, const LineMap &m, long L0,
                                           int lt, int pitch, bool contig) {
    const int n = m.n, total = lt * n;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        int l, e;
        if (contig) {
            l = idx / n;
            e = idx - l * n;
        } else {
            e = idx / lt;
            l = idx - e * lt;
        }
        const long L = L0 + l;
        if (L < m.nlines)
            g[line_start(m, L) + (long)e * m.es] = t[l * pitch + e];
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fft_pass(c2_t<T> *data, LineMap m, Radices rp, const c2_t<T> *__restrict__ tw, int inv, int lt,
           int contig) {
    using V = c2_t<T>;
    extern __shared__ unsigned char smem[];
    const int pitch = m.n + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    const long L0 = (long)blockIdx.x * lt;
    load_tile(data, A, m, L0, lt, pitch, contig != 0);
    __syncthreads();
    V *r = fft_tile(A, B, lt, m.n, pitch, rp, tw, inv);
    store_tile(data, r, m, L0, lt, pitch, contig != 0);
}

// r2c along z: real lines of length n (dense, line L at L*n) -> first n/2+1
// coefficients (line L at L*H).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fft_r2c_z(const T *__restrict__ in, c2_t<T> *__restrict__ out, long nlines, int n, Radices rp,
            const c2_t<T> *__restrict__ tw, int lt) {
    using V = c2_t<T>;
    extern __shared__ unsigned char smem[];
    const int pitch = n + 1, H = n / 2 + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    const long L0 = (long)blockIdx.x * lt;
    for (int idx = threadIdx.x; idx < lt * n; idx += blockDim.x) {
        const int l = idx / n, e = idx - l * n;
        if (L0 + l < nlines) {
            V v;
            v.x = in[(L0 + l) * n + e];
            v.y = T(0);
            A[l * pitch + e] = v;
        }
    }
    __syncthreads();
    V *r = fft_tile(A, B, lt, n, pitch, rp, tw, 0);
    for (int idx = threadIdx.x; idx < lt * H; idx += blockDim.x) {
        const int l = idx / H, e = idx - l * H;
        if (L0 + l < nlines)
            out[(L0 + l) * H + e] = r[l * pitch + e];
    }
}

// c2r along z: n/2+1 Hermitian coefficients per line -> n real values.  The
// full line is rebuilt as X[n-k] = conj(X[k]) and the real part of the
// backward transform kept, which is FFTW's c2r on Hermitian input.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fft_c2r_z(const c2_t<T> *__restrict__ in, T *__restrict__ out, long nlines, int n, Radices rp,
            const c2_t<T> *__restrict__ tw, int lt) {
    using V = c2_t<T>;
    extern __shared__ unsigned char smem[];
    const int pitch = n + 1, H = n / 2 + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    const long L0 = (long)blockIdx.x * lt;
    for (int idx = threadIdx.x; idx < lt * n; idx += blockDim.x) {
        const int l = idx / n, e = idx - l * n;
        if (L0 + l < nlines) {
            V v;
            if (e < H) {
                v = in[(L0 + l) * H + e];
            } else {
                v = in[(L0 + l) * H + (n - e)];
                v.y = -v.y;
            }
            A[l * pitch + e] = v;
        }
    }
    __syncthreads();
    V *r = fft_tile(A, B, lt, n, pitch, rp, tw, 1);
    for (int idx = threadIdx.x; idx < lt * n; idx += blockDim.x) {
        const int l = idx / n, e = idx - l * n;
        if (L0 + l < nlines)
            out[(L0 + l) * n + e] = r[l * pitch + e].x;
    }
}

// Fused x pass: forward along x, multiply, backward along x.
// Lines are c = j*W + k (W = M for c2c, M/2+1 for r2c), element e = n_x.
// nout == 1: out0 = G * X.  nout == 3: outd = i (2 pi/L) Dn[n_d] G * X.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_convolve_x(const c2_t<T> *in, c2_t<T> *out0, c2_t<T> *out1, c2_t<T> *out2, LineMap m, int W,
             Radices rp, const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
             const double *__restrict__ Dn, double twopi_L, int nout, int lt) {
    using V = c2_t<T>;
    extern __shared__ unsigned char smem[];
    const int n = m.n, pitch = n + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    V *Cb = B + lt * pitch;
    const long L0 = (long)blockIdx.x * lt;
    load_tile(in, A, m, L0, lt, pitch, false);
    __syncthreads();
    V *X = fft_tile(A, B, lt, n, pitch, rp, tw, 0);
    V *other = (X == A) ? B : A;
    const int M = n;
    // X <- G * X (FP64 product, rounded to T)
    for (int idx = threadIdx.x; idx < lt * n; idx += blockDim.x) {
        const int e = idx / lt, l = idx - e * lt;
        const long L = L0 + l;
        if (L < m.nlines) {
            const int j = (int)(L / W), k = (int)(L - (long)j * W);
            const double g = G[((long)e * M + j) * M + k];
            V v = X[l * pitch + e];
            v.x = (T)((double)v.x * g);
            v.y = (T)((double)v.y * g);
            X[l * pitch + e] = v;
        }
    }
    __syncthreads();
    if (nout == 1) {
        V *r = fft_tile(X, other, lt, n, pitch, rp, tw, 1);
        store_tile(out0, r, m, L0, lt, pitch, false);
        return;
    }
    V *outs[3] = {out0, out1, out2};
    for (int d = 0; d < 3; d++) {
        for (int idx = threadIdx.x; idx < lt * n; idx += blockDim.x) {
            const int e = idx / lt, l = idx - e * lt;
            const long L = L0 + l;
            if (L < m.nlines) {
                const int j = (int)(L / W), k = (int)(L - (long)j * W);
                const int nd = (d == 0) ? e : ((d == 1) ? j : k);
                const double f = twopi_L * Dn[nd];
                const V v = X[l * pitch + e];
                V w;
                w.x = (T)(-f * (double)v.y);
                w.y = (T)(f * (double)v.x);
                Cb[l * pitch + e] = w;
            }
        }
        __syncthreads();
        V *r = fft_tile(Cb, other, lt, n, pitch, rp, tw, 1);
        store_tile(outs[d], r, m, L0, lt, pitch, false);
        __syncthreads();
    }
}

// Adjoint of the ik gradient step: x-axis forward pass of three spectra,
// out = sum_d (-i (2 pi / L) Dn[n_d] G) X_d, x-axis backward pass.  The
// multiplier of the forward gradient, i (2 pi / L) Dn G, is odd and purely
// imaginary, so its transpose as a real operator is its negative.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_divergence_x(const c2_t<T> *in0, const c2_t<T> *in1, const c2_t<T> *in2, c2_t<T> *out, LineMap m,
               int W, Radices rp, const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
               const double *__restrict__ Dn, double twopi_L, int lt) {
    using V = c2_t<T>;
    extern __shared__ unsigned char smem[];
    const int n = m.n, pitch = n + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    V *Cb = B + lt * pitch;
    const long L0 = (long)blockIdx.x * lt;
    const V *ins[3] = {in0, in1, in2};
    const int M = n;
    for (int d = 0; d < 3; d++) {
        load_tile(ins[d], A, m, L0, lt, pitch, false);
        __syncthreads();
        const V *X = fft_tile(A, B, lt, n, pitch, rp, tw, 0);
        for (int idx = threadIdx.x; idx < lt * n; idx += blockDim.x) {
            const int e = idx / lt, l = idx - e * lt;
            const long L = L0 + l;
            if (L < m.nlines) {
                const int j = (int)(L / W), k = (int)(L - (long)j * W);
                const int nd = (d == 0) ? e : ((d == 1) ? j : k);
                const double f = twopi_L * Dn[nd] * G[((long)e * M + j) * M + k];
                const V v = X[l * pitch + e];
                V w;
                w.x = (T)(f * (double)v.y);
                w.y = (T)(-f * (double)v.x);
                if (d > 0)
                    w = cadd(w, Cb[l * pitch + e]);
                Cb[l * pitch + e] = w;
            }
        }
        __syncthreads();
    }
    V *r = fft_tile(Cb, A, lt, n, pitch, rp, tw, 1);
    store_tile(out, r, m, L0, lt, pitch, false);
}

bool make_radices(int n, Radices &rp) {
    rp.n = 0;
    static const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
    for (int r : order) {
        while (n % r == 0) {
            if (rp.n >= kMaxPasses)
                return false;
            rp.r[rp.n++] = r;
            n /= r;
        }
    }
    return n == 1;
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

// lines per workgroup: as many as fit nbuf LDS buffers, capped so that a
// strided row of the tile spans at most 256 bytes
template <typename T> int tile_lines(int n, int nbuf) {
    const size_t line = (size_t)(n + 1) * sizeof(c2_t<T>);
    int lt = (int)(kLdsBudget / (nbuf * line));
    const int cap = 256 / (int)sizeof(c2_t<T>);
    if (lt > cap)
        lt = cap;
    if (lt < 1)
        lt = 1;
    return lt;
}

template <typename T>
int pass(p3mk_fft *f, c2_t<T> *data, const LineMap &m, int inv, bool contig, p3mk_queue *q) {
    const int lt = tile_lines<T>(m.n, 2);
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
    const int lt = tile_lines<T>(M, 2);
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
    const int lt = tile_lines<T>(M, 2);
    const size_t lds = 2 * (size_t)lt * (M + 1) * sizeof(V);
    hipLaunchKernelGGL(k_fft_c2r_z<T>, dim3((unsigned)((nl + lt - 1) / lt)), dim3(kThreads), lds,
                       p3mk_stream(q), (const V *)spec, (T *)mesh, nl, M, f->rp, tw, lt);
    return p3mk_launch_check("fft c2r");
}

template <typename T>
int convolve(p3mk_fft *f, int nout, const double *G, const double *Dn, void *spec,
             void *const out[3], p3mk_queue *q) {
    using V = c2_t<T>;
    const int M = f->M;
    const int W = f->g.r2c ? f->H : M;
    const long nl = (long)M * W;
    LineMap x = {M, nl, nl, 0, nl};
    const int nbuf = (nout == 3) ? 3 : 2;
    const int lt = tile_lines<T>(M, nbuf);
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
    const int lt = tile_lines<T>(M, 3);
    const size_t lds = (size_t)3 * lt * (M + 1) * sizeof(V);
    const c2_t<T> *tw = sizeof(T) == 8 ? (const V *)f->tw64 : (const V *)f->tw32;
    const double twopi_L = 2.0 * 3.14159265358979323846 / f->g.L;
    hipLaunchKernelGGL(k_divergence_x<T>, dim3((unsigned)((nl + lt - 1) / lt)), dim3(kThreads), lds,
                       p3mk_stream(q), (const V *)in[0], (const V *)in[1], (const V *)in[2],
                       (V *)out, x, W, f->rp, tw, G, Dn, twopi_L, lt);
    return p3mk_launch_check("divergence");
}

}  // namespace

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
