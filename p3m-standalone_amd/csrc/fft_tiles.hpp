// Batched 1D FFT tiles in LDS shared by the single-system and the batched
// FFT kernels: Stockham radix passes (radices 2,3,4,5,7,8,11,13) over `lt`
// lines held in LDS, tile load/store, and the per-tile bodies of the axis
// passes (the caller supplies the tile's first line L0 and the dynamic LDS).
//
// Conventions are FFTW's (the reference's transforms, p3m-ik.c:84-88 etc.):
// forward X[k] = sum x[n] exp(-2 pi i nk/N), backward with +, unnormalised.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace p3m {
namespace fft {


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

// ---- register-resident passes for power-of-two lines --------------------
//
// A line of N points is held by N / 8 threads, eight values each: thread t
// owns elements t + (N / 8) s, s = 0..7, which are exactly the inputs of its
// butterflies in every Stockham pass (radix R: butterflies t + (N / 8) m,
// m < 8 / R) and, after the last pass, its outputs in natural order.  A
// pass computes in registers; between passes the outputs go through one LDS
// line (positions of the autosort Stockham, as butterfly() writes them) and
// back.  The radix sequence is make_radices', so the results equal the LDS
// tile passes' bit for bit.
template <int N> struct P2Plan;
template <> struct P2Plan<16> { static constexpr int np = 2, r[3] = {8, 2, 1}; };
template <> struct P2Plan<32> { static constexpr int np = 2, r[3] = {8, 4, 1}; };
template <> struct P2Plan<64> { static constexpr int np = 2, r[3] = {8, 8, 1}; };
template <> struct P2Plan<128> { static constexpr int np = 3, r[3] = {8, 8, 2}; };
template <> struct P2Plan<256> { static constexpr int np = 3, r[3] = {8, 8, 4}; };

// one pass: v holds the inputs t + (N / 8) s, returns the outputs in the
// same slots; slot m + r (8 / R) is output r of butterfly m
template <int N, int R, typename V>
__device__ __forceinline__ void p2_pass(V v[8], int t, int Ns, const V *__restrict__ tw, int inv) {
    constexpr int TPL = N / 8, B = 8 / R;
#pragma unroll
    for (int m = 0; m < B; m++) {
        const int j = t + m * TPL;
        const int k = j & (Ns - 1);
        V x[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            x[r] = v[m + r * B];
        if (Ns > 1) {
            const int step = k * (N / (Ns * R));
#pragma unroll
            for (int r = 1; r < R; r++)
                x[r] = cmul(x[r], twid(tw, r * step, inv));
        }
        Dft<R, V>::run(x, tw, N, inv);
#pragma unroll
        for (int r = 0; r < R; r++)
            v[m + r * B] = x[r];
    }
}

// the outputs of a pass to their autosort positions in the LDS line
template <int N, int R, typename V>
__device__ __forceinline__ void p2_write(V *line, const V v[8], int t, int Ns) {
    constexpr int TPL = N / 8, B = 8 / R;
#pragma unroll
    for (int m = 0; m < B; m++) {
        const int j = t + m * TPL;
        const int base = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
        for (int r = 0; r < R; r++)
            line[base + r * Ns] = v[m + r * B];
    }
}

template <int N, typename V>
__device__ __forceinline__ void p2_read(const V *line, V v[8], int t) {
    constexpr int TPL = N / 8;
#pragma unroll
    for (int s = 0; s < 8; s++)
        v[s] = line[t + TPL * s];
}

// all passes of the plan on the registers of thread t; `line` is its line's
// LDS row (N + 1 values), shared by the N / 8 threads of the line
template <int N, typename V>
__device__ __forceinline__ void p2_fft(V v[8], V *line, int t, const V *__restrict__ tw, int inv) {
    using Pl = P2Plan<N>;
    p2_pass<N, Pl::r[0]>(v, t, 1, tw, inv);
    p2_write<N, Pl::r[0]>(line, v, t, 1);
    __syncthreads();
    p2_read<N>(line, v, t);
    __syncthreads();
    p2_pass<N, Pl::r[1]>(v, t, Pl::r[0], tw, inv);
    if constexpr (Pl::np == 3) {
        p2_write<N, Pl::r[1]>(line, v, t, Pl::r[0]);
        __syncthreads();
        p2_read<N>(line, v, t);
        __syncthreads();
        p2_pass<N, Pl::r[2]>(v, t, Pl::r[0] * Pl::r[1], tw, inv);
    }
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
__device__ __forceinline__ void
fft_pass_tile(c2_t<T> *data, LineMap m, Radices rp, const c2_t<T> *__restrict__ tw, int inv, int lt,
           int contig,
    long L0, unsigned char *smem) {
    using V = c2_t<T>;
    const int pitch = m.n + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    load_tile(data, A, m, L0, lt, pitch, contig != 0);
    __syncthreads();
    V *r = fft_tile(A, B, lt, m.n, pitch, rp, tw, inv);
    store_tile(data, r, m, L0, lt, pitch, contig != 0);
}

// r2c along z: real lines of length n (dense, line L at L*n) -> first n/2+1
// coefficients (line L at L*H).
template <typename T>
__device__ __forceinline__ void
fft_r2c_z_tile(const T *__restrict__ in, c2_t<T> *__restrict__ out, long nlines, int n, Radices rp,
            const c2_t<T> *__restrict__ tw, int lt,
    long L0, unsigned char *smem) {
    using V = c2_t<T>;
    const int pitch = n + 1, H = n / 2 + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
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
__device__ __forceinline__ void
fft_c2r_z_tile(const c2_t<T> *__restrict__ in, T *__restrict__ out, long nlines, int n, Radices rp,
            const c2_t<T> *__restrict__ tw, int lt,
    long L0, unsigned char *smem) {
    using V = c2_t<T>;
    const int pitch = n + 1, H = n / 2 + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
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
__device__ __forceinline__ void
convolve_x_tile(const c2_t<T> *in, c2_t<T> *out0, c2_t<T> *out1, c2_t<T> *out2, LineMap m, int W,
             Radices rp, const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
             const double *__restrict__ Dn, double twopi_L, int nout, int lt,
    long L0, unsigned char *smem) {
    using V = c2_t<T>;
    const int n = m.n, pitch = n + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    V *Cb = B + lt * pitch;
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
__device__ __forceinline__ void
divergence_x_tile(const c2_t<T> *in0, const c2_t<T> *in1, const c2_t<T> *in2, c2_t<T> *out, LineMap m,
               int W, Radices rp, const c2_t<T> *__restrict__ tw, const double *__restrict__ G,
               const double *__restrict__ Dn, double twopi_L, int lt,
    long L0, unsigned char *smem) {
    using V = c2_t<T>;
    const int n = m.n, pitch = n + 1;
    V *A = reinterpret_cast<V *>(smem);
    V *B = A + lt * pitch;
    V *Cb = B + lt * pitch;
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

inline bool make_radices(int n, Radices &rp) {
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


// lines per workgroup: as many as fit nbuf LDS buffers, capped so that a
// strided row of the tile spans at most 256 bytes
template <typename T> __host__ __device__ inline int tile_lines(int n, int nbuf) {
    const size_t line = (size_t)(n + 1) * sizeof(c2_t<T>);
    int lt = (int)(kLdsBudget / (nbuf * line));
    const int cap = 256 / (int)sizeof(c2_t<T>);
    if (lt > cap)
        lt = cap;
    if (lt < 1)
        lt = 1;
    return lt;
}

// lines per tile of the launches: the largest power of two up to 16 that
// tile_lines allows.  Sixteen lines keep a strided row of the tile aligned
// and let several workgroups share a CU (at M = 128 the FP32 FFT + multiply
// stage takes 0.14 ms with 16 lines against 0.28 ms with the 31 that fill
// 64 KiB; FP64: 8 lines, 0.19 against 0.27 ms with 15).
template <typename T> __host__ __device__ inline int pow2_tile_lines(int n, int nbuf) {
    const int fit = tile_lines<T>(n, nbuf);
    int lt = 1;
    while (2 * lt <= fit && 2 * lt <= 16)
        lt *= 2;
    return lt;
}

// lines per tile of the batched passes, whose meshes (16-48 points per line)
// give short lines: the largest power of two that tile_lines allows and
// that keeps a workgroup at most kBatchTilePoints points
constexpr int kBatchTilePoints = 2048;
template <typename T> __host__ __device__ inline int batch_tile_lines(int n, int nbuf) {
    const int fit = tile_lines<T>(n, nbuf);
    int lt = 1;
    while (2 * lt <= fit && 2 * lt * n <= kBatchTilePoints)
        lt *= 2;
    return lt;
}

// last index s in [0, count) with prefix[s] <= v (prefix[0] == 0)
template <typename I>
__host__ __device__ inline int prefix_slot(const I *prefix, int count, long v) {
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if ((long)prefix[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

}  // namespace fft
}  // namespace p3m
