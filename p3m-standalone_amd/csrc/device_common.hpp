// Shared device helpers for the P3M kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "bspline_coeffs.hpp"

namespace p3m {

// Resolution of the reference's weight tables: the offset of a particle from
// its nearest mesh point is quantised to 1/(2*65536) of a mesh spacing
// (interpol.h:22, interpol.c:53-57, charge-assign.c:94-97).
constexpr int kMaxInterpol = 65536;
constexpr double kInvTableStep = 1.0 / (2.0 * kMaxInterpol);

template <typename T> struct C2;
template <> struct C2<double> { using type = double2; };
template <> struct C2<float> { using type = float2; };
template <typename T> using c2_t = typename C2<T>::type;

template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) {
    V r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename V> __device__ __forceinline__ V csub(V a, V b) {
    V r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}

// One period wrap (charge-assign.c:37-39): valid for ind in [-M, 2M).
__device__ __forceinline__ int wrap1(int ind, int M) {
    return ind + (ind < 0) * M - (ind >= M) * M;
}

// Position -> (wrapped nearest mesh point, quantised offset), rounding step
// for step as the reference's FP64 code does (charge-assign.c:153-156):
//   pos = x*Hi - pos_shift + 0.5*ii;  nmp = floor(pos + 0.5);
//   arg = floor((pos - nmp + 0.5) * 131072);  xq = (arg - 65536) / 131072.
// Contraction into FMA is disabled here: a fused x*Hi - pos_shift rounds
// differently and would move particles across table bins.
__device__ __forceinline__ void locate(double r, double Hi, double pos_shift, double half_ii,
                                       int M, int &base, double &xq) {
#pragma clang fp contract(off)
    double pos = r * Hi;
    pos = pos - pos_shift;
    pos = pos + half_ii;
    const double f = floor(pos + 0.5);
    const int nmp = (int)f;
    // equals wrap1(nmp, M) for positions in [0, L); a full modulo keeps any
    // other input inside the mesh
    int b = nmp % M;
    base = b < 0 ? b + M : b;
    const double t = (pos - f) + 0.5;
    const int arg = (int)floor(t * (2.0 * kMaxInterpol));
    xq = (double)(arg - kMaxInterpol) * kInvTableStep;
}

// Weights (and optionally derivative weights) of the P stencil points along
// one axis.
template <int P, typename T>
__device__ __forceinline__ void axis_weights(double r, double Hi, double half_ii, int M,
                                             int &base, T *w) {
    double xq;
    locate(r, Hi, (double)((P - 1) / 2), half_ii, M, base, xq);
    Bspline<P, T>::w((T)xq, w);
}
template <int P, typename T>
__device__ __forceinline__ void axis_weights_d(double r, double Hi, double half_ii, int M,
                                               int &base, T *w, T *dw) {
    double xq;
    locate(r, Hi, (double)((P - 1) / 2), half_ii, M, base, xq);
    Bspline<P, T>::w((T)xq, w);
    Bspline<P, T>::dw((T)xq, dw);
}

// Error-free transformation of a sum (Knuth's TwoSum): s + e == a + b exactly.
template <typename T> __device__ __forceinline__ void two_sum(T a, T b, T &s, T &e) {
#pragma clang fp contract(off)
    s = a + b;
    const T bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated accumulator: value = hi + lo.
template <typename T> struct CompSum {
    T hi, lo;
    __device__ __forceinline__ CompSum() : hi(0), lo(0) {}
    __device__ __forceinline__ void add(T v) {
        T s, e;
        two_sum(hi, v, s, e);
        hi = s;
        lo += e;
    }
    __device__ __forceinline__ double value() const { return (double)hi + (double)lo; }
};

}  // namespace p3m
