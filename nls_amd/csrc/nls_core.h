// nls_core.h -- device-side building blocks shared by the kernels of
// libnls_amd: complex load/store types, the pointwise right-hand side and
// the RK4 stage bookkeeping.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nls_stencil.h"

namespace nls {

// BC_SLAB (2-D fused kernel only): one row slab of a grid that is periodic
// along its rows, with the 4R ghost rows above and below supplied in the
// buffers; x wraps, y reads the ghost rows.
enum { BC_REF = 0, BC_PERIODIC = 1, BC_SLAB = 2 };

template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using cplx_t = typename cvec<T>::type;

template <typename T> __device__ __forceinline__ cplx_t<T> mk(T a, T b) {
    cplx_t<T> r; r.x = a; r.y = b; return r;
}

// Coefficients of the model actually read by the core (reference c(k)).
template <typename T> struct Coeffs {
    T c3, c4, c5, c6;        // gain, loss, self-interaction, reservoir coupling
    T c11, c12, c13, c14;    // reservoir: d/dt, pumping, decay, depletion
    T d12, d13, d14;         // c12/c11, c13/c11, c14/c11 (formed in double, rounded once); 0 when c11 = 0
};

template <typename T> inline Coeffs<T> make_coeffs(const double *c) {
    Coeffs<T> r;
    r.c3 = (T)c[2]; r.c4 = (T)c[3]; r.c5 = (T)c[4]; r.c6 = (T)c[5];
    r.c11 = (T)c[10]; r.c12 = (T)c[11]; r.c13 = (T)c[12]; r.c14 = (T)c[13];
    const bool dyn = c[10] != 0.0;
    r.d12 = dyn ? (T)(c[11] / c[10]) : (T)0;
    r.d13 = dyn ? (T)(c[12] / c[10]) : (T)0;
    r.d14 = dyn ? (T)(c[13] / c[10]) : (T)0;
    return r;
}

// Right-hand side at one point, given the stage input (y, yn), pumping P and
// the Laplacian of y.  Mirrors hamiltonian / revervoir of the reference
// (nls/nls.f90:570-581, :637-647):
//   r      = c12 P / (c13 + c14 |y|^2)   (adiabatic)   or yn (dynamic)
//   k.re   = (c3 r - c4) y.re + (c5 |y|^2 + c6 r) y.im - (L y).im
//   k.im   = (c3 r - c4) y.im - (c5 |y|^2 + c6 r) y.re + (L y).re
//   kn     = (c12 P - (c13 + c14 |y|^2) yn) / c11        (dynamic only)
// The reservoir's 1/c11 is folded into its coefficients once
// (kn = d12 P - (d13 + d14 |y|^2) yn), and each psi component is two fused
// multiply-adds; both regroupings move results by rounding only.
//
// rhs_point_cp takes the pumping term already formed as cP = c12 P
// (adiabatic) or d12 P (dynamic), so kernels that keep a point's pumping in
// a register across the four stages form it once (pump_term).
template <typename T, bool DYN> __device__ __forceinline__ T pump_term(const Coeffs<T> &c, T P) {
    return DYN ? c.d12 * P : c.c12 * P;
}

template <typename T, bool DYN>
__device__ __forceinline__ void rhs_point_cp(const Coeffs<T> &c, T cP, cplx_t<T> y, T yn, cplx_t<T> lap,
                                             cplx_t<T> &k, T &kn) {
    const T s = y.x * y.x + y.y * y.y;
    const T r = DYN ? yn : cP / (c.c13 + c.c14 * s);
    const T g = c.c3 * r - c.c4;
    const T q = c.c5 * s + c.c6 * r;
    k.x = fma(g, y.x, fma(q, y.y, -lap.y));
    k.y = fma(g, y.y, fma(-q, y.x, lap.x));
    if (DYN) kn = fma(-fma(c.d14, s, c.d13), yn, cP);
}

// The same right-hand side with the Laplacian given as its neighbour part S
// (lap = wc y + S, wc the centre weight): the centre term joins the
// point's own factor, q' = c5 |y|^2 + c6 r - wc, so
//   k.re = (c3 r - c4) y.re + q' y.im - S.im,  k.im = (c3 r - c4) y.im - q' y.re + S.re,
// which saves the two products wc y.re, wc y.im.  Equal to rhs_point_cp up to
// rounding (q' rounds at the scale of wc, as wc y does in the sum).
template <typename T, bool DYN>
__device__ __forceinline__ void rhs_point_centre(const Coeffs<T> &c, T cP, T wc, cplx_t<T> y, T yn, cplx_t<T> S,
                                                 cplx_t<T> &k, T &kn) {
    const T s = y.x * y.x + y.y * y.y;
    const T r = DYN ? yn : cP / (c.c13 + c.c14 * s);
    const T g = c.c3 * r - c.c4;
    const T q = fma(c.c5, s, fma(c.c6, r, -wc));
    k.x = fma(g, y.x, fma(q, y.y, -S.y));
    k.y = fma(g, y.y, fma(-q, y.x, S.x));
    if (DYN) kn = fma(-fma(c.d14, s, c.d13), yn, cP);
}

template <typename T, bool DYN>
__device__ __forceinline__ void rhs_point(const Coeffs<T> &c, T P, cplx_t<T> y, T yn, cplx_t<T> lap,
                                          cplx_t<T> &k, T &kn) {
    rhs_point_cp<T, DYN>(c, pump_term<T, DYN>(c, P), y, yn, lap, k, kn);
}

// Cartesian Laplacian of a point with value c from its neighbours at
// distance d = 1..R (weights w_d of nls_stencil.h).
//   double precision: the stencil sum as written, the centre weight times c
//     (lap_centre) plus w_d times the sum of the neighbours at distance d;
//   single precision: the centre weight is folded into the neighbour terms
//     (the weights of a stencil sum to zero).  For d = 1 the differences to c
//     are formed first -- exact when the values are close, as on a smooth
//     field on a fine grid -- so the largest weights (w_1 ~ 1/h^2) no longer
//     amplify the rounding of c; for d >= 2 the neighbour sum less 2c (1-D) or
//     4c (2-D).  This keeps the single-precision Laplacian accurate where the
//     stencil sum cancels to a small result.
// The two forms are equal up to rounding.
template <typename T> __device__ __forceinline__ T lap_centre(T wc, T c) {
    return sizeof(T) == 8 ? wc * c : (T)0;
}

template <typename T> __device__ __forceinline__ T lap_pair(T w, T c, T a0, T a1, int d) {
    if (sizeof(T) == 8) return w * (a0 + a1);
    if (d == 1) return w * ((a0 - c) + (a1 - c));
    return w * ((a0 + a1) - (T)2 * c);
}

template <typename T> __device__ __forceinline__ T lap_quad(T w, T c, T a0, T a1, T b0, T b1, int d) {
    if (sizeof(T) == 8) return w * ((a0 + a1) + (b0 + b1));
    if (d == 1) return w * (((a0 - c) + (a1 - c)) + ((b0 - c) + (b1 - c)));
    return w * (((a0 + a1) + (b0 + b1)) - (T)4 * c);
}

// Periodic index wrap for offsets that may exceed one period.
__device__ __forceinline__ int wrap_index(int g, int n) {
    while (g < 0) g += n;
    while (g >= n) g -= n;
    return g;
}

// Periodic index wrap for -n <= g < 2n as one three-way unsigned minimum:
// of g, g + n and g - n exactly one lies in [0, n) and the others, as
// unsigned values, are larger.
__device__ __forceinline__ int wrap_near3(int g, int n) {
    const unsigned a = (unsigned)g, b = (unsigned)(g + n), c = (unsigned)(g - n);
    return (int)min(min(a, b), c);
}

// Linear index gy * n + gx of a point of an n x n grid whose index fits 32
// bits (so gy, n < 2^24): one 24-bit multiply-add.  Written as a 32-bit
// product the compiler emits a 64-bit multiply-add whose high addend half
// it may take from any register -- in the fused kernels, one a ring-table
// load is still filling, so the first field loads waited for that load.
__device__ __forceinline__ unsigned grid_index(int gy, int n, int gx) {
    return __umul24((unsigned)gy, (unsigned)n) + (unsigned)gx;
}

// Periodic index wrap for -n <= g < 2n.
__device__ __forceinline__ int wrap_near(int g, int n) {
    g += g < 0 ? n : 0;
    g -= g >= n ? n : 0;
    return g;
}

}  // namespace nls
