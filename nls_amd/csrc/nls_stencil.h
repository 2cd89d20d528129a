// nls_stencil.h -- finite-difference stencil weights of the time-stepping
// core (shared by host and device code).
//
// Second-derivative rows s2 and first-derivative rows s1 of width
// order = 2R+1, as functions of the offset d = -R..R:
//   order 3: s2 = (1, -2, 1)/h^2                 s1 = (-1, 0, 1)/(2h)
//   order 5: s2 = (-1, 16, -30, 16, -1)/(12h^2)  s1 = (1, -8, 0, 8, -1)/(12h)
//   order 7: s2 = (2, -27, 270, -490, 270, -27, 2)/(180h^2)
//            s1 = (-1, 9, -45, 0, 45, -9, 1)/(60h)
// These are the interior rows of make_laplacian_o3/o5/o7 (row2d / row1d in
// reference nls/nls.f90:148-149, :179-180, :213-214; the reference stores
// row1d in band order, i.e. reversed in d).  The 2-D operators of
// make_laplacian_2d_o3/o5/o7 (nls.f90:297-385) are s2 along each axis, so
// their middle band has twice the 1-D centre weight.
//
// The 1-D "ref" operator is the axisymmetric radial Laplacian
// d2/dr2 + (1/r) d/dr on r_k = k*h, k = 0..n-1 (make_laplacian,
// nls.f90:278-295).  Written as a dense row of the operator:
//   row 0    : 2 * (s2 folded by the mirror psi(-r) = psi(r)); the first
//              derivative row is cleared (nls.f90:93-107), the factor 2 is
//              the r -> 0 limit of (1/r) d/dr -> d2/dr2 (nls.f90:154-155,
//              :187-189, :235-238);
//   row k >= 1: s2 + s1/(k*h), with taps that fall on r < 0 folded back onto
//              r > 0 by the same mirror (the corrections of nls.f90:185-186,
//              :219-253) and taps beyond r_{n-1} dropped (nls.f90:75-90).
//
// All weights are scalar functions of (order, d): no arrays, so device code
// built from them needs no private (scratch) memory.
#pragma once

#if defined(__HIPCC__)
#define NLS_HD __host__ __device__
#else
#define NLS_HD
#endif

namespace nls {

// s2(d): second-derivative weight at offset d.
NLS_HD inline double stencil_s2(int order, int d, double h) {
    const int a = d < 0 ? -d : d;
    if (order == 3) return (a == 0 ? -2.0 : 1.0) / (1.0 * (h * h));
    if (order == 5) return (a == 0 ? -30.0 : a == 1 ? 16.0 : -1.0) / (12.0 * (h * h));
    return (a == 0 ? -490.0 : a == 1 ? 270.0 : a == 2 ? -27.0 : 2.0) / (180.0 * (h * h));
}

// s1(d): first-derivative weight at offset d (odd in d).
NLS_HD inline double stencil_s1(int order, int d, double h) {
    if (d == 0) return 0.0;
    const int a = d < 0 ? -d : d;
    const double sg = d < 0 ? -1.0 : 1.0;
    if (order == 3) return sg * 1.0 / (2.0 * h);
    if (order == 5) return sg * (a == 1 ? 8.0 : -1.0) / (12.0 * h);
    return sg * (a == 1 ? 45.0 : a == 2 ? -9.0 : 1.0) / (60.0 * h);
}

// Weight of the 1-D radial operator in row i on column i + (k - R),
// k = 0..order-1; zero for columns outside [0, n).
NLS_HD inline double radial_weight(int i, int k, int n, int order, double h) {
    const int R = (order - 1) / 2;
    const int j = i + k - R;
    if (j < 0 || j >= n) return 0.0;
    double w = 0.0;
    for (int d = -R; d <= R; ++d) {
        int c = i + d;
        if (c < 0) c = -c;                    // mirror at r = 0
        if (c != j) continue;
        w += (i == 0) ? 2.0 * stencil_s2(order, d, h)
                      : stencil_s2(order, d, h) + stencil_s1(order, d, h) / ((double)i * h);
    }
    return w;
}

}  // namespace nls
