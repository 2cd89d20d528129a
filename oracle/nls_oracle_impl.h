/*
 * nls_oracle_impl.h -- body of the CPU oracle, included once per precision by
 * nls_oracle.c with REAL (float or double) and SFX (symbol suffix) defined.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran core (nls/nls.f90) used as the checker for the HIP path and as the
 * "port" CPU baseline.  Nothing in the product links or calls it.
 *
 * Array conventions (identical to the reference's memory layout):
 *   - complex arrays are interleaved (re, im) pairs of REAL;
 *   - 2-D arrays are n*n, flat index f = row*n + col with col contiguous.  In
 *     the reference's column-major u(i,j), i is contiguous: col = i-1,
 *     row = j-1, so the two layouts coincide byte for byte;
 *   - banded matrices use BLAS general-band storage, Fortran shape (m, n):
 *     element (p, j) (1-based) at band[(j-1)*m + (p-1)], holding A(p+j-ku-1, j).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* Fortran-style 1-based accessor for an (m, n) column-major band. */
#define BAND(mat, m, p, j) (mat)[((j) - 1) * (m) + ((p) - 1)]

/* make_banded_matrix, nls/nls.f90:61-91.  Loop order kept (left triangle,
 * body, right triangle) so that small n, where the parts overlap, gives the
 * same array as the reference. */
int FN(oracle_make_banded_matrix)(int n, int m, const REAL *row, REAL *mat)
{
    int j, p, k = (m - 1) / 2, l = k + 1, r = n - k;
    if (n < 1 || m < 1 || (m % 2) == 0) return -1;
    for (j = 1; j <= k; ++j) {
        for (p = k + 2 - j; p <= m; ++p) BAND(mat, m, p, j) = row[p - 1];
        for (p = 1; p <= k + 1 - j; ++p) BAND(mat, m, p, j) = 0;
    }
    for (j = l; j <= r; ++j)
        for (p = 1; p <= m; ++p) BAND(mat, m, p, j) = row[p - 1];
    for (j = 1; j <= k; ++j) {
        for (p = 1; p <= m - j; ++p) BAND(mat, m, p, n - k + j) = row[p - 1];
        for (p = m - j + 1; p <= m; ++p) BAND(mat, m, p, n - k + j) = 0;
    }
    return 0;
}

/* clear_first_row_of_derivative, nls/nls.f90:93-107 */
static void FN(clear_first_row)(int n, int m, REAL *L1)
{
    int i;
    (void)n;
    for (i = 1; i <= (m + 1) / 2; ++i) BAND(L1, m, i, (m + 1) / 2 - i + 1) = 0;
}

/* divide_derivative_on_radius, nls/nls.f90:109-130: band element (i, j) lies
 * on matrix row k = i + j - (m+3)/2 (0-based), divided by r_k = k*h. */
static void FN(divide_on_radius)(int n, int m, REAL h, REAL *L1)
{
    int i, j, row;
    for (i = 1; i <= m; ++i)
        for (j = 1; j <= n; ++j) {
            row = i + j - (m + 3) / 2;
            if (row > 0) BAND(L1, m, i, j) = BAND(L1, m, i, j) / ((REAL)row * h);
        }
}

/* make_laplacian_o3/o5/o7, nls/nls.f90:132-259: the axisymmetric radial
 * operator d2/dr2 + (1/r) d/dr on r_k = k*h, k = 0..n-1, as a band of width
 * m = order.  L1 receives the first-derivative part (row 0 cleared, rows k>0
 * divided by k*h), L2 the second-derivative part with the r = 0 mirror
 * weights; the operator is L1 + L2.  L1/L2 may be NULL. */
int FN(oracle_laplacian_1d_parts)(int n, int order, REAL h, REAL *L1out, REAL *L2out)
{
    int m = order, p;
    REAL dx1, dx2, row1d[7], row2d[7];
    REAL *L1, *L2;
    if (!(order == 3 || order == 5 || order == 7) || n < order) return -1;
    L1 = (REAL *)calloc((size_t)m * n, sizeof(REAL));
    L2 = (REAL *)calloc((size_t)m * n, sizeof(REAL));
    if (!L1 || !L2) { free(L1); free(L2); return -2; }
    if (order == 3) {                               /* nls.f90:145-160 */
        static const int a[3] = {1, 0, -1}, b[3] = {1, -2, 1};
        dx1 = 2 * h; dx2 = 1 * (h * h);
        for (p = 0; p < 3; ++p) { row1d[p] = (REAL)a[p] / dx1; row2d[p] = (REAL)b[p] / dx2; }
        FN(oracle_make_banded_matrix)(n, m, row1d, L1);
        FN(oracle_make_banded_matrix)(n, m, row2d, L2);
        BAND(L2, m, 2, 1) = 2 * BAND(L2, m, 2, 1);
        BAND(L2, m, 1, 2) = 4 * BAND(L2, m, 1, 2);
    } else if (order == 5) {                        /* nls.f90:176-194 */
        static const int a[5] = {-1, 8, 0, -8, 1}, b[5] = {-1, 16, -30, 16, -1};
        dx1 = 12 * h; dx2 = 12 * (h * h);
        for (p = 0; p < 5; ++p) { row1d[p] = (REAL)a[p] / dx1; row2d[p] = (REAL)b[p] / dx2; }
        FN(oracle_make_banded_matrix)(n, m, row1d, L1);
        FN(oracle_make_banded_matrix)(n, m, row2d, L2);
        BAND(L1, m, 3, 2) = BAND(L1, m, 3, 2) + (REAL)1.0 / dx1;
        BAND(L2, m, 3, 2) = BAND(L2, m, 3, 2) - (REAL)1.0 / dx2;
        BAND(L2, m, 3, 1) = 2 * BAND(L2, m, 3, 1);
        BAND(L2, m, 2, 2) = 4 * BAND(L2, m, 2, 2);
        BAND(L2, m, 1, 3) = 4 * BAND(L2, m, 1, 3);
    } else {                                        /* nls.f90:210-258 */
        static const int a[7] = {1, -9, 45, 0, -45, 9, -1};
        static const int b[7] = {2, -27, 270, -490, 270, -27, 2};
        dx1 = 60 * h; dx2 = 180 * (h * h);
        for (p = 0; p < 7; ++p) { row1d[p] = (REAL)a[p] / dx1; row2d[p] = (REAL)b[p] / dx2; }
        FN(oracle_make_banded_matrix)(n, m, row1d, L1);
        FN(oracle_make_banded_matrix)(n, m, row2d, L2);
        BAND(L1, m, 4, 2) = BAND(L1, m, 4, 2) + (REAL)9.0 / dx1;
        BAND(L1, m, 3, 3) = BAND(L1, m, 3, 3) - (REAL)1.0 / dx1;
        BAND(L1, m, 5, 2) = BAND(L1, m, 5, 2) - (REAL)1.0 / dx1;
        BAND(L2, m, 4, 1) = 2 * BAND(L2, m, 4, 1);
        BAND(L2, m, 3, 2) = 4 * BAND(L2, m, 3, 2);
        BAND(L2, m, 2, 3) = 4 * BAND(L2, m, 2, 3);
        BAND(L2, m, 1, 4) = 4 * BAND(L2, m, 1, 4);
        BAND(L2, m, 4, 2) = BAND(L2, m, 4, 2) - (REAL)27.0 / dx2;
        BAND(L2, m, 3, 3) = BAND(L2, m, 3, 3) + (REAL)2.0 / dx2;
        BAND(L2, m, 5, 2) = BAND(L2, m, 5, 2) + (REAL)2.0 / dx2;
    }
    FN(clear_first_row)(n, m, L1);
    FN(divide_on_radius)(n, m, h, L1);
    if (L1out) memcpy(L1out, L1, sizeof(REAL) * (size_t)m * n);
    if (L2out) memcpy(L2out, L2, sizeof(REAL) * (size_t)m * n);
    free(L1); free(L2);
    return 0;
}

/* make_laplacian, nls/nls.f90:278-295: op = L1 + L2. */
int FN(oracle_make_laplacian)(int n, int order, REAL h, REAL *op)
{
    size_t i, sz = (size_t)order * (size_t)n;
    REAL *L1 = (REAL *)malloc(sizeof(REAL) * sz), *L2 = (REAL *)malloc(sizeof(REAL) * sz);
    int rc;
    if (!L1 || !L2) { free(L1); free(L2); return -2; }
    rc = FN(oracle_laplacian_1d_parts)(n, order, h, L1, L2);
    if (rc == 0) for (i = 0; i < sz; ++i) op[i] = L1[i] + L2[i];
    free(L1); free(L2);
    return rc;
}

/* gbmv with kl = ku = klu, trans = 'N', beta = 1 as called by rgbmv
 * (nls/nls.f90:530-541): y := alpha*A*x + y, restricted to the n x n matrix.
 * Column-oriented like the reference BLAS: for each column j,
 * temp = alpha*x(j) and y(i) += temp*A(i,j). */
int FN(oracle_gbmv)(int n, int klu, REAL alpha, const REAL *band, const REAL *x, REAL *y)
{
    int j, i, m = 2 * klu + 1;
    for (j = 1; j <= n; ++j) {
        REAL temp = alpha * x[j - 1];
        int ilo = j - klu < 1 ? 1 : j - klu, ihi = j + klu > n ? n : j + klu;
        for (i = ilo; i <= ihi; ++i) y[i - 1] += temp * BAND(band, m, klu + 1 + i - j, j);
    }
    return 0;
}

/* Second-derivative weights of the Cartesian stencils (centre at index R). */
static int FN(cart_weights)(int order, REAL h, REAL *w)
{
    int d;
    if (order == 3) {           /* row2d of nls.f90:149 */
        static const int c[3] = {1, -2, 1};
        for (d = 0; d < 3; ++d) w[d] = (REAL)c[d] / (1 * (h * h));
    } else if (order == 5) {    /* row2d of nls.f90:180; middle of :340 is twice the centre */
        static const int c[5] = {-1, 16, -30, 16, -1};
        for (d = 0; d < 5; ++d) w[d] = (REAL)c[d] / (12 * (h * h));
    } else if (order == 7) {    /* row2d of nls.f90:214; middle of :371 */
        static const int c[7] = {2, -27, 270, -490, 270, -27, 2};
        for (d = 0; d < 7; ++d) w[d] = (REAL)c[d] / (180 * (h * h));
    } else return -1;
    return 0;
}

/* y += alpha * L x for the chosen operator.
 *  dim 1, bc 0 ("ref"): radial band of nls.f90:278 applied as rgbmv;
 *  dim 2, bc 0 ("ref"): block-banded Cartesian Laplacian with zero values
 *    outside the grid, applied in the block order of rbbmv_o3/o5/o7
 *    (nls.f90:408-508): off-diagonal blocks and the middle band;
 *  bc 1 ("periodic"): the same interior Cartesian stencils with wrap-around.
 * band is the 1-D radial band (dim 1, bc 0 only). */
static void FN(apply_laplacian)(int dim, int n, int order, int bc, REAL h, const REAL *band,
                                REAL alpha, const REAL *x, REAL *y)
{
    int R = (order - 1) / 2, i, j, d;
    REAL w[7];
    if (dim == 1 && bc == 0) { FN(oracle_gbmv)(n, R, alpha, band, x, y); return; }
    FN(cart_weights)(order, h, w);
    if (dim == 1) {
        for (i = 0; i < n; ++i) {
            REAL acc = y[i];
            for (d = -R; d <= R; ++d) acc += (alpha * x[((i + d) % n + n) % n]) * w[R + d];
            y[i] = acc;
        }
        return;
    }
    /* 2-D.  Block order of the reference: o3 = (-1, 0, +1), o5 = (+2, +1, 0,
     * -1, -2), o7 = (-3, -2, -1, 0, +1, +2, +3); block 0 is the middle band
     * along the contiguous axis whose centre is twice the 1-D centre. */
    {
        static const int ord3[3] = {-1, 0, 1}, ord5[5] = {2, 1, 0, -1, -2};
        static const int ord7[7] = {-3, -2, -1, 0, 1, 2, 3};
        const int *ord = order == 3 ? ord3 : order == 5 ? ord5 : ord7;
        int b;
        #pragma omp parallel for private(j, b, d) schedule(static)
        for (i = 0; i < n; ++i) {               /* i: row (slow axis) */
            for (b = 0; b < order; ++b) {
                int off = ord[b];
                if (off != 0) {
                    int src = i + off;
                    if (bc == 0 && (src < 0 || src >= n)) continue;
                    src = ((src % n) + n) % n;
                    for (j = 0; j < n; ++j) y[i * n + j] += (alpha * x[src * n + j]) * w[R + off];
                } else {
                    /* middle band over the contiguous axis, column-oriented
                     * like gbmv: contributions to y(j) arrive in order of the
                     * source column jj = j-R .. j+R */
                    REAL *yr = y + (size_t)i * n;
                    const REAL *xr = x + (size_t)i * n;
                    int jj;
                    for (j = 0; j < n; ++j) {
                        REAL acc = yr[j];
                        for (jj = j - R; jj <= j + R; ++jj) {
                            REAL wt = (jj == j) ? 2 * w[R] : w[R + (jj - j)];
                            int s = jj;
                            if (bc == 0 && (s < 0 || s >= n)) continue;
                            s = ((s % n) + n) % n;
                            acc += (alpha * xr[s]) * wt;
                        }
                        yr[j] = acc;
                    }
                }
            }
        }
    }
}

/* The right-hand side F(psi, n) -- hamiltonian + revervoir of
 * nls/nls.f90:570-650 (1-D) and :829-870 (2-D), with the reservoir either
 * algebraic (c11 == 0, as in the reference, nls.f90:580) or a dynamic field:
 *   dpsi = (c3 r - c4) psi - i (c5 |psi|^2 + c6 r) psi + i L psi
 *   r    = c12 P / (c13 + c14 |psi|^2)             (adiabatic)
 *        = n                                       (dynamic)
 *   dn   = (c12 P - (c13 + c14 |psi|^2) n) / c11   (dynamic only)
 * coeffs holds the reference's 23 coefficients, 0-based (c_k = coeffs[k-1]).
 */
static void FN(rhs)(int dim, int n, int order, int bc, REAL h, const REAL *band, const REAL *P,
                    const double *coeffs, const REAL *psi, const REAL *nres, REAL *dpsi, REAL *dn,
                    REAL *ur, REAL *ui, REAL *vr, REAL *vi)
{
    size_t N = dim == 1 ? (size_t)n : (size_t)n * n, k;
    const REAL c3 = (REAL)coeffs[2], c4 = (REAL)coeffs[3], c5 = (REAL)coeffs[4], c6 = (REAL)coeffs[5];
    const REAL c11 = (REAL)coeffs[10], c12 = (REAL)coeffs[11], c13 = (REAL)coeffs[12], c14 = (REAL)coeffs[13];
    const int dynamic = coeffs[10] != 0.0;
    for (k = 0; k < N; ++k) {
        REAL a = psi[2 * k], b = psi[2 * k + 1];
        REAL s = a * a + b * b;
        REAL r = dynamic ? nres[k] : c12 * P[k] / (c13 + c14 * s);
        ur[k] = a; ui[k] = b;
        vr[k] = (c3 * r - c4) * a + (c5 * s + c6 * r) * b;
        vi[k] = (c3 * r - c4) * b - (c5 * s + c6 * r) * a;
        if (dynamic) dn[k] = (c12 * P[k] - (c13 + c14 * s) * nres[k]) / c11;
    }
    FN(apply_laplacian)(dim, n, order, bc, h, band, (REAL)-1.0, ui, vr);
    FN(apply_laplacian)(dim, n, order, bc, h, band, (REAL)+1.0, ur, vi);
    for (k = 0; k < N; ++k) { dpsi[2 * k] = vr[k]; dpsi[2 * k + 1] = vi[k]; }
}

static int FN(check_args)(int dim, int n, int order, int bc)
{
    if (dim != 1 && dim != 2) return -1;
    if (!(order == 3 || order == 5 || order == 7)) return -1;
    if (bc != 0 && bc != 1) return -1;
    if (n < order) return -1;
    return 0;
}

/* One evaluation of F, exposed for tests (hamiltonian / hamiltonian_2d). */
int FN(oracle_rhs)(int dim, int n, int order, int bc, REAL dx, const REAL *P, const double *coeffs,
                   const REAL *psi, const REAL *nres, REAL *dpsi, REAL *dn)
{
    size_t N = dim == 1 ? (size_t)n : (size_t)n * n;
    REAL *band = NULL, *tmp;
    if (FN(check_args)(dim, n, order, bc)) return -1;
    tmp = (REAL *)malloc(sizeof(REAL) * 4 * N);
    if (!tmp) return -2;
    if (dim == 1 && bc == 0) {
        band = (REAL *)malloc(sizeof(REAL) * (size_t)order * n);
        FN(oracle_make_laplacian)(n, order, dx, band);
    }
    FN(rhs)(dim, n, order, bc, dx, band, P, coeffs, psi, nres, dpsi, dn, tmp, tmp + N, tmp + 2 * N, tmp + 3 * N);
    free(band); free(tmp);
    return 0;
}

/* Classical RK4, runge_kutta / runge_kutta_2d (nls/nls.f90:705-734,
 * :873-901), advancing psi (and n when c11 != 0) in place by iters steps:
 *   k1 = F(u), k2 = F(u + k1*dt/2), k3 = F(u + k2*dt/2), k4 = F(u + k3*dt/1)
 *   u  = u + (k1 + 2*k2 + 2*k3 + k4)*dt/6
 * with the reference's left-to-right evaluation order of each expression. */
int FN(oracle_rk4)(int dim, int n, int order, int bc, REAL dt, REAL dx, int iters, const REAL *P,
                   const double *coeffs, REAL *psi, REAL *nres)
{
    size_t N = dim == 1 ? (size_t)n : (size_t)n * n, k;
    const int dynamic = coeffs[10] != 0.0;
    REAL *band = NULL, *w, *k1, *k2, *k3, *k4, *y, *m1, *m2, *m3, *m4, *yn, *tmp;
    int it, rc = 0;
    if (FN(check_args)(dim, n, order, bc) || iters < 0) return -1;
    if (dynamic && !nres) return -1;
    w = (REAL *)calloc(2 * N * 5 + N * 5 + 4 * N, sizeof(REAL));
    if (!w) return -2;
    k1 = w; k2 = k1 + 2 * N; k3 = k2 + 2 * N; k4 = k3 + 2 * N; y = k4 + 2 * N;
    m1 = y + 2 * N; m2 = m1 + N; m3 = m2 + N; m4 = m3 + N; yn = m4 + N; tmp = yn + N;
    if (dim == 1 && bc == 0) {
        band = (REAL *)malloc(sizeof(REAL) * (size_t)order * n);
        if (!band) { free(w); return -2; }
        rc = FN(oracle_make_laplacian)(n, order, dx, band);
    }
    for (it = 0; it < iters && rc == 0; ++it) {
        /* k1 = F(u + 0.*dt/2) */
        for (k = 0; k < 2 * N; ++k) y[k] = psi[k] + (REAL)0.0 * dt / 2;
        if (dynamic) for (k = 0; k < N; ++k) yn[k] = nres[k];
        FN(rhs)(dim, n, order, bc, dx, band, P, coeffs, y, yn, k1, m1, tmp, tmp + N, tmp + 2 * N, tmp + 3 * N);
        /* k2 = F(u + k1*dt/2) */
        for (k = 0; k < 2 * N; ++k) y[k] = psi[k] + k1[k] * dt / 2;
        if (dynamic) for (k = 0; k < N; ++k) yn[k] = nres[k] + m1[k] * dt / 2;
        FN(rhs)(dim, n, order, bc, dx, band, P, coeffs, y, yn, k2, m2, tmp, tmp + N, tmp + 2 * N, tmp + 3 * N);
        /* k3 = F(u + k2*dt/2) */
        for (k = 0; k < 2 * N; ++k) y[k] = psi[k] + k2[k] * dt / 2;
        if (dynamic) for (k = 0; k < N; ++k) yn[k] = nres[k] + m2[k] * dt / 2;
        FN(rhs)(dim, n, order, bc, dx, band, P, coeffs, y, yn, k3, m3, tmp, tmp + N, tmp + 2 * N, tmp + 3 * N);
        /* k4 = F(u + k3*dt/1) */
        for (k = 0; k < 2 * N; ++k) y[k] = psi[k] + k3[k] * dt / 1;
        if (dynamic) for (k = 0; k < N; ++k) yn[k] = nres[k] + m3[k] * dt / 1;
        FN(rhs)(dim, n, order, bc, dx, band, P, coeffs, y, yn, k4, m4, tmp, tmp + N, tmp + 2 * N, tmp + 3 * N);
        /* u = u + (k1 + 2*k2 + 2*k3 + k4)*dt/6 */
        for (k = 0; k < 2 * N; ++k) psi[k] = psi[k] + (k1[k] + 2 * k2[k] + 2 * k3[k] + k4[k]) * dt / 6;
        if (dynamic)
            for (k = 0; k < N; ++k) nres[k] = nres[k] + (m1[k] + 2 * m2[k] + 2 * m3[k] + m4[k]) * dt / 6;
    }
    free(band); free(w);
    return rc;
}

/* chemical_potential_1d (nls/nls.f90:921-948): order 5, one RHS evaluation
 * u = F(u0), r_i = (i-1)*dx, M = sum conj(u0) u0 r, E = i sum conj(u0) u r,
 * mu = E/M (complex, returned as (re, im)). */
int FN(oracle_chemical_potential_1d)(REAL dx, int n, const REAL *P, const double *coeffs,
                                     const REAL *u0, REAL *mu)
{
    double c[23];
    REAL *u = (REAL *)malloc(sizeof(REAL) * 2 * (size_t)n);
    REAL Mr = 0, Mi = 0, Er = 0, Ei = 0, den;
    int i, rc;
    if (!u) return -2;
    memcpy(c, coeffs, sizeof c);
    c[10] = 0.0;                    /* the reference's hamiltonian is adiabatic */
    rc = FN(oracle_rhs)(1, n, 5, 0, dx, P, c, u0, NULL, u, NULL);
    if (rc) { free(u); return rc; }
    for (i = 0; i < n; ++i) {
        REAL r = ((REAL)(i + 1) - (REAL)1.0) * dx;
        REAL ar = u0[2 * i], ai = u0[2 * i + 1];
        REAL br = u0[2 * i] * r, bi = u0[2 * i + 1] * r;           /* u0*r */
        REAL cr = u[2 * i] * r, ci = u[2 * i + 1] * r;             /* u*r  */
        Mr += ar * br + ai * bi; Mi += ar * bi - ai * br;          /* conj(a)*b */
        Er += ar * cr + ai * ci; Ei += ar * ci - ai * cr;
    }
    /* E = i * (Er + i Ei) = -Ei + i Er ; mu = E / M */
    { REAL xr = -Ei, xi = Er; den = Mr * Mr + Mi * Mi;
      mu[0] = (xr * Mr + xi * Mi) / den; mu[1] = (xi * Mr - xr * Mi) / den; }
    free(u);
    return 0;
}

/* chemical_potential_2d (nls/nls.f90:950-971): unweighted sums; the result
 * is assigned to a real, so only Re(E/M) survives. */
int FN(oracle_chemical_potential_2d)(REAL dx, int n, const REAL *P, const double *coeffs,
                                     const REAL *u0, REAL *mu)
{
    double c[23];
    size_t N = (size_t)n * n, i;
    REAL *u = (REAL *)malloc(sizeof(REAL) * 2 * N);
    REAL Mr = 0, Mi = 0, Er = 0, Ei = 0;
    int rc;
    if (!u) return -2;
    memcpy(c, coeffs, sizeof c);
    c[10] = 0.0;
    rc = FN(oracle_rhs)(2, n, 5, 0, dx, P, c, u0, NULL, u, NULL);
    if (rc) { free(u); return rc; }
    for (i = 0; i < N; ++i) {
        REAL ar = u0[2 * i], ai = u0[2 * i + 1];
        Mr += ar * ar + ai * ai; Mi += ar * ai - ai * ar;
        Er += ar * u[2 * i] + ai * u[2 * i + 1]; Ei += ar * u[2 * i + 1] - ai * u[2 * i];
    }
    { REAL xr = -Ei, xi = Er, den = Mr * Mr + Mi * Mi;
      mu[0] = (xr * Mr + xi * Mi) / den; }
    free(u);
    return 0;
}

#undef BAND
#undef FN
#undef CAT
#undef CAT_
