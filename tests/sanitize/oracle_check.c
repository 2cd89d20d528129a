/*
 * oracle_check.c -- drives every function of the CPU oracle (oracle/
 * nls_oracle.c, compiled into this program) on small inputs, so that a build
 * with AddressSanitizer and UndefinedBehaviorSanitizer sees all its array
 * indexing: band builders at every order and the smallest sizes, the band
 * apply, 1-D and 2-D right-hand sides and RK4 on both boundary families and
 * both reservoir modes, the chemical potential.  Test infrastructure only.
 * Exit status 0 when every call succeeded and every result is finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/nls_oracle.c"

static int bad = 0;

static void expect_finite(const char *what, const double *a, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(a[i])) { fprintf(stderr, "%s: non-finite at %zu\n", what, i); bad = 1; return; }
}

static double urand(unsigned *s) { *s = *s * 1103515245u + 12345u; return ((*s >> 8) & 0xffff) / 65536.0; }

int main(void) {
    unsigned seed = 7;
    double c[23] = {0};
    c[2] = 0.5; c[3] = 0.5; c[4] = 1.0; c[5] = 4.0; c[11] = 1.0; c[12] = 1.5; c[13] = 1.0;
    for (int order = 3; order <= 7; order += 2) {
        for (int n = order; n <= order + 9; ++n) {
            double *op = calloc((size_t)order * n, sizeof(double));
            double *L1 = calloc((size_t)order * n, sizeof(double)), *L2 = calloc((size_t)order * n, sizeof(double));
            float *opf = calloc((size_t)order * n, sizeof(float));
            if (oracle_make_laplacian_f64(n, order, 0.1, op) || oracle_make_laplacian_f32(n, order, 0.1f, opf) ||
                oracle_laplacian_1d_parts_f64(n, order, 0.1, L1, L2)) bad = 1;
            expect_finite("make_laplacian", op, (size_t)order * n);
            double *row = calloc(order, sizeof(double)), *mat = calloc((size_t)order * n, sizeof(double));
            for (int k = 0; k < order; ++k) row[k] = k + 1;
            if (oracle_make_banded_matrix_f64(n, order, row, mat)) bad = 1;
            double *x = calloc(n, sizeof(double)), *y = calloc(n, sizeof(double));
            for (int i = 0; i < n; ++i) x[i] = urand(&seed);
            if (oracle_gbmv_f64(n, (order - 1) / 2, 1.0, op, x, y)) bad = 1;
            expect_finite("gbmv", y, n);
            free(op); free(L1); free(L2); free(opf); free(row); free(mat); free(x); free(y);
        }
        for (int dim = 1; dim <= 2; ++dim)
            for (int bc = 0; bc <= 1; ++bc)
                for (int dyn = 0; dyn <= 1; ++dyn) {
                    /* the reference's 2-D order 7 is defined at n = 7 only */
                    const int n = dim == 2 && order == 7 && bc == 0 ? 7 : (dim == 1 ? 24 : 11);
                    const size_t N = dim == 1 ? (size_t)n : (size_t)n * n;
                    double *P = malloc(N * sizeof(double)), *psi = malloc(2 * N * sizeof(double));
                    double *nres = malloc(N * sizeof(double)), *dpsi = malloc(2 * N * sizeof(double));
                    double *dn = malloc(N * sizeof(double));
                    float *Pf = malloc(N * sizeof(float)), *psif = malloc(2 * N * sizeof(float));
                    float *nf = malloc(N * sizeof(float));
                    for (size_t i = 0; i < N; ++i) {
                        P[i] = 1.0 + urand(&seed); Pf[i] = (float)P[i];
                        nres[i] = 0.1 * urand(&seed); nf[i] = (float)nres[i];
                        psi[2 * i] = 0.3 + 0.1 * urand(&seed); psi[2 * i + 1] = 0.1 * urand(&seed);
                        psif[2 * i] = (float)psi[2 * i]; psif[2 * i + 1] = (float)psi[2 * i + 1];
                    }
                    c[10] = dyn ? 1.0 : 0.0;
                    if (oracle_rhs_f64(dim, n, order, bc, 0.2, P, c, psi, dyn ? nres : NULL, dpsi, dyn ? dn : NULL) ||
                        oracle_rk4_f64(dim, n, order, bc, 1e-3, 0.2, 5, P, c, psi, dyn ? nres : NULL) ||
                        oracle_rk4_f32(dim, n, order, bc, 1e-3f, 0.2f, 5, Pf, c, psif, dyn ? nf : NULL))
                        bad = 1;
                    expect_finite("rk4", psi, 2 * N);
                    if (order == 5 && bc == 0 && !dyn) {
                        double mu[2] = {0, 0};
                        if (dim == 1 ? oracle_chemical_potential_1d_f64(0.2, n, P, c, psi, mu)
                                     : oracle_chemical_potential_2d_f64(0.2, n, P, c, psi, mu))
                            bad = 1;
                        expect_finite("chemical_potential", mu, dim == 1 ? 2 : 1);
                    }
                    free(P); free(psi); free(nres); free(dpsi); free(dn); free(Pf); free(psif); free(nf);
                }
    }
    if (bad) fprintf(stderr, "oracle_check: FAILED\n");
    else printf("oracle_check: ok\n");
    return bad;
}
