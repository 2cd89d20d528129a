// nls_helpers.hip -- host drop-ins of the reference module's operator
// builders and BLAS-2 wrappers: make_laplacian (nls/nls.f90:278-295),
// make_banded_matrix (:61-91),
// make_laplacian_2d (:297-403), rbbmv (:408-527), rgbmv (:530-541),
// revervoir / revervoir_2d (:570-581, :829-839).  These build or apply the
// operator arrays on the host, as the reference does; the solvers never call
// them (their kernels hold the stencils).  Single precision, with the
// reference's evaluation order, so results match the reference's arithmetic.
#include <cstring>
#include <vector>

#include "nls_amd.h"
#include "nls_internal.h"
#include "nls_stencil.h"

using namespace nls;

namespace {

// y += sign A x for the band A stored as op(2 klu + 1, n): the column sweep of
// the reference BLAS sgbmv ('N', beta = 1), y(i) += (sign x(j)) A(i, j).
void band_apply(const float *x, float *y, float sign, const float *op, int klu, int n) {
    const int lda = 2 * klu + 1;
    for (int j = 0; j < n; ++j) {
        if (x[j] == 0.0f) continue;
        const float t = sign * x[j];
        const int i0 = j - klu > 0 ? j - klu : 0, i1 = j + klu < n - 1 ? j + klu : n - 1;
        for (int i = i0; i <= i1; ++i) y[i] += t * op[(size_t)j * lda + (klu + i - j)];
    }
}

// make_banded_matrix on one row of m coefficients: interior columns carry the
// row, the first and last (m - 1)/2 columns drop the entries outside the matrix.
void banded(int n, int m, const float *row, float *mat) {
    const int k = (m - 1) / 2;
    for (int j = 0; j < k; ++j)
        for (int b = 0; b < m; ++b) mat[(size_t)j * m + b] = b >= k - j ? row[b] : 0.0f;
    for (int j = k; j < n - k; ++j)
        for (int b = 0; b < m; ++b) mat[(size_t)j * m + b] = row[b];
    for (int q = 1; q <= k; ++q) {
        const int j = n - k + q - 1;
        for (int b = 0; b < m; ++b) mat[(size_t)j * m + b] = b < m - q ? row[b] : 0.0f;
    }
}

}  // namespace

extern "C" {

int nls_make_banded_matrix(int n, int m, const float *row, float *mat) {
    if (!row || !mat) return fail(NLS_EINVAL, "NULL argument");
    if (m < 1 || m % 2 == 0) return fail(NLS_EINVAL, "m must be odd and positive (got %d)", m);
    if (n < m) return fail(NLS_EINVAL, "n=%d is smaller than the row length %d", n, m);
    banded(n, m, row, mat);
    return NLS_OK;
}

int nls_make_laplacian_2d(int n, int m, float h, float *blocks, int *orders) {
    if (!blocks || !orders) return fail(NLS_EINVAL, "NULL argument");
    if (m != 3 && m != 5 && m != 7) return fail(NLS_EINVAL, "m must be 3, 5 or 7 (got %d)", m);
    if (n < m) return fail(NLS_EINVAL, "n=%d is smaller than the stencil width %d", n, m);
    // The reference's interior weights of d^2/dx^2 + d^2/dy^2 over its
    // denominator (nls.f90:313-315, :339-343, :368-374), in single precision.
    // Order 7 is declared blocks(m, 2m - 1) there (:360), valid only at
    // n = 7; this builds the evident (n, 2m - 1) blocks at every n.
    static const float mid3[3] = {1, -4, 1}, mid5[5] = {-1, 16, -60, 16, -1}, mid7[7] = {2, -27, 270, -980, 270, -27, 2};
    const float *mid = m == 3 ? mid3 : m == 5 ? mid5 : mid7;
    const float hh = h * h;
    const float dx2 = m == 3 ? 1.0f * hh : m == 5 ? 12.0f * hh : 180.0f * hh;
    const int R = (m - 1) / 2;
    float row[7];
    for (int b = 0; b < m; ++b) row[b] = mid[b] / dx2;
    for (int d = 1; d <= R; ++d) {
        const float v = mid[R - d] / dx2;          // the off-diagonal blocks: one diagonal each
        for (int j = 0; j < n; ++j) {
            blocks[(size_t)(R - d) * n + j] = v;
            blocks[(size_t)(R + m - 1 + d) * n + j] = v;
        }
    }
    banded(n, m, row, blocks + (size_t)R * n);
    for (int q = 0; q < m; ++q) orders[q] = q == R ? R : 0;
    return NLS_OK;
}

int nls_rgbmv(const float *x, float *u, float sign, const float *op, int klu, int n) {
    if (!x || !u || !op) return fail(NLS_EINVAL, "NULL argument");
    if (klu < 0 || n < 1) return fail(NLS_EINVAL, "bad klu=%d or n=%d", klu, n);
    band_apply(x, u, sign, op, klu, n);
    return NLS_OK;
}

int nls_rbbmv(const float *x, float *y, float sign, const float *blocks, const int *ms, int m, int n) {
    if (!x || !y || !blocks || !ms) return fail(NLS_EINVAL, "NULL argument");
    if (m != 3 && m != 5 && m != 7) return fail(NLS_EINVAL, "m must be 3, 5 or 7 (got %d)", m);
    if (n < 1) return fail(NLS_EINVAL, "bad n=%d", n);
    // Block q (offset q - R lines) applies the band at its column of blocks
    // to grid line i + q - R of x and adds to line i of y, in the order the
    // reference's rbbmv_o3/o5/o7 visit them (nls.f90:420-431, :446-465,
    // :480-506): order 5 from the last block to the first, 3 and 7 the other
    // way round.
    const int R = (m - 1) / 2;
    for (int t = 0; t < m; ++t) {
        const int q = m == 5 ? m - 1 - t : t;
        const int off = q - R;
        const int col = q <= R ? q : R + m - 1 + (q - R);
        const float *band = blocks + (size_t)col * n;
        for (int i = 0; i < n; ++i) {
            const int l = i + off;
            if (l >= 0 && l < n) band_apply(x + (size_t)l * n, y + (size_t)i * n, sign, band, ms[q], n);
        }
    }
    return NLS_OK;
}

static int reservoir(const float *pumping, const float *coeffs, const float *u_sqr, float *r, size_t count) {
    if (!pumping || !coeffs || !u_sqr || !r) return fail(NLS_EINVAL, "NULL argument");
    for (size_t i = 0; i < count; ++i) r[i] = coeffs[11] * pumping[i] / (coeffs[12] + coeffs[13] * u_sqr[i]);
    return NLS_OK;
}

int nls_revervoir(const float *pumping, const float *coeffs, const float *u_sqr, float *r, int n) {
    if (n < 1) return fail(NLS_EINVAL, "bad n=%d", n);
    return reservoir(pumping, coeffs, u_sqr, r, (size_t)n);
}

int nls_revervoir_2d(const float *pumping, const float *coeffs, const float *u_sqr, float *r, int n) {
    if (n < 1) return fail(NLS_EINVAL, "bad n=%d", n);
    return reservoir(pumping, coeffs, u_sqr, r, (size_t)n * n);
}

// make_laplacian (nls.f90:278-295): the 1-D radial operator in BLAS band
// storage, built in double precision from the closed-form row weights.
int nls_make_laplacian(int n, int order, double h, double *op) {
    if (!op) return fail(NLS_EINVAL, "NULL argument");
    if (order != 3 && order != 5 && order != 7) return fail(NLS_EINVAL, "order must be 3, 5 or 7");
    if (n < order) return fail(NLS_EINVAL, "n smaller than the stencil width");
    const int R = (order - 1) / 2;
    std::vector<double> rows((size_t)n * order);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < order; ++k) rows[(size_t)i * order + k] = radial_weight(i, k, n, order, h);
    // band storage: op(p, j) = A(I, j) with I = p + j - R - 1 (1-based)
    for (int j = 1; j <= n; ++j)
        for (int pp = 1; pp <= order; ++pp) {
            const int I = pp + j - R - 1;
            op[(size_t)(j - 1) * order + (pp - 1)] =
                (I >= 1 && I <= n) ? rows[(size_t)(I - 1) * order + (2 * R + 1 - pp)] : 0.0;
        }
    return NLS_OK;
}

// clear_first_row_of_derivative (nls.f90:93-107) and
// divide_derivative_on_radius (:109-130), in place on a first-derivative band
// L1(m, n) in BLAS band storage (column-major, L1(i, j) at L1[(j-1) m + i-1]),
// as make_laplacian_o3/o5/o7 apply them: zero the band entries of the first
// (m+1)/2 entries that lie on matrix row 0, then divide every entry on matrix
// row k = i + j - (m+3)/2 > 0 by r_k = k h, in single precision.
int nls_clear_first_row_of_derivative(int n, int m, float *L1) {
    if (!L1) return fail(NLS_EINVAL, "NULL argument");
    if (m < 1 || m % 2 == 0 || n < (m + 1) / 2) return fail(NLS_EINVAL, "bad band shape (m=%d, n=%d)", m, n);
    for (int i = 1; i <= (m + 1) / 2; ++i) L1[(size_t)((m + 1) / 2 - i) * m + (i - 1)] = 0.0f;
    return NLS_OK;
}

int nls_divide_derivative_on_radius(int n, int m, float h, float *L1) {
    if (!L1) return fail(NLS_EINVAL, "NULL argument");
    if (m < 1 || m % 2 == 0 || n < 1) return fail(NLS_EINVAL, "bad band shape (m=%d, n=%d)", m, n);
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const int row = i + j - (m + 3) / 2;
            if (row > 0) L1[(size_t)(j - 1) * m + (i - 1)] = L1[(size_t)(j - 1) * m + (i - 1)] / ((float)row * h);
        }
    return NLS_OK;
}

}  // extern "C"
