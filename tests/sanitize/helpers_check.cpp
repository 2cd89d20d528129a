// helpers_check.cpp -- drives the host code of libnls_amd's operator builders
// and BLAS-2 drop-ins (nls_amd/csrc/nls_helpers.hip) and its status reporting
// (nls_status.hip) on every order, the smallest sizes and the argument errors,
// so that a build with AddressSanitizer and UndefinedBehaviorSanitizer sees
// all of their indexing.  No GPU call is made.  Test infrastructure only.
// Exit status 0 when every call behaved and every result is finite.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nls_amd.h"

static int bad = 0;

template <typename V> static void finite(const char *what, const V &a) {
    for (size_t i = 0; i < a.size(); ++i)
        if (!std::isfinite((double)a[i])) { std::fprintf(stderr, "%s: non-finite at %zu\n", what, i); bad = 1; return; }
}

int main() {
    for (int m = 3; m <= 7; m += 2) {
        for (int n = m; n <= m + 9; ++n) {
            std::vector<double> op((size_t)m * n);
            if (nls_make_laplacian(n, m, 0.1, op.data())) bad = 1;
            finite("make_laplacian", op);
            std::vector<float> row(m), mat((size_t)m * n);
            for (int k = 0; k < m; ++k) row[k] = (float)(k + 1);
            if (nls_make_banded_matrix(n, m, row.data(), mat.data())) bad = 1;
            if (nls_clear_first_row_of_derivative(n, m, mat.data())) bad = 1;
            if (nls_divide_derivative_on_radius(n, m, 0.1f, mat.data())) bad = 1;
            finite("derivative band", mat);
            std::vector<float> blocks((size_t)n * (2 * m - 1)), x((size_t)n * n, 1.0f), y((size_t)n * n, 0.0f);
            std::vector<int> orders(m);
            if (nls_make_laplacian_2d(n, m, 0.1f, blocks.data(), orders.data())) bad = 1;
            if (nls_rbbmv(x.data(), y.data(), 1.0f, blocks.data(), orders.data(), m, n)) bad = 1;
            finite("rbbmv", y);
            std::vector<float> band((size_t)m * n), u(n, 0.0f), xv(n, 0.5f);
            for (size_t i = 0; i < band.size(); ++i) band[i] = (float)op[i];
            if (nls_rgbmv(xv.data(), u.data(), -1.0f, band.data(), (m - 1) / 2, n)) bad = 1;
            finite("rgbmv", u);
        }
    }
    std::vector<float> coeffs(23, 0.0f), P(64, 2.0f), s(64, 0.25f), r(64);
    coeffs[11] = 1.0f; coeffs[12] = 1.5f; coeffs[13] = 1.0f;
    if (nls_revervoir(P.data(), coeffs.data(), s.data(), r.data(), 64)) bad = 1;
    if (nls_revervoir_2d(P.data(), coeffs.data(), s.data(), r.data(), 8)) bad = 1;
    finite("revervoir", r);
    // argument errors come back as codes with a message, nothing is written
    std::vector<double> op(3 * 2);
    if (nls_make_laplacian(2, 3, 0.1, op.data()) != NLS_EINVAL) bad = 1;
    if (nls_make_laplacian(16, 4, 0.1, op.data()) != NLS_EINVAL) bad = 1;
    if (std::strlen(nls_last_error()) == 0) bad = 1;
    if (nls_make_banded_matrix(4, 2, nullptr, nullptr) != NLS_EINVAL) bad = 1;
    int a = 0, b = 0, c = 0;
    if (nls_version(&a, &b, &c) || a != 0 || b != 2 || c != 0) bad = 1;
    if (bad) std::fprintf(stderr, "helpers_check: FAILED\n");
    else std::printf("helpers_check: ok\n");
    return bad;
}
