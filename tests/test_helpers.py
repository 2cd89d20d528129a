"""The reference module's operator builders and BLAS-2 wrappers through the
C ABI (host code, no GPU): make_banded_matrix, make_laplacian_2d, rgbmv,
rbbmv, revervoir.  Pinned by the reference's own known-answer tests
(test/test_nls.f95:50-59, :166-184) and, for the 2-D operator, by the
oracle's independent 2-D right-hand side."""
import numpy as np
import pytest

from nls_amd import native

nls = native.nls


def test_make_banded_matrix_kat():
    mat = nls.make_banded_matrix(7, [1, 2, 3, 4, 5])
    ans = np.array([[0, 0, 3, 4, 5], [0, 2, 3, 4, 5], [1, 2, 3, 4, 5], [1, 2, 3, 4, 5],
                    [1, 2, 3, 4, 5], [1, 2, 3, 4, 0], [1, 2, 3, 0, 0]], np.float32).T
    assert mat.shape == (5, 7) and mat.dtype == np.float32
    np.testing.assert_array_equal(mat, ans)


@pytest.mark.parametrize("m", (1, 3, 7))
def test_make_banded_matrix_matches_oracle(oracle, m):
    row = np.arange(1, m + 1, dtype=np.float32) * 0.5
    np.testing.assert_array_equal(nls.make_banded_matrix(11, row), oracle.make_banded_matrix(11, m, row, np.float32))


def test_rgbmv_identity_kat():
    op = nls.make_banded_matrix(7, [1, 2, 3, 4, 5])
    z = np.array([[3, 4, 5, 0, 0, 0, 0], [2, 3, 4, 5, 0, 0, 0], [1, 2, 3, 4, 5, 0, 0],
                  [0, 1, 2, 3, 4, 5, 0], [0, 0, 1, 2, 3, 4, 5], [0, 0, 0, 1, 2, 3, 4],
                  [0, 0, 0, 0, 1, 2, 3]], np.float32)
    for i in range(7):
        e = np.zeros(7, np.float32)
        e[i] = 1.0
        y = np.zeros(7, np.float32)
        assert nls.rgbmv(e, y, 1.0, op, 2) is None          # intent(inout): y is updated in place
        np.testing.assert_array_equal(y, z[i])


def test_rgbmv_matches_oracle_on_a_random_band(oracle):
    rng = np.random.default_rng(3)
    n, klu = 40, 3
    op = rng.standard_normal((2 * klu + 1, n)).astype(np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    y0 = rng.standard_normal(n).astype(np.float32)
    y = y0.copy()
    nls.rgbmv(x, y, -0.75, op, klu)
    ref = oracle.gbmv(n, klu, -0.75, op, x, y0)
    np.testing.assert_allclose(y, ref, rtol=1e-5, atol=1e-5)
    with pytest.raises(ValueError):                          # intent(inout) needs a float32 array
        nls.rgbmv(x, list(y0), 1.0, op, klu)


@pytest.mark.parametrize("m", (3, 5, 7))
def test_make_laplacian_2d_and_rbbmv_apply_the_2d_operator(oracle, m):
    """rbbmv with make_laplacian_2d's blocks is the 2-D Laplacian with zero
    outside the grid: the imaginary part of the oracle's right-hand side of
    a real field with every other term switched off."""
    n, h = 23, 0.3
    blocks, orders = nls.make_laplacian_2d(n, m, h)
    assert blocks.shape == (n, 2 * m - 1) and list(orders) == [0] * ((m - 1) // 2) + [(m - 1) // 2] + [0] * ((m - 1) // 2)
    x = np.random.default_rng(m).standard_normal((n, n))
    y = np.zeros(n * n, np.float32)
    nls.rbbmv(x.ravel(), y, 1.0, blocks, orders)
    c = np.zeros(23)
    c[12] = 1.0                                          # r = c12 P / c13 = 0
    F, _ = oracle.rhs(2, n, m, 0, h, np.zeros((n, n)), c, x)
    np.testing.assert_allclose(y.reshape(n, n), F.imag, rtol=1e-4, atol=1e-4 * np.abs(F.imag).max())
    # the order-specific names are the same procedure
    y5 = np.zeros(n * n, np.float32)
    getattr(nls, "rbbmv_o%d" % m)(x.ravel(), y5, 1.0, blocks, orders)
    np.testing.assert_array_equal(y5, y)
    b2, o2 = getattr(nls, "make_laplacian_2d_o%d" % m)(n, h)
    np.testing.assert_array_equal(b2, blocks)


def test_revervoir_is_the_algebraic_reservoir_of_the_rhs(oracle):
    """On a uniform periodic field with only c3 = 1 among the psi terms, the
    right-hand side is r psi: r from the oracle against revervoir."""
    n = 16
    c = np.zeros(23)
    c[2], c[11], c[12], c[13] = 1.0, 0.8, 1.5, 0.7
    for amp in (0.3, 1.1):
        psi = np.full(n, amp + 0.2j)
        P = np.full(n, 2.5)
        F, _ = oracle.rhs(1, n, 5, 1, 0.1, P, c, psi)
        r = nls.revervoir(P, c, np.abs(psi) ** 2)
        np.testing.assert_allclose(r, (F / psi).real, rtol=1e-6)
        r2 = nls.revervoir_2d(np.full((4, 4), 2.5), c, np.full((4, 4), abs(amp + 0.2j) ** 2))
        np.testing.assert_allclose(r2, r[0], rtol=1e-6)


def test_operator_arguments_checked_on_the_host():
    n = 16
    c = np.zeros(23)
    u = np.ones(n, np.complex64)
    with pytest.raises(ValueError):                       # a band of width 9 is beyond the kernels
        nls.hamiltonian(np.ones(n), c, u, np.zeros((9, n), np.float32), 4)
    blocks, orders = nls.make_laplacian_2d(n, 5, 0.2)
    bad = blocks.copy()
    bad[:, 0] *= 2.0                                      # far-left block differs from far-right
    with pytest.raises(ValueError):
        nls.runge_kutta_2d(1e-3, 0.0, np.ones((n, n)), bad, orders, 1, np.ones((n, n)), c)
    with pytest.raises(ValueError):
        nls.hamiltonian_2d(np.ones((n, n)), c, np.ones((n, n)), blocks, np.array([0, 1, 1, 1, 0]))


@pytest.mark.parametrize("m", (3, 5, 7))
def test_derivative_band_steps(m):
    """clear_first_row_of_derivative and divide_derivative_on_radius
    (nls.f90:93-130) in place on a band of ones, against the closed form:
    band entry (i, j) lies on matrix row k = i + j - (m+3)/2; row 0 becomes
    0, rows k > 0 become 1/(k h) (one single-precision division), entries
    outside the matrix (k < 0) are left alone."""
    from nls_amd.native import nls
    n, h = 12, np.float32(0.3)
    L1 = np.ones((m, n), np.float32, order="F")
    out = nls.clear_first_row_of_derivative(n, m, L1)
    assert out is L1
    nls.divide_derivative_on_radius(n, m, float(h), L1)
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            k = i + j - (m + 3) // 2
            want = np.float32(1.0) if k < 0 else (np.float32(0.0) if k == 0 else np.float32(1.0) / (np.float32(k) * h))
            assert L1[i - 1, j - 1] == want, (i, j, k)
    # without a band: f2py's fresh zero-filled intent(out) array
    assert not nls.clear_first_row_of_derivative(n, m).any()
