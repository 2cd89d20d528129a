"""The reference's hamiltonian / runge_kutta family (nls.f90:621-650,
:705-734, :841-901) through the C ABI, on the GPU, with the operator given
as the reference's arrays.

Pins: the oracle's FP64 right-hand side and RK4 with the same operators
(built by make_laplacian / make_laplacian_2d); the drop-in solvers, which
build the same operators themselves; and, for an arbitrary (non-symmetric)
band, the closed form of RK4 on a linear system, u_N = G(i dt A)^N u0 with
G(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 and A the dense matrix of the band --
a transposed or shifted band fails it.
"""
import numpy as np
import pytest

from nls_amd import configs, native

pytestmark = pytest.mark.gpu
nls = native.nls


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def coeffs():
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 2.0, 0.5, 1.0, 2.0
    return c


def dense_of_band(op, klu):
    m, n = op.shape
    A = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - klu), min(n, j + klu + 1)):
            A[i, j] = op[klu + i - j, j]
    return A


@pytest.mark.parametrize("order", (3, 5, 7))
def test_hamiltonian_matches_oracle(gpu, oracle, order):
    n, h = 96, 0.1
    x = np.arange(n) * h
    P = 3.0 * np.exp(-x ** 2 / 13.6)
    u = (0.3 * np.exp(-x ** 2 / 8.0) + 0.05j * np.cos(x)).astype(np.complex64)
    op = nls.make_laplacian(n, order, h)
    v = nls.hamiltonian(P, coeffs(), u, op, (order - 1) // 2)
    F, _ = oracle.rhs(1, n, order, 0, h, P, coeffs(), u)
    assert v.dtype == np.complex64 and v.shape == (n,)
    assert rel(v, F) <= 1e-5, rel(v, F)


def test_hamiltonian_of_an_arbitrary_band(gpu):
    """Linear terms only: v = i A u for a random non-symmetric band."""
    rng = np.random.default_rng(4)
    n, klu = 50, 2
    op = rng.standard_normal((2 * klu + 1, n)).astype(np.float32)
    u = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    c = np.zeros(23)
    c[12] = 1.0
    v = nls.hamiltonian(np.zeros(n), c, u, op, klu)
    A = dense_of_band(op.astype(np.float64), klu)
    assert rel(v, 1j * (A @ u.astype(np.complex128))) <= 1e-5


@pytest.mark.parametrize("order", (3, 5, 7))
def test_runge_kutta_matches_oracle_and_solve_nls(gpu, oracle, order):
    n, h, dt, steps = 128, 0.1, 5e-4, 400
    x = np.arange(n) * h
    P = 3.0 * np.exp(-x ** 2 / 13.6)
    u0 = 0.1 * np.ones(n)
    op = nls.make_laplacian(n, order, h)
    u = nls.runge_kutta(dt, 0.0, u0, op, steps, P, coeffs())
    assert u.dtype == np.complex64 and u.shape == (n,)
    o64, _ = oracle.rk4(1, n, order, 0, dt, h, steps, P, coeffs(), u0)
    assert rel(u, o64) <= 1e-4, rel(u, o64)
    assert rel(u, nls.solve_nls(dt, h, order, steps, P, coeffs(), u0)) <= 1e-5


def test_runge_kutta_on_an_arbitrary_band_is_the_rk4_polynomial(gpu):
    rng = np.random.default_rng(6)
    n, order, dt, steps = 40, 5, 2e-3, 50
    klu = (order - 1) // 2
    op = (rng.standard_normal((order, n)) * 3.0).astype(np.float32)
    u0 = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    c = np.zeros(23)
    c[12] = 1.0                                     # every term but the operator is zero
    u = nls.runge_kutta(dt, 0.0, u0, op, steps, np.zeros(n), c)
    Z = 1j * dt * dense_of_band(op.astype(np.float64), klu)
    G = np.eye(n) + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24
    exact = np.linalg.matrix_power(G, steps) @ u0.astype(np.complex128)
    assert rel(u, exact) <= 1e-5, rel(u, exact)


@pytest.mark.parametrize("order", (3, 5, 7))
def test_2d_family_matches_oracle_and_solve_nls_2d(gpu, oracle, order):
    n, h, dt, steps = 40, 0.2, 1e-3, 200
    X, Y, _ = configs.grid_2d(n, n * h)
    P = 2.5 * np.exp(-(X ** 2 + Y ** 2) / 10.0)
    u0 = (0.1 + 0.02 * np.exp(-(X ** 2 + Y ** 2) / 4.0)).astype(np.complex64)
    c = configs.model_coeffs(dynamic=False)
    blocks, orders = nls.make_laplacian_2d(n, order, h)
    v = nls.hamiltonian_2d(P, c, u0, blocks, orders)
    F, _ = oracle.rhs(2, n, order, 0, h, P, c, u0)
    assert rel(v, F) <= 1e-5, rel(v, F)
    u = nls.runge_kutta_2d(dt, 0.0, u0, blocks, orders, steps, P, c)
    o64, _ = oracle.rk4(2, n, order, 0, dt, h, steps, P, c, u0)
    assert rel(u, o64) <= 1e-4, rel(u, o64)
    assert rel(u, nls.solve_nls_2d(dt, h, order, steps, P, c, u0)) <= 1e-5


def test_dropins_ignore_c11_like_the_reference(gpu):
    """The reference never reads coeffs(11) (nls.f90:570-581): the drop-in
    solve gives the same result whatever it holds."""
    n, h, dt, steps = 64, 0.1, 1e-3, 100
    x = np.arange(n) * h
    P = 3.0 * np.exp(-x ** 2 / 13.6)
    c0 = coeffs()
    c1 = c0.copy()
    c1[10] = 2.0
    a = nls.solve_nls(dt, h, 5, steps, P, c0, 0.1 * np.ones(n))
    b = nls.solve_nls(dt, h, 5, steps, P, c1, 0.1 * np.ones(n))
    np.testing.assert_array_equal(a, b)
