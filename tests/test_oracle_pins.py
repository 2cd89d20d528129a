"""Pins of the CPU oracle (oracle/), so that it can serve as the checker.

Sources of truth, in order:
  * known-answer tests held by the reference: make_banded_matrix
    (test/test_nls.f95:50-59), rgbmv (:166-184), make_laplacian_o5 (:102-108,
    whose golden encodes L1 + L2/2 where the code builds L1 + L2 -- used with
    that factor restored) and the dense construction of nls/nls_test.py:30-41;
  * the reference Fortran itself, compiled in place (oracle/_ref) and, where
    that is absent, the vectors it wrote (tests/golden/ref_fixtures.npz);
  * mathematics: plane-wave eigenvalues of the stencils, the RK4
    amplification polynomial, an independent ODE integrator on uniform
    states, steady states, the norm balance and the radial Laplacian of a
    Gaussian.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "ref_fixtures.npz")


def band_to_dense(op, n):
    m = op.shape[0]
    R = (m - 1) // 2
    A = np.zeros((n, n))
    for j in range(n):
        for p in range(m):
            i = p + j - R
            if 0 <= i < n:
                A[i, j] = op[p, j]
    return A


# ---------------------------------------------------------------- KATs ------

def test_make_banded_matrix_kat(oracle):
    mat = oracle.make_banded_matrix(7, 5, [1, 2, 3, 4, 5], dtype=np.float32)
    ans = np.array([[0, 0, 3, 4, 5], [0, 2, 3, 4, 5], [1, 2, 3, 4, 5], [1, 2, 3, 4, 5],
                    [1, 2, 3, 4, 5], [1, 2, 3, 4, 0], [1, 2, 3, 0, 0]], np.float32).T
    np.testing.assert_array_equal(mat, ans)


def test_rgbmv_identity_kat(oracle):
    op = oracle.make_banded_matrix(7, 5, [1, 2, 3, 4, 5], dtype=np.float32)
    z = np.array([[3, 4, 5, 0, 0, 0, 0], [2, 3, 4, 5, 0, 0, 0], [1, 2, 3, 4, 5, 0, 0],
                  [0, 1, 2, 3, 4, 5, 0], [0, 0, 1, 2, 3, 4, 5], [0, 0, 0, 1, 2, 3, 4],
                  [0, 0, 0, 0, 1, 2, 3]], np.float32)
    for i in range(7):
        e = np.zeros(7, np.float32)
        e[i] = 1.0
        y = oracle.gbmv(7, 2, 1.0, op, e, np.zeros(7, np.float32), dtype=np.float32)
        np.testing.assert_array_equal(y, z[i])


def test_make_laplacian_o5_kat_with_factor_restored(oracle):
    golden = np.array([
        [0.0, 0.0, -1666.66666667, -1250.0, -833.33333333, -694.44444444, -625.0],
        [0.0, 26666.66666667, 13333.33333333, 10000.0, 8888.88888889, 8333.33333333, 8000.0],
        [-25000.0, -12083.33333333, -12500.0, -12500.0, -12500.0, -12500.0, -12500.0],
        [0.0, 3333.33333333, 4444.44444444, 5000.0, 5333.33333333, 5555.55555555, 0.0],
        [0.0, -138.88888888, -208.33333333, -250.0, -277.77777777, 0.0, 0.0]])
    L1, L2 = oracle.laplacian_1d_parts(7, 5, 0.01)
    np.testing.assert_allclose(L1 + 0.5 * L2, golden, rtol=1e-8, atol=1e-6)
    # the code's operator is L1 + L2 (nls.f90:194): golden entries differ
    op = oracle.make_laplacian(7, 5, 0.01)
    assert op[2, 0] == pytest.approx(-50000.0) and op[2, 1] == pytest.approx(-25000.0)


def test_dense_construction_of_nls_test(oracle):
    n, h = 8, 0.1
    r = np.arange(0, h * n - 1e-10, h)
    D1 = sum(np.diag(np.full(n - abs(k), v), k) for v, k in zip((1, -8, 0, 8, -1), (-2, -1, 0, 1, 2)))
    D2 = sum(np.diag(np.full(n - abs(k), v), k) for v, k in zip((-1, 16, -30, 16, -1), (-2, -1, 0, 1, 2)))
    r[0] = 1
    D1[0, :] = 0
    D1[1, 1] += 1
    D2[0, :3] = [-60, 64, -4]
    D2[1, 1] += -1
    D = D2 / (12 * h ** 2) + np.diag(1.0 / r).dot(D1) / (12 * h)
    A = band_to_dense(oracle.make_laplacian(n, 5, h), n)
    np.testing.assert_allclose(A, D, rtol=1e-12, atol=1e-9)


# ------------------------------------------------- the reference itself ----

def _golden():
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixtures missing")
    return np.load(GOLDEN, allow_pickle=False)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_oracle_matches_reference_golden_vectors(oracle):
    g = _golden()
    keys = sorted(k[:-len("_meta")] for k in g.files if k.endswith("_meta"))
    assert len(keys) == 6
    for key in keys:
        dim, n, order, steps, dt, dx = g[key + "_meta"]
        dim, n, order, steps = int(dim), int(n), int(order), int(steps)
        P, u0, c = g[key + "_P"], g[key + "_u0"], g[key + "_coeffs"]
        o64, _ = oracle.rk4(dim, n, order, 0, dt, dx, steps, P, c, u0)
        o32, _ = oracle.rk4(dim, n, order, 0, np.float32(dt), np.float32(dx), steps, P, c, u0, dtype=np.float32)
        assert rel(o64, g[key + "_u_ref64"]) <= 1e-12, key
        assert rel(o32, g[key + "_u_ref32"]) <= 1e-5, key
    for order in (3, 5, 7):
        np.testing.assert_array_equal(oracle.make_laplacian(16, order, np.float32(0.1), dtype=np.float32),
                                      g["laplacian_n16_o%d_ref32" % order])
        np.testing.assert_array_equal(oracle.make_laplacian(16, order, 0.1), g["laplacian_n16_o%d_ref64" % order])
    for dim, n in ((1, 24), (2, 12)):
        key = "mu_d%d_n%d" % (dim, n)
        mu = oracle.chemical_potential(dim, 0.2, n, g[key + "_P"], g[key + "_coeffs"], g[key + "_u0"])
        assert abs(mu - complex(g[key + "_ref64"])) <= 1e-12 * max(1.0, abs(mu))
    assert tuple(g["version"]) == (0, 2, 0)


def test_oracle_matches_compiled_reference_live(oracle):
    if not oracle.RefLib.available(64):
        pytest.skip("oracle/_ref not built")
    R64, R32 = oracle.RefLib(64), oracle.RefLib(32)
    rng = np.random.default_rng(5)
    for dim, n, order in ((1, 48, 5), (1, 40, 7), (2, 24, 3), (2, 20, 5)):
        shape = (n,) if dim == 1 else (n, n)
        P = 1.0 + rng.random(shape)
        u0 = 0.3 + 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        c = np.zeros(23)
        c[2], c[3], c[4], c[5], c[11], c[12], c[13] = 0.4, 0.6, 1.2, 3.0, 0.8, 1.1, 0.9
        o, _ = oracle.rk4(dim, n, order, 0, 5e-4, 0.15, 300, P, c, u0)
        assert rel(o, R64.solve(dim, 5e-4, 0.15, n, order, 300, P, c, u0)) <= 1e-12
        o32, _ = oracle.rk4(dim, n, order, 0, np.float32(5e-4), np.float32(0.15), 300, P, c, u0, dtype=np.float32)
        assert rel(o32, R32.solve(dim, 5e-4, 0.15, n, order, 300, P, c, u0)) <= 1e-5
    h = R32.hamiltonian_1d(np.ones(24), c, 0.5 * np.ones(24), 5, 0.1)
    oh, _ = oracle.rhs(1, 24, 5, 0, np.float32(0.1), np.ones(24), c, 0.5 * np.ones(24), dtype=np.float32)
    np.testing.assert_allclose(oh, h, rtol=1e-5, atol=1e-5)


# ------------------------------------------------------------ mathematics --

def stencil_symbol(order, theta):
    """h^2 * eigenvalue of the 1-D second-derivative stencil on exp(i k x), theta = k h."""
    if order == 3:
        return -4.0 * np.sin(theta / 2) ** 2
    if order == 5:
        return (-2 * np.cos(2 * theta) + 32 * np.cos(theta) - 30) / 12.0
    return (4 * np.cos(3 * theta) - 54 * np.cos(2 * theta) + 540 * np.cos(theta) - 490) / 180.0


def linear_coeffs():
    c = np.zeros(23)
    c[12] = 1.0          # r = c12 P / (c13 + ...) = 0: only i L psi remains
    return c


@pytest.mark.parametrize("order", (3, 5, 7))
@pytest.mark.parametrize("dim", (1, 2))
def test_plane_wave_eigenvalue(oracle, dim, order):
    n, h = 32, 0.3
    kx, ky = 3, 5
    x = np.arange(n) * h
    L = n * h
    if dim == 1:
        psi = np.exp(2j * np.pi * kx * x / L)
        lam = stencil_symbol(order, 2 * np.pi * kx / n) / h ** 2
    else:
        psi = np.exp(2j * np.pi * (ky * x[:, None] + kx * x[None, :]) / L)
        lam = (stencil_symbol(order, 2 * np.pi * kx / n) + stencil_symbol(order, 2 * np.pi * ky / n)) / h ** 2
    P = np.zeros(psi.shape)
    F, _ = oracle.rhs(dim, n, order, 1, h, P, linear_coeffs(), psi)
    np.testing.assert_allclose(F, 1j * lam * psi, rtol=0, atol=1e-11 * abs(lam))


@pytest.mark.parametrize("order", (3, 5, 7))
def test_rk4_amplification_polynomial(oracle, order):
    n, h, dt, N = 64, 0.2, 2e-3, 250
    k = 7
    x = np.arange(n) * h
    psi0 = np.exp(2j * np.pi * k * x / (n * h))
    lam = stencil_symbol(order, 2 * np.pi * k / n) / h ** 2
    z = 1j * lam * dt
    G = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    psi, _ = oracle.rk4(1, n, order, 1, dt, h, N, np.zeros(n), linear_coeffs(), psi0)
    np.testing.assert_allclose(psi, G ** N * psi0, rtol=0, atol=1e-12)


def uniform_ode(c, P):
    def f(t, y):
        psi = y[0] + 1j * y[1]
        nn = y[2]
        s = abs(psi) ** 2
        dpsi = (c[2] * nn - c[3]) * psi - 1j * (c[4] * s + c[5] * nn) * psi
        dn = (c[11] * P - (c[12] + c[13] * s) * nn) / c[10]
        return [dpsi.real, dpsi.imag, dn]
    return f


def test_uniform_state_matches_independent_integrator(oracle):
    from scipy.integrate import solve_ivp
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[10], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 4.0, 1.0, 1.0, 1.5, 1.0
    n, P0, dt, steps = 8, 2.0, 1e-3, 1500
    psi0 = (0.3 + 0.1j) * np.ones((n, n))
    sol = solve_ivp(uniform_ode(c, P0), (0, dt * steps), [0.3, 0.1, 0.2], rtol=1e-13, atol=1e-14, method="DOP853")
    psi, nres = oracle.rk4(2, n, 5, 1, dt, 0.5, steps, P0 * np.ones((n, n)), c, psi0, 0.2 * np.ones((n, n)))
    yT = sol.y[:, -1]
    np.testing.assert_allclose(psi, (yT[0] + 1j * yT[1]) * np.ones((n, n)), rtol=1e-10)
    np.testing.assert_allclose(nres, yT[2] * np.ones((n, n)), rtol=1e-10)


def test_dynamic_and_adiabatic_steady_states(oracle):
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[10], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 4.0, 1.0, 1.0, 1.5, 1.0
    n, P0 = 4, 2.0
    psi, nres = oracle.rk4(1, n, 3, 1, 1e-2, 1.0, 20000, P0 * np.ones(n), c, 0.3 * np.ones(n), np.zeros(n))
    nstar = c[3] / c[2]
    sstar = (c[11] * P0 / nstar - c[12]) / c[13]
    np.testing.assert_allclose(nres, nstar, rtol=1e-6)
    np.testing.assert_allclose(abs(psi) ** 2, sstar, rtol=1e-6)
    c[10] = 0.0
    psi, _ = oracle.rk4(1, n, 3, 1, 1e-2, 1.0, 20000, P0 * np.ones(n), c, 0.3 * np.ones(n))
    np.testing.assert_allclose(abs(psi) ** 2, (c[2] * c[11] * P0 / c[3] - c[12]) / c[13], rtol=1e-6)


@pytest.mark.parametrize("bc", (0, 1))
def test_norm_balance(oracle, bc):
    """sum Re(conj(psi) F) = sum (c3 r - c4)|psi|^2 when L is symmetric (2-D)."""
    rng = np.random.default_rng(9)
    n = 20
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    P = 1.0 + rng.random((n, n))
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[11], c[12], c[13] = 0.7, 0.3, 1.1, 2.0, 1.3, 0.9, 1.7
    F, _ = oracle.rhs(2, n, 5, bc, 0.2, P, c, psi)
    s = abs(psi) ** 2
    r = c[11] * P / (c[12] + c[13] * s)
    lhs = np.sum((psi.conj() * F).real)
    assert lhs == pytest.approx(np.sum((c[2] * r - c[3]) * s), rel=1e-12)


@pytest.mark.parametrize("order", (3, 5, 7))
def test_radial_operator_on_gaussian(oracle, order):
    """d2/dr2 + (1/r) d/dr of exp(-r^2) is (4 r^2 - 4) exp(-r^2); the FD error
    shrinks at the stencil's order, including at r = 0 and the first rows."""
    errs = []
    for h in (0.05, 0.025):
        n = int(round(6.0 / h))
        r = np.arange(n) * h
        f = np.exp(-r ** 2)
        A = band_to_dense(oracle.make_laplacian(n, order, h), n)
        exact = (4 * r ** 2 - 4) * f
        errs.append(np.abs(A.dot(f) - exact)[: int(3.0 / h)].max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > order - 1.5, (errs, rate)
