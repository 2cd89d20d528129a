"""Split-step Fourier scheme.

CPU: pins of the numpy oracle (oracle/ssf_oracle.py) -- the free
Schrodinger flow of a Gaussian (closed form; fixes the transform conventions
and the sign of the phase), uniform states against an independent ODE
integrator, norm conservation of the unitary case and second-order
convergence of the Strang splitting.
GPU: the HIP row/column passes and the 1-D line kernel against the oracle
(FP64 relative L2 <= 1e-10, FP32 <= 1e-4), through the C ABI.
"""
import numpy as np
import pytest

from oracle import ssf_oracle as S

TOL64, TOL32 = 1e-10, 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def free_coeffs():
    c = np.zeros(23)
    c[12] = 1.0          # r = 0: the pointwise flow is the identity
    return c


def gaussian_free(x, s, t):
    """psi_t = i psi_xx from exp(-x^2/(2 s^2)): (s/s_t) exp(-x^2/(2 s_t^2)), s_t^2 = s^2 + 2 i t."""
    st2 = s * s + 2j * t
    return np.sqrt(s * s / st2) * np.exp(-x ** 2 / (2 * st2))


@pytest.mark.parametrize("dim", (1, 2))
def test_oracle_free_flow_closed_form(dim):
    n, L, s, dt, steps = 256, 60.0, 1.5, 0.01, 50
    x = -L / 2 + L * np.arange(n) / n
    if dim == 1:
        psi0 = gaussian_free(x, s, 0.0)
        exact = gaussian_free(x, s, dt * steps)
    else:
        psi0 = gaussian_free(x[:, None], s, 0.0) * gaussian_free(x[None, :], s, 0.0)
        exact = gaussian_free(x[:, None], s, dt * steps) * gaussian_free(x[None, :], s, dt * steps)
    psi, _ = S.ssf(dim, n, L / n, dt, steps, np.zeros(psi0.shape), free_coeffs(), psi0)
    assert rel(psi, exact) <= 1e-12


def test_oracle_uniform_state_matches_independent_integrator():
    from scipy.integrate import solve_ivp
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[10], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 4.0, 1.0, 1.0, 1.5, 1.0
    P0, dt, steps, n = 2.0, 1e-3, 800, 16

    def f(t, y):
        psi = y[0] + 1j * y[1]
        s = abs(psi) ** 2
        d = (c[2] * y[2] - c[3]) * psi - 1j * (c[4] * s + c[5] * y[2]) * psi
        return [d.real, d.imag, (c[11] * P0 - (c[12] + c[13] * s) * y[2]) / c[10]]
    sol = solve_ivp(f, (0, dt * steps), [0.3, 0.1, 0.2], rtol=1e-13, atol=1e-14, method="DOP853")
    psi, nn = S.ssf(2, n, 0.5, dt, steps, P0 * np.ones((n, n)), c, (0.3 + 0.1j) * np.ones((n, n)),
                    0.2 * np.ones((n, n)))
    yT = sol.y[:, -1]
    np.testing.assert_allclose(psi, yT[0] + 1j * yT[1], rtol=1e-10)
    np.testing.assert_allclose(nn, yT[2], rtol=1e-10)


def test_oracle_norm_conserved_when_unitary():
    c = np.zeros(23)
    c[4], c[5], c[11], c[12], c[13] = 1.0, 2.0, 1.0, 1.0, 0.5     # no gain or loss: c3 = c4 = 0
    n, L = 128, 20.0
    x = -L / 2 + L * np.arange(n) / n
    rng = np.random.default_rng(3)
    psi0 = np.exp(-x ** 2) * (1 + 0.1 * rng.standard_normal(n)) + 0j
    P = 1.0 + np.exp(-x ** 2)
    psi, _ = S.ssf(1, n, L / n, 1e-3, 200, P, c, psi0)
    assert abs(np.sum(abs(psi) ** 2) / np.sum(abs(psi0) ** 2) - 1.0) <= 1e-11


def test_oracle_strang_is_second_order():
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 4.0, 1.0, 1.5, 1.0
    n, L, T = 64, 16.0, 0.4
    x = -L / 2 + L * np.arange(n) / n
    psi0 = 0.5 * np.exp(-x ** 2 / 4) + 0.2j * np.exp(-(x - 1) ** 2)
    P = 3.0 * np.exp(-x ** 2 / 8)
    ref, _ = S.ssf(1, n, L / n, T / 640, 640, P, c, psi0)
    errs = [rel(S.ssf(1, n, L / n, T / k, k, P, c, psi0)[0], ref) for k in (20, 40)]
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


# ------------------------------------------------------------------ GPU ----

def ssf_case(dim, n, dyn, seed=4):
    from nls_amd import configs
    rng = np.random.default_rng(seed)
    if dim == 1:
        x, h = configs.grid_1d(n, 0.1 * n)
        r2 = x ** 2
        shape = (n,)
    else:
        X, Y, h = configs.grid_2d(n, 0.125 * n)
        r2 = X ** 2 + Y ** 2
        shape = (n, n)
    psi0 = 0.3 * np.exp(-r2 / 6.0) + configs.noise(rng, shape, 1e-2)
    P = 2.5 * np.exp(-r2 / 10.0)
    c = configs.model_coeffs(dynamic=dyn)
    n0 = 0.1 * np.exp(-r2 / 5.0) if dyn else None
    return h, P, psi0, n0, c


# every FFT length class: radix-8 stages with a radix-2 or radix-4 tail or
# none (double precision), radix-4 with or without a radix-2 tail (single),
# several rows / columns per workgroup (n < 2048) and one (n = 2048)
CASES = [(1, 16), (1, 64), (1, 512), (1, 2048), (1, 4096),
         (2, 16), (2, 32), (2, 64), (2, 128), (2, 256), (2, 512), (2, 1024)]


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n", CASES)
@pytest.mark.parametrize("dyn", (False, True))
def test_gpu_ssf_f64_matches_oracle(gpu, dim, n, dyn):
    from nls_amd import native
    h, P, psi0, n0, c = ssf_case(dim, n, dyn)
    steps = 40 if n <= 512 else 10
    dt = 2e-3
    desc = native.ProblemDesc.make(dim, n, 5, dt, h, c, bc="periodic", dtype="f64", scheme="ssf")
    psi, nres = native.solve_host(desc, steps, P, psi0, n0)
    opsi, onres = S.ssf(dim, n, h, dt, steps, P, c, psi0, n0)
    assert rel(psi, opsi) <= TOL64, rel(psi, opsi)
    if dyn:
        assert rel(nres, onres) <= TOL64


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n", [(1, 32), (1, 1024), (1, 4096), (2, 16), (2, 64), (2, 256), (2, 2048)])
def test_gpu_ssf_f32_within_tolerance(gpu, dim, n):
    from nls_amd import native
    h, P, psi0, n0, c = ssf_case(dim, n, True)
    steps = 20 if n < 2048 else 4
    desc = native.ProblemDesc.make(dim, n, 5, 2e-3, h, c, bc="periodic", dtype="f32", scheme="ssf")
    psi, nres = native.solve_host(desc, steps, P, psi0, n0)
    opsi, onres = S.ssf(dim, n, h, 2e-3, steps, P, c, psi0, n0)
    assert rel(psi, opsi) <= TOL32
    assert rel(nres, onres) <= TOL32


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,dyn", [("f64", True), ("f64", False), ("f32", False)])
def test_gpu_ssf_2048(gpu, dtype, dyn):
    """2048-point lines take their own row and column kernels (first and last
    FFT stages against HBM, the turn between transforms in registers): every
    storage type and reservoir mode against the oracle."""
    from nls_amd import native
    h, P, psi0, n0, c = ssf_case(2, 2048, dyn)
    desc = native.ProblemDesc.make(2, 2048, 5, 2e-3, h, c, bc="periodic", dtype=dtype, scheme="ssf")
    psi, nres = native.solve_host(desc, 3, P, psi0, n0)
    opsi, onres = S.ssf(2, 2048, h, 2e-3, 3, P, c, psi0, n0)
    tol = TOL64 if dtype == "f64" else TOL32
    assert rel(psi, opsi) <= tol
    if dyn:
        assert rel(nres, onres) <= tol


def test_ssf_argument_errors():
    from nls_amd import native
    c = np.zeros(23)
    bad = native.ProblemDesc.make(2, 96, 5, 1e-3, 0.1, c, bc="periodic", dtype="f64", scheme="ssf")
    with pytest.raises(ValueError):
        native.solve_host(bad, 1, np.zeros((96, 96)), np.zeros((96, 96)))
    big = native.ProblemDesc.make(2, 4096, 5, 1e-3, 0.1, c, bc="periodic", dtype="f32", scheme="ssf")
    with pytest.raises(ValueError):
        native.solve_host(big, 0, np.zeros((4096, 4096), np.float32), np.zeros((4096, 4096), np.complex64))
