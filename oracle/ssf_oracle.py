"""Split-step Fourier oracle -- TEST INFRASTRUCTURE ONLY (numpy, double).

The reference has no split-step scheme; this is the scheme as specified for
the build, written with numpy.fft so that the GPU's LDS FFTs and fused passes
are checked against a library transform.  One step (Strang splitting) is

    N(dt/2) -> L(dt) -> N(dt/2)

  L(dt): psi_hat(k) *= exp(-i |k|^2 dt), k the spectral wavenumbers of the
         periodic grid (exact flow of psi_t = i lap psi);
  N(t):  one classical RK4 step of size t of the pointwise system
           psi' = (c3 r - c4) psi - i (c5 |psi|^2 + c6 r) psi
           n'   = (c12 P - (c13 + c14 |psi|^2) n) / c11        (c11 != 0)
         with r = n (dynamic) or c12 P / (c13 + c14 |psi|^2) (c11 == 0) --
         the pointwise part of the reference's right-hand side
         (nls/nls.f90:570-581, :637-644).
Only tests/ import this module.
"""
import numpy as np


def _f(psi, n, P, c, dyn):
    s = (psi.conj() * psi).real
    r = n if dyn else c[11] * P / (c[12] + c[13] * s)
    dpsi = (c[2] * r - c[3]) * psi - 1j * (c[4] * s + c[5] * r) * psi
    dn = (c[11] * P - (c[12] + c[13] * s) * n) / c[10] if dyn else None
    return dpsi, dn


def pointwise_rk4(psi, n, P, c, tau, dyn):
    k1, m1 = _f(psi, n, P, c, dyn)
    k2, m2 = _f(psi + k1 * tau / 2, n + m1 * tau / 2 if dyn else None, P, c, dyn)
    k3, m3 = _f(psi + k2 * tau / 2, n + m2 * tau / 2 if dyn else None, P, c, dyn)
    k4, m4 = _f(psi + k3 * tau, n + m3 * tau if dyn else None, P, c, dyn)
    psi = psi + (k1 + 2 * k2 + 2 * k3 + k4) * tau / 6
    if dyn:
        n = n + (m1 + 2 * m2 + 2 * m3 + m4) * tau / 6
    return psi, n


def wavenumbers(n, dx):
    return 2.0 * np.pi * np.fft.fftfreq(n, d=dx)


def ssf(dim, n, dx, dt, steps, P, coeffs, psi0, n0=None):
    """Advance (psi, n) by `steps` Strang steps on the periodic grid."""
    c = np.asarray(coeffs, np.float64)
    dyn = c[10] != 0.0
    psi = np.asarray(psi0, np.complex128).copy()
    P = np.asarray(P, np.float64)
    nn = (np.zeros(psi.shape) if n0 is None else np.asarray(n0, np.float64).copy()) if dyn else None
    k = wavenumbers(n, dx)
    if dim == 1:
        phase = np.exp(-1j * k ** 2 * dt)
        fwd, inv = np.fft.fft, np.fft.ifft
    else:
        phase = np.exp(-1j * (k[:, None] ** 2 + k[None, :] ** 2) * dt)
        fwd, inv = np.fft.fft2, np.fft.ifft2
    for _ in range(steps):
        psi, nn = pointwise_rk4(psi, nn, P, c, dt / 2, dyn)
        psi = inv(phase * fwd(psi))
        psi, nn = pointwise_rk4(psi, nn, P, c, dt / 2, dyn)
    return psi, nn
