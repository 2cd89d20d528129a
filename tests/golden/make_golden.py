"""Write tests/golden/ref_fixtures.npz: input/output vectors of the reference
Fortran core itself (oracle/_ref, compiled in place from nls/nls.f90 by
oracle/ref.mk), so the oracle can be checked against the reference on
machines where the reference tree is absent.

Run in the build container after `make -C oracle ref`:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def case_inputs(dim, n, seed):
    rng = np.random.default_rng(seed)
    if dim == 1:
        x = np.arange(n) * 0.2
        P = 2.5 * np.exp(-x ** 2 / 6.0) + 0.5 * np.exp(-(x - 3.0) ** 2)
        u0 = 0.2 * np.exp(-x ** 2 / 8.0) + 1e-2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    else:
        x = np.linspace(-1.6, 1.6, n)
        r2 = x[:, None] ** 2 + x[None, :] ** 2
        P = 2.5 * np.exp(-r2 / 2.0)
        u0 = 0.2 * np.exp(-r2 / 3.0) + 1e-2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 4.0, 1.0, 1.5, 1.0
    return P, u0, c


def main():
    out = {}
    refs = {b: O.RefLib(b) for b in (32, 64)}
    cases = [(1, 32, 3), (1, 32, 5), (1, 32, 7), (2, 16, 3), (2, 16, 5), (2, 7, 7)]
    dt, dx, steps = 1e-3, 0.2, 200
    for k, (dim, n, order) in enumerate(cases):
        P, u0, c = case_inputs(dim, n, 100 + k)
        key = "solve_d%d_n%d_o%d" % (dim, n, order)
        out[key + "_P"] = P
        out[key + "_u0"] = u0
        out[key + "_coeffs"] = c
        out[key + "_meta"] = np.array([dim, n, order, steps, dt, dx])
        for b, R in refs.items():
            out[key + "_u_ref%d" % b] = R.solve(dim, dt, dx, n, order, steps, P, c, u0)
    for order in (3, 5, 7):
        for b, R in refs.items():
            out["laplacian_n16_o%d_ref%d" % (order, b)] = R.make_laplacian(16, order, 0.1)
    for dim, n in ((1, 24), (2, 12)):
        P, u0, c = case_inputs(dim, n, 7)
        key = "mu_d%d_n%d" % (dim, n)
        out[key + "_P"], out[key + "_u0"], out[key + "_coeffs"] = P, u0, c
        for b, R in refs.items():
            out[key + "_ref%d" % b] = np.array(R.chemical_potential(dim, 0.2, n, P, c, u0))
    out["version"] = np.array(refs[32].version())
    np.savez_compressed(os.path.join(HERE, "ref_fixtures.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
