"""Write tests/golden/ref_layout_model2d.mat and ref_layout_model2d.npz.

The .mat file holds a stored 2-D solution in the reference's own layout: the
keys AbstractModel.store writes (nls/model.py:257-285: model as
str(type(self)), date, dim, dimlesses, init_solution, num_iters, num_nodes,
order, originals, pumping, spatial_step, time_step) and the keys
Solution.store appends (:500-517: elapsed_time, solution, version), nothing
else.  Its solution is the reference Fortran's own single-precision result
(oracle/_ref, solve_nls_2d) after N steps from init_solution; the .npz holds
the reference's uninterrupted 2N-step result from the same start, so a test
can resume the stored solution for N steps and compare.

Run in the build container after `make -C oracle ref`:
    python tests/golden/make_mat_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

N_STEPS = 150
ORIGINALS = {"R": 0.0242057488654, "gamma": 0.0242057488654, "g": 0.00162178517398,
             "tilde_g": 0.0169440242057, "gamma_R": 0.242057488654}


def coefficients(o):
    """The reference's physical -> dimensionless map (nls/model.py:138-160)."""
    c = np.zeros(23)
    phi0 = np.sqrt(o["gamma"] / (2.0 * o["g"]))
    n0 = 2.0 / (o["R"] * phi0)
    c[0] = c[1] = c[2] = c[3] = c[4] = 1.0
    c[5] = 4.0 * o["tilde_g"] / o["R"]
    c[11] = 1.0 / (n0 * o["gamma_R"])
    c[12] = 1.0
    c[13] = o["R"] * phi0 ** 2 / o["gamma_R"]
    return c


def main():
    from scipy.io import savemat
    n, dx, dt, order = 16, 0.5, 1e-3, 3
    x = np.linspace(-n * dx / 2.0, n * dx / 2.0, n)        # the reference's 2-D grid (model.py:220-232)
    X, Y = np.meshgrid(x, x)
    pumping = 20.0 * np.exp(-(X ** 2 + Y ** 2) / (2.0 * 1.5 ** 2))
    rng = np.random.default_rng(2024)
    u0 = 0.1 * np.ones((n, n)) + 1e-3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    c = coefficients(ORIGINALS)
    R = O.RefLib(32)
    u_n = R.solve(2, dt, dx, n, order, N_STEPS, pumping, c, u0)
    u_2n = R.solve(2, dt, dx, n, order, 2 * N_STEPS, pumping, c, u0)
    mat = {"model": "<class 'nls.model.Model2D'>", "date": "2026-01-01T00:00:00", "dim": 2, "dimlesses": c,
           "init_solution": u0, "num_iters": N_STEPS, "num_nodes": n, "order": order, "originals": ORIGINALS,
           "pumping": pumping, "spatial_step": dx, "time_step": dt,
           "elapsed_time": 0.01, "solution": u_n, "version": np.array([0, 2, 0])}
    savemat(os.path.join(HERE, "ref_layout_model2d.mat"), mat)
    np.savez_compressed(os.path.join(HERE, "ref_layout_model2d.npz"), u_n=u_n, u_2n=u_2n, steps=N_STEPS)
    print("wrote ref_layout_model2d.mat/.npz")


if __name__ == "__main__":
    main()
