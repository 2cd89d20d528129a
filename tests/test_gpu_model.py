"""End to end through the reference's API: Problem().model(...).solve() with
the reference's own example and benchmark settings (examples/solve1d.py:21-39,
benchmark/benchmark.py:21-38, shortened), checked against the oracle."""
import numpy as np
import pytest

from nls_amd import configs
from nls_amd import model as M
from nls_amd import pumping as PU

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_example_solve1d(gpu, oracle):
    m = M.Problem().model(model="1d", dx=1.0e-1, dt=1.0e-3, t0=0.0, u0=0.1, order=5, num_nodes=400,
                          num_iters=3000, pumping=PU.GaussianRingPumping1D(power=20.0, radius=10.0, variation=3.14))
    sol = m.solve()
    u = sol.getSolution()
    assert u.dtype == np.complex64 and u.shape == (400,)
    o64, _ = oracle.rk4(1, 400, 5, 0, 1e-3, 0.1, 3000, m.getPumping(), m.getCoefficients(), m.getInitialSolution())
    assert rel(u, o64) <= 1e-4
    assert sol.getElapsedTime() > 0.0
    mu = m.getChemicalPotential(sol)
    omu = oracle.chemical_potential(1, 0.1, 400, m.getPumping(), m.getCoefficients(), u.astype(np.complex128))
    assert abs(mu - omu) <= 1e-4 * max(1.0, abs(omu))
    assert np.isfinite(sol.getDampingIntegral())


def test_benchmark_model_2d(gpu, oracle):
    m = M.Problem().model(model="2d", dx=2.0e-1, dt=2.0e-3, t0=0.0, u0=0.1, order=5, num_nodes=64,
                          num_iters=1000, pumping=PU.GaussianPumping2D(power=15.0, variation=3.14))
    sol = m.solve()
    o64, _ = oracle.rk4(2, 64, 5, 0, 2e-3, 0.2, 1000, m.getPumping(), m.getCoefficients(), m.getInitialSolution())
    assert rel(sol.getSolution(), o64) <= 1e-4
    mu = m.getChemicalPotential(sol)
    omu = oracle.chemical_potential(2, 0.2, 64, m.getPumping(), m.getCoefficients(),
                                    sol.getSolution().astype(np.complex128))
    assert abs(mu - omu) <= 1e-4 * max(1.0, abs(omu))


def test_extended_model_f64_periodic_dynamic(gpu, oracle):
    inst = configs.line256()
    m = M.Problem().model(model="1d", dx=inst.dx, dt=inst.dt, order=5, num_nodes=inst.n, num_iters=inst.steps,
                          u0=inst.psi0, pumping=PU.GridPumping(inst.P), coeffs=inst.coeffs, bc="periodic",
                          dtype="f64", n0=inst.n0)
    sol = m.solve()
    opsi, onres = oracle.rk4(1, inst.n, 5, 1, inst.dt, inst.dx, inst.steps, inst.P, inst.coeffs, inst.psi0, inst.n0)
    assert rel(sol.getSolution(), opsi) <= 1e-10
    assert rel(sol.getReservoir(), onres) <= 1e-10
