"""The Python API of the reference (nls/model.py, nls/solver.py,
nls/pumping.py) as kept by nls_amd: factory defaults, coefficient
derivation, pumping grids and functors, the solver's call into the native
layer, Solution accessors and .mat round trips.  Host only; the native call
is replaced by a recorder."""
import numpy as np
import pytest

from nls_amd import model as M
from nls_amd import pumping as PU
from nls_amd import solver as S


def test_default_coefficients_golden():
    """model.py:56-65 defaults through model.py:138-160."""
    m = M.Problem().model(model="1d", num_nodes=16, pumping=PU.GaussianPumping1D())
    c = m.getCoefficients()
    assert c[5] == pytest.approx(2.79999999998678, rel=1e-13)
    assert c[11] == pytest.approx(0.13658959117711406, rel=1e-13)
    assert c[13] == pytest.approx(0.7462686567172462, rel=1e-13)
    assert c[10] == 0.0 and c[14] == 0.0
    assert list(c[:5]) == [1.0] * 5
    assert m.getCharacteristicScale("phi") == pytest.approx(2.7317918235422813, rel=1e-13)
    assert m.getCharacteristicScale("n") == pytest.approx(30.245716122215104, rel=1e-13)
    assert m.getCharacteristicScale("nope") is None


def test_factory_dispatch_and_defaults():
    P = M.Problem()
    m1 = P.model(model="default", pumping=PU.GaussianPumping1D())
    assert isinstance(m1, M.Model1D) and isinstance(m1.getSolver(), S.Solver1D)
    assert (m1.dx, m1.dt, m1.order, m1.num_nodes, m1.num_iters) == (0.1, 1e-3, 5, 1000, 100000)
    np.testing.assert_array_equal(m1.getInitialSolution(), 0.1 * np.ones(1000))
    m2 = P.model(model="2d")
    assert isinstance(m2, M.Model2D) and isinstance(m2.getSolver(), S.Solver2D)
    assert (m2.order, m2.num_nodes, m2.num_iters) == (3, 40, 1000)
    assert m2.getInitialSolution().shape == (40, 40)
    with pytest.raises(Exception):
        P.model(model="3d")


def test_callable_u0_sampled_on_linspace():
    m = M.Problem().model(model="1d", num_nodes=11, dx=0.5, u0=lambda x: x ** 2, pumping=PU.GaussianPumping1D())
    np.testing.assert_allclose(m.getInitialSolution(), np.linspace(0.0, 5.5, 11) ** 2)


def test_pumping_grids():
    m1 = M.Problem().model(model="1d", num_nodes=9, dx=0.5, pumping=PU.GaussianPumping1D(power=2.0, variation=1.0))
    x = np.linspace(0.0, 4.5, 9)            # spacing n*dx/(n-1), model.py:222-224
    np.testing.assert_allclose(m1.getPumping(), 2.0 * np.exp(-x ** 2 / 2.0))
    m2 = M.Problem().model(model="2d", num_nodes=8, dx=0.5, pumping=PU.GaussianPumping2D(variation=1.5))
    x = np.linspace(-2.0, 2.0, 8)
    gx, gy = np.meshgrid(x, x)
    np.testing.assert_allclose(m2.getPumping(), np.exp(-(gx ** 2 + gy ** 2) / (2 * 1.5 ** 2)))
    m3 = M.Problem().model(model="2d", num_nodes=8, dx=0.5, bc="periodic", pumping=PU.GaussianPumping2D())
    np.testing.assert_allclose(m3.getGrid(), -2.0 + 0.5 * np.arange(8))


def test_pumping_functors():
    r = np.linspace(0, 10, 41)
    ring = PU.GaussianRingPumping2D(power=3.0, radius=4.0, variation=1.5)
    val = ring(r, np.zeros_like(r))
    np.testing.assert_allclose(val, 3.0 * (np.exp(-(r - 4) ** 2 / 4.5) + np.exp(-(r + 4) ** 2 / 4.5)))
    a, b = PU.GaussianPumping1D(2.0, variation=1.0), PU.GaussianPumping1D(1.0, x0=1.0, variation=2.0)
    np.testing.assert_allclose((a + b)(r), a(r) + b(r))
    np.testing.assert_allclose((a - b)(r), a(r) - b(r))
    s = a + b
    s.setPower(0.5)
    np.testing.assert_allclose(s(r), 0.5 * (a(r) + b(r)))
    rect = PU.RectangularPumping1D(power=7.0, x0=2.0, width=1.0)
    np.testing.assert_array_equal(rect(np.array([1.4, 1.5, 2.0, 2.5, 2.6])), [0, 7, 7, 7, 0])
    assert PU.AbstractPumping(power=5.0).power == 1.0          # pumping.py:15-16
    assert PU.GaussianElipticPumping2D(power=5.0).power == 1.0
    with pytest.raises(Exception):
        PU.AbstractPumping()(r)
    with pytest.raises(Exception):
        PU.OpMulPumping(a, b)(r)
    single = PU.RingPumping2D(power=3.0, radius=6.0, width=1.0)
    np.testing.assert_allclose(single(r, 0 * r), 3.0 * np.exp(-(r - 6) ** 2 / 2.0))
    grid = PU.GridPumping(np.arange(4.0))
    np.testing.assert_array_equal(grid(), np.arange(4.0))


class Recorder(object):
    def __init__(self):
        self.calls = []

    def solve_nls(self, *args):
        self.calls.append(("solve_nls", args))
        return np.asarray(args[-1], np.complex64) * 2

    def solve_nls_2d(self, *args):
        self.calls.append(("solve_nls_2d", args))
        return np.asarray(args[-1], np.complex64) * 3


def test_solver_calls_native_in_reference_order(monkeypatch):
    rec = Recorder()
    monkeypatch.setattr(S, "nls", rec)
    m = M.Problem().model(model="1d", num_nodes=12, dx=0.2, dt=1e-3, order=7, num_iters=50,
                          pumping=PU.GaussianPumping1D())
    sol = m.solve(num_iters=77)            # solver.py:22-23 override
    name, args = rec.calls[-1]
    assert name == "solve_nls"
    dt, dx, order, iters, pumping, coeffs, u0 = args
    assert (dt, dx, order, iters) == (1e-3, 0.2, 7, 77)
    np.testing.assert_array_equal(pumping, m.getPumping())
    np.testing.assert_array_equal(coeffs, m.getCoefficients())
    np.testing.assert_allclose(sol.getSolution(), 2 * m.getInitialSolution(), rtol=1e-7)
    assert m.getNumberOfIterations() == 77
    assert sol.getElapsedTime() >= 0.0 and sol.getModel() is m
    m2 = M.Problem().model(model="2d", num_nodes=6)
    sol2 = m2.solve()
    assert rec.calls[-1][0] == "solve_nls_2d"
    np.testing.assert_allclose(sol2.getSolution(), 3 * m2.getInitialSolution(), rtol=1e-7)


def test_solution_accessors_and_damping_integral():
    m = M.Problem().model(model="1d", num_nodes=10, dx=0.5, pumping=PU.GaussianPumping1D(power=2.0))
    sol = M.Solution(m, 0.5 * np.ones(10, np.complex64))
    np.testing.assert_allclose(sol.getDensity(), 0.25)
    c = m.getCoefficients()
    res = c[11] * m.getPumping() / (c[12] + c[13] * 0.25)
    np.testing.assert_allclose(sol.getReservoir(), res)
    r = np.linspace(0, 5.0, 10)
    assert sol.getDampingIntegral() == pytest.approx(2 * np.pi * np.sum((res - 1.0) * 0.25 * r * 0.5))
    m2 = M.Problem().model(model="2d", num_nodes=5, dx=0.5)
    sol2 = M.Solution(m2, np.ones((5, 5)))
    res2 = sol2.getReservoir()
    assert sol2.getDampingIntegral() == pytest.approx(np.sum((res2 - 1.0) * 0.25))


def test_store_restore_round_trip(tmp_path):
    m = M.Problem().model(model="2d", num_nodes=6, dx=0.3, dt=2e-3, order=5, bc="periodic", dtype="f64",
                          pumping=PU.GaussianPumping2D(power=2.0), n0=np.full((6, 6), 0.25))
    sol = M.Solution(m, (1 + 2j) * np.ones((6, 6)))
    sol.setReservoir(np.full((6, 6), 0.5))
    sol.setElapsedTime(1.5)
    f = str(tmp_path / "s.mat")
    sol.store(f, label="lab", desc="run")
    m2 = M.Problem().model(filename=f)
    assert isinstance(m2, M.Model2D)
    assert (m2.num_nodes, m2.order, m2.dx, m2.dt, m2.bc, m2.dtype) == (6, 5, 0.3, 2e-3, "periodic", "f64")
    np.testing.assert_allclose(m2.getPumping(), m.getPumping())
    np.testing.assert_allclose(m2.getCoefficients(), m.getCoefficients())
    np.testing.assert_allclose(m2.getInitialReservoir(), 0.25)
    s2 = M.Solution(m2).restore(f)
    np.testing.assert_allclose(s2.getSolution(), (1 + 2j) * np.ones((6, 6)))
    np.testing.assert_allclose(s2.getReservoir(), 0.5)
    assert s2.getElapsedTime() == 1.5
    m1 = M.Problem().model(model="1d", num_nodes=7, pumping=PU.GaussianPumping1D())
    f1 = str(tmp_path / "m1.mat")
    m1.store(f1)
    r1 = M.Problem().model(filename=f1)
    assert isinstance(r1, M.Model1D) and r1.getInitialSolution().shape == (7,)


def test_reference_mode_selection():
    m = M.Problem().model(model="1d", num_nodes=8, pumping=PU.GaussianPumping1D())
    assert m.getSolver()._reference_mode(m.getCoefficients())
    m = M.Problem().model(model="1d", num_nodes=8, dtype="f64", pumping=PU.GaussianPumping1D())
    assert not m.getSolver()._reference_mode(m.getCoefficients())
    c = np.zeros(23)
    c[10] = 1.0
    m = M.Problem().model(model="1d", num_nodes=8, coeffs=c, pumping=PU.GaussianPumping1D())
    assert not m.getSolver()._reference_mode(m.getCoefficients())
