"""Parameter sweeps solved as one batch (nls_amd.sweep): the worker's job
(schedulr/workr.py:41-60) and the batch against each model's own solve."""
import numpy as np
import pytest

from nls_amd import model as M
from nls_amd import pumping as PU
from nls_amd import sweep


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_workr_model_is_the_workers_job():
    m = sweep.workr_model(3.5)
    assert isinstance(m, M.Model2D)
    assert (m.getNumberOfNodes(), m.getApproximationOrder(), m.getNumberOfIterations()) == (200, 5, 2000)
    assert (m.getSpatialStep(), m.getTimeStep()) == (0.2, 2e-3)
    assert m.pumping.power == 3.5 and m.pumping.variation == 5.0
    np.testing.assert_array_equal(m.getInitialSolution(), 0.1 * np.ones((200, 200)))
    # the worker's physical parameters are the reference's defaults: the same coefficients
    np.testing.assert_array_equal(m.getCoefficients(), M.Problem().model(model="2d").getCoefficients())
    assert sweep.workr_model(1.0, num_iters=7, num_nodes=32).getNumberOfNodes() == 32


def test_solve_models_rejects_mixed_families():
    a = sweep.workr_model(1.0, num_nodes=16, num_iters=2)
    with pytest.raises(ValueError):
        sweep.solve_models([])
    for b in (sweep.workr_model(1.0, num_nodes=16, order=3),
              sweep.workr_model(1.0, num_nodes=16, dtype="f64"),
              sweep.workr_model(1.0, num_nodes=16, bc="periodic"),
              M.Problem().model(model="1d", num_nodes=16)):
        with pytest.raises(ValueError):
            sweep.solve_models([a, b])
    with pytest.raises(ValueError):
        sweep.solve_models([sweep.workr_model(1.0, num_nodes=16, bc="periodic", scheme="ssf")])


@pytest.mark.gpu
def test_gpu_power_sweep_equals_each_solve(gpu):
    """The worker's job list at four powers (grid and steps as the worker's,
    fewer steps, one model stopping early): each batched result equals the
    model's own model.solve() bit for bit (the same fused kernel, alone and in
    a batch)."""
    powers = (1.0, 2.5, 4.0, 15.0)
    models = [sweep.workr_model(p, num_iters=200) for p in powers]
    models[2].setNumberOfIterations(77)
    sols = sweep.solve_models(models)
    for m, s in zip(models, sols):
        alone = sweep.workr_model(m.pumping.power, num_iters=m.getNumberOfIterations()).solve().getSolution()
        got = s.getSolution()
        assert got.dtype == np.complex64 and got.shape == (200, 200)
        assert np.isfinite(got).all()
        np.testing.assert_array_equal(got, alone)
        assert s.getModel() is m and s.getElapsedTime() > 0.0
    # the powers give different results
    assert rel(sols[0].getSolution(), sols[3].getSolution()) > 1e-3


@pytest.mark.gpu
def test_gpu_sweep_against_oracle(gpu, oracle):
    """Double precision, dynamic reservoir and periodic boundaries, mixed sizes
    and step counts, each member against the FP64 oracle at 1e-10."""
    c = M.Problem().model(model="2d").getCoefficients().copy()
    c[10] = 1.0
    models = []
    for n, p, it in ((48, 3.0, 40), (64, 8.0, 25), (96, 5.0, 40)):
        x = -n * 0.2 / 2 + 0.2 * np.arange(n)
        u0 = 0.1 * np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / 8.0) + 0j
        models.append(M.Problem().model(model="2d", dx=0.2, dt=2e-3, order=5, num_nodes=n, num_iters=it, u0=u0,
                                        pumping=PU.GaussianPumping2D(power=p, variation=5.0), coeffs=c,
                                        bc="periodic", dtype="f64", n0=np.zeros((n, n))))
    sols = sweep.solve_models(models)
    for m, s in zip(models, sols):
        n = m.getNumberOfNodes()
        opsi, onres = oracle.rk4(2, n, 5, 1, 2e-3, 0.2, m.getNumberOfIterations(), m.getPumping(), c,
                                 m.getInitialSolution(), np.zeros((n, n)))
        assert rel(s.getSolution(), opsi) <= 1e-10
        assert rel(s.getReservoir(), onres) <= 1e-10


@pytest.mark.gpu
def test_gpu_sweep_1d_radial_equals_each_solve(gpu):
    """1-D radial models (the reference's default model family) of different
    sizes, powers and step counts in one batch: each equals its own
    model.solve() bit for bit (the resident engine alone and in the batch)."""
    models = [M.Problem().model(model="1d", num_nodes=n, num_iters=it, dx=0.1, dt=1e-3,
                                pumping=PU.GaussianRingPumping1D(power=p, radius=5.0, variation=3.14))
              for n, p, it in ((200, 2.0, 300), (400, 5.0, 301), (1000, 20.0, 150))]
    sols = sweep.solve_models(models)
    for m, s in zip(models, sols):
        alone = M.Problem().model(model="1d", num_nodes=m.getNumberOfNodes(), num_iters=m.getNumberOfIterations(),
                                  dx=0.1, dt=1e-3, pumping=m.pumping).solve().getSolution()
        assert s.getSolution().shape == (m.getNumberOfNodes(),)
        assert np.isfinite(s.getSolution()).all()
        np.testing.assert_array_equal(s.getSolution(), alone)
