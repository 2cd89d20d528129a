"""Parameter sweeps: many models of one family solved together in one batch.

The reference solves a sweep one job per worker process
(schedulr/workr.py:39-66): each job builds Problem().model(model='2d',
dx=0.2, dt=2e-3, order=5, num_nodes=200, num_iters=2000,
pumping=GaussianPumping2D(power=p, variation=5.0), original_params={...})
from the job's pumping power and calls model.solve().  Here the sweep's
models are the members of one batch (``nls_amd.batch.Batch`` over
``nls_batch_*``), advanced together on the GPU, each for its own
num_iters; every model gets back the ``Solution`` its own model.solve()
returns, with the same arithmetic per point (a member's result does not
depend on the rest of the batch).
"""
from time import time

import numpy as np

from .model import Problem, Solution
from .pumping import GaussianPumping2D

# The worker's physical parameters and grid (schedulr/workr.py:41-60).
WORKR_ORIGINALS = {
    "R": 0.0242057488654,
    "gamma": 0.0242057488654,
    "g": 0.00162178517398,
    "tilde_g": 0.0169440242057,
    "gamma_R": 0.242057488654,
}


def workr_model(power, **overrides):
    """The model of one job of schedulr/workr.py:41-60 at pumping power
    `power`; keyword arguments replace the worker's values (num_iters,
    num_nodes, dtype, ...)."""
    kw = dict(model="2d", dx=2.0e-1, dt=2.0e-3, t0=0.0, u0=0.1, order=5, num_nodes=200, num_iters=2000,
              pumping=GaussianPumping2D(power=power, variation=5.0), original_params=dict(WORKR_ORIGINALS),
              dimless_params={})
    kw.update(overrides)
    return Problem().model(**kw)


class _Member(object):
    """A model as a batch member (n, dx, dt, coeffs, P, psi0, n0).

    A model in the reference's own mode (single precision, 'ref'
    boundaries, RK4, algebraic reservoir) solves through the drop-ins,
    whose dt, dx and coefficients are single precision (real(sp),
    nls.f90:800-804); its member takes them rounded the same way, so that
    it steps exactly as its model.solve() does."""

    def __init__(self, model):
        self.n = int(model.getNumberOfNodes())
        self.dx = float(model.getSpatialStep())
        self.dt = float(model.getTimeStep())
        self.coeffs = np.asarray(model.getCoefficients(), np.float64)
        if model.getSolver()._reference_mode(self.coeffs):
            self.dx, self.dt = float(np.float32(self.dx)), float(np.float32(self.dt))
            self.coeffs = self.coeffs.astype(np.float32).astype(np.float64)
        self.P = np.asarray(model.getPumping())
        self.psi0 = np.asarray(model.getInitialSolution())
        self.n0 = model.getInitialReservoir()


def _family(model):
    return (model.getInitialSolution().ndim, int(model.getApproximationOrder()), getattr(model, "bc", "ref"),
            getattr(model, "dtype", "f32"), bool(np.asarray(model.getCoefficients())[10] != 0.0))


def solve_models(models, engine="auto", device=None, graphs=True):
    """Solve `models` together; returns one ``Solution`` per model, in order.

    The models share dimension, approximation order, boundary family
    ('ref' or 'periodic'), dtype and reservoir mode (algebraic when
    coeffs[10] == 0, else dynamic), and use the RK4 scheme; each keeps its
    own grid size, steps, coefficients, pumping, initial state and
    num_iters.  `engine` is the batch engine ('auto', 'stage', 'fused',
    'resident', 'hybrid').  Raises ValueError for an empty list or mixed
    families.  A solution's elapsed time is the whole batch's wall-clock
    time."""
    from .batch import Batch
    models = list(models)
    if not models:
        raise ValueError("solve_models: no models")
    fam = _family(models[0])
    for m in models:
        if _family(m) != fam:
            raise ValueError("solve_models: models differ in dimension, order, boundary, dtype or reservoir mode")
        if getattr(m, "scheme", "rk4") != "rk4":
            raise ValueError("solve_models: the batch engine runs the RK4 scheme only")
    dim, order, bc, dtype, _ = fam
    elapsed = -time()
    batch = Batch([_Member(m) for m in models], dim=dim, order=order, bc=bc, dtype=dtype, engine=engine,
                  device=device, graphs=graphs)
    try:
        iters = [int(m.getNumberOfIterations()) for m in models]
        batch.set_steps(iters)
        batch.advance(max(iters))
        out = []
        for i, m in enumerate(models):
            psi, nres = batch.member(i)
            s = Solution(m)
            s.setSolution(psi)
            if nres is not None:
                s.setReservoir(nres)
            out.append(s)
    finally:
        batch.close()
    elapsed += time()
    for s in out:
        s.setElapsedTime(elapsed)
    return out


def power_sweep(powers, engine="auto", device=None, **overrides):
    """The worker's job list for the given pumping powers, solved as one
    batch: a list of (power, Solution)."""
    models = [workr_model(p, **overrides) for p in powers]
    return list(zip(powers, solve_models(models, engine=engine, device=device)))
