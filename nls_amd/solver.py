"""Solver facade (the reference's ``nls/solver.py``): a template method that
collects (dt, dx, order, iters, pumping, coeffs, u0) from the model in the
reference's argument order, times the native call with wall-clock time and
wraps the result in a ``Solution`` (solver.py:21-38).

A model in the reference's own mode (single precision, 'ref' boundaries,
RK4, algebraic reservoir) goes through the f2py-shaped drop-ins
``nls.solve_nls`` / ``nls.solve_nls_2d`` exactly as the reference does
(solver.py:66, :80); any other model goes through ``native.solve_host`` with
its dtype / boundary family / scheme, which also returns the reservoir field
when it is dynamic.
"""
from time import time

import numpy as np

from . import native
from .native import nls


class AbstractSolver(object):

    def __init__(self, model):
        self.model = model
        self.elapsed_time = 0.0
        self.solution = None

    def __call__(self, num_iters=None):
        from .model import Solution
        if num_iters:
            self.model.setNumberOfIterations(num_iters)
        self.elapsed_time = -time()
        self.solution = Solution(self.model)
        result = self.solve(
            self.model.getTimeStep(),
            self.model.getSpatialStep(),
            self.model.getApproximationOrder(),
            self.model.getNumberOfIterations(),
            self.model.getPumping(),
            self.model.getCoefficients(),
            self.model.getInitialSolution())
        if isinstance(result, tuple):
            psi, nres = result
            self.solution.setSolution(psi)
            self.solution.setReservoir(nres)
        else:
            self.solution.setSolution(result)
        self.elapsed_time += time()
        self.solution.setElapsedTime(self.elapsed_time)
        return self.solution

    def solve(self, *args, **kwargs):
        raise Exception("AbstractSolver: native solver routine is not passed!")

    def chemicalPotential(self, solution, *args, **kwargs):
        return self.chemicalPotentialRoutine(
            self.model.getSpatialStep(),
            self.model.getPumping(),
            self.model.getCoefficients(),
            solution)

    def chemicalPotentialRoutine(self, *args, **kwargs):
        raise Exception("AbstractSolver: native solver routine is not passed!")

    # -- the extended path ------------------------------------------------
    def _reference_mode(self, coeffs):
        m = self.model
        return (getattr(m, "dtype", "f32") == "f32" and getattr(m, "bc", "ref") == "ref"
                and getattr(m, "scheme", "rk4") == "rk4" and getattr(m, "engine", "auto") == "auto"
                and np.asarray(coeffs)[10] == 0.0)

    def _solve_extended(self, dim, dt, dx, order, iters, pumping, coeffs, u0):
        m = self.model
        engine = getattr(m, "engine", "auto")
        desc = native.ProblemDesc.make(dim, m.getNumberOfNodes(), order, dt, dx, coeffs, bc=m.bc,
                                       dtype=m.dtype, scheme=m.scheme,
                                       engine="auto" if engine == "tuned" else engine)
        if engine == "tuned":
            from .autotune import Tuner
            desc = Tuner().tuned(desc)
        psi, nres = native.solve_host(desc, iters, pumping, u0, m.getInitialReservoir())
        return (psi, nres) if nres is not None else psi


class Solver1D(AbstractSolver):
    """1-D solver (solver.py:58-69)."""

    def solve(self, dt, dx, order, iters, pumping, coeffs, u0):
        if self._reference_mode(coeffs):
            return nls.solve_nls(dt, dx, order, iters, pumping, coeffs, u0)
        return self._solve_extended(1, dt, dx, order, iters, pumping, coeffs, u0)

    def chemicalPotentialRoutine(self, *args, **kwargs):
        return nls.chemical_potential_1d(*args, **kwargs)


class Solver2D(AbstractSolver):
    """2-D solver (solver.py:72-83)."""

    def solve(self, dt, dx, order, iters, pumping, coeffs, u0):
        if self._reference_mode(coeffs):
            return nls.solve_nls_2d(dt, dx, order, iters, pumping, coeffs, u0)
        return self._solve_extended(2, dt, dx, order, iters, pumping, coeffs, u0)

    def chemicalPotentialRoutine(self, *args, **kwargs):
        return nls.chemical_potential_2d(*args, **kwargs)
