"""Problem / Model / Solution -- the reference's user-facing API
(``nls/model.py``) in Python 3, backed by libnls_amd on the GPU.

Kept as in the reference: the factory ``Problem().model(**kwargs)`` and its
defaults (model.py:32-117), the derivation of the 23 dimensionless
coefficients from the physical parameters (model.py:125-160), the pumping
grids of ``getPumping`` (model.py:220-232; note the spacing n*dx/(n-1) of
``linspace``), ``model.solve(num_iters)`` and the ``Solution`` accessors
(model.py:337-380), and ``.mat`` store/restore (model.py:257-314,
:500-532).

Added keyword arguments of ``Problem().model``:
  dtype  'f32' (default: the reference's precision) or 'f64'
  bc     'ref' (default: the reference's radial 1-D / zero-outside 2-D
         operators) or 'periodic' (Cartesian stencils with wrap-around; the
         pumping and u0 grids are then x_k = -n dx/2 + k dx)
  scheme 'rk4' (default) or 'ssf' (split-step Fourier, periodic only)
  engine RK4 engine: 'auto' (default), 'stage', 'fused', 'resident' or
         'tuned' (the fastest on this device, measured once and cached by
         nls_amd.autotune)
  coeffs an explicit 23-entry coefficient vector replacing the derived one
  n0     initial reservoir density when coeffs[10] != 0 (default zero)
Plotting (model.py:391-498) is not part of this package.
"""
from datetime import datetime
from types import FunctionType

import numpy as np

from .pumping import GaussianPumping, GridPumping

DEFAULT_ORIGINALS = {
    "R": 0.0242057488654,
    "gamma": 0.0242057488654,
    "g": 0.00162178517398,
    "tilde_g": 0.0169440242057,
    "gamma_R": 0.242057488654,
}

HBAR = 6.61e-34
M_E = 9.1e-31
M_0 = 1.0e-5 * M_E


def version():
    """(0, 2, 0): the reference module's version (nls/version.py:5-11, nls.f90:29-37)."""
    return 0, 2, 0


# The .mat "model" field as the reference writes it, str(type(self)) of its
# classes in package nls (nls/model.py:265), and the strings its Problem.model
# dispatch accepts (:67-72): files written here restore in the reference, and
# the reference's files restore here.
REF_MODEL_NAMES = {1: "<class 'nls.model.Model1D'>", 2: "<class 'nls.model.Model2D'>"}


def _model_kind(name):
    """1 or 2 for a model name Problem.model accepts, else None: '1d',
    'default', '2d', str() of this module's classes or of the reference's,
    and the bare class names."""
    name = str(name).strip()
    if name in ("1d", "default", str(Model1D), REF_MODEL_NAMES[1], "Model1D"):
        return 1
    if name in ("2d", str(Model2D), REF_MODEL_NAMES[2], "Model2D"):
        return 2
    return None


class Problem(object):
    """Factory of models (model.py:27-117)."""

    def model(self, *args, **kwargs):
        if "filename" in kwargs:
            return self.modelFromFile(kwargs["filename"])
        kwargs.pop("params", None)
        kwargs["model"] = kwargs.get("model", "default")
        originals = dict(kwargs.get("original_params") or {})
        for k, v in DEFAULT_ORIGINALS.items():
            originals.setdefault(k, v)
        kwargs["original_params"] = originals
        kind = _model_kind(kwargs["model"])
        if kind == 1:
            return self.fabricateModel1D(*args, **kwargs)
        if kind == 2:
            return self.fabricateModel2D(*args, **kwargs)
        raise Exception("Unknown model passed!")

    def modelFromFile(self, filename):
        """model.py:74-84: a default model of the file's kind, then restore."""
        from scipy.io import loadmat
        mat = loadmat(filename)
        if "model" not in mat:
            raise ValueError("%s holds no model" % filename)
        return self.model(model=str(mat["model"][0])).restore(filename)

    @staticmethod
    def _grid(kwargs, n):
        dx = kwargs["dx"]
        if kwargs.get("bc", "ref") == "periodic":
            return -n * dx / 2.0 + dx * np.arange(n)
        return np.linspace(0.0, dx * n, n)

    def fabricateModel1D(self, *args, **kwargs):
        kwargs.setdefault("dx", 1.0e-1)
        kwargs.setdefault("dt", 1.0e-3)
        kwargs.setdefault("t0", 0.0)
        kwargs.setdefault("u0", 1.0e-1)
        kwargs.setdefault("order", 5)
        kwargs.setdefault("pumping", GaussianPumping())
        kwargs.setdefault("num_nodes", 1000)
        kwargs.setdefault("num_iters", 100000)
        if isinstance(kwargs["u0"], (int, float, complex)):
            kwargs["u0"] = kwargs["u0"] * np.ones(kwargs["num_nodes"])
        elif isinstance(kwargs["u0"], FunctionType):
            kwargs["u0"] = kwargs["u0"](self._grid(kwargs, kwargs["num_nodes"]))
        return Model1D(**kwargs)

    def fabricateModel2D(self, *args, **kwargs):
        kwargs.setdefault("dx", 1.0e-1)
        kwargs.setdefault("dt", 1.0e-3)
        kwargs.setdefault("t0", 0.0)
        kwargs.setdefault("u0", 1.0e-1)
        kwargs.setdefault("order", 3)
        kwargs.setdefault("pumping", GaussianPumping())
        kwargs.setdefault("num_nodes", 40)
        kwargs.setdefault("num_iters", 1000)
        if isinstance(kwargs["u0"], (int, float, complex)):
            kwargs["u0"] = kwargs["u0"] * np.ones((kwargs["num_nodes"], kwargs["num_nodes"]))
        return Model2D(**kwargs)


class AbstractModel(object):
    """Model state and the coefficient derivation (model.py:120-314)."""

    dim = None

    def __init__(self, *args, **kwargs):
        self.dt = kwargs["dt"]
        self.dx = kwargs["dx"]
        self.order = kwargs["order"]
        self.num_nodes = kwargs["num_nodes"]
        self.num_iters = kwargs["num_iters"]
        self.pumping = kwargs["pumping"]
        self.init_sol = np.asarray(kwargs["u0"])
        self.originals = kwargs["original_params"]
        self.verbose = bool(kwargs.get("verbose"))
        self.dtype = kwargs.get("dtype", "f32")
        self.bc = kwargs.get("bc", "ref")
        self.scheme = kwargs.get("scheme", "rk4")
        self.engine = kwargs.get("engine", "auto")
        self.init_res = kwargs.get("n0")
        self.solver = None
        self.coeffs = self.deriveCoefficients(self.originals)
        if kwargs.get("coeffs") is not None:
            c = np.zeros(23)
            cc = np.asarray(kwargs["coeffs"], np.float64).ravel()
            c[:min(23, cc.size)] = cc[:23]
            self.coeffs = c

    @staticmethod
    def deriveCoefficients(o):
        """model.py:138-160: dimensionless coefficients from R, gamma, g, tilde_g, gamma_R."""
        c = np.zeros(23)
        phi0 = np.sqrt(o["gamma"] / (2.0 * o["g"]))
        t0 = phi0
        n0 = 2.0 / (o["R"] * t0)
        c[0] = 1.0
        c[1] = 1.0
        c[2] = 1.0
        c[3] = 1.0
        c[4] = 1.0
        c[5] = 4.0 * o["tilde_g"] / o["R"]
        c[10] = 0.0
        c[11] = 1.0 / (n0 * o["gamma_R"])
        c[12] = 1.0
        c[13] = o["R"] * phi0 ** 2 / o["gamma_R"]
        c[14] = 0.0
        return c

    def __repr__(self):
        from pprint import pformat
        return pformat({"dt": self.dt, "dx": self.dx, "order": self.order, "num_nodes": self.num_nodes,
                        "num_iters": self.num_iters, "pumping": self.pumping, "originals": self.originals,
                        "dtype": self.dtype, "bc": self.bc, "scheme": self.scheme,
                        "engine": self.engine}) + "\n" + str(self.coeffs)

    def getApproximationOrder(self):
        return self.order

    def getCharacteristicScale(self, scale):
        o = self.originals
        phi0 = np.sqrt(o["gamma"] / (2.0 * o["g"]))
        t0 = phi0
        x0 = np.sqrt(HBAR * t0 / (2 * M_0))
        n0 = 2.0 / (o["R"] * t0)
        return {"x": x0, "t": t0, "n": n0, "phi": phi0}.get(scale)

    def getChemicalPotential(self, solution):
        if isinstance(solution, Solution):
            solution = solution.getSolution()
        self.mu = self.solver.chemicalPotential(solution)
        return self.mu

    def getCoefficients(self):
        return self.coeffs

    def getInitialSolution(self):
        return self.init_sol

    def getInitialReservoir(self):
        return self.init_res

    def getNumberOfIterations(self):
        return self.num_iters

    def getNumberOfNodes(self):
        return self.num_nodes

    def getParticleNumber(self, method="simps"):
        """Simpson integral of |psi|^2 over the last solution with step dx
        (model.py:217-218; the reference reads a never-set attribute)."""
        from scipy.integrate import simpson
        sol = self.solver.solution.getSolution()
        return simpson((sol.conj() * sol).real, dx=self.dx)

    def getGrid(self):
        n, dx = self.num_nodes, self.dx
        if self.bc == "periodic":
            return -n * dx / 2.0 + dx * np.arange(n)
        if self.init_sol.ndim == 1:
            return np.linspace(0.0, n * dx, n)
        return np.linspace(-n * dx / 2.0, n * dx / 2.0, n)

    def getPumping(self):
        x = self.getGrid()
        if self.init_sol.ndim == 1:
            return self.pumping(*np.meshgrid(x))
        return self.pumping(*np.meshgrid(x, x))

    def getSpatialStep(self):
        return self.dx

    def getSolver(self):
        return self.solver

    def getTimeStep(self):
        return self.dt

    def setNumberOfIterations(self, num_iters):
        self.num_iters = num_iters

    def setPumping(self, pumping):
        self.pumping = pumping

    def setInitialSolution(self, solution):
        self.init_sol = np.asarray(solution)

    def setInitialReservoir(self, reservoir):
        self.init_res = None if reservoir is None else np.asarray(reservoir)

    def solve(self, num_iters=None):
        return self.solver(num_iters)

    def store(self, filename=None, label=None, desc=None, date=None):
        """Write the model to a .mat file (model.py:257-285) plus the fields
        of this package (dtype, bc, scheme, engine, initial reservoir)."""
        from scipy.io import savemat
        date = date if date else datetime.now()
        date = date.replace(microsecond=0).isoformat()
        filename = filename if filename else date + ".mat"
        mat = {"model": REF_MODEL_NAMES[self.init_sol.ndim], "date": date, "dim": self.init_sol.ndim,
               "dimlesses": self.coeffs, "init_solution": self.init_sol, "num_iters": self.num_iters,
               "num_nodes": self.num_nodes, "order": self.order, "originals": self.originals,
               "pumping": self.getPumping(), "spatial_step": self.dx, "time_step": self.dt,
               "dtype": self.dtype, "bc": self.bc, "scheme": self.scheme, "engine": self.engine}
        if self.init_res is not None:
            mat["init_reservoir"] = self.init_res
        if desc:
            mat["desc"] = desc
        if label:
            mat["label"] = label
        savemat(filename, mat)

    def restore(self, filename):
        """Read a model written by ``store`` (model.py:287-314)."""
        from scipy.io import loadmat
        mat = loadmat(filename)
        dim = int(np.asarray(mat["dim"]).ravel()[0])
        init = mat["init_solution"]
        pump = mat["pumping"]
        if dim == 1:
            init, pump = init[0, :], pump[0, :]
        self.coeffs = np.asarray(mat["dimlesses"]).ravel().astype(np.float64)
        self.init_sol = init
        self.num_nodes = int(np.asarray(mat["num_nodes"]).ravel()[0])
        self.num_iters = int(np.asarray(mat["num_iters"]).ravel()[0])
        self.order = int(np.asarray(mat["order"]).ravel()[0])
        self.pumping = GridPumping(pump)
        self.dx = float(np.asarray(mat["spatial_step"]).ravel()[0])
        self.dt = float(np.asarray(mat["time_step"]).ravel()[0])
        o = mat["originals"][0, 0]
        self.originals = {name: float(np.asarray(o[name]).ravel()[0]) for name in o.dtype.names}
        for key in ("dtype", "bc", "scheme", "engine"):
            if key in mat:
                setattr(self, key, str(np.asarray(mat[key]).ravel()[0]))
        if "init_reservoir" in mat:
            r = mat["init_reservoir"]
            self.init_res = r[0, :] if dim == 1 else r
        if "desc" in mat:
            self.desc = str(mat["desc"][0])
        if "label" in mat:
            self.label = str(mat["label"][0])
        return self


class Model1D(AbstractModel):
    """1-D model: axisymmetric radial grid for bc='ref' (model.py:317-324)."""

    dim = 1

    def __init__(self, *args, **kwargs):
        super(Model1D, self).__init__(*args, **kwargs)
        from .solver import Solver1D
        self.solver = Solver1D(self)


class Model2D(AbstractModel):
    """2-D model on an n x n grid (model.py:327-334)."""

    dim = 2

    def __init__(self, *args, **kwargs):
        super(Model2D, self).__init__(*args, **kwargs)
        from .solver import Solver2D
        self.solver = Solver2D(self)


class Solution(object):
    """Result of a solve (model.py:337-536, plotting excluded)."""

    def __init__(self, model, solution=None, verbose=False):
        self.elapsed_time = 0.0
        self.model = model
        self.solution = solution
        self.reservoir = None
        self.verbose = verbose

    def getDampingIntegral(self):
        """Rectangle-rule integral of (n - 1)|psi|^2 (model.py:350-365): 2 pi r dr
        weighted on the 1-D radial grid, dx^2 in 2-D."""
        reservoir = self.getReservoir()
        density = self.getDensity()
        length = self.model.getSpatialStep()
        if self.solution.ndim == 1:
            nodes = self.model.getNumberOfNodes()
            radius = np.linspace(0, nodes * length, nodes)
            return 2 * np.pi * np.sum((reservoir - 1.0) * density * radius * length)
        return np.sum((reservoir - 1.0) * density * length ** 2)

    def getDensity(self):
        return (self.solution.conj() * self.solution).real

    def getElapsedTime(self):
        return self.elapsed_time

    def getModel(self):
        return self.model

    def getReservoir(self):
        """The reservoir density: the integrated field for a dynamic reservoir,
        else c12 P / (c13 + c14 |psi|^2) (model.py:376-380)."""
        if self.reservoir is not None:
            return self.reservoir
        c = self.model.coeffs
        return c[11] * self.model.getPumping() / (c[12] + c[13] * self.getDensity())

    def getSolution(self):
        return self.solution

    def setElapsedTime(self, seconds):
        self.elapsed_time = seconds

    def setSolution(self, solution):
        self.solution = solution

    def setReservoir(self, reservoir):
        self.reservoir = reservoir

    def visualize(self, *args, **kwargs):
        raise NotImplementedError("plotting is not part of nls_amd")

    def store(self, filename=None, label=None, desc=None, date=None):
        """Model fields plus solution, elapsed time and version (model.py:500-517)."""
        from scipy.io import loadmat, savemat
        date = datetime.now() if date is None else date
        filename = filename if filename else date.replace(microsecond=0).isoformat() + ".mat"
        self.model.store(filename, label, desc, date)
        mat = loadmat(filename)
        mat = {k: v for k, v in mat.items() if not k.startswith("__")}
        mat.update({"elapsed_time": self.elapsed_time, "solution": self.solution, "version": version()})
        if self.reservoir is not None:
            mat["reservoir"] = self.reservoir
        savemat(filename, mat)

    def restore(self, filename):
        from scipy.io import loadmat
        mat = loadmat(filename)
        sol = mat["solution"]
        dim = int(np.asarray(mat["dim"]).ravel()[0])
        self.solution = sol[0, :] if dim == 1 else sol
        self.elapsed_time = float(np.asarray(mat["elapsed_time"]).ravel()[0])
        if "reservoir" in mat:
            r = mat["reservoir"]
            self.reservoir = r[0, :] if dim == 1 else r
        return self

    def report(self):
        message = "Elapsed in {0} seconds with {1} iteration on {2} grid nodes."
        print(message.format(self.elapsed_time, self.model.getNumberOfIterations(), self.model.getNumberOfNodes()))
