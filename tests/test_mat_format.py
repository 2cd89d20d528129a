"""The .mat store/restore format against the reference's layout
(nls/model.py:257-314, :500-532).

tests/golden/ref_layout_model2d.mat holds exactly the keys the reference's
AbstractModel.store and Solution.store write, with its model field as the
reference writes it (str(type(self)) of nls.model.Model2D) and a solution
computed by the reference Fortran (tests/golden/make_mat_fixture.py)."""
import os

import numpy as np
from scipy.io import loadmat

from nls_amd import model as M
from nls_amd import pumping as PU

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURE = os.path.join(GOLD, "ref_layout_model2d.mat")
# keys the reference's store() methods write (nls/model.py:263-276, :505-509)
REF_MODEL_KEYS = {"model", "date", "dim", "dimlesses", "init_solution", "num_iters", "num_nodes", "order",
                  "originals", "pumping", "spatial_step", "time_step"}
REF_SOLUTION_KEYS = {"elapsed_time", "solution", "version"}
# model strings the reference's Problem.model dispatch accepts when the
# package is imported as nls (nls/model.py:67-72)
REF_ACCEPTS = {1: ("1d", "default", "<class 'nls.model.Model1D'>"), 2: ("2d", "<class 'nls.model.Model2D'>")}


def test_fixture_has_the_reference_layout():
    mat = loadmat(FIXTURE)
    assert {k for k in mat if not k.startswith("__")} == REF_MODEL_KEYS | REF_SOLUTION_KEYS
    assert str(mat["model"][0]) in REF_ACCEPTS[2]


def test_restore_reference_file():
    m = M.Problem().model(filename=FIXTURE)
    mat = loadmat(FIXTURE)
    assert isinstance(m, M.Model2D)
    assert m.getNumberOfNodes() == 16 and m.getNumberOfIterations() == 150 and m.order == 3
    assert m.getSpatialStep() == 0.5 and m.getTimeStep() == 1e-3
    np.testing.assert_array_equal(m.getCoefficients(), mat["dimlesses"].ravel())
    np.testing.assert_array_equal(m.getInitialSolution(), mat["init_solution"])
    np.testing.assert_array_equal(m.getPumping(), mat["pumping"])
    assert set(m.originals) == {"R", "gamma", "g", "tilde_g", "gamma_R"}
    sol = M.Solution(m).restore(FIXTURE)
    np.testing.assert_array_equal(sol.getSolution(), mat["solution"])
    assert sol.getElapsedTime() == 0.01


def test_store_writes_what_the_reference_restores(tmp_path):
    for kind, shape in ((1, (32,)), (2, (8, 8))):
        pump = PU.GaussianPumping1D() if kind == 1 else PU.GaussianPumping2D()
        m = M.Problem().model(model="%dd" % kind, num_nodes=shape[0], u0=0.1, num_iters=7, pumping=pump)
        path = str(tmp_path / ("m%d.mat" % kind))
        m.store(path)
        mat = loadmat(path)
        assert REF_MODEL_KEYS <= set(mat)
        # the reference's dispatch accepts the string, and so does this one
        assert str(mat["model"][0]) in REF_ACCEPTS[kind]
        back = M.Problem().model(filename=path)
        assert isinstance(back, M.Model1D if kind == 1 else M.Model2D)
        assert back.getNumberOfIterations() == 7
        np.testing.assert_array_equal(back.getInitialSolution(), m.getInitialSolution())


def test_dispatch_accepts_older_and_own_names():
    for name in ("Model2D", str(M.Model2D), "<class 'nls.model.Model2D'>", "2d"):
        assert isinstance(M.Problem().model(model=name), M.Model2D)
    for name in ("Model1D", str(M.Model1D), "<class 'nls.model.Model1D'>", "1d", "default"):
        assert isinstance(M.Problem().model(model=name), M.Model1D)
