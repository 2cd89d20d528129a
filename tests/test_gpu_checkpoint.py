"""Checkpoint and resume: a solve of 2N steps equals a solve of N steps whose
state is saved (to a .mat file through the model's store/restore, or kept on
the host) and resumed for N more steps -- bitwise, on every engine.  The
resumed state carries the dynamic reservoir."""
import numpy as np
import pytest

from nls_amd import configs, native
from nls_amd import model as M
from nls_amd import pumping as PU
from nls_amd.batch import Batch

pytestmark = pytest.mark.gpu


def case(dim, n, seed=4):
    rng = np.random.default_rng(seed)
    if dim == 1:
        x, h = configs.grid_1d(n, 0.1 * n)
        r2 = x ** 2
        shape = (n,)
    else:
        X, Y, h = configs.grid_2d(n, 10.0)
        r2 = X ** 2 + Y ** 2
        shape = (n, n)
    psi0 = 0.3 * np.exp(-r2 / 4.0) + configs.noise(rng, shape, 1e-2)
    return h, 2.0 * np.exp(-r2 / 6.0), psi0, 0.2 * np.exp(-r2 / 3.0), configs.model_coeffs(dynamic=True)


ENGINES = [(1, 300, "stage"), (1, 300, "resident"), (2, 40, "stage"), (2, 40, "fused"), (2, 40, "resident"),
           (2, 130, "fused")]


@pytest.mark.parametrize("dim,n,engine", ENGINES)
@pytest.mark.parametrize("dtype", ("f32", "f64"))
def test_resume_equals_uninterrupted(gpu, dim, n, engine, dtype):
    h, P, psi0, n0, c = case(dim, n)
    dt = 5e-4 if dim == 1 else 1e-3
    d = native.ProblemDesc.make(dim, n, 5, dt, h, c, bc="periodic", dtype=dtype, engine=engine)
    N = 150                                       # 2N crosses the step-block graph threshold
    full_psi, full_n = native.solve_host(d, 2 * N, P, psi0, n0)
    half_psi, half_n = native.solve_host(d, N, P, psi0, n0)
    res_psi, res_n = native.solve_host(d, N, P, half_psi, half_n)
    np.testing.assert_array_equal(res_psi, full_psi)
    np.testing.assert_array_equal(res_n, full_n)


@pytest.mark.parametrize("engine", ("fused", "stage", "resident"))
def test_batch_resume_equals_uninterrupted(gpu, engine):
    ms = [m for m in configs.sweep512(count=12, steps=40) if m.n == 64][:4]
    a = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine=engine)
    a.advance(40)
    b = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine=engine)
    b.advance(17)
    b.advance(23)
    for i in range(len(ms)):
        pa, na = a.member(i)
        pb, nb = b.member(i)
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(na, nb)
    a.close()
    b.close()


def test_model_checkpoint_through_mat_file(gpu, tmp_path):
    n, N = 48, 120
    h, P, psi0, n0, c = case(2, n)

    def model(u0, r0, iters):
        return M.Problem().model(model="2d", dx=h, dt=1e-3, order=5, num_nodes=n, num_iters=iters, u0=u0,
                                 pumping=PU.GridPumping(P), coeffs=c, bc="periodic", dtype="f64", n0=r0)

    full = model(psi0, n0, 2 * N).solve()
    first = model(psi0, n0, N).solve()
    # checkpoint: the state after N steps becomes the initial state of a stored model
    ck = model(first.getSolution(), first.getReservoir(), N)
    path = str(tmp_path / "checkpoint.mat")
    ck.store(path)
    resumed = model(psi0, n0, 1)
    resumed.restore(path)
    assert resumed.num_iters == N and resumed.dtype == "f64" and resumed.bc == "periodic"
    second = resumed.solve()
    np.testing.assert_array_equal(second.getSolution(), full.getSolution())
    np.testing.assert_array_equal(second.getReservoir(), full.getReservoir())


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_resume_reference_checkpoint_matches_reference(gpu):
    """A checkpoint the reference wrote (its .mat layout, its own FP32
    solution after N steps; tests/golden/make_mat_fixture.py) resumed the
    reference's way (setInitialSolution, then solve: tools/check.py:34-36)
    for N more steps equals the reference's uninterrupted 2N-step solve to
    FP32 rounding (relative L2 <= 1e-5)."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    m = M.Problem().model(filename=os.path.join(gold, "ref_layout_model2d.mat"))
    sol = M.Solution(m).restore(os.path.join(gold, "ref_layout_model2d.mat"))
    ref = np.load(os.path.join(gold, "ref_layout_model2d.npz"))
    m.setInitialSolution(sol.getSolution())
    resumed = m.solve().getSolution()
    assert resumed.dtype == np.complex64
    assert _rel(resumed, ref["u_2n"]) <= 1e-5, _rel(resumed, ref["u_2n"])


@pytest.mark.parametrize("dim,n", [(1, 300), (2, 48)])
def test_mat_resume_matches_oracle_uninterrupted(gpu, oracle, tmp_path, dim, n):
    """Double precision, dynamic reservoir: N steps, a checkpoint of psi and
    the reservoir through the model's .mat file, N more steps -- against the
    oracle's uninterrupted 2N-step solve (relative L2 <= 1e-10)."""
    N = 140
    h, P, psi0, n0, c = case(dim, n)
    dt = 5e-4 if dim == 1 else 1e-3
    pump = PU.GridPumping(P)

    def model(u0, r0, iters):
        return M.Problem().model(model="%dd" % dim, dx=h, dt=dt, order=5, num_nodes=n, num_iters=iters, u0=u0,
                                 pumping=pump, coeffs=c, bc="periodic", dtype="f64", n0=r0)

    first = model(psi0, n0, N).solve()
    ck = model(first.getSolution(), first.getReservoir(), N)
    path = str(tmp_path / "ck.mat")
    ck.store(path)
    second = M.Problem().model(filename=path).solve()
    opsi, onres = oracle.rk4(dim, n, 5, 1, dt, h, 2 * N, P, c, psi0, n0)
    assert _rel(second.getSolution(), opsi) <= 1e-10, _rel(second.getSolution(), opsi)
    assert _rel(second.getReservoir(), onres) <= 1e-10
