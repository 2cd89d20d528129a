"""GPU parity of the RK4 finite-difference path against the oracle.

Bar: FP64 on the GPU matches the FP64 oracle to relative L2 <= 1e-10 after
the full step count; FP32 on the GPU matches the FP64 oracle to <= 1e-4.
Every call goes through the C ABI of libnls_amd.so.
"""
import itertools

import numpy as np
import pytest

from nls_amd import configs, native

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def small_case(dim, n, bc, dyn, seed=7):
    rng = np.random.default_rng(seed)
    if dim == 1:
        if bc == "ref":
            x = np.arange(n) * 0.1
        else:
            x, _ = configs.grid_1d(n, 12.0)
        r2 = x ** 2
        shape = (n,)
    else:
        X, Y, _ = configs.grid_2d(n, 8.0)
        r2 = X ** 2 + Y ** 2
        shape = (n, n)
    psi0 = 0.3 * np.exp(-r2 / 4.0) + configs.noise(rng, shape, 1e-2)
    P = 2.0 * np.exp(-r2 / 6.0)
    c = configs.model_coeffs(dynamic=dyn)
    n0 = 0.2 * np.exp(-r2 / 3.0) if dyn else None
    return P, psi0, n0, c


# Every plain-path RK4 engine: per-stage kernels, the fused one-launch-per-step
# kernel (2-D), and the resident kernel (these sizes fit on chip; it is the
# default for them)
CASES = [(dim, bc, dyn, order, engine)
         for dim, bc, dyn, order in itertools.product((1, 2), ("ref", "periodic"), (False, True), (3, 5, 7))
         for engine in (("stage", "resident") if dim == 1 else ("stage", "fused", "resident"))]


@pytest.mark.parametrize("dim,bc,dyn,order,engine", CASES)
def test_rk4_f64_matches_oracle(gpu, oracle, dim, bc, dyn, order, engine):
    n = 96 if dim == 1 else 40
    if engine == "resident" and dim == 2 and order == 7:
        # no double-precision 2-D order-7 resident variant (it would spill):
        # the library refuses the request instead of running another engine
        P, psi0, n0, c = small_case(dim, n, bc, dyn)
        desc = native.ProblemDesc.make(dim, n, order, 1e-3, 0.2, c, bc=bc, dtype="f64", engine=engine)
        with pytest.raises(ValueError, match="resident engine"):
            native.solve_host(desc, 10, P, psi0, n0)
        return
    dx = 0.1 if dim == 1 else 0.2
    P, psi0, n0, c = small_case(dim, n, bc, dyn)
    steps = 300
    dt = 2e-3 if dim == 1 else 1e-3
    if dim == 1 and bc == "ref":
        dt = 5e-4          # the radial operator's 1/r rows are stiffer
    desc = native.ProblemDesc.make(dim, n, order, dt, dx, c, bc=bc, dtype="f64", engine=engine)
    psi, nres = native.solve_host(desc, steps, P, psi0, n0)
    opsi, onres = oracle.rk4(dim, n, order, 0 if bc == "ref" else 1, dt, dx, steps, P, c, psi0, n0)
    assert np.all(np.isfinite(psi))
    assert rel(psi, opsi) <= TOL64, rel(psi, opsi)
    if dyn:
        assert rel(nres, onres) <= TOL64


@pytest.mark.parametrize("dim,bc,dyn,order,engine", CASES)
def test_rk4_f32_within_tolerance_of_f64_oracle(gpu, oracle, dim, bc, dyn, order, engine):
    n = 96 if dim == 1 else 40
    dx = 0.1 if dim == 1 else 0.2
    P, psi0, n0, c = small_case(dim, n, bc, dyn)
    steps = 300
    dt = 5e-4 if (dim == 1 and bc == "ref") else (2e-3 if dim == 1 else 1e-3)
    desc = native.ProblemDesc.make(dim, n, order, dt, dx, c, bc=bc, dtype="f32", engine=engine)
    psi, nres = native.solve_host(desc, steps, P, psi0, n0)
    opsi, onres = oracle.rk4(dim, n, order, 0 if bc == "ref" else 1, dt, dx, steps, P, c, psi0, n0)
    assert rel(psi, opsi) <= TOL32, rel(psi, opsi)
    if dyn:
        assert rel(nres, onres) <= TOL32


def test_dropin_solve_nls_matches_oracle_f32(gpu, oracle):
    """nls.solve_nls with the reference's argument list (solver.py:66)."""
    n, dx, dt, steps = 128, 0.1, 1e-3, 500
    x = np.arange(n) * dx
    P = 3.0 * np.exp(-x ** 2 / 13.6)
    c = np.zeros(23)
    c[2], c[3], c[4], c[5], c[11], c[12], c[13] = 0.5, 0.5, 1.0, 2.0, 0.5, 1.0, 2.0
    u0 = 0.1 * np.ones(n)
    u = native.nls.solve_nls(dt, dx, 5, steps, P, c, u0)
    assert u.dtype == np.complex64 and u.shape == (n,)
    o32, _ = oracle.rk4(1, n, 5, 0, dt, dx, steps, P, c, u0, dtype=np.float32)
    o64, _ = oracle.rk4(1, n, 5, 0, dt, dx, steps, P, c, u0)
    assert rel(u, o64) <= TOL32
    assert rel(u, o32) <= 1e-5


def test_dropin_solve_nls_2d_matches_oracle_f32(gpu, oracle):
    n, dx, dt, steps = 48, 0.2, 1e-3, 400
    X, Y, _ = configs.grid_2d(n, n * dx)
    P = 2.5 * np.exp(-(X ** 2 + Y ** 2) / 10.0)
    c = configs.model_coeffs(dynamic=False)
    u0 = 0.1 * np.ones((n, n))
    u = native.nls.solve_nls_2d(dt, dx, 5, steps, P, c, u0)
    assert u.dtype == np.complex64 and u.shape == (n, n)
    o64, _ = oracle.rk4(2, n, 5, 0, dt, dx, steps, P, c, u0)
    assert rel(u, o64) <= TOL32


def test_line256_full_f64(gpu, oracle):
    """The 1-D workload in full: N=256 periodic, 2000 steps of 1e-3."""
    inst = configs.line256()
    desc = native.ProblemDesc.make(1, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic", dtype="f64")
    psi, nres = native.solve_host(desc, inst.steps, inst.P, inst.psi0, inst.n0)
    opsi, onres = oracle.rk4(1, inst.n, 5, 1, inst.dt, inst.dx, inst.steps, inst.P, inst.coeffs, inst.psi0, inst.n0)
    assert rel(psi, opsi) <= TOL64
    assert rel(nres, onres) <= TOL64


@pytest.mark.parametrize("engine", ("resident", "stage"))
def test_line256_full_f32(gpu, oracle, engine):
    """Single precision on the 1-D workload in full against the FP64 oracle.
    The per-stage and resident kernels form the Laplacian from differences to
    the centre, so the stencil's large weights do not amplify the rounding of
    the state: the error stays within 3e-6 (the direct stencil sum gives 7e-6)."""
    inst = configs.line256()
    desc = native.ProblemDesc.make(1, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic", dtype="f32",
                                   engine=engine)
    psi, nres = native.solve_host(desc, inst.steps, inst.P, inst.psi0, inst.n0)
    opsi, onres = oracle.rk4(1, inst.n, 5, 1, inst.dt, inst.dx, inst.steps, inst.P, inst.coeffs, inst.psi0, inst.n0)
    assert rel(psi, opsi) <= 3e-6, rel(psi, opsi)
    assert rel(nres, onres) <= 3e-6


def test_square256_f64_first_1000_steps(gpu, oracle):
    """The 256^2 workload for 1000 of its 10^4 steps (the oracle's share)."""
    inst = configs.square256()
    steps = 1000
    oracle.set_threads(8)
    desc = native.ProblemDesc.make(2, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic", dtype="f64")
    psi, nres = native.solve_host(desc, steps, inst.P, inst.psi0, inst.n0)
    opsi, onres = oracle.rk4(2, inst.n, 5, 1, inst.dt, inst.dx, steps, inst.P, inst.coeffs, inst.psi0, inst.n0)
    assert rel(psi, opsi) <= TOL64
    assert rel(nres, onres) <= TOL64
    desc32 = native.ProblemDesc.make(2, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic", dtype="f32")
    psi32, _ = native.solve_host(desc32, steps, inst.P, inst.psi0, inst.n0)
    assert rel(psi32, opsi) <= TOL32


@pytest.mark.parametrize("n,dyn", [(1152, True), (640, True), (512, False)])
def test_fused_plain_f64_tile_forms(gpu, oracle, n, dyn):
    """The plain fused solve of one grid in double precision, in each tile
    form (on the 256 compute units of MI355X): 1152^2 in 512-thread 64 x 16
    tiles, two workgroups a CU (its 648 64 x 32 tiles take three rounds);
    640^2 in 1024-thread 64 x 32 tiles (200 of them); 512^2 in 1024-thread
    64 x 16 tiles (256).  Each against the oracle; the small grids above
    cover the 64 x 16 form at every order and boundary."""
    steps = 12
    oracle.set_threads(8)
    X, Y, h = configs.grid_2d(n, 0.15 * n)
    r2 = X ** 2 + Y ** 2
    rng = np.random.default_rng(n)
    psi0 = 0.3 * np.exp(-r2 / 200.0) + configs.noise(rng, (n, n), 1e-2)
    P = 2.0 * np.exp(-r2 / 300.0)
    n0 = 0.2 * np.exp(-r2 / 100.0) if dyn else None
    c = configs.model_coeffs(dynamic=dyn)
    desc = native.ProblemDesc.make(2, n, 5, 1e-3, h, c, bc="periodic", dtype="f64", engine="fused")
    psi, nres = native.solve_host(desc, steps, P, psi0, n0)
    opsi, onres = oracle.rk4(2, n, 5, 1, 1e-3, h, steps, P, c, psi0, n0)
    assert np.all(np.isfinite(psi))
    assert rel(psi, opsi) <= TOL64, rel(psi, opsi)
    if dyn:
        assert rel(nres, onres) <= TOL64


@pytest.mark.parametrize("dim", (1, 2))
def test_dropin_chemical_potential(gpu, oracle, dim):
    n = 64 if dim == 1 else 32
    dx = 0.1
    rng = np.random.default_rng(3)
    shape = (n,) if dim == 1 else (n, n)
    u0 = 0.5 + 0.1 * rng.standard_normal(shape) + 0.1j * rng.standard_normal(shape)
    P = 1.0 + rng.random(shape)
    c = configs.model_coeffs(dynamic=False)
    if dim == 1:
        mu = native.nls.chemical_potential_1d(dx, P, c, u0)
        ref = oracle.chemical_potential(1, dx, n, P, c, u0)
    else:
        mu = native.nls.chemical_potential_2d(dx, P, c, u0)
        ref = oracle.chemical_potential(2, dx, n, P, c, u0)
    assert abs(mu - ref) <= 1e-4 * max(1.0, abs(ref))


_SQ2048 = {}


def _square2048_oracle(oracle, steps=10):
    """The oracle's first `steps` steps of the 2048^2 workload (computed once)."""
    import os
    if steps not in _SQ2048:
        inst = configs.square2048()
        oracle.set_threads(min(16, os.cpu_count() or 1))
        opsi, onres = oracle.rk4(2, inst.n, 5, 1, inst.dt, inst.dx, steps, inst.P, inst.coeffs, inst.psi0, inst.n0)
        _SQ2048[steps] = (inst, opsi, onres)
    inst, opsi, onres = _SQ2048[steps]
    return inst, steps, opsi, onres


@pytest.mark.parametrize("engine", ("fused", "stage"))
def test_square2048_f64_first_steps(gpu, oracle, engine):
    """The 2048^2 workload itself (periodic on [-64,64)^2, ring pumping P0 = 3,
    r0 = 20, w = 2, dt = 2.5e-4, dynamic reservoir, order 5) for its first 10
    steps on each 2-D RK4 engine of the plain path: the fused kernel in the
    form this launch takes on MI355X (512-thread 64 x 16 tiles, two
    workgroups a CU: 4096 64 x 32 tiles are more than two rounds of 256 CUs)
    and the per-stage kernels, against the oracle in double precision
    (<= 1e-10); single precision on the same engine within 1e-4 of it."""
    inst, steps, opsi, onres = _square2048_oracle(oracle)
    desc = native.ProblemDesc.make(2, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic", dtype="f64",
                                   engine=engine)
    psi, nres = native.solve_host(desc, steps, inst.P, inst.psi0, inst.n0)
    assert np.all(np.isfinite(psi))
    assert rel(psi, opsi) <= TOL64, rel(psi, opsi)
    assert rel(nres, onres) <= TOL64, rel(nres, onres)
    desc32 = native.ProblemDesc.make(2, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic", dtype="f32",
                                     engine=engine)
    psi32, nres32 = native.solve_host(desc32, steps, inst.P, inst.psi0, inst.n0)
    assert rel(psi32, opsi) <= TOL32, rel(psi32, opsi)
    assert rel(nres32, onres) <= TOL32


def test_f64_fused_holds_grids_past_4_gib(gpu):
    """Double precision keeps 64-bit addresses in the fused kernel, so it
    holds a 16384^2 grid (its complex array is 4 GiB; single precision, with
    32-bit buffer offsets, hands such grids to the per-stage engine): two
    steps on device arrays equal the per-stage engine's to rounding."""
    import torch
    n, dx = 16384, 0.01
    x = (torch.arange(n, dtype=torch.float64, device="cuda") - n / 2) * dx
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    P = 2.0 * torch.exp(-r2 / 6.0)
    psi0 = (0.3 * torch.exp(-r2 / 4.0)).to(torch.complex128)
    n0 = 0.2 * torch.exp(-r2 / 3.0)
    del r2
    c = configs.model_coeffs(dynamic=True)
    out = {}
    for engine in ("fused", "stage"):
        d = native.ProblemDesc.make(2, n, 5, 1e-5, dx, c, bc="periodic", dtype="f64", engine=engine)
        psi, nres = psi0.clone(), n0.clone()
        ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device="cuda")
        native.advance(d, 2, P, psi, nres, ws)
        torch.cuda.synchronize()
        del ws
        out[engine] = (psi, nres)
    for a, b in zip(out["fused"], out["stage"]):
        err = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        assert err <= 1e-12, err
    assert float(torch.linalg.norm(out["fused"][0] - psi0)) > 0


@pytest.mark.parametrize("dim", (1, 2))
def test_unstable_dt_is_reported(gpu, dim):
    """A dt far above the RK4 stability limit (|lambda| dt <= 2 sqrt 2) blows
    the solution up.  The extended host entry point reports it
    (NLS_ENONFINITE -> FloatingPointError); the drop-in keeps the reference's
    behaviour (nls.f90:725-733): it returns the non-finite state silently."""
    n = 64 if dim == 1 else 32
    P, psi0, n0, c = small_case(dim, n, "periodic", True)
    dx = 0.1
    lam = dim * (16.0 / 3.0) / dx ** 2                 # largest |eigenvalue| of the order-5 Laplacian
    dt = 20.0 * 2.0 * np.sqrt(2.0) / lam
    desc = native.ProblemDesc.make(dim, n, 5, dt, dx, c, bc="periodic", dtype="f64")
    with pytest.raises(FloatingPointError):
        native.solve_host(desc, 400, P, psi0, n0)
    stable = native.ProblemDesc.make(dim, n, 5, 0.1 * dt / 20.0, dx, c, bc="periodic", dtype="f64")
    psi, _ = native.solve_host(stable, 50, P, psi0, n0)
    assert np.all(np.isfinite(psi))
    fn = native.nls.solve_nls if dim == 1 else native.nls.solve_nls_2d
    u = fn(dt, dx, 5, 400, P, configs.model_coeffs(dynamic=False), psi0)
    assert not np.all(np.isfinite(u))


def test_stage_f32_nontemporal_grid(gpu, oracle):
    """A single-precision grid past the per-stage kernels' non-temporal-store
    threshold (2304^2 periodic, dynamic reservoir: 276 MB of stage arrays,
    above 256 MB) for 4 steps on the per-stage engine, against the FP64
    oracle (the FP64 form of that store path runs in
    test_square2048_f64_first_steps)."""
    n, dx, dt, steps = 2304, 0.0625, 2.5e-4, 4
    x = -n * dx / 2 + dx * np.arange(n)
    r = np.sqrt(x[:, None] ** 2 + x[None, :] ** 2)
    P = configs.ring(r, 3.0, 20.0, 2.0)
    c = configs.model_coeffs(dynamic=True)
    psi0 = (0.1 * np.exp(-r ** 2 / 400.0) + configs.noise(np.random.default_rng(5), (n, n))).astype(np.complex128)
    n0 = np.zeros((n, n))
    oracle.set_threads(8)
    opsi, onres = oracle.rk4(2, n, 5, 1, dt, dx, steps, P, c, psi0, n0)
    desc = native.ProblemDesc.make(2, n, 5, dt, dx, c, bc="periodic", dtype="f32", engine="stage")
    psi, nres = native.solve_host(desc, steps, P, psi0, n0)
    assert np.all(np.isfinite(psi))
    assert rel(psi, opsi) <= TOL32, rel(psi, opsi)
    assert rel(nres, onres) <= TOL32, rel(nres, onres)
