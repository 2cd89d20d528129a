"""Input fields generated on the device (nls_fill_profile / nls_fill_state).

CPU: the host mirror reproduces the reference's functors on the model's grids
(it is what the device values are checked against), argument errors.
GPU: device values equal the host mirror to within the last bit of exp/sqrt
(FP64 relative 4e-16, FP32 one rounding); the noise is reproducible,
independent of the launch, and standard normal; served requests whose inputs
are generated on the device give the same result as the same inputs
uploaded."""
import ctypes as C

import numpy as np
import pytest

from nls_amd import fields, native
from nls_amd import model as M
from nls_amd import pumping as PU


def reference_pumping(functor, dim, n, dx, bc="ref"):
    """The model's own sampling (nls/model.py:220-232)."""
    if dim == 1:
        m = M.Problem().model(model="1d", dx=dx, num_nodes=n, pumping=functor, bc=bc)
    else:
        m = M.Problem().model(model="2d", dx=dx, num_nodes=n, pumping=functor, bc=bc)
    return m.getPumping()


FUNCTORS = [
    (1, PU.GaussianPumping1D(power=3.0, x0=1.0, variation=2.0)),
    (1, PU.GaussianRingPumping1D(power=20.0, radius=10.0, variation=3.14)),
    (2, PU.GaussianPumping2D(power=15.0, x0=0.5, y0=-1.0, variation=3.14)),
    (2, PU.GaussianRingPumping2D(power=2.0, x0=0.3, y0=0.2, variation=1.5, radius=4.0)),
    (2, PU.RingPumping2D(power=3.0, radius=6.0, width=1.0)),
    (2, PU.GaussianPumping2D(power=1.0, variation=2.0) + PU.RingPumping2D(power=2.0, radius=3.0, width=0.5)),
]


@pytest.mark.parametrize("dim,functor", FUNCTORS)
@pytest.mark.parametrize("bc", ("ref", "periodic"))
def test_host_mirror_equals_reference_functors(dim, functor, bc):
    n, dx = (400, 0.1) if dim == 1 else (64, 0.2)
    ref = reference_pumping(functor, dim, n, dx, bc)
    grid = fields.Grid.of_model(n, dx, dim, bc)
    terms = fields.from_functor(functor, dim)
    assert terms is not None

    def ev(ts):
        (t,) = ts
        if t[0] == "tree":
            _, node, lhs, rhs = t
            return node.power * (ev(lhs) + ev(rhs)) if isinstance(node, PU.OpSumPumping) else \
                node.power * (ev(lhs) - ev(rhs))
        scale, prof = t
        v = fields.profile_numpy(prof, dim, n, grid)
        return v if scale == 1.0 else scale * v
    np.testing.assert_allclose(ev(terms), ref, rtol=4e-16, atol=0)


def test_grids_match_numpy_constructions():
    n, dx = 33, 0.37
    np.testing.assert_array_equal(fields.Grid("linspace", 0.0, n * dx).numpy(n), np.linspace(0.0, n * dx, n))
    np.testing.assert_array_equal(fields.Grid("step", -n * dx / 2.0, dx).numpy(n), -n * dx / 2.0 + dx * np.arange(n))
    L = 32.0
    np.testing.assert_array_equal(fields.Grid("fraction", -L / 2, L).numpy(n), -L / 2 + L * np.arange(n) / n)


def test_noise_mirror_is_standard_normal():
    z = fields.noise_numpy(200000, np.sqrt(2.0), seed=11)     # unit variance per component
    assert abs(z.real.mean()) < 0.01 and abs(z.imag.mean()) < 0.01
    assert abs(z.real.std() - 1.0) < 0.01 and abs(z.imag.std() - 1.0) < 0.01
    assert abs(np.corrcoef(z.real, z.imag)[0, 1]) < 0.01
    z2 = fields.noise_numpy(10, np.sqrt(2.0), seed=11)
    np.testing.assert_array_equal(z[:10], z2)


def test_unsupported_functors_are_refused():
    assert fields.from_functor(PU.GridPumping(np.ones(8)), 1) is None
    assert fields.from_functor(PU.RectangularPumping1D(), 1) is None


def test_fill_argument_errors():
    L = native.load()
    p, g = fields.Profile("gaussian").desc(), fields.Grid("step", 0, 1).desc()
    buf = C.c_void_p(1)
    assert L.nls_fill_profile(C.byref(p), 3, 16, C.byref(g), 0, buf, None) == native.NLS_EINVAL
    bad = fields.ProfileDesc(9, 0, 1, 0, 0, 1, 0)
    assert L.nls_fill_profile(C.byref(bad), 1, 16, C.byref(g), 0, buf, None) == native.NLS_EINVAL
    badg = fields.GridDesc(7, 0, 0, 1)
    assert L.nls_fill_profile(C.byref(p), 1, 16, C.byref(badg), 0, buf, None) == native.NLS_EINVAL
    z = fields.Profile("gaussian", width=0.0).desc()
    assert L.nls_fill_state(C.byref(z), 0.0, 1, 1, 16, C.byref(g), 0, buf, None) == native.NLS_EINVAL


# ---- GPU ------------------------------------------------------------------

PROFILES = [fields.Profile("constant", 2.5), fields.Profile("gaussian", 3.0, 0.5, -1.0, 2.0),
            fields.Profile("gaussian_ring", 2.0, 0.3, 0.2, 1.5, 4.0), fields.Profile("ring", 3.0, 0.0, 1.0, 0.7, 5.0)]
GRIDS = [fields.Grid("linspace", -8.0, 8.0), fields.Grid("step", -6.4, 0.2), fields.Grid("fraction", -16.0, 32.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n", [(1, 1000), (2, 96)])
@pytest.mark.parametrize("dtype", ("f64", "f32"))
def test_device_profiles_equal_host_mirror(gpu, dim, n, dtype):
    import torch
    rt = torch.float64 if dtype == "f64" else torch.float32
    for prof in PROFILES:
        for grid in GRIDS:
            out = torch.empty((n,) if dim == 1 else (n, n), dtype=rt, device="cuda:0")
            fields.fill_profile(out, prof, dim, n, grid)
            got = out.cpu().numpy()
            want = fields.profile_numpy(prof, dim, n, grid, "f64")
            if dtype == "f64":
                np.testing.assert_allclose(got, want, rtol=4.5e-16, atol=1e-300)
            else:
                # one rounding of a double that may differ in its last bit
                np.testing.assert_allclose(got, want.astype(np.float32), rtol=1.2e-7, atol=1e-38)


@pytest.mark.gpu
def test_device_state_noise(gpu):
    import torch
    n = 256
    grid = fields.Grid("fraction", -16.0, 32.0)
    prof = fields.Profile("gaussian", 0.1, width=np.sqrt(2.0))
    a = torch.empty((n, n), dtype=torch.complex128, device="cuda:0")
    b = torch.empty((n, n), dtype=torch.complex128, device="cuda:0")
    fields.fill_state(a, prof, 1e-3, 12345, 2, n, grid)
    fields.fill_state(b, prof, 1e-3, 12345, 2, n, grid)
    torch.cuda.synchronize()
    A, B = a.cpu().numpy(), b.cpu().numpy()
    np.testing.assert_array_equal(A, B)                                  # reproducible
    want = fields.state_numpy(prof, 1e-3, 12345, 2, n, grid)
    assert np.max(np.abs(A - want)) <= 1e-15                              # the host mirror, to libm's last bits
    z = (A - fields.profile_numpy(prof, 2, n, grid)) * (np.sqrt(2.0) / 1e-3)
    assert abs(z.real.std() - 1.0) < 0.02 and abs(z.imag.std() - 1.0) < 0.02
    c = torch.empty((n, n), dtype=torch.complex128, device="cuda:0")
    fields.fill_state(c, prof, 1e-3, 12346, 2, n, grid)
    assert not np.array_equal(c.cpu().numpy(), A)                        # another seed, another draw


@pytest.mark.gpu
@pytest.mark.parametrize("dim,functor", FUNCTORS)
def test_model_pumping_on_device(gpu, dim, functor):
    n, dx = (400, 0.1) if dim == 1 else (64, 0.2)
    ref = reference_pumping(functor, dim, n, dx)
    got = fields.pumping_on_device(functor, dim, n, fields.Grid.of_model(n, dx, dim), "f64").cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=4.5e-16, atol=1e-300)


@pytest.mark.gpu
def test_served_device_inputs_equal_uploaded_inputs(gpu):
    """The same requests served twice: inputs generated on the device, and the
    device-generated arrays downloaded and given as host arrays."""
    import torch
    from nls_amd import serving
    rng = np.random.default_rng(4)
    reqs = [serving.make_request(i, 0.0, rng, step_range=(50, 200), device_inputs=True) for i in range(8)]
    reqs = [r for r in reqs if r.dim == 1 or r.n <= 192]
    eng = serving.Engine(budget_bytes=1 << 30, target_block_ms=2.0, min_block=8, max_block=64)
    done = eng.run(reqs)
    eng.close()
    twins = []
    for r in reqs:
        shape = (r.n,) if r.dim == 1 else (r.n, r.n)
        P = torch.empty(shape, dtype=torch.float32, device="cuda:0")
        psi = torch.empty(shape, dtype=torch.complex64, device="cuda:0")
        fields.fill_profile(P, r.spec["pump"], r.dim, r.n, r.spec["grid"])
        fields.fill_state(psi, r.spec["state"], r.spec["noise"], r.spec["seed"], r.dim, r.n, r.spec["grid"])
        twins.append(serving.Request(r.rid, 0.0, r.dim, r.n, r.steps, r.dt, r.dx, r.coeffs, P.cpu().numpy(),
                                     psi.cpu().numpy(), np.zeros(shape)))
    eng = serving.Engine(budget_bytes=1 << 30, target_block_ms=2.0, min_block=8, max_block=64)
    done2 = eng.run(twins)
    eng.close()
    by = {r.rid: r for r in done2}
    assert len(done) == len(reqs) == len(done2)
    for r in done:
        np.testing.assert_array_equal(r.psi, by[r.rid].psi)
        np.testing.assert_array_equal(r.nres, by[r.rid].nres)
