"""Observables on the device (nls_observables, nls_batch_observables):
norm sum |psi|^2 dV, reservoir population sum n dV, peak |psi|^2 and the
non-finite count, accumulated in double precision.

Pins: closed forms for uniform states (which fix the volume element,
2 pi r dr with r = k dx on the 1-D radial grid), numpy's FP64 sums of random
states, planted non-finite points, and -- for the batch reduction -- the
single-instance reduction of each member and bitwise independence of a
member's values from the rest of the batch.
"""
import math

import numpy as np
import pytest

from nls_amd import configs, native
from nls_amd.batch import Batch

pytestmark = pytest.mark.gpu


def dev_state(psi, nres, dtype):
    import torch
    ct = torch.complex128 if dtype == "f64" else torch.complex64
    rt = torch.float64 if dtype == "f64" else torch.float32
    d = torch.device("cuda:0")
    return torch.as_tensor(psi, dtype=ct, device=d), torch.as_tensor(nres, dtype=rt, device=d)


@pytest.mark.parametrize("dtype", ("f64", "f32"))
@pytest.mark.parametrize("dim,bc", [(1, "ref"), (1, "periodic"), (2, "periodic"), (2, "ref")])
def test_uniform_state_closed_forms(gpu, dtype, dim, bc):
    n, dx, a, n0 = 48, 0.25, 0.75, 1.5
    shape = (n,) if dim == 1 else (n, n)
    psi, nres = dev_state(np.full(shape, a + 0j), np.full(shape, n0), dtype)
    desc = native.ProblemDesc.make(dim, n, 5, 1e-3, dx, configs.model_coeffs(dynamic=True), bc=bc, dtype=dtype)
    o = native.observables(desc, psi, nres)
    if dim == 1 and bc == "ref":
        vol = 2.0 * math.pi * dx * dx * n * (n - 1) / 2.0    # sum over k of 2 pi (k dx) dx
    else:
        vol = n * dx if dim == 1 else n * n * dx * dx
    assert o["norm"] == pytest.approx(a * a * vol, rel=1e-12)
    assert o["population"] == pytest.approx(n0 * vol, rel=1e-12)
    assert o["peak"] == pytest.approx(a * a, rel=1e-12) and o["nonfinite"] == 0


def test_random_state_matches_numpy_sums_and_counts_nonfinite(gpu):
    rng = np.random.default_rng(2)
    n, dx = 100, 0.1
    psi_h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    n_h = rng.random((n, n))
    psi, nres = dev_state(psi_h, n_h, "f64")
    desc = native.ProblemDesc.make(2, n, 5, 1e-3, dx, configs.model_coeffs(dynamic=True), bc="periodic", dtype="f64")
    o = native.observables(desc, psi, nres)
    s = np.abs(psi_h) ** 2
    assert o["norm"] == pytest.approx(s.sum() * dx * dx, rel=1e-12)
    assert o["population"] == pytest.approx(n_h.sum() * dx * dx, rel=1e-12)
    # the peak to rounding: |psi|^2 here goes through hypot, on the device re^2 + im^2
    assert o["peak"] == pytest.approx(s.max(), rel=1e-14) and o["nonfinite"] == 0
    psi_h[3, 7] = np.nan
    n_h[50, 1] = np.inf
    psi, nres = dev_state(psi_h, n_h, "f64")
    o = native.observables(desc, psi, nres)
    ok = np.isfinite(np.abs(psi_h)) & np.isfinite(n_h)
    assert o["nonfinite"] == 2
    assert o["norm"] == pytest.approx((np.abs(psi_h[ok]) ** 2).sum() * dx * dx, rel=1e-12)


def members(sizes, dim=2):
    out = []
    for k, n in enumerate(sizes):
        if dim == 2:
            out.append(configs.square(n, 0.25 * n, 1.0 + k, 2.0, 1.0, 40 + k, 1e-3, 0, "o%d" % k))
        else:
            x, h = configs.grid_1d(n, 0.1 * n)
            rng = np.random.default_rng(50 + k)
            out.append(configs.Instance("l%d" % k, 1, n, h, 1e-3, 0, np.exp(-x ** 2), 0.3 * np.exp(-x ** 2 / 4.0)
                                        + configs.noise(rng, x.shape, 1e-2), 0.1 * np.ones(n),
                                        configs.model_coeffs(dynamic=True)))
    return out


@pytest.mark.parametrize("dtype", ("f64", "f32"))
@pytest.mark.parametrize("dim,bc,engine", [(2, "periodic", "fused"), (2, "ref", "stage"), (1, "ref", "stage"),
                                           (1, "periodic", "resident")])
def test_batch_observables_match_each_member_alone(gpu, dtype, dim, bc, engine):
    import torch
    sizes = (48, 64, 100) if dim == 2 else (300, 512, 1000)
    ms = members(sizes, dim)
    b = Batch(ms, dim=dim, bc=bc, dtype=dtype, engine=engine)
    b.advance(30)
    torch.cuda.synchronize()
    obs = b.observables()
    for i, m in enumerate(ms):
        psi_h, n_h = b.member(i)
        psi, nres = dev_state(psi_h, n_h, dtype)
        desc = native.ProblemDesc.make(dim, m.n, 5, m.dt, m.dx, m.coeffs, bc=bc, dtype=dtype)
        one = native.observables(desc, psi, nres)
        assert obs[i, 0] == pytest.approx(one["norm"], rel=1e-12)
        assert obs[i, 1] == pytest.approx(one["population"], rel=1e-12)
        assert obs[i, 2] == pytest.approx(one["peak"], rel=1e-14) and obs[i, 3] == 0
    # a member's values do not depend on the rest of the batch
    alone = Batch(ms[1:2], dim=dim, bc=bc, dtype=dtype, engine=engine)
    alone.advance(30)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(alone.observables()[0], obs[1])
    b.close()
    alone.close()


def test_batch_observables_count_nonfinite_per_member(gpu):
    import torch
    ms = members((48, 64))
    b = Batch(ms, dtype="f32", engine="fused")
    b.psi[b.offsets[1] + 5] = float("nan")
    torch.cuda.synchronize()
    obs = b.observables()
    assert obs[0, 3] == 0 and obs[1, 3] == 1
    b.close()
