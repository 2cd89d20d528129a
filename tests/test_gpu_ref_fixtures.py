"""The GPU path against vectors written by the reference Fortran itself.

tests/golden/ref_fixtures.npz holds inputs and outputs of nls/nls.f90 compiled
as written (FP32, BLAS sgbmv) and with every real made double (FP64, dgbmv);
tests/golden/make_golden.py wrote it.  Here the GPU runs the same inputs
through the C ABI, with no oracle in between:

* the single-precision drop-ins (solve_nls, solve_nls_2d: nls.f90:797, :903)
  against the FP32 reference, relative L2 <= 1e-5;
* the double-precision path on the reference's boundaries with its
  algebraic reservoir (nls_solve_host, every RK4 engine that holds the case)
  against the FP64 reference, relative L2 <= 1e-10;
* the chemical potential (nls.f90:921-971): the FP32 drop-ins within 1e-5 of
  the FP32 reference, the FP64 extended entry point within 1e-12 of the FP64
  reference.
"""
import os

import numpy as np
import pytest

from nls_amd import native

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_fixtures.npz"))
SOLVES = sorted({k[: -len("_meta")] for k in FIX.files if k.startswith("solve_") and k.endswith("_meta")})


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def case(key):
    dim, n, order, steps, dt, dx = FIX[key + "_meta"]
    return (int(dim), int(n), int(order), int(steps), float(dt), float(dx), FIX[key + "_P"], FIX[key + "_u0"],
            FIX[key + "_coeffs"])


def test_fixture_cases_present():
    assert len(SOLVES) == 6
    assert tuple(FIX["version"]) == native.nls.version()


@pytest.mark.parametrize("key", SOLVES)
def test_dropin_f32_matches_reference_f32(gpu, key):
    dim, n, order, steps, dt, dx, P, u0, c = case(key)
    fn = native.nls.solve_nls if dim == 1 else native.nls.solve_nls_2d
    u = fn(dt, dx, order, steps, P, c, u0)
    ref = FIX[key + "_u_ref32"]
    assert u.dtype == np.complex64 and u.shape == ref.shape
    assert rel(u, ref) <= 1e-5, rel(u, ref)


def engines(dim, order):
    if dim == 1:
        return ("stage", "resident")
    # the resident engine has no double-precision 2-D order-7 variant
    return ("stage", "fused") if order == 7 else ("stage", "fused", "resident")


@pytest.mark.parametrize("key,engine", [(k, e) for k in SOLVES for e in engines(case(k)[0], case(k)[2])])
def test_f64_matches_reference_f64(gpu, key, engine):
    dim, n, order, steps, dt, dx, P, u0, c = case(key)
    assert c[10] == 0.0                     # the reference's algebraic reservoir
    desc = native.ProblemDesc.make(dim, n, order, dt, dx, c, bc="ref", dtype="f64", engine=engine)
    psi, _ = native.solve_host(desc, steps, P, u0)
    ref = FIX[key + "_u_ref64"]
    assert rel(psi, ref) <= 1e-10, rel(psi, ref)


@pytest.mark.parametrize("dim,n", [(1, 24), (2, 12)])
def test_chemical_potential_matches_reference(gpu, dim, n):
    import torch
    key = "mu_d%d_n%d" % (dim, n)
    P, u0, c = FIX[key + "_P"], FIX[key + "_u0"], FIX[key + "_coeffs"]
    dx = 0.2
    if dim == 1:
        mu32 = native.nls.chemical_potential_1d(dx, P, c, u0)
    else:
        mu32 = native.nls.chemical_potential_2d(dx, P, c, u0)
    ref32 = complex(FIX[key + "_ref32"])
    assert abs(complex(mu32) - ref32) <= 1e-5 * abs(ref32)
    # double precision: the extended entry point (order 5, reference boundaries)
    dev = torch.device("cuda:0")
    desc = native.ProblemDesc.make(dim, n, 5, 1e-3, dx, c, bc="ref", dtype="f64")
    mu = native.chemical_potential(desc, torch.from_numpy(np.ascontiguousarray(P)).to(dev),
                                   torch.from_numpy(np.ascontiguousarray(u0)).to(dev))
    ref64 = complex(FIX[key + "_ref64"])
    if dim == 2:
        mu = mu.real                        # the reference returns the real part (nls.f90:955)
    assert abs(mu - ref64) <= 1e-12 * abs(ref64), (mu, ref64)
