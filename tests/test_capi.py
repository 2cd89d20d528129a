"""The C ABI: library loads, exports every declared symbol, validates
arguments on the host and builds the reference's band operator.  No GPU
compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from nls_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names.update(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nls_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = native.load()
    names = declared_symbols()
    assert "nls_solve_nls" in names and "nls_advance" in names
    for name in names:
        assert hasattr(L, name), name
    assert set(native.EXPORTS) <= set(names)


def test_version_matches_reference_module():
    assert native.nls.version() == (0, 2, 0)     # nls.f90:34-36


@pytest.mark.parametrize("order", (3, 5, 7))
@pytest.mark.parametrize("n", (7, 16, 33))
def test_make_laplacian_matches_oracle(oracle, n, order):
    op = native.nls.make_laplacian(n, order, 0.1)
    ref = oracle.make_laplacian(n, order, 0.1)
    assert op.shape == (order, n)
    L = native.load()
    op64 = np.zeros((n, order))
    assert L.nls_make_laplacian(n, order, 0.1, op64.ctypes.data_as(C.c_void_p)) == 0
    np.testing.assert_allclose(op64.T, ref, rtol=1e-13, atol=1e-12 * np.abs(ref).max())


def test_argument_errors_raise_valueerror():
    with pytest.raises(ValueError):
        native.nls.solve_nls(1e-3, 0.1, 4, 10, np.ones(16), np.zeros(23), np.ones(16))   # bad order
    with pytest.raises(ValueError):
        native.nls.solve_nls(1e-3, 0.1, 5, 10, np.ones(16), np.zeros(23), np.ones(15))   # u0 shape
    with pytest.raises(ValueError):
        native.nls.solve_nls(1e-3, 0.1, 5, 10, np.ones(4), np.zeros(23), np.ones(4))     # n < order
    with pytest.raises(ValueError):
        native.nls.solve_nls_2d(1e-3, 0.1, 5, 10, np.ones((8, 8)), np.zeros(22), np.ones((8, 8)))
    d = native.ProblemDesc.make(2, 64, 5, 1e-3, 0.1, np.zeros(23), bc="ref", scheme="ssf")
    with pytest.raises(ValueError):
        native.solve_host(d, 1, np.ones((64, 64)), np.ones((64, 64)))                      # SSF needs periodic


def test_workspace_bytes():
    # per-stage engine: two stage slopes and the accumulator per field
    c = np.zeros(23)
    d = native.ProblemDesc.make(2, 256, 5, 1e-3, 0.1, c, bc="periodic", dtype="f64", engine="stage")
    w_adi = native.workspace_bytes(d)
    assert w_adi >= 3 * 256 * 256 * 16
    c[10] = 1.0
    d = native.ProblemDesc.make(2, 256, 5, 1e-3, 0.1, c, bc="periodic", dtype="f64", engine="stage")
    assert native.workspace_bytes(d) >= w_adi + 3 * 256 * 256 * 8
    # fused engine (the 2-D default above 64 x 64): the other half of the state ping-pong
    for engine in ("auto", "fused"):
        d = native.ProblemDesc.make(2, 256, 5, 1e-3, 0.1, c, bc="periodic", dtype="f64", engine=engine)
        assert 256 * 256 * 24 <= native.workspace_bytes(d) < 256 * 256 * 24 + 1024


def test_engine_field():
    c = np.zeros(23)
    with pytest.raises(ValueError):       # the fused engine is 2-D only
        native.solve_host(native.ProblemDesc.make(1, 64, 5, 1e-3, 0.1, c, engine="fused"), 1,
                          np.ones(64), np.ones(64))
    with pytest.raises(ValueError):       # the resident engine holds 2-D n <= 64 only
        native.solve_host(native.ProblemDesc.make(2, 128, 5, 1e-3, 0.1, c, engine="resident"), 1,
                          np.ones((128, 128)), np.ones((128, 128)))
    with pytest.raises(ValueError):
        native.solve_host(native.ProblemDesc.make(2, 16, 5, 1e-3, 0.1, c, engine=7), 1,
                          np.ones((16, 16)), np.ones((16, 16)))
    with pytest.raises(ValueError):       # the hybrid engine splits batches; one instance takes one engine
        native.solve_host(native.ProblemDesc.make(2, 16, 5, 1e-3, 0.1, c, engine="hybrid"), 1,
                          np.ones((16, 16)), np.ones((16, 16)))
    assert C.sizeof(native.ProblemDesc) == 8 * 4 + 2 * 8 + 23 * 8


def test_slab_arguments_on_host():
    # every call below fails its host-side checks before any device work
    L = native.load()
    assert [L.nls_slab_halo_rows(o) for o in (3, 5, 7)] == [4, 8, 12]
    assert L.nls_slab_halo_rows(4) == native.NLS_EINVAL
    c = np.zeros(23)
    c[10] = 1.0
    a, b, q = C.c_void_p(0x1000), C.c_void_p(0x2000), C.c_void_p(0x3000)
    p = native.ProblemDesc.make(2, 64, 5, 1e-3, 0.1, c, bc="periodic")
    step = lambda d, rows, *bufs: L.nls_slab_step(C.byref(d), rows, *bufs, None)
    assert step(p, 7, q, a, a, b, b) == native.NLS_EINVAL                 # fewer rows than the 8 ghost rows
    assert step(p, 32, q, a, a, a, b) == native.NLS_EINVAL                # psi in place
    assert step(p, 32, q, a, None, b, b) == native.NLS_EINVAL             # no reservoir input
    for bad in (native.ProblemDesc.make(2, 64, 5, 1e-3, 0.1, c, bc="ref"),
                native.ProblemDesc.make(1, 64, 5, 1e-3, 0.1, c, bc="periodic"),
                native.ProblemDesc.make(2, 64, 5, 1e-3, 0.1, c, bc="periodic", scheme="ssf")):
        assert step(bad, 32, q, a, a, b, b) == native.NLS_EUNSUPPORTED
