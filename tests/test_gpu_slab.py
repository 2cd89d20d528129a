"""Row slabs of one periodic grid (nls_slab_step), all held on one device.

A grid split into row slabs with 4R ghost rows, exchanged once per step, must
give the whole-grid fused solve bit for bit: one step at a point depends only
on the state within 4R rows, which the ghost rows supply, and every point is
computed by the same arithmetic in either form.  The whole-grid fused path is
itself pinned to the FP64 oracle; one case checks the slabs against the
oracle directly as well.  Shapes cover an equal split, uneven slabs whose
heights are not a multiple of the kernel's tile height, a single slab (its own
neighbour), and a grid narrower than a tile plus its halo (far wrap in x).
"""
import numpy as np
import pytest

from nls_amd import configs, native
from nls_amd.slab import SlabGrid

pytestmark = pytest.mark.gpu


def case(n, order, dtype, dyn, seed=11):
    import torch
    rng = np.random.default_rng(seed)
    X, Y, h = configs.grid_2d(n, 0.08 * n)
    r2 = X ** 2 + Y ** 2
    psi0 = 0.3 * np.exp(-r2 / 4.0) + configs.noise(rng, (n, n), 1e-2)
    P = 2.0 * np.exp(-r2 / 6.0)
    n0 = 0.2 * np.exp(-r2 / 3.0) if dyn else None
    c = configs.model_coeffs(dynamic=dyn)
    desc = native.ProblemDesc.make(2, n, order, 1e-3, h, c, bc="periodic", dtype=dtype, engine="fused")
    rt = torch.float64 if dtype == "f64" else torch.float32
    ct = torch.complex128 if dtype == "f64" else torch.complex64
    dev = torch.device("cuda:0")
    t = lambda a, ty: None if a is None else torch.as_tensor(a, dtype=ty, device=dev).clone()
    return desc, t(P, rt), (lambda: (t(psi0, ct), t(n0, rt))), (P, psi0, n0, c, h)


def whole_grid(desc, Pt, state, steps):
    import torch
    psi, nres = state()
    ws = torch.empty(max(1, native.workspace_bytes(desc)), dtype=torch.uint8, device=psi.device)
    native.advance(desc, steps, Pt, psi, nres, ws)
    torch.cuda.synchronize()
    return psi, nres


@pytest.mark.parametrize("n,order,dtype,dyn,rows", [
    (256, 5, "f64", True, 4),
    (256, 5, "f32", True, [64, 64, 64, 64]),
    (200, 7, "f32", True, [72, 64, 64]),
    (200, 3, "f64", False, [72, 64, 64]),
    (130, 5, "f32", True, [40, 50, 40]),
    (96, 5, "f32", True, 1),
    (48, 5, "f64", True, [24, 24]),
    (48, 3, "f32", False, [8, 8, 16, 16]),
    (640, 5, "f64", True, [320, 320]),        # whole grid in 64 x 32 tiles, slabs in 64 x 16
    (640, 5, "f32", True, [320, 320]),        # the same
    (1152, 5, "f64", True, [576, 576]),       # whole grid in 512-thread 64 x 16 tiles, slabs in 64 x 32
])
def test_slabs_equal_whole_grid(gpu, n, order, dtype, dyn, rows):
    import torch
    desc, Pt, state, _ = case(n, order, dtype, dyn)
    steps = 20
    ref_psi, ref_n = whole_grid(desc, Pt, state, steps)
    psi0, n0 = state()
    g = SlabGrid(desc, rows, Pt, psi0, n0)
    g.step(steps)
    psi, nres = g.gather()
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(psi)).all()
    assert torch.equal(psi, ref_psi), float((psi - ref_psi).abs().max())
    if dyn:
        assert torch.equal(nres, ref_n)


def test_slabs_match_oracle_f64(gpu, oracle):
    n, order, steps = 64, 5, 100
    desc, Pt, state, (P, psi0, n0, c, h) = case(n, order, "f64", True)
    p, q = state()
    g = SlabGrid(desc, [24, 16, 24], Pt, p, q)
    g.step(steps)
    psi, nres = g.gather()
    opsi, onres = oracle.rk4(2, n, order, 1, 1e-3, h, steps, P, c, psi0, n0)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    assert rel(psi.cpu().numpy(), opsi) <= 1e-10
    assert rel(nres.cpu().numpy(), onres) <= 1e-10


def test_slab_arguments(gpu):
    desc, Pt, state, _ = case(64, 5, "f32", True)
    psi, nres = state()
    with pytest.raises(ValueError):
        SlabGrid(desc, [4, 60], Pt, psi, nres)          # a slab thinner than its 8 ghost rows
    with pytest.raises(ValueError):
        SlabGrid(desc, [32, 16], Pt, psi, nres)         # heights do not cover the grid
    g = SlabGrid(desc, 2, Pt, psi, nres)
    with pytest.raises(ValueError):                     # in place: input and output the same buffer
        native.slab_step(desc, 32, g.P[0], g.cur[0], g.ncur[0], g.cur[0], g.nalt[0])
    assert g.ghost_bytes_per_step() == 2 * 8 * 64 * (8 + 4)
