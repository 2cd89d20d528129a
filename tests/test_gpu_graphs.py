"""Step-block graphs of nls_advance.

A solve of >= 128 steps on the fused or per-stage engine replays a captured
block of 64 steps as one graph launch.  The replayed result must equal the
result of issuing every launch directly (bitwise), on every replay -- the
test replays the cached graph several times with direct work and stream
synchronisations in between -- and match the FP64 oracle.
"""
import numpy as np
import pytest

from nls_amd import configs, native

pytestmark = pytest.mark.gpu


def setup(dim, n, dtype, engine, dyn=True, seed=5):
    import torch
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
    P = 2.0 * np.exp(-r2 / 6.0)
    n0 = 0.2 * np.exp(-r2 / 3.0)
    c = configs.model_coeffs(dynamic=dyn)
    dt = 1e-3 if dim == 2 else 5e-4
    kw = dict(bc="periodic", dtype=dtype, engine=engine)
    d_graph = native.ProblemDesc.make(dim, n, 5, dt, h, c, **kw)
    d_eager = native.ProblemDesc.make(dim, n, 5, dt, h, c, graphs=False, **kw)
    rt = torch.float64 if dtype == "f64" else torch.float32
    ct = torch.complex128 if dtype == "f64" else torch.complex64
    dev = torch.device("cuda:0")
    Pt = torch.as_tensor(P, dtype=rt, device=dev)
    state = lambda: (torch.as_tensor(psi0, dtype=ct, device=dev).clone(),
                     torch.as_tensor(n0, dtype=rt, device=dev).clone())
    ws = torch.empty(max(1, native.workspace_bytes(d_graph)), dtype=torch.uint8, device=dev)
    ws2 = torch.empty(max(1, native.workspace_bytes(d_eager)), dtype=torch.uint8, device=dev)
    return d_graph, d_eager, Pt, state, ws, ws2, (P, psi0, n0, c, dt, h)


@pytest.mark.parametrize("dim,n,engine", [(2, 48, "fused"), (2, 40, "stage"), (1, 300, "stage"), (2, 130, "fused")])
@pytest.mark.parametrize("dtype", ("f64", "f32"))
def test_graph_replays_equal_direct_launches(gpu, dim, n, engine, dtype):
    import torch
    d_graph, d_eager, Pt, state, ws, ws2, _ = setup(dim, n, dtype, engine)
    steps = 200                                   # 3 graph blocks + 8 direct steps
    psi_e, n_e = state()
    native.advance(d_eager, steps, Pt, psi_e, n_e, ws2)
    torch.cuda.synchronize()
    ref_psi, ref_n = psi_e.cpu().numpy(), n_e.cpu().numpy()
    for rep in range(3):                          # first call captures, later ones hit the cache
        psi_g, n_g = state()
        native.advance(d_graph, steps, Pt, psi_g, n_g, ws)
        # direct work on the same stream between the replays
        psi_x, n_x = state()
        native.advance(d_eager, 5, Pt, psi_x, n_x, ws2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(psi_g.cpu().numpy(), ref_psi)
        np.testing.assert_array_equal(n_g.cpu().numpy(), ref_n)
        # the same buffers again: the state continues from where it stopped
        native.advance(d_graph, steps, Pt, psi_g, n_g, ws)
        native.advance(d_eager, steps, Pt, psi_e, n_e, ws2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(psi_g.cpu().numpy(), psi_e.cpu().numpy())
        psi_e, n_e = state()
        native.advance(d_eager, steps, Pt, psi_e, n_e, ws2)
        torch.cuda.synchronize()
    native.release()


def test_graph_solve_matches_oracle(gpu, oracle):
    import torch
    d_graph, _, Pt, state, ws, _, (P, psi0, n0, c, dt, h) = setup(2, 48, "f64", "fused")
    psi, nres = state()
    native.advance(d_graph, 300, Pt, psi, nres, ws)
    torch.cuda.synchronize()
    opsi, onres = oracle.rk4(2, 48, 5, 1, dt, h, 300, P, c, psi0, n0)
    rel = np.linalg.norm(psi.cpu().numpy() - opsi) / np.linalg.norm(opsi)
    assert rel <= 1e-10, rel
    reln = np.linalg.norm(nres.cpu().numpy() - onres) / np.linalg.norm(onres)
    assert reln <= 1e-10, reln


def test_graph_cache_eviction(gpu):
    """More distinct buffer sets than cache entries: every solve stays right."""
    import torch
    d_graph, d_eager, Pt, state, ws, ws2, _ = setup(2, 40, "f32", "fused")
    psi_e, n_e = state()
    native.advance(d_eager, 130, Pt, psi_e, n_e, ws2)
    torch.cuda.synchronize()
    ref = psi_e.cpu().numpy()
    for k in range(10):
        psi, nres = state()
        native.advance(d_graph, 130, Pt, psi, nres, ws)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(psi.cpu().numpy(), ref)
    native.release()


def test_graph_follows_stage_tile_height(gpu):
    """The cached graph of a per-stage solve is keyed by the tile height: after
    nls_tune_stage_tile the next solve runs kernels of the new height (same
    results, since the tile shape does not change a point's arithmetic)."""
    import torch
    d_graph, d_eager, Pt, state, ws, ws2, _ = setup(2, 96, "f64", "stage")
    old = native.stage_tile("f64")
    outs = []
    try:
        for rows in (8, 32, 16):
            native.stage_tile("f64", rows)
            psi, nres = state()
            native.advance(d_graph, 130, Pt, psi, nres, ws)
            torch.cuda.synchronize()
            outs.append(psi.cpu().numpy())
    finally:
        native.stage_tile("f64", old)
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])
    native.release()


def batch_members(dtype_dim=2):
    out = []
    for k, n in enumerate((48, 64, 80)):
        inst = configs.square(n, 0.2 * n, 2.0 + k, 3.0, 1.0, 30 + k, 1e-3, 0, "b%d" % k)
        out.append(inst)
    return out


@pytest.mark.parametrize("engine", ("fused", "stage", "hybrid"))
@pytest.mark.parametrize("dtype", ("f64", "f32"))
def test_batch_graph_replays_equal_direct_launches(gpu, engine, dtype):
    """A batch replays a captured 64-step block when every member's budget
    covers the call: bitwise equal to direct launches, on every replay.  A
    hybrid batch (48^2 and 64^2 resident, 80^2 fused) passes the graph
    setting to its fused part."""
    import torch
    from nls_amd.batch import Batch
    ms = batch_members()
    steps = 200                                   # 3 graph blocks + 8 direct steps
    eager = Batch(ms, dtype=dtype, engine=engine, graphs=False)
    eager.advance(steps)
    torch.cuda.synchronize()
    ref = [eager.member(i) for i in range(len(ms))]
    g = Batch(ms, dtype=dtype, engine=engine, graphs=True)
    g.prepare(steps)                              # captures the block now; every call below replays it
    for rep in range(3):
        g.reset()
        g.advance(steps)
        g.psi.mul_(1.0)                           # direct work on the same stream between replays
        torch.cuda.synchronize()
        for i in range(len(ms)):
            psi, nres = g.member(i)
            np.testing.assert_array_equal(psi, ref[i][0])
            np.testing.assert_array_equal(nres, ref[i][1])
    eager.close()
    g.close()


@pytest.mark.parametrize("engine", ("fused", "stage", "hybrid"))
def test_batch_budgets_inside_a_call_are_honoured(gpu, engine):
    """Budgets that end inside a call: the launches are issued directly, the
    step index of each is exact, and every member stops at its budget."""
    import torch
    from nls_amd.batch import Batch
    ms = batch_members()
    budgets = [200, 90, 150]
    g = Batch(ms, dtype="f32", engine=engine, graphs=True)
    g.set_steps(budgets)
    g.advance(200)
    torch.cuda.synchronize()
    for i, b in enumerate(budgets):
        one = Batch([ms[i]], dtype="f32", engine=engine, graphs=False)
        one.advance(b)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(g.member(i)[0], one.member(0)[0])
        one.close()
    g.close()
