"""GPU parity of the batch engines (nls_batch_* through the C ABI).

Every member of a batch must match the oracle run on that member alone:
FP64 to relative L2 <= 1e-10, FP32 to <= 1e-4 against the FP64 oracle; a
member's result must not depend on what else is in the batch (bitwise).
"""
import numpy as np
import pytest

from nls_amd import configs
from nls_amd.batch import Batch

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-10, 1e-4


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def members_2d(sizes, dyn=True, seed=11, ref=False, spacing=None):
    out = []
    for i, n in enumerate(sizes):
        rng = np.random.default_rng(seed + i)
        L = 12.0 + 4.0 * (i % 3) if spacing is None else spacing * n
        X, Y, h = configs.grid_2d(n, L)
        r = np.sqrt(X ** 2 + Y ** 2)
        psi0 = 0.2 * np.exp(-r ** 2 / 5.0) + configs.noise(rng, r.shape, 1e-2)
        P = configs.ring(r, rng.uniform(1.0, 4.0), rng.uniform(1.0, 3.0), rng.uniform(0.5, 1.5))
        c = configs.model_coeffs(dynamic=dyn, g=rng.uniform(0.5, 1.5))
        n0 = 0.1 * np.exp(-r ** 2 / 4.0)
        out.append(configs.Instance("m%d" % i, 2, n, h, rng.choice([5e-4, 1e-3]), 0, P, psi0, n0, c))
    return out


def members_1d(sizes, dyn=True, seed=21, ref=False):
    out = []
    for i, n in enumerate(sizes):
        rng = np.random.default_rng(seed + i)
        if ref:
            h = 0.1
            x = np.arange(n) * h
        else:
            x, h = configs.grid_1d(n, 0.1 * n)
        psi0 = 0.3 * np.exp(-x ** 2 / 4.0) + configs.noise(rng, x.shape, 1e-2)
        P = 2.0 * np.exp(-x ** 2 / rng.uniform(4.0, 10.0))
        c = configs.model_coeffs(dynamic=dyn)
        out.append(configs.Instance("l%d" % i, 1, n, h, 5e-4, 0, P, psi0, 0.1 * np.exp(-x ** 2), c))
    return out


def oracle_member(oracle, m, steps, bc, dim=2, order=5):
    return oracle.rk4(dim, m.n, order, 0 if bc == "ref" else 1, m.dt, m.dx, steps, m.P, m.coeffs, m.psi0, m.n0)


def test_hybrid_members_equal_their_engine_alone(gpu):
    """The hybrid engine runs the members that fit one workgroup (2-D n <= 64)
    on the resident engine on a stream of its own and the others on the fused
    engine: each member equals the same member batched alone on its engine,
    bit for bit, with step budgets, and so do its FP64 observables.  AUTO
    picks the hybrid engine for such a mix (same results)."""
    ms = members_2d([64, 128, 40, 96, 64, 33, 200], seed=7)
    budgets = [60, 60, 25, 60, 0, 41, 60]
    mix = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine="hybrid")
    assert mix.tiles == sum(((m.n + 63) // 64) * ((m.n + 15) // 16) for m in ms if m.n > 64)
    mix.set_steps(budgets)
    mix.advance(60)
    auto = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine="auto")
    auto.set_steps(budgets)
    auto.advance(60)
    obs = mix.observables()
    for i, m in enumerate(ms):
        alone = Batch([m], dim=2, order=5, bc="periodic", dtype="f32", engine="resident" if m.n <= 64 else "fused")
        alone.set_steps([budgets[i]])
        alone.advance(60)
        a, an = alone.member(0)
        b, bn = mix.member(i)
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(an, bn)
        np.testing.assert_array_equal(auto.member(i)[0], b)
        np.testing.assert_array_equal(alone.observables()[0], obs[i])
        if budgets[i] == 0:
            np.testing.assert_array_equal(b, m.psi0.astype(np.complex64))
        alone.close()
    auto.close()
    mix.close()


RAGGED = [131, 333, 200, 257, 190, 300, 150, 221, 289, 177, 310, 243, 162, 275, 199, 321]


@pytest.mark.parametrize("dtype,bc,rows", [("f64", "periodic", 16), ("f64", "ref", 32), ("f32", "periodic", 32)])
def test_fused_tile_forms_on_ragged_grids(gpu, oracle, dtype, bc, rows):
    """Grid sides that are not multiples of the tile (131 .. 333) in the forms
    a launch of many tiles takes on MI355X (572 64 x 32 tiles, more than two
    rounds of 256 CUs): double precision periodic in 512-thread 64 x 16 tiles,
    reflecting in 64 x 32 tiles, against the oracle; single precision in
    64 x 32 tiles against each member alone (64 x 16 tiles), bit for bit."""
    steps = 20
    ms = members_2d(RAGGED, seed=50, spacing=0.15)
    b = Batch(ms, dim=2, order=5, bc=bc, dtype=dtype, engine="fused")
    assert b.tiles == sum(((n + 63) // 64) * ((n + rows - 1) // rows) for n in RAGGED)
    b.advance(steps)
    oracle.set_threads(8)
    for i in (0, 1, 3, 8, 15):
        psi, nres = b.member(i)
        assert np.all(np.isfinite(psi))
        if dtype == "f64":
            opsi, onres = oracle_member(oracle, ms[i], steps, bc)
            assert rel(psi, opsi) <= TOL64, (i, ms[i].n, rel(psi, opsi))
            assert rel(nres, onres) <= TOL64
        else:
            alone = Batch([ms[i]], dim=2, order=5, bc=bc, dtype=dtype, engine="fused")
            alone.advance(steps)
            np.testing.assert_array_equal(alone.member(0)[0], psi)
            np.testing.assert_array_equal(alone.member(0)[1], nres)
            alone.close()
    b.close()


@pytest.mark.parametrize("sizes,dtype,order,engine", [([32, 64, 48], "f32", 5, "resident"),
                                                     ([96, 128], "f32", 5, "fused"),
                                                     ([32, 96], "f64", 7, "fused")])
def test_hybrid_with_one_part_is_that_engine(gpu, sizes, dtype, order, engine):
    """A hybrid batch whose members all fit the resident engine, or none do
    (double precision has no 2-D order-7 resident variant), is a batch of the
    one engine that holds them: bitwise the same results."""
    ms = members_2d(sizes, seed=3)
    a = Batch(ms, dim=2, order=order, bc="periodic", dtype=dtype, engine="hybrid")
    b = Batch(ms, dim=2, order=order, bc="periodic", dtype=dtype, engine=engine)
    assert a.tiles == b.tiles
    a.advance(40)
    b.advance(40)
    for i in range(len(ms)):
        assert np.all(np.isfinite(a.member(i)[0]))
        np.testing.assert_array_equal(a.member(i)[0], b.member(i)[0])
    a.close()
    b.close()


@pytest.mark.parametrize("bc", ("ref", "periodic"))
def test_hybrid_1d_f64_matches_oracle(gpu, oracle, bc):
    """1-D lines of up to 4096 points on the resident engine, longer ones on
    the per-stage engine, in one hybrid batch."""
    ms = members_1d([300, 5000, 4096, 6000], ref=(bc == "ref"))
    b = Batch(ms, dim=1, order=5, bc=bc, dtype="f64", engine="hybrid")
    b.advance(150)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        opsi, onres = oracle.rk4(1, m.n, 5, 0 if bc == "ref" else 1, m.dt, m.dx, 150, m.P, m.coeffs, m.psi0, m.n0)
        assert rel(psi, opsi) <= TOL64, (i, rel(psi, opsi))
        assert rel(nres, onres) <= TOL64
    b.close()


@pytest.mark.parametrize("dtype,bc,count,rows", [("f32", "periodic", 20, 32), ("f64", "periodic", 20, 16),
                                                 ("f64", "periodic", 12, 32), ("f64", "ref", 20, 32)])
@pytest.mark.parametrize("order", (3, 5))
def test_fused_tile_forms_agree(gpu, oracle, dtype, bc, count, rows, order):
    """The fused kernel's tile form is chosen per launch against the 256
    compute units of MI355X: 1024-thread 64 x 16 tiles when they fit one a
    CU (each member alone here: 64 of them), else 1024-thread 64 x 32 tiles,
    except in double precision on periodic grids when those take more than
    two rounds of the CUs (20 members: 640), which runs 512-thread 64 x 16
    tiles, two workgroups a CU.  A point's update is the same arithmetic in every form:
    each member of the batch equals the member alone bit for bit.  In double
    precision a member of the batch also matches the oracle."""
    steps = 30
    ms = members_2d([256] * count, seed=40, spacing=0.15)
    mix = Batch(ms, dim=2, order=order, bc=bc, dtype=dtype, engine="fused")
    assert mix.tiles == count * 4 * (256 // rows)
    mix.advance(steps)
    for i in (0, count - 1):
        alone = Batch([ms[i]], dim=2, order=order, bc=bc, dtype=dtype, engine="fused")
        assert alone.tiles == 4 * 16                  # 64 x 16 tiles
        alone.advance(steps)
        a, an = alone.member(0)
        b, bn = mix.member(i)
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(an, bn)
        alone.close()
    if dtype == "f64":
        psi, nres = mix.member(0)
        opsi, onres = oracle_member(oracle, ms[0], steps, bc, order=order)
        assert rel(psi, opsi) <= TOL64, rel(psi, opsi)
        assert rel(nres, onres) <= TOL64
    mix.close()


@pytest.mark.parametrize("engine", ("stage", "fused", "hybrid"))
@pytest.mark.parametrize("bc", ("periodic", "ref"))
@pytest.mark.parametrize("dyn", (True, False))
def test_batch_2d_f64_matches_oracle(gpu, oracle, engine, bc, dyn):
    ms = members_2d([16, 40, 64, 72, 100, 33], dyn=dyn)
    steps = 150
    b = Batch(ms, dim=2, order=5, bc=bc, dtype="f64", engine=engine)
    b.advance(steps)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        opsi, onres = oracle_member(oracle, m, steps, bc)
        assert rel(psi, opsi) <= TOL64, (i, m.n, rel(psi, opsi))
        if dyn:
            assert rel(nres, onres) <= TOL64
    b.close()


@pytest.mark.parametrize("order", (3, 7))
def test_batch_fused_orders(gpu, oracle, order):
    ms = members_2d([48, 64], dyn=True)
    b = Batch(ms, dim=2, order=order, bc="periodic", dtype="f64", engine="fused")
    b.advance(100)
    for i, m in enumerate(ms):
        psi, _ = b.member(i)
        opsi, _ = oracle.rk4(2, m.n, order, 1, m.dt, m.dx, 100, m.P, m.coeffs, m.psi0, m.n0)
        assert rel(psi, opsi) <= TOL64


@pytest.mark.parametrize("bc", ("ref", "periodic"))
def test_batch_1d_f64_matches_oracle(gpu, oracle, bc):
    ms = members_1d([40, 256, 300, 1000], ref=(bc == "ref"))
    b = Batch(ms, dim=1, order=5, bc=bc, dtype="f64", engine="stage")
    b.advance(200)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        opsi, onres = oracle_member(oracle, m, 200, bc, dim=1)
        assert rel(psi, opsi) <= TOL64, (i, rel(psi, opsi))


@pytest.mark.parametrize("engine", ("stage", "fused", "hybrid"))
def test_step_budgets(gpu, oracle, engine):
    ms = members_2d([32, 64, 48, 80], dyn=True)
    budgets = [0, 37, 100, 120]
    b = Batch(ms, dim=2, order=5, bc="periodic", dtype="f64", engine=engine)
    b.set_steps(budgets)
    b.advance(120)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        if budgets[i] == 0:
            np.testing.assert_array_equal(psi, m.psi0)
            continue
        opsi, onres = oracle_member(oracle, m, budgets[i], "periodic")
        assert rel(psi, opsi) <= TOL64, (i, rel(psi, opsi))
        assert rel(nres, onres) <= TOL64


def test_fused_f32_sweep_members_within_tolerance(gpu, oracle):
    ms = configs.sweep512(count=6, steps=400)
    b = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine="fused")
    b.advance(400)
    oracle.set_threads(8)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        opsi, onres = oracle_member(oracle, m, 400, "periodic")
        assert rel(psi, opsi) <= TOL32, (i, m.n, rel(psi, opsi))
        assert rel(nres, onres) <= TOL32


@pytest.mark.parametrize("engine", ("stage", "fused"))
def test_member_alone_equals_member_in_mix(gpu, engine):
    ms = configs.sweep512(count=8, steps=50)
    mix = Batch(ms, dim=2, order=5, bc="periodic", dtype="f32", engine=engine)
    mix.set_steps([50, 20, 50, 50, 7, 50, 50, 33])
    mix.advance(50)
    for i in (0, 3, 6):
        alone = Batch([ms[i]], dim=2, order=5, bc="periodic", dtype="f32", engine=engine)
        alone.advance(50)
        a, an = alone.member(0)
        b, bn = mix.member(i)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(an, bn)
        alone.close()
    mix.close()


@pytest.mark.parametrize("bc", ("periodic", "ref"))
@pytest.mark.parametrize("dyn", (True, False))
@pytest.mark.parametrize("order", (3, 5))
def test_resident_2d_f64_matches_oracle(gpu, oracle, bc, dyn, order):
    ms = members_2d([16, 40, 64, 33], dyn=dyn)
    b = Batch(ms, dim=2, order=order, bc=bc, dtype="f64", engine="resident")
    b.advance(150)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        opsi, onres = oracle.rk4(2, m.n, order, 0 if bc == "ref" else 1, m.dt, m.dx, 150, m.P, m.coeffs, m.psi0,
                                 m.n0)
        assert rel(psi, opsi) <= TOL64, (i, m.n, rel(psi, opsi))
        if dyn:
            assert rel(nres, onres) <= TOL64
    b.close()


@pytest.mark.parametrize("bc", ("ref", "periodic"))
@pytest.mark.parametrize("order", (3, 5, 7))
def test_resident_1d_f64_matches_oracle(gpu, oracle, bc, order):
    ms = members_1d([40, 256, 1000, 4096], ref=(bc == "ref"))
    b = Batch(ms, dim=1, order=order, bc=bc, dtype="f64", engine="resident")
    b.advance(200)
    for i, m in enumerate(ms):
        psi, nres = b.member(i)
        opsi, onres = oracle.rk4(1, m.n, order, 0 if bc == "ref" else 1, m.dt, m.dx, 200, m.P, m.coeffs, m.psi0,
                                 m.n0)
        assert rel(psi, opsi) <= TOL64, (i, rel(psi, opsi))
        assert rel(nres, onres) <= TOL64


def test_resident_budgets_and_invariance(gpu, oracle):
    ms = configs.sweep512(count=12, steps=60)
    small = [m for m in ms if m.n == 64][:4]
    b = Batch(small, dim=2, order=5, bc="periodic", dtype="f32", engine="resident")
    b.set_steps([60, 0, 17, 59])
    b.advance(60)
    np.testing.assert_array_equal(b.member(1)[0], small[1].psi0.astype(np.complex64))
    for i in (0, 2, 3):
        alone = Batch([small[i]], dim=2, order=5, bc="periodic", dtype="f32", engine="resident")
        alone.advance([60, 0, 17, 59][i])
        np.testing.assert_array_equal(alone.member(0)[0], b.member(i)[0])
        alone.close()
    opsi, _ = oracle.rk4(2, 64, 5, 1, small[2].dt, small[2].dx, 17, small[2].P, small[2].coeffs, small[2].psi0,
                         small[2].n0)
    assert rel(b.member(2)[0], opsi) <= TOL32
    b.close()


@pytest.mark.parametrize("dtype", ("f32", "f64"))
@pytest.mark.parametrize("dim", (1, 2))
@pytest.mark.parametrize("steps", (61, 129))
def test_hybrid_odd_step_counts(gpu, dtype, dim, steps):
    """An advance of an odd number of steps (61: direct launches; 129: two
    64-step graph replays and one direct step) leaves the fused or per-stage
    part's state in the workspace.  Only that part's own points may go back:
    the resident members sit between them in the packed arrays and are
    stepped concurrently on the batch's side stream.  Every member equals the
    same member alone on its engine, bit for bit, after one such call and
    after a second one."""
    if dim == 2:
        ms = members_2d([64, 128, 40, 96, 64, 33, 200], seed=17, spacing=0.15)
        rest, fits = "fused", (lambda m: m.n <= 64)
    else:
        ms = members_1d([300, 5000, 4096, 6000, 1000], seed=31)
        rest, fits = "stage", (lambda m: m.n <= 4096)
    assert any(fits(m) for m in ms) and not all(fits(m) for m in ms)
    mix = Batch(ms, dim=dim, order=5, bc="periodic", dtype=dtype, engine="hybrid")
    alone = [Batch([m], dim=dim, order=5, bc="periodic", dtype=dtype, engine="resident" if fits(m) else rest)
             for m in ms]
    for call in range(2):
        mix.advance(steps)
        for i, a in enumerate(alone):
            a.advance(steps)
            psi, nres = mix.member(i)
            apsi, anres = a.member(0)
            assert np.all(np.isfinite(apsi)), (call, i)
            np.testing.assert_array_equal(apsi, psi, err_msg="call %d member %d" % (call, i))
            np.testing.assert_array_equal(anres, nres, err_msg="call %d member %d" % (call, i))
    for a in alone:
        a.close()
    mix.close()
