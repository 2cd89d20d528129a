"""The continuous-batching request loop (nls_amd/serving.py).

CPU: the pool allocator.  GPU: requests served together -- arriving at
different times, with different sizes, dimensionalities and step counts --
give exactly the result each gives when solved alone, and that result is
within FP32 tolerance of the FP64 oracle."""
import numpy as np
import pytest

from nls_amd import serving


def test_pool_allocator_first_fit_and_merge():
    a = serving.PoolAllocator(64 * 10)
    x = a.alloc(100)          # 128
    y = a.alloc(64)
    z = a.alloc(200)          # 256
    assert (x, y, z) == (0, 128, 192)
    assert a.alloc(64 * 10) is None
    a.release(y, 64)
    assert a.alloc(50) == 128
    a.release(128, 64)
    a.release(x, 100)
    a.release(z, 200)
    assert a.free == [(0, 640)]
    assert a.alloc(640) == 0


def test_stable_dt():
    h = 0.125
    assert serving.stable_dt(2, h) == pytest.approx(0.5 * 2 * np.sqrt(2) / (2 * 16 / 3 / h ** 2))


def test_requests_follow_the_workload_recipe():
    reqs = serving.poisson_requests(200.0, 1.0, seed=3)
    assert 120 < len(reqs) < 300
    dims = [r.dim for r in reqs]
    assert 0.3 < np.mean(np.array(dims) == 1) < 0.7
    for r in reqs:
        assert r.n % 64 == 0
        if r.dim == 1:
            assert 256 <= r.n <= 4096
        else:
            assert 64 <= r.n <= 512
        assert 1000 <= r.steps <= 100000
    assert all(b.arrival > a.arrival for a, b in zip(reqs, reqs[1:]))


@pytest.mark.gpu
def test_result_alone_equals_result_in_the_mix(gpu, oracle):
    from nls_amd.batch import Batch
    rng = np.random.default_rng(8)
    reqs = []
    for i in range(10):
        r = serving.make_request(i, 0.0 if i < 6 else 0.05 * i, rng, step_range=(40, 400))
        if r.dim == 2 and r.n > 192:
            continue
        reqs.append(r)
    eng = serving.Engine(budget_bytes=1 << 30, target_block_ms=2.0, min_block=8, max_block=64)
    done = eng.run(reqs)
    eng.close()
    assert len(done) == len(reqs)
    for r in done:
        eng_name = eng.groups[eng.group_of(r)].engine
        alone = Batch([r], dim=r.dim, order=5, bc="periodic", dtype="f32", engine=eng_name)
        alone.advance(r.steps)
        psi, nres = alone.member(0)
        alone.close()
        np.testing.assert_array_equal(r.psi, psi)
        np.testing.assert_array_equal(r.nres, nres)
    # and the served FP32 result is within tolerance of the FP64 oracle
    r = min(done, key=lambda q: q.points)
    opsi, _ = oracle.rk4(r.dim, r.n, 5, 1, r.dt, r.dx, r.steps, r.P, r.coeffs, r.psi0, r.n0)
    assert np.linalg.norm(r.psi - opsi) / np.linalg.norm(opsi) <= 1e-4
    assert all(r.latency > 0 for r in done)
    # every result carries its FP64 observables, from the state it returns
    for r in done:
        dV = r.dx if r.dim == 1 else r.dx * r.dx
        s2 = np.abs(r.psi.astype(np.complex128)) ** 2
        assert r.observables["nonfinite"] == 0
        assert np.isfinite(r.observables["energy"])
        assert r.observables["norm"] == pytest.approx(s2.sum() * dV, rel=1e-12)
        assert r.observables["population"] == pytest.approx(r.nres.astype(np.float64).sum() * dV, rel=1e-12)


def _request_2d(rid, n, steps, seed, dt=None, length=32.0, arrival=0.0):
    from nls_amd import configs
    rng = np.random.default_rng(seed)
    X, Y, h = configs.grid_2d(n, length)
    r = np.sqrt(X ** 2 + Y ** 2)
    P = configs.ring(r, rng.uniform(1.0, 4.0), rng.uniform(3.0, 6.0), rng.uniform(0.5, 1.5))
    psi0 = 0.1 * np.exp(-r ** 2 / 4.0) + configs.noise(rng, r.shape)
    return serving.Request(rid, arrival, 2, n, steps, serving.stable_dt(2, h) if dt is None else dt, h,
                           configs.C_MODEL.copy(), P, psi0, np.zeros(r.shape))


def _request_1d(rid, n, steps, seed, arrival=0.0):
    from nls_amd import configs
    rng = np.random.default_rng(seed)
    x, h = configs.grid_1d(n, 64.0)
    P = rng.uniform(1.0, 4.0) * np.exp(-x ** 2 / (2 * rng.uniform(1.0, 3.0) ** 2))
    psi0 = 0.1 * np.exp(-x ** 2 / 4.0) + configs.noise(rng, x.shape)
    return serving.Request(rid, arrival, 1, n, steps, serving.stable_dt(1, h), h, configs.C_MODEL.copy(), P, psi0,
                           np.zeros(x.shape))


@pytest.mark.gpu
def test_large_grid_and_long_request_match_oracle(gpu, oracle):
    """The request loop in double precision with the sizes of the serving
    mix's large pool: a 512 x 512 request (fused pool) beside smaller ones,
    and a 1-D request of 1300 steps, both with budgets ending inside a block
    (blocks of at most 256 steps); each result against the oracle's solve of
    the same request (relative L2 <= 1e-10)."""
    import os
    reqs = [_request_2d(0, 512, 300, seed=1), _request_1d(1, 1024, 1300, seed=2),
            _request_2d(2, 192, 700, seed=3, arrival=0.01), _request_2d(3, 64, 500, seed=4)]
    eng = serving.Engine(budget_bytes=2 << 30, dtype="f64", target_block_ms=2.0, min_block=8, max_block=256)
    done = eng.run(reqs)
    eng.close()
    assert len(done) == len(reqs) and all(r.error is None for r in done)
    assert eng.blocks > 6
    oracle.set_threads(min(16, os.cpu_count() or 1))
    for r in done:
        if r.rid not in (0, 1):
            continue
        opsi, onres = oracle.rk4(r.dim, r.n, 5, 1, r.dt, r.dx, r.steps, r.P, r.coeffs, r.psi0, r.n0)
        err = np.linalg.norm(r.psi - opsi) / np.linalg.norm(opsi)
        assert err <= 1e-10, (r.rid, err)
        assert np.linalg.norm(r.nres - onres) / np.linalg.norm(onres) <= 1e-10


@pytest.mark.gpu
def test_blown_up_request_is_evicted_and_others_unchanged(gpu):
    """A request whose dt is far above the RK4 stability limit blows up; the
    loop evicts it once its FP64 observables count non-finite points and
    returns it with error NLS_ENONFINITE, while the healthy requests beside it
    in the same pool finish bit for bit as when served without it."""
    from nls_amd import native

    def healthy():
        return [_request_2d(0, 128, 600, seed=11), _request_2d(1, 96, 400, seed=12), _request_1d(2, 512, 500, seed=13)]

    bad = _request_2d(9, 128, 5000, seed=14)
    bad.dt = 40.0 * bad.dt
    kw = dict(budget_bytes=1 << 30, target_block_ms=2.0, min_block=8, max_block=64)
    eng = serving.Engine(**kw)
    done = {r.rid: r for r in eng.run(healthy() + [bad])}
    eng.close()
    ref_eng = serving.Engine(**kw)
    ref = {r.rid: r for r in ref_eng.run(healthy())}
    ref_eng.close()
    b = done[9]
    assert b.error == native.NLS_ENONFINITE and b.psi is None
    assert b.observables["nonfinite"] > 0
    assert b.remaining > 0                      # evicted before its steps ran out
    for rid in (0, 1, 2):
        assert done[rid].error is None
        np.testing.assert_array_equal(done[rid].psi, ref[rid].psi)
        np.testing.assert_array_equal(done[rid].nres, ref[rid].nres)


@pytest.mark.gpu
def test_batch_rate_runs_every_pool(gpu):
    """The batch engine's rate on a set of requests' grids (the reference
    the loop's saturated rate is reported against) runs the three pools'
    batches and counts their work."""
    reqs = [_request_2d(0, 128, 100, seed=21), _request_2d(1, 48, 100, seed=22), _request_1d(2, 1024, 100, seed=23),
            _request_2d(3, 256, 100, seed=24)]
    rate = serving.batch_rate(reqs, steps=32)
    assert np.isfinite(rate) and rate > 1e6
