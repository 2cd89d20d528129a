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
