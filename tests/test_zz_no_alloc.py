"""nls_advance_static: nls_advance without allocation (its workspace is a
device array the library holds, launches issued directly).  Through the C
ABI, two calls in a row -- the second on a smaller problem, so the array
still holds the first call's larger contents -- each equal, bit for bit, to
nls_advance on the same inputs with a workspace of the caller's; problems
whose workspace exceeds the array are refused."""
import numpy as np
import pytest

from nls_amd import configs, native

pytestmark = pytest.mark.gpu


def inputs(dim, n, dtype, seed):
    rng = np.random.default_rng(seed)
    if dim == 1:
        x, h = configs.grid_1d(n, 0.05 * n)
        r2 = x ** 2
        shape = (n,)
    else:
        X, Y, h = configs.grid_2d(n, 0.1 * n)
        r2 = X ** 2 + Y ** 2
        shape = (n, n)
    psi0 = 0.3 * np.exp(-r2 / 8.0) + configs.noise(rng, shape, 1e-2)
    P = 2.0 * np.exp(-r2 / 10.0)
    n0 = 0.2 * np.exp(-r2 / 6.0)
    return h, P, psi0, n0


def solve(dim, n, dtype, engine, scheme, steps, static, seed):
    import torch
    dev = torch.device("cuda:0")
    h, P, psi0, n0 = inputs(dim, n, dtype, seed)
    c = configs.model_coeffs(dynamic=True)
    dt = 2e-4 if scheme == "rk4" else 1e-3
    d = native.ProblemDesc.make(dim, n, 5, dt, h, c, bc="periodic", dtype=dtype, engine=engine, scheme=scheme)
    rt = torch.float64 if dtype == "f64" else torch.float32
    ct = torch.complex128 if dtype == "f64" else torch.complex64
    Pd = torch.as_tensor(np.ascontiguousarray(P), dtype=rt, device=dev)
    psi = torch.as_tensor(np.ascontiguousarray(psi0), dtype=ct, device=dev)
    nres = torch.as_tensor(np.ascontiguousarray(n0), dtype=rt, device=dev)
    if static:
        native.advance_static(d, steps, Pd, psi, nres)
    else:
        ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device=dev)
        native.advance(d, steps, Pd, psi, nres, ws)
    torch.cuda.synchronize()
    return psi.cpu().numpy(), nres.cpu().numpy()


# (dim, first n, second n, dtype, engine, scheme): each engine of the plain path
CASES = [(2, 512, 200, "f32", "fused", "rk4"), (2, 256, 96, "f64", "fused", "rk4"),
         (2, 300, 130, "f64", "stage", "rk4"), (1, 6000, 5000, "f32", "stage", "rk4"),
         (1, 4096, 700, "f64", "resident", "rk4"), (2, 64, 40, "f32", "resident", "rk4"),
         (2, 512, 128, "f32", "auto", "ssf"), (1, 2048, 256, "f64", "auto", "ssf"),
         (2, 2048, 1024, "f32", "auto", "ssf"), (1, 4096, 1024, "f64", "auto", "ssf")]


@pytest.mark.parametrize("dim,n1,n2,dtype,engine,scheme", CASES)
def test_static_equals_advance_on_two_calls(gpu, dim, n1, n2, dtype, engine, scheme):
    steps = 131                    # odd, and past the 128 steps from which nls_advance replays graphs
    for k, n in enumerate((n1, n2)):
        psi_s, n_s = solve(dim, n, dtype, engine, scheme, steps, True, seed=k)
        psi_a, n_a = solve(dim, n, dtype, engine, scheme, steps, False, seed=k)
        assert np.all(np.isfinite(psi_s)), (n, k)
        np.testing.assert_array_equal(psi_s, psi_a, err_msg="call %d, n=%d" % (k + 1, n))
        np.testing.assert_array_equal(n_s, n_a, err_msg="call %d, n=%d" % (k + 1, n))


def test_static_refuses_problems_beyond_the_array(gpu):
    import torch
    dev = torch.device("cuda:0")
    n = 2048                       # per-stage engine, FP64: three state-sized buffers, 302 MB
    d = native.ProblemDesc.make(2, n, 5, 1e-4, 0.05, configs.model_coeffs(), bc="periodic", dtype="f64",
                                engine="stage")
    assert native.workspace_bytes(d) > native.STATIC_WORKSPACE_BYTES
    P = torch.zeros((n, n), dtype=torch.float64, device=dev)
    psi = torch.zeros((n, n), dtype=torch.complex128, device=dev)
    nres = torch.zeros((n, n), dtype=torch.float64, device=dev)
    with pytest.raises(ValueError, match="holds at most"):
        native.advance_static(d, 1, P, psi, nres)
