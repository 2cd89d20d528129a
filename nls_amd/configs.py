"""Synthetic inputs for the standard workloads of the time-stepping core.

Model (all workloads): i psi_t = -lap psi + g|psi|^2 psi + 2 g_R n psi
+ (i/2)(R n - gamma_c) psi and n_t = P - (gamma_R + R|psi|^2) n with g = 1,
g_R = 2, R = 1, gamma_c = 1, gamma_R = 1.5.  In the reference's coefficient
vector (nls.f90:643-647 for psi, :580 for the reservoir) that is c3 = R/2,
c4 = gamma_c/2, c5 = g, c6 = 2 g_R, c11 = 1, c12 = 1, c13 = gamma_R,
c14 = R.

Initial state: psi0 = 0.1 exp(-r^2/4) + seeded complex Gaussian noise of
total std 1e-3 (each component 1e-3/sqrt(2)); n0 = 0.

Workloads (name: grid, domain, pumping, dt x steps):
  line256   1-D N=256 on [-10,10) periodic, Gaussian P = 2 exp(-x^2/8),
            dt = 1e-3 x 2000, seed 0
  square256 2-D 256^2 on [-16,16)^2 periodic, ring P = 3 exp(-(r-6)^2/2),
            dt = 5e-4 x 10^4, seed 1
  square2048 2-D 2048^2 on [-64,64)^2 periodic, ring P0 = 3, r0 = 20, w = 2,
            dt = 2.5e-4 x 2*10^4, seed 2
  sweep512  512 independent 2-D instances on [-16,16)^2, side drawn from
            {64, 128, 256} with p = 0.5/0.3/0.2, ring P0 ~ U[0.5, 5],
            r0 ~ U[3, 8], w ~ U[0.5, 2], seed 1000 + i, dt = 5e-4 x 5000
All use the order-5 stencil.
"""
import numpy as np

C_MODEL = np.zeros(23)
C_MODEL[2], C_MODEL[3], C_MODEL[4], C_MODEL[5] = 0.5, 0.5, 1.0, 4.0
C_MODEL[10], C_MODEL[11], C_MODEL[12], C_MODEL[13] = 1.0, 1.0, 1.5, 1.0


def model_coeffs(g=1.0, g_R=2.0, R=1.0, gamma_c=1.0, gamma_R=1.5, dynamic=True):
    """Coefficient vector of the model above (c11 = 0 gives the algebraic
    reservoir limit, which the reference runs unchanged)."""
    c = np.zeros(23)
    c[2], c[3], c[4], c[5] = R / 2.0, gamma_c / 2.0, g, 2.0 * g_R
    c[10], c[11], c[12], c[13] = (1.0 if dynamic else 0.0), 1.0, gamma_R, R
    return c


def grid_1d(n, length=20.0):
    x = -length / 2 + length * np.arange(n) / n
    return x, length / n


def grid_2d(n, length=32.0):
    x = -length / 2 + length * np.arange(n) / n
    return x[:, None], x[None, :], length / n


def noise(rng, shape, std=1e-3):
    xi = rng.standard_normal((2,) + tuple(shape))
    return (std / np.sqrt(2.0)) * (xi[0] + 1j * xi[1])


def ring(r, P0, r0, w):
    return P0 * np.exp(-(r - r0) ** 2 / (2.0 * w * w))


class Instance(object):
    """One solve: grid, inputs and stepping."""

    def __init__(self, name, dim, n, dx, dt, steps, P, psi0, n0, coeffs, order=5, bc="periodic", params=None):
        self.name, self.dim, self.n, self.dx, self.dt, self.steps = name, dim, n, dx, dt, steps
        self.P, self.psi0, self.n0, self.coeffs = P, psi0, n0, coeffs
        self.order, self.bc = order, bc
        self.params = params or {}

    @property
    def points(self):
        return self.n if self.dim == 1 else self.n * self.n

    @property
    def work(self):
        return self.points * self.steps


def line256(seed=0):
    x, h = grid_1d(256, 20.0)
    rng = np.random.default_rng(seed)
    psi0 = 0.1 * np.exp(-x ** 2 / 4.0) + noise(rng, x.shape)
    P = 2.0 * np.exp(-x ** 2 / 8.0)
    return Instance("line256", 1, 256, h, 1e-3, 2000, P, psi0, np.zeros_like(x), C_MODEL.copy())


def square(n, length, P0, r0, w, seed, dt, steps, name):
    X, Y, h = grid_2d(n, length)
    rng = np.random.default_rng(seed)
    r2 = X ** 2 + Y ** 2
    psi0 = 0.1 * np.exp(-r2 / 4.0) + noise(rng, r2.shape)
    P = ring(np.sqrt(r2), P0, r0, w)
    return Instance(name, 2, n, h, dt, steps, P, psi0, np.zeros(r2.shape), C_MODEL.copy(),
                    params={"P0": P0, "r0": r0, "w": w})


def square256(seed=1):
    return square(256, 32.0, 3.0, 6.0, 1.0, seed, 5e-4, 10000, "square256")


def square2048(seed=2):
    return square(2048, 128.0, 3.0, 20.0, 2.0, seed, 2.5e-4, 20000, "square2048")


def sweep512(count=512, steps=5000, dt=5e-4, base_seed=1000):
    """The batched parameter sweep: list of Instance (2-D, periodic)."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(base_seed + i)
        side = int(rng.choice([64, 128, 256], p=[0.5, 0.3, 0.2]))
        P0 = rng.uniform(0.5, 5.0)
        r0 = rng.uniform(3.0, 8.0)
        w = rng.uniform(0.5, 2.0)
        X, Y, h = grid_2d(side, 32.0)
        r2 = X ** 2 + Y ** 2
        psi0 = 0.1 * np.exp(-r2 / 4.0) + noise(rng, r2.shape)
        P = ring(np.sqrt(r2), P0, r0, w)
        out.append(Instance("sweep512[%d]" % i, 2, side, h, dt, steps, P, psi0, np.zeros(r2.shape),
                            C_MODEL.copy(), params={"P0": P0, "r0": r0, "w": w, "seed": base_seed + i}))
    return out
