"""Input fields generated on the device (nls_fill_profile / nls_fill_state).

A pumping profile or an initial state given by parameters -- the reference's
functors (nls/pumping.py:113-178) on the model's grids (nls/model.py:220-232)
-- is sampled straight into device memory, so a solve needs no host-to-device
upload of its inputs.  ``profile_numpy`` / ``state_numpy`` evaluate the same
formulas on the host in the same order (they are what the device values are
checked against and what host-side callers use when they need the arrays).

  Grid(kind, a, b)     kind "linspace" (numpy.linspace(a, b, n)), "step"
                       (a + b k) or "fraction" (a + (b k) / n)
  Profile(kind, amplitude, x0, y0, width, radius)
                       kind "constant", "gaussian", "gaussian_ring" (the
                       reference's two-Gaussian ring) or "ring"
"""
import ctypes as C

import numpy as np

from . import native

GRID_KIND = {"linspace": 0, "step": 1, "fraction": 2}
PROFILE_KIND = {"constant": 0, "gaussian": 1, "gaussian_ring": 2, "ring": 3}


GridDesc = native.GridDesc
ProfileDesc = native.ProfileDesc


class Grid(object):
    def __init__(self, kind, a, b):
        assert kind in GRID_KIND
        self.kind, self.a, self.b = kind, float(a), float(b)

    def desc(self):
        return GridDesc(GRID_KIND[self.kind], 0, self.a, self.b)

    @classmethod
    def of_model(cls, n, dx, dim, bc="ref"):
        """The grid of nls_amd.model.AbstractModel.getGrid (model.py:220-232)."""
        if bc == "periodic":
            return cls("step", -n * dx / 2.0, dx)
        if dim == 1:
            return cls("linspace", 0.0, n * dx)
        return cls("linspace", -n * dx / 2.0, n * dx / 2.0)

    def numpy(self, n):
        k = np.arange(n)
        if self.kind == "linspace":
            return np.linspace(self.a, self.b, n)
        if self.kind == "step":
            return self.a + self.b * k.astype(np.float64)
        return self.a + (self.b * k.astype(np.float64)) / n


class Profile(object):
    def __init__(self, kind, amplitude=1.0, x0=0.0, y0=0.0, width=1.0, radius=0.0):
        assert kind in PROFILE_KIND
        self.kind = kind
        self.amplitude, self.x0, self.y0, self.width, self.radius = (float(amplitude), float(x0), float(y0),
                                                                     float(width), float(radius))

    def desc(self):
        return ProfileDesc(PROFILE_KIND[self.kind], 0, self.amplitude, self.x0, self.y0, self.width, self.radius)

    def numpy(self, x, y=None):
        """The profile at points x (1-D) or (x, y), in the device's order."""
        a, w2 = self.amplitude, 2.0 * (self.width * self.width)
        x = np.asarray(x, np.float64)
        if self.kind == "constant":
            return np.full(np.broadcast(x, x if y is None else y).shape, a)
        if self.kind == "gaussian":
            dxx = x - self.x0
            dyy = (0.0 if y is None else np.asarray(y, np.float64)) - self.y0
            return a * np.exp(-(dxx * dxx + dyy * dyy) / w2)
        if y is None:
            rho = x if self.kind == "gaussian_ring" else np.abs(x - self.x0)
        else:
            dxx, dyy = x - self.x0, np.asarray(y, np.float64) - self.y0
            rho = np.sqrt(dxx * dxx + dyy * dyy)
        if self.kind == "gaussian_ring":
            u, v = rho - self.radius, rho - (-self.radius)
            return 1.0 * (a * np.exp(-(u * u + 0.0) / w2) + a * np.exp(-(v * v + 0.0) / w2))
        u = rho - self.radius
        return a * np.exp(-(u * u) / w2)


def profile_numpy(profile, dim, n, grid, dtype="f64"):
    x = grid.numpy(n)
    if dim == 1:
        v = profile.numpy(x)
    else:
        X, Y = np.meshgrid(x, x)
        v = profile.numpy(X, Y)
    return v.astype(np.float64 if dtype == "f64" else np.float32)


_M64 = (1 << 64) - 1


def _uniform53(z):
    """splitmix64 finaliser of z, as a double in (0, 1] (vectorised, uint64)."""
    z = (z + np.uint64(0x9e3779b97f4a7c15)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)) & np.uint64(_M64)
    z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53


def noise_numpy(count, noise, seed):
    """The device's noise for elements 0..count-1 (complex128)."""
    with np.errstate(over="ignore"):
        i = np.arange(count, dtype=np.uint64)
        s = np.uint64(seed)
        u1 = _uniform53(s + np.uint64(2) * i + np.uint64(1))
        u2 = _uniform53(s + np.uint64(2) * i + np.uint64(2))
    r = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    sc = noise / 1.4142135623730951
    return sc * (r * np.cos(th)) + 1j * (sc * (r * np.sin(th)))


def state_numpy(profile, noise, seed, dim, n, grid, dtype="f64"):
    base = profile_numpy(profile, dim, n, grid, "f64")
    psi = base.astype(np.complex128)
    if noise:
        z = noise_numpy(base.size, noise, seed).reshape(base.shape)
        psi = (base + z.real) + 1j * z.imag
    return psi.astype(np.complex128 if dtype == "f64" else np.complex64)


# ---- device ----------------------------------------------------------------

def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def fill_profile(out, profile, dim, n, grid, stream=None):
    """Sample profile into the real torch tensor `out` (n or n*n elements)."""
    dtype = 1 if out.dtype.itemsize == 8 else 0
    assert out.numel() == (n if dim == 1 else n * n) and out.is_contiguous()
    p, g = profile.desc(), grid.desc()
    native._check(native.load().nls_fill_profile(C.byref(p), int(dim), int(n), C.byref(g), dtype,
                                                 C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


def fill_state(out, profile, noise, seed, dim, n, grid, stream=None):
    """Sample profile + seeded noise into the complex torch tensor `out`."""
    dtype = 1 if out.dtype.itemsize == 16 else 0
    assert out.numel() == (n if dim == 1 else n * n) and out.is_contiguous()
    p, g = profile.desc(), grid.desc()
    native._check(native.load().nls_fill_state(C.byref(p), float(noise), C.c_uint64(int(seed)), int(dim), int(n),
                                               C.byref(g), dtype, C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


def from_functor(pumping, dim):
    """Terms [(scale, Profile)] whose scaled sum is the reference functor
    `pumping` (nls_amd.pumping), or None if it has a part that is not a
    parametric profile (GridPumping, elliptic or rectangular profiles)."""
    from . import pumping as PU
    if isinstance(pumping, PU.GaussianRingPumping1D) and dim == 1:
        g = pumping.lhs
        return [(pumping.power, Profile("gaussian_ring", g.power, 0.0, 0.0, g.variation, g.x0))]
    if isinstance(pumping, PU.GaussianRingPumping2D) and dim == 2:
        g = pumping.pumping.lhs
        return [(pumping.pumping.power, Profile("gaussian_ring", g.power, pumping.x0, pumping.y0, g.variation, g.x0))]
    if isinstance(pumping, PU.RingPumping2D):
        return [(1.0, Profile("ring", pumping.power, pumping.x0, pumping.y0, pumping.width, pumping.radius))]
    if isinstance(pumping, PU.GaussianPumping):
        return [(1.0, Profile("gaussian", pumping.power, pumping.x0, pumping.y0, pumping.variation))]
    if isinstance(pumping, (PU.OpSumPumping, PU.OpSubPumping)):
        lhs, rhs = from_functor(pumping.lhs, dim), from_functor(pumping.rhs, dim)
        if lhs is None or rhs is None:
            return None
        return [("tree", pumping, lhs, rhs)]
    return None


def pumping_on_device(pumping, dim, n, grid, dtype="f64", device=None):
    """The functor sampled on the grid, evaluated on the device (a torch
    tensor, shape (n,) or (n, n)); sums and differences are combined in the
    functor's own order, power * (lhs +/- rhs)."""
    import torch
    terms = from_functor(pumping, dim)
    if terms is None:
        raise ValueError("pumping %r is not a parametric profile" % (pumping,))
    dev = device or torch.device("cuda:0")
    rt = torch.float64 if dtype == "f64" else torch.float32
    shape = (n,) if dim == 1 else (n, n)

    def ev(ts):
        (t,) = ts
        if t[0] == "tree":
            _, node, lhs, rhs = t
            a, b = ev(lhs), ev(rhs)
            from . import pumping as PU
            return node.power * (a + b) if isinstance(node, PU.OpSumPumping) else node.power * (a - b)
        scale, prof = t
        out = torch.empty(shape, dtype=torch.float64, device=dev)
        fill_profile(out, prof, dim, n, grid)
        return out if scale == 1.0 else scale * out
    return ev(terms).to(rt)
