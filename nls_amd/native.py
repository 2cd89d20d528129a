"""ctypes boundary to ``libnls_amd.so`` (declared in ``include/nls_amd.h``).

``nls`` mirrors the f2py module object the reference imports as
``from .native import nls`` (reference ``nls/solver.py:9``): same procedure
names, positional argument order and optional trailing ``n`` as f2py derives
them from ``nls/nls.f90`` (e.g. ``u = nls.solve_nls(dt, dx, order, iters,
pumping, coeffs, u0[, n])``).  Inputs are cast like f2py casts them
(float32 / complex64, contiguous copies), and a shape mismatch raises
``ValueError`` as f2py does.  The work runs on the GPU; there is no CPU
fallback -- a missing library raises ``ImportError``.

The extended entry points (``solve_host``, ``advance``, ``rhs``,
``observables``, ``chemical_potential``) take an :class:`ProblemDesc` and
cover double precision, periodic grids, the dynamic reservoir and the
split-step Fourier scheme.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnls_amd.so")

NLS_OK, NLS_EINVAL, NLS_EUNSUPPORTED, NLS_EHIP, NLS_ENONFINITE, NLS_ENOMEM = 0, -1, -2, -3, -4, -5
BC = {"ref": 0, "periodic": 1}
DTYPE = {"f32": 0, "f64": 1}
SCHEME = {"rk4": 0, "ssf": 1}

EXPORTS = (
    "nls_last_error", "nls_version", "nls_solve_nls", "nls_solve_nls_1d", "nls_solve_nls_2d",
    "nls_chemical_potential_1d", "nls_chemical_potential_2d", "nls_make_laplacian",
    "nls_solve_host", "nls_workspace_bytes", "nls_advance", "nls_rhs",
    "nls_chemical_potential", "nls_observables", "nls_energy",
    "nls_batch_create", "nls_batch_destroy", "nls_batch_points", "nls_batch_tiles",
    "nls_batch_workspace_bytes", "nls_batch_set_steps", "nls_batch_advance", "nls_tune_stage_tile",
    "nls_release", "nls_fill_profile", "nls_fill_state", "nls_slab_halo_rows", "nls_slab_step",
    "nls_make_banded_matrix", "nls_make_laplacian_2d", "nls_rgbmv", "nls_rbbmv", "nls_revervoir",
    "nls_revervoir_2d", "nls_hamiltonian", "nls_hamiltonian_2d", "nls_runge_kutta", "nls_runge_kutta_2d",
    "nls_batch_set_flags", "nls_batch_observables", "nls_batch_prepare", "nls_batch_energy",
    "nls_batch_update", "nls_batch_observables_async", "nls_batch_energy_async", "nls_advance_static",
    "nls_clear_first_row_of_derivative", "nls_divide_derivative_on_radius", "nls_tune_fused_form",
)
ENGINE = {"auto": -1, "stage": 0, "fused": 1, "resident": 2, "hybrid": 3}
FLAG_EAGER = 1
FLAG_F64_ARITH = 2


class NativeError(RuntimeError):
    """A HIP/runtime failure reported by libnls_amd."""


class GridDesc(C.Structure):
    """Mirror of ``nls_grid`` (include/nls_amd.h)."""
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("a", C.c_double), ("b", C.c_double)]


class ProfileDesc(C.Structure):
    """Mirror of ``nls_profile`` (include/nls_amd.h)."""
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("amplitude", C.c_double), ("x0", C.c_double),
                ("y0", C.c_double), ("width", C.c_double), ("radius", C.c_double)]


class ProblemDesc(C.Structure):
    """Mirror of ``nls_problem`` (include/nls_amd.h)."""
    _fields_ = [("dim", C.c_int32), ("n", C.c_int32), ("order", C.c_int32), ("bc", C.c_int32),
                ("dtype", C.c_int32), ("scheme", C.c_int32), ("engine", C.c_int32), ("flags", C.c_int32),
                ("dt", C.c_double), ("dx", C.c_double), ("coeffs", C.c_double * 23)]

    @classmethod
    def make(cls, dim, n, order, dt, dx, coeffs, bc="ref", dtype="f32", scheme="rk4", engine="auto",
             graphs=True, f64_arith=False):
        """engine: RK4 engine of a plain solve -- "auto" (resident when the
        instance fits on chip, else fused for 2-D and stage for 1-D),
        "stage" (one launch per RK4 stage), "fused" (one launch per step,
        2-D only) or "resident" (1-D n <= 4096, 2-D n <= 64).  graphs=False
        sets NLS_FLAG_EAGER: no step-block graph replay.  f64_arith=True
        sets NLS_FLAG_F64_ARITH: a single-precision 2-D periodic problem
        stepped with double-precision arithmetic on its float fields."""
        p = cls()
        p.dim, p.n, p.order = int(dim), int(n), int(order)
        p.bc = BC[bc] if isinstance(bc, str) else int(bc)
        p.dtype = DTYPE[dtype] if isinstance(dtype, str) else int(dtype)
        p.scheme = SCHEME[scheme] if isinstance(scheme, str) else int(scheme)
        p.engine = ENGINE[engine] if isinstance(engine, str) else int(engine)
        p.flags = (0 if graphs else FLAG_EAGER) | (FLAG_F64_ARITH if f64_arith else 0)
        p.dt, p.dx = float(dt), float(dx)
        c = np.zeros(23)
        cc = np.asarray(coeffs, np.float64).ravel()
        c[:min(23, cc.size)] = cc[:23]
        p.coeffs[:] = list(c)
        return p

    @property
    def dynamic(self):
        return self.coeffs[10] != 0.0

    @property
    def points(self):
        return self.n if self.dim == 1 else self.n * self.n

    @property
    def real_dtype(self):
        return np.float64 if self.dtype == 1 else np.float32

    @property
    def complex_dtype(self):
        return np.complex128 if self.dtype == 1 else np.complex64

    @property
    def shape(self):
        return (self.n,) if self.dim == 1 else (self.n, self.n)


class Member(C.Structure):
    """Mirror of ``nls_member`` (include/nls_amd.h)."""
    _fields_ = [("n", C.c_int32), ("reserved", C.c_int32), ("offset", C.c_int64), ("dt", C.c_double),
                ("dx", C.c_double), ("coeffs", C.c_double * 23)]


_lib = None


def load():
    """Load libnls_amd.so; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libnls_amd.so is not built (python -m nls_amd.build)")
    L = C.CDLL(LIB_PATH)
    vp, i, f, d, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
    pp = C.POINTER(ProblemDesc)
    sig = {
        "nls_last_error": ([], C.c_char_p),
        "nls_version": ([vp, vp, vp], i),
        "nls_solve_nls": ([f, f, i, i, i, vp, vp, vp, vp], i),
        "nls_solve_nls_1d": ([f, f, i, i, i, vp, vp, vp, vp], i),
        "nls_solve_nls_2d": ([f, f, i, i, i, vp, vp, vp, vp], i),
        "nls_chemical_potential_1d": ([f, i, vp, vp, vp, vp], i),
        "nls_chemical_potential_2d": ([f, i, vp, vp, vp, vp], i),
        "nls_make_laplacian": ([i, i, d, vp], i),
        "nls_solve_host": ([pp, i, vp, vp, vp], i),
        "nls_workspace_bytes": ([pp], sz),
        "nls_advance": ([pp, i, vp, vp, vp, vp, sz, vp], i),
        "nls_advance_static": ([pp, i, vp, vp, vp, vp], i),
        "nls_release": ([], i),
        "nls_slab_halo_rows": ([i], i),
        "nls_make_banded_matrix": ([i, i, vp, vp], i),
        "nls_clear_first_row_of_derivative": ([i, i, vp], i),
        "nls_divide_derivative_on_radius": ([i, i, f, vp], i),
        "nls_make_laplacian_2d": ([i, i, f, vp, vp], i),
        "nls_rgbmv": ([vp, vp, f, vp, i, i], i),
        "nls_rbbmv": ([vp, vp, f, vp, vp, i, i], i),
        "nls_revervoir": ([vp, vp, vp, vp, i], i),
        "nls_revervoir_2d": ([vp, vp, vp, vp, i], i),
        "nls_hamiltonian": ([vp, vp, vp, vp, vp, i, i], i),
        "nls_hamiltonian_2d": ([vp, vp, vp, vp, vp, vp, i, i], i),
        "nls_runge_kutta": ([f, f, vp, vp, i, i, i, vp, vp, vp], i),
        "nls_runge_kutta_2d": ([f, f, vp, i, vp, vp, i, i, vp, vp, vp], i),
        "nls_slab_step": ([pp, i, vp, vp, vp, vp, vp, vp], i),
        "nls_fill_profile": ([C.POINTER(ProfileDesc), i, i, C.POINTER(GridDesc), i, vp, vp], i),
        "nls_fill_state": ([C.POINTER(ProfileDesc), d, C.c_uint64, i, i, C.POINTER(GridDesc), i, vp, vp], i),
        "nls_rhs": ([pp, vp, vp, vp, vp, vp, vp], i),
        "nls_chemical_potential": ([pp, vp, vp, vp, vp, vp], i),
        "nls_observables": ([pp, vp, vp, vp, vp], i),
        "nls_energy": ([pp, vp, vp, vp, vp, vp], i),
        "nls_batch_create": ([i, i, i, i, i, C.POINTER(Member), i, C.POINTER(vp)], i),
        "nls_batch_destroy": ([vp], i),
        "nls_batch_points": ([vp], sz),
        "nls_batch_tiles": ([vp], i),
        "nls_batch_workspace_bytes": ([vp], sz),
        "nls_batch_set_steps": ([vp, vp, vp], i),
        "nls_batch_advance": ([vp, i, vp, vp, vp, vp, sz, vp], i),
        "nls_batch_set_flags": ([vp, i], i),
        "nls_batch_observables": ([vp, vp, vp, vp, vp], i),
        "nls_batch_energy": ([vp, vp, vp, vp, vp, vp], i),
        "nls_batch_prepare": ([vp, i, vp, vp, vp, vp, sz], i),
        "nls_batch_update": ([vp, C.POINTER(Member), i, vp, vp], i),
        "nls_batch_observables_async": ([vp, vp, vp, vp, vp], i),
        "nls_batch_energy_async": ([vp, vp, vp, vp, vp, vp], i),
        "nls_tune_stage_tile": ([i, i], i),
        "nls_tune_fused_form": ([i, i], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc):
    if rc == NLS_OK:
        return
    msg = load().nls_last_error().decode(errors="replace")
    if rc in (NLS_EINVAL, NLS_EUNSUPPORTED):
        raise ValueError(msg)
    if rc == NLS_ENONFINITE:
        raise FloatingPointError(msg)
    raise NativeError("libnls_amd error %d: %s" % (rc, msg))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _arr(x, dtype, name, pos, shape=None, order="C"):
    """f2py-style conversion: cast, make contiguous (order "F" for the
    reference's 2-D operator arrays, whose memory the C ABI reads
    column-major), check the fixed shape."""
    try:
        a = np.array(x, dtype=dtype, copy=True, order=order)
    except (TypeError, ValueError) as e:
        raise ValueError("failed in converting %s argument `%s' to C/Fortran array: %s" % (pos, name, e))
    if shape is not None:
        if a.ndim == 0:
            a = np.full(shape, a, dtype=dtype)
        if a.shape != tuple(shape):
            raise ValueError("%s argument `%s' has shape %s, expected %s" % (pos, name, a.shape, tuple(shape)))
    return a


class _NlsModule(object):
    """The f2py-shaped namespace ``nls`` (procedures of module nls, nls.f90:13-24)."""

    @staticmethod
    def version():
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(load().nls_version(C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    @staticmethod
    def _solve(fn, dim, dt, dx, order, iters, pumping, coeffs, u0, n):
        P = _arr(pumping, np.float32, "pumping", "5th")
        if n is None:
            n = P.shape[0]
        shape = (n,) if dim == 1 else (n, n)
        P = _arr(P, np.float32, "pumping", "5th", shape)
        co = _arr(coeffs, np.float32, "coeffs", "6th", (23,))
        u0 = _arr(u0, np.complex64, "u0", "7th", shape)
        u = np.empty(shape, np.complex64)
        _check(fn(float(dt), float(dx), int(n), int(order), int(iters), _ptr(P), _ptr(co), _ptr(u0), _ptr(u)))
        return u

    def solve_nls(self, dt, dx, order, iters, pumping, coeffs, u0, n=None):
        """u = solve_nls(dt, dx, order, iters, pumping, coeffs, u0[, n]) -- nls.f90:797."""
        return self._solve(load().nls_solve_nls, 1, dt, dx, order, iters, pumping, coeffs, u0, n)

    def solve_nls_1d(self, dt, dx, order, iters, pumping, coeffs, u0, n=None):
        """u = solve_nls_1d(...) -- nls.f90:815."""
        return self._solve(load().nls_solve_nls_1d, 1, dt, dx, order, iters, pumping, coeffs, u0, n)

    def solve_nls_2d(self, dt, dx, order, iters, pumping, coeffs, u0, n=None):
        """u = solve_nls_2d(dt, dx, order, iters, pumping, coeffs, u0[, n]) -- nls.f90:903."""
        return self._solve(load().nls_solve_nls_2d, 2, dt, dx, order, iters, pumping, coeffs, u0, n)

    def chemical_potential_1d(self, dx, pumping, coeffs, u0, n=None):
        """mu = chemical_potential_1d(dx, pumping, coeffs, u0[, n]) -- nls.f90:921 (complex)."""
        P = _arr(pumping, np.float32, "pumping", "2nd")
        n = P.shape[0] if n is None else n
        P = _arr(P, np.float32, "pumping", "2nd", (n,))
        co = _arr(coeffs, np.float32, "coeffs", "3rd", (23,))
        u0 = _arr(u0, np.complex64, "u0", "4th", (n,))
        mu = np.zeros(1, np.complex64)
        _check(load().nls_chemical_potential_1d(float(dx), int(n), _ptr(P), _ptr(co), _ptr(u0), _ptr(mu)))
        return mu[0]

    def chemical_potential_2d(self, dx, pumping, coeffs, u0, n=None):
        """mu = chemical_potential_2d(dx, pumping, coeffs, u0[, n]) -- nls.f90:950 (real)."""
        P = _arr(pumping, np.float32, "pumping", "2nd")
        n = P.shape[0] if n is None else n
        P = _arr(P, np.float32, "pumping", "2nd", (n, n))
        co = _arr(coeffs, np.float32, "coeffs", "3rd", (23,))
        u0 = _arr(u0, np.complex64, "u0", "4th", (n, n))
        mu = np.zeros(1, np.float32)
        _check(load().nls_chemical_potential_2d(float(dx), int(n), _ptr(P), _ptr(co), _ptr(u0), _ptr(mu)))
        return mu[0]

    @staticmethod
    def make_laplacian(n, m, h):
        """op = make_laplacian(n, m, h) -- nls.f90:278, band (m, n) like the Fortran op(m, n)."""
        op = np.zeros((int(n), int(m)), np.float64)   # C (n, m) == Fortran (m, n)
        _check(load().nls_make_laplacian(int(n), int(m), float(h), _ptr(op)))
        return op.T.astype(np.float32)

    # The module's other public procedures (nls.f90:13-24).  Argument lists
    # follow f2py's rules: intent(out) arrays are returned, intent(inout)
    # arrays (rgbmv's u, rbbmv's y) are updated in place and must be float32
    # contiguous arrays, dimension arguments that an input array's shape
    # fixes are optional and come last.

    @staticmethod
    def _band_inplace(L1, n, m):
        """The band L1(m, n) as a Fortran-ordered float32 array: the caller's,
        updated in place when it already is one; a new one otherwise."""
        if L1 is None:
            # f2py allocates an intent(out) array afresh, zero-filled: the
            # procedure then works on zeros (parity unpinned: no build of the
            # reference's extension here)
            return np.zeros((int(m), int(n)), np.float32, order="F")
        a = L1 if (isinstance(L1, np.ndarray) and L1.dtype == np.float32 and L1.flags.f_contiguous) else \
            _arr(L1, np.float32, "l1", "3rd", None, order="F")
        if a.shape != (int(m), int(n)):
            raise ValueError("l1 has shape %s, expected %s" % (a.shape, (int(m), int(n))))
        return a

    def clear_first_row_of_derivative(self, n, m, L1=None):
        """l1 = clear_first_row_of_derivative(n, m[, l1]) -- nls.f90:93-107.  The
        Fortran procedure works in place on make_laplacian's first-derivative
        band; given a Fortran-ordered float32 l1 of shape (m, n) it does so here
        too, and returns it."""
        a = self._band_inplace(L1, n, m)
        _check(load().nls_clear_first_row_of_derivative(int(n), int(m), _ptr(a)))
        return a

    def divide_derivative_on_radius(self, n, m, h, L1=None):
        """l1 = divide_derivative_on_radius(n, m, h[, l1]) -- nls.f90:109-130, in
        place on l1 as clear_first_row_of_derivative."""
        a = self._band_inplace(L1, n, m)
        _check(load().nls_divide_derivative_on_radius(int(n), int(m), float(h), _ptr(a)))
        return a

    def make_laplacian_o3(self, n, h):
        return self.make_laplacian(n, 3, h)

    def make_laplacian_o5(self, n, h):
        return self.make_laplacian(n, 5, h)

    def make_laplacian_o7(self, n, h):
        return self.make_laplacian(n, 7, h)

    @staticmethod
    def make_banded_matrix(n, row, m=None):
        """mat = make_banded_matrix(n, row[, m]) -- nls.f90:61, mat(m, n)."""
        r = _arr(row, np.float32, "row", "2nd")
        m = r.shape[0] if m is None else int(m)
        r = _arr(r, np.float32, "row", "2nd", (m,))
        mat = np.zeros((m, int(n)), np.float32, order="F")
        _check(load().nls_make_banded_matrix(int(n), m, _ptr(r), _ptr(mat)))
        return mat

    @staticmethod
    def make_laplacian_2d(n, m, h):
        """blocks, orders = make_laplacian_2d(n, m, h) -- nls.f90:388, blocks(n, 2m-1)."""
        n, m = int(n), int(m)
        blocks = np.zeros((n, 2 * m - 1), np.float32, order="F")
        orders = np.zeros(m, np.int32)
        _check(load().nls_make_laplacian_2d(n, m, float(h), _ptr(blocks), _ptr(orders)))
        return blocks, orders

    def make_laplacian_2d_o3(self, n, h):
        return self.make_laplacian_2d(n, 3, h)

    def make_laplacian_2d_o5(self, n, h):
        return self.make_laplacian_2d(n, 5, h)

    def make_laplacian_2d_o7(self, n, h):
        return self.make_laplacian_2d(n, 7, h)

    @staticmethod
    def _inout(y, name, pos, size):
        if not (isinstance(y, np.ndarray) and y.dtype == np.float32 and y.size == size and
                (y.flags.c_contiguous or y.flags.f_contiguous)):
            raise ValueError("%s argument `%s' must be a contiguous float32 array of %d elements (intent(inout))"
                             % (pos, name, size))
        return y

    def rgbmv(self, x, u, sign, op, klu, n=None):
        """rgbmv(x, u, sign, op, klu[, n]) -- nls.f90:530: u += sign A x in place, A = op(2 klu + 1, n)."""
        xa = _arr(x, np.float32, "x", "1st")
        n = xa.shape[0] if n is None else int(n)
        xa = _arr(xa, np.float32, "x", "1st", (n,))
        ua = self._inout(u, "u", "2nd", n)
        oa = _arr(op, np.float32, "op", "4th", (2 * int(klu) + 1, n), order="F")
        _check(load().nls_rgbmv(_ptr(xa), _ptr(ua), float(sign), _ptr(oa), int(klu), n))

    def rbbmv(self, x, y, sign, blocks, ms, m=None, n=None):
        """rbbmv(x, y, sign, blocks, ms[, m, n]) -- nls.f90:510: y += sign B x in place on n*n points."""
        msa = _arr(ms, np.int32, "ms", "5th")
        m = msa.shape[0] if m is None else int(m)
        ba = _arr(blocks, np.float32, "blocks", "4th", order="F")
        n = ba.shape[0] if n is None else int(n)
        ba = _arr(ba, np.float32, "blocks", "4th", (n, 2 * m - 1), order="F")
        msa = _arr(msa, np.int32, "ms", "5th", (m,))
        xa = _arr(np.ravel(x), np.float32, "x", "1st", (n * n,))
        ya = self._inout(y, "y", "2nd", n * n)
        _check(load().nls_rbbmv(_ptr(xa), _ptr(ya), float(sign), _ptr(ba), _ptr(msa), m, n))

    def rbbmv_o3(self, x, y, sign, blocks, ms, n=None):
        return self.rbbmv(x, y, sign, blocks, ms, 3, n)

    def rbbmv_o5(self, x, y, sign, blocks, ms, n=None):
        return self.rbbmv(x, y, sign, blocks, ms, 5, n)

    def rbbmv_o7(self, x, y, sign, blocks, ms, n=None):
        return self.rbbmv(x, y, sign, blocks, ms, 7, n)

    @staticmethod
    def _reservoir(fn, dim, pumping, coeffs, u_sqr, n):
        P = _arr(pumping, np.float32, "pumping", "1st")
        n = P.shape[0] if n is None else int(n)
        shape = (n,) if dim == 1 else (n, n)
        P = _arr(P, np.float32, "pumping", "1st", shape)
        co = _arr(coeffs, np.float32, "coeffs", "2nd", (23,))
        us = _arr(u_sqr, np.float32, "u_sqr", "3rd", shape)
        r = np.empty(shape, np.float32)
        _check(fn(_ptr(P), _ptr(co), _ptr(us), _ptr(r), n))
        return r

    def revervoir(self, pumping, coeffs, u_sqr, n=None):
        """r = revervoir(pumping, coeffs, u_sqr[, n]) -- nls.f90:570."""
        return self._reservoir(load().nls_revervoir, 1, pumping, coeffs, u_sqr, n)

    def revervoir_2d(self, pumping, coeffs, u_sqr, n=None):
        """r = revervoir_2d(pumping, coeffs, u_sqr[, n]) -- nls.f90:829."""
        return self._reservoir(load().nls_revervoir_2d, 2, pumping, coeffs, u_sqr, n)

    def hamiltonian(self, pumping, coeffs, u, op, klu, n=None):
        """v = hamiltonian(pumping, coeffs, u, op, klu[, n]) -- nls.f90:621 (on the GPU)."""
        P = _arr(pumping, np.float32, "pumping", "1st")
        n = P.shape[0] if n is None else int(n)
        P = _arr(P, np.float32, "pumping", "1st", (n,))
        co = _arr(coeffs, np.float32, "coeffs", "2nd", (23,))
        ua = _arr(u, np.complex64, "u", "3rd", (n,))
        oa = _arr(op, np.float32, "op", "4th", (2 * int(klu) + 1, n), order="F")
        v = np.empty(n, np.complex64)
        _check(load().nls_hamiltonian(_ptr(P), _ptr(co), _ptr(ua), _ptr(v), _ptr(oa), int(klu), n))
        return v

    def hamiltonian_2d(self, pumping, coeffs, u, blocks, orders, order=None, n=None):
        """v = hamiltonian_2d(pumping, coeffs, u, blocks, orders[, order, n]) -- nls.f90:841 (on the GPU)."""
        oa = _arr(orders, np.int32, "orders", "5th")
        order = oa.shape[0] if order is None else int(order)
        P = _arr(pumping, np.float32, "pumping", "1st")
        n = P.shape[0] if n is None else int(n)
        P = _arr(P, np.float32, "pumping", "1st", (n, n))
        co = _arr(coeffs, np.float32, "coeffs", "2nd", (23,))
        ua = _arr(u, np.complex64, "u", "3rd", (n, n))
        ba = _arr(blocks, np.float32, "blocks", "4th", (n, 2 * order - 1), order="F")
        oa = _arr(oa, np.int32, "orders", "5th", (order,))
        v = np.empty((n, n), np.complex64)
        _check(load().nls_hamiltonian_2d(_ptr(P), _ptr(co), _ptr(ua), _ptr(v), _ptr(ba), _ptr(oa), order, n))
        return v

    def runge_kutta(self, dt, t0, u0, op, iters, pumping, coeffs, n=None, order=None):
        """u = runge_kutta(dt, t0, u0, op, iters, pumping, coeffs[, n, order]) -- nls.f90:705 (on the GPU)."""
        ua = _arr(u0, np.complex64, "u0", "3rd")
        n = ua.shape[0] if n is None else int(n)
        ua = _arr(ua, np.complex64, "u0", "3rd", (n,))
        oa = _arr(op, np.float32, "op", "4th", order="F")
        order = oa.shape[0] if order is None else int(order)
        oa = _arr(oa, np.float32, "op", "4th", (order, n), order="F")
        P = _arr(pumping, np.float32, "pumping", "6th", (n,))
        co = _arr(coeffs, np.float32, "coeffs", "7th", (23,))
        u = np.empty(n, np.complex64)
        _check(load().nls_runge_kutta(float(dt), float(t0), _ptr(ua), _ptr(oa), n, order, int(iters), _ptr(u),
                                      _ptr(P), _ptr(co)))
        return u

    def runge_kutta_2d(self, dt, t0, u0, blocks, orders, iters, pumping, coeffs, n=None, order=None):
        """u = runge_kutta_2d(dt, t0, u0, blocks, orders, iters, pumping, coeffs[, n, order]) -- nls.f90:873."""
        ua = _arr(u0, np.complex64, "u0", "3rd")
        n = ua.shape[0] if n is None else int(n)
        ua = _arr(ua, np.complex64, "u0", "3rd", (n, n))
        oa = _arr(orders, np.int32, "orders", "5th")
        order = oa.shape[0] if order is None else int(order)
        oa = _arr(oa, np.int32, "orders", "5th", (order,))
        ba = _arr(blocks, np.float32, "blocks", "4th", (n, 2 * order - 1), order="F")
        P = _arr(pumping, np.float32, "pumping", "7th", (n, n))
        co = _arr(coeffs, np.float32, "coeffs", "8th", (23,))
        u = np.empty((n, n), np.complex64)
        _check(load().nls_runge_kutta_2d(float(dt), float(t0), _ptr(ua), n, _ptr(ba), _ptr(oa), order, int(iters),
                                         _ptr(u), _ptr(P), _ptr(co)))
        return u


nls = _NlsModule()


# ---------------------------------------------------------------------------
# Extended entry points
# ---------------------------------------------------------------------------

def solve_host(desc, iters, pumping, psi0, n0=None):
    """Host-memory solve (nls_solve_host).  Returns (psi, nres or None)."""
    rt, ct, shape = desc.real_dtype, desc.complex_dtype, desc.shape
    P = _arr(pumping, rt, "pumping", "3rd", shape)
    psi = _arr(psi0, ct, "psi0", "4th", shape)
    nres = None
    if desc.dynamic:
        nres = np.zeros(shape, rt) if n0 is None else _arr(n0, rt, "n0", "5th", shape)
    _check(load().nls_solve_host(C.byref(desc), int(iters), _ptr(P), _ptr(psi),
                                 _ptr(nres) if nres is not None else None))
    return psi, nres


def stage_tile(dtype, rows=0):
    """Get (rows=0) or set the tile height of the 2-D stage kernels for dtype."""
    r = load().nls_tune_stage_tile(DTYPE[dtype] if isinstance(dtype, str) else int(dtype), int(rows))
    if r < 0:
        _check(r)
    return r


FUSED_FORMS = {"auto": 0, "tall": 1, "short": 2, "half": 3, "tall512": 4}


def fused_form(dtype, form=None):
    """Get (form=None) or set the fused kernel's tile form for dtype:
    "auto" (per launch), "tall", "short", "half" or "tall512"
    (nls_tune_fused_form)."""
    code = -1 if form is None else FUSED_FORMS[form]
    r = load().nls_tune_fused_form(DTYPE[dtype] if isinstance(dtype, str) else int(dtype), code)
    if r < 0:
        _check(r)
    return {v: k for k, v in FUSED_FORMS.items()}[r]


def workspace_bytes(desc):
    return int(load().nls_workspace_bytes(C.byref(desc)))


def _dp(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def advance(desc, iters, pumping, psi, nres, workspace, stream=None):
    """Device solve (nls_advance) on torch tensors; psi/nres advanced in place.

    ``stream`` is a torch.cuda.Stream (default: the current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    _check(load().nls_advance(C.byref(desc), int(iters), _dp(pumping), _dp(psi), _dp(nres),
                              _dp(workspace), workspace.numel() * workspace.element_size() if workspace is not None else 0,
                              C.c_void_p(s.cuda_stream)))


STATIC_WORKSPACE_BYTES = 64 << 20      # NLS_STATIC_WORKSPACE_BYTES


def advance_static(desc, iters, pumping, psi, nres, stream=None):
    """Device solve without allocation (nls_advance_static): the library's
    own per-device workspace array, direct launches only."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    _check(load().nls_advance_static(C.byref(desc), int(iters), _dp(pumping), _dp(psi), _dp(nres),
                                     C.c_void_p(s.cuda_stream)))


def release():
    """Free the calling thread's cached step-block graphs (nls_release)."""
    _check(load().nls_release())


def slab_halo_rows(order):
    """Ghost rows on each side of a row slab (nls_slab_halo_rows): 2 (order - 1)."""
    h = load().nls_slab_halo_rows(int(order))
    if h < 0:
        _check(h)
    return h


def slab_step(desc, rows, pumping, psi_in, nres_in, psi_out, nres_out, stream=None):
    """One RK4 step of a row slab (nls_slab_step) on torch tensors of
    (rows + 2 h) x n points; writes the slab's own rows of psi_out/nres_out."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    _check(load().nls_slab_step(C.byref(desc), int(rows), _dp(pumping), _dp(psi_in), _dp(nres_in), _dp(psi_out),
                                _dp(nres_out), C.c_void_p(s.cuda_stream)))


def rhs(desc, pumping, psi, nres, dpsi, dn, stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    _check(load().nls_rhs(C.byref(desc), _dp(pumping), _dp(psi), _dp(nres), _dp(dpsi), _dp(dn),
                          C.c_void_p(s.cuda_stream)))


def chemical_potential(desc, pumping, psi, nres=None, stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    mu = (C.c_double * 2)()
    _check(load().nls_chemical_potential(C.byref(desc), _dp(pumping), _dp(psi), _dp(nres), mu,
                                         C.c_void_p(s.cuda_stream)))
    return complex(mu[0], mu[1])


def energy(desc, pumping, psi, nres=None, stream=None):
    """Energy of a device state for the conservative part of the right-hand
    side (nls_energy): {"energy", "kinetic", "interaction", "reservoir"},
    accumulated in double precision.  The kinetic term uses the Laplacian the
    problem's scheme evolves: the order-p finite-difference stencil for RK4,
    the spectral operator (sum of |k|^2 |psi_hat|^2) for split-step (SSF)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    out = (C.c_double * 4)()
    _check(load().nls_energy(C.byref(desc), _dp(pumping), _dp(psi), _dp(nres), out, C.c_void_p(s.cuda_stream)))
    return {"energy": out[0], "kinetic": out[1], "interaction": out[2], "reservoir": out[3]}


def observables(desc, psi, nres=None, stream=None):
    """(norm, population, peak density, non-finite count) of a device state."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    out = (C.c_double * 4)()
    _check(load().nls_observables(C.byref(desc), _dp(psi), _dp(nres), out, C.c_void_p(s.cuda_stream)))
    return {"norm": out[0], "population": out[1], "peak": out[2], "nonfinite": int(out[3])}
