"""CPU oracle for the time-stepping core of nls -- TEST INFRASTRUCTURE ONLY.

Thin ctypes access to two checkers:

* ``liboracle.so`` (built from ``nls_oracle.c`` in this directory): a plain-C
  restatement of the reference Fortran core ``nls/nls.f90`` in float and
  double, plus the periodic-boundary and dynamic-reservoir extensions;
* ``_ref/libnls_ref32.so`` / ``_ref/libnls_ref64.so`` (built by ``ref.mk``
  from the reference source where it lies): the reference Fortran itself,
  called through its module-procedure symbols with Fortran by-reference
  arguments.  Optional: absent on machines without the reference tree.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only to check or time against it.

Conventions: complex arrays are numpy complex64/complex128; 2-D fields are
C-ordered (n, n) arrays whose flat memory equals the reference's
column-major ``u(n, n)`` (col contiguous).  ``coeffs`` is the reference's
23-entry coefficient vector, 0-based (Fortran ``c(k)`` is ``coeffs[k-1]``).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

BC_REF, BC_PERIODIC = 0, 1

_lib = None


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        for sfx, real in (("_f64", C.c_double), ("_f32", C.c_float)):
            f = getattr(L, "oracle_rk4" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, real, real, C.c_int,
                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            f = getattr(L, "oracle_rhs" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, real, C.c_void_p, C.c_void_p,
                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            f = getattr(L, "oracle_laplacian_1d_parts" + sfx)
            f.argtypes = [C.c_int, C.c_int, real, C.c_void_p, C.c_void_p]
            f = getattr(L, "oracle_make_laplacian" + sfx)
            f.argtypes = [C.c_int, C.c_int, real, C.c_void_p]
            f = getattr(L, "oracle_make_banded_matrix" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
            f = getattr(L, "oracle_gbmv" + sfx)
            f.argtypes = [C.c_int, C.c_int, real, C.c_void_p, C.c_void_p, C.c_void_p]
            for d in ("1d", "2d"):
                f = getattr(L, "oracle_chemical_potential_%s%s" % (d, sfx))
                f.argtypes = [real, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _types(dtype):
    dtype = np.dtype(dtype)
    if dtype in (np.dtype(np.float64), np.dtype(np.complex128)):
        return "_f64", np.float64, np.complex128
    if dtype in (np.dtype(np.float32), np.dtype(np.complex64)):
        return "_f32", np.float32, np.complex64
    raise ValueError("dtype must be single or double precision")


def set_threads(n):
    return lib().oracle_set_threads(int(n))


def _shape(dim, n):
    return (n,) if dim == 1 else (n, n)


def rk4(dim, n, order, bc, dt, dx, iters, P, coeffs, psi0, n0=None, dtype=np.float64):
    """Advance (psi, n) by ``iters`` RK4 steps (runge_kutta[_2d], nls.f90:705-734, :873-901).

    Returns (psi, nres); nres is None for the adiabatic reservoir (c11 == 0).
    """
    sfx, rt, ct = _types(dtype)
    coeffs = np.ascontiguousarray(coeffs, np.float64)
    psi = np.ascontiguousarray(np.asarray(psi0).reshape(_shape(dim, n)), ct).copy()
    P = np.ascontiguousarray(np.asarray(P).reshape(_shape(dim, n)), rt)
    dynamic = coeffs[10] != 0.0
    nres = None
    if dynamic:
        nres = (np.zeros(_shape(dim, n), rt) if n0 is None
                else np.ascontiguousarray(np.asarray(n0).reshape(_shape(dim, n)), rt).copy())
    rc = getattr(lib(), "oracle_rk4" + sfx)(dim, n, order, bc, dt, dx, int(iters), _vp(P),
                                           _vp(coeffs), _vp(psi), _vp(nres) if dynamic else None)
    if rc:
        raise ValueError("oracle_rk4 rejected its arguments (rc=%d)" % rc)
    return psi, nres


def rhs(dim, n, order, bc, dx, P, coeffs, psi, nres=None, dtype=np.float64):
    """One RHS evaluation F(psi, n) (hamiltonian[_2d], nls.f90:621-650, :841-870)."""
    sfx, rt, ct = _types(dtype)
    coeffs = np.ascontiguousarray(coeffs, np.float64)
    psi = np.ascontiguousarray(np.asarray(psi).reshape(_shape(dim, n)), ct)
    P = np.ascontiguousarray(np.asarray(P).reshape(_shape(dim, n)), rt)
    dpsi = np.zeros_like(psi)
    dynamic = coeffs[10] != 0.0
    dn = np.zeros(_shape(dim, n), rt) if dynamic else None
    nr = np.ascontiguousarray(nres, rt) if dynamic else None
    rc = getattr(lib(), "oracle_rhs" + sfx)(dim, n, order, bc, dx, _vp(P), _vp(coeffs), _vp(psi),
                                           _vp(nr) if dynamic else None, _vp(dpsi),
                                           _vp(dn) if dynamic else None)
    if rc:
        raise ValueError("oracle_rhs rejected its arguments (rc=%d)" % rc)
    return dpsi, dn


def laplacian_1d_parts(n, order, h, dtype=np.float64):
    """(L1, L2) bands of make_laplacian_o3/o5/o7 (nls.f90:132-259), each (order, n)
    indexed like the Fortran op(p, j) (0-based)."""
    sfx, rt, _ = _types(dtype)
    L1 = np.zeros((n, order), rt)
    L2 = np.zeros((n, order), rt)
    rc = getattr(lib(), "oracle_laplacian_1d_parts" + sfx)(n, order, h, _vp(L1), _vp(L2))
    if rc:
        raise ValueError("bad order/n")
    return L1.T.copy(), L2.T.copy()


def make_laplacian(n, order, h, dtype=np.float64):
    sfx, rt, _ = _types(dtype)
    op = np.zeros((n, order), rt)
    rc = getattr(lib(), "oracle_make_laplacian" + sfx)(n, order, h, _vp(op))
    if rc:
        raise ValueError("bad order/n")
    return op.T.copy()


def make_banded_matrix(n, m, row, dtype=np.float64):
    """make_banded_matrix (nls.f90:61-91) -> (m, n) band."""
    sfx, rt, _ = _types(dtype)
    row = np.ascontiguousarray(row, rt)
    mat = np.full((n, m), np.nan, rt)
    rc = getattr(lib(), "oracle_make_banded_matrix" + sfx)(n, m, _vp(row), _vp(mat))
    if rc:
        raise ValueError("bad m/n")
    return mat.T.copy()


def gbmv(n, klu, alpha, band, x, y, dtype=np.float64):
    """y + alpha*A*x for the (2klu+1, n) band A (rgbmv, nls.f90:530-541)."""
    sfx, rt, _ = _types(dtype)
    b = np.ascontiguousarray(np.asarray(band, rt).T)
    x = np.ascontiguousarray(x, rt)
    y = np.ascontiguousarray(y, rt).copy()
    getattr(lib(), "oracle_gbmv" + sfx)(n, klu, alpha, _vp(b), _vp(x), _vp(y))
    return y


def chemical_potential(dim, dx, n, P, coeffs, u0, dtype=np.float64):
    """chemical_potential_1d/2d (nls.f90:921-971): complex in 1-D, real in 2-D."""
    sfx, rt, ct = _types(dtype)
    coeffs = np.ascontiguousarray(coeffs, np.float64)
    P = np.ascontiguousarray(np.asarray(P).reshape(_shape(dim, n)), rt)
    u0 = np.ascontiguousarray(np.asarray(u0).reshape(_shape(dim, n)), ct)
    mu = np.zeros(2, rt)
    rc = getattr(lib(), "oracle_chemical_potential_%dd%s" % (dim, sfx))(dx, n, _vp(P), _vp(coeffs),
                                                                       _vp(u0), _vp(mu))
    if rc:
        raise ValueError("bad arguments")
    return complex(mu[0], mu[1]) if dim == 1 else float(mu[0])


# ---------------------------------------------------------------------------
# The reference Fortran itself (oracle/_ref, built by ref.mk).
# ---------------------------------------------------------------------------

class RefLib(object):
    """The compiled reference module ``nls`` (nls/nls.f90) called through its
    LLVM-flang module-procedure symbols (``_QMnlsP<name>``); every argument is
    passed by reference as Fortran expects.  ``bits`` is 32 (as written) or 64
    (same source compiled with an 8-byte real kind)."""

    def __init__(self, bits=32):
        self.bits = bits
        path = os.path.join(REF_DIR, "libnls_ref%d.so" % bits)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = C.CDLL(path)
        self.rt = np.float32 if bits == 32 else np.float64
        self.ct = np.complex64 if bits == 32 else np.complex128

    @staticmethod
    def available(bits=32):
        return os.path.exists(os.path.join(REF_DIR, "libnls_ref%d.so" % bits))

    def _r(self, x):
        return np.array([x], self.rt)

    @staticmethod
    def _i(x):
        return C.byref(C.c_int(int(x)))

    def version(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self.L._QMnlsPversion(C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def make_banded_matrix(self, n, m, row):
        row = np.ascontiguousarray(row, self.rt)
        mat = np.zeros((n, m), self.rt)
        self.L._QMnlsPmake_banded_matrix(self._i(n), self._i(m), _vp(row), _vp(mat))
        return mat.T.copy()

    def make_laplacian(self, n, order, h):
        op = np.zeros((n, order), self.rt)
        hh = self._r(h)
        self.L._QMnlsPmake_laplacian(self._i(n), self._i(order), _vp(hh), _vp(op))
        return op.T.copy()

    def solve(self, dim, dt, dx, n, order, iters, P, coeffs, u0):
        """solve_nls (nls.f90:797) or solve_nls_2d (:903)."""
        shape = _shape(dim, n)
        P = np.ascontiguousarray(np.asarray(P).reshape(shape), self.rt)
        co = np.ascontiguousarray(coeffs, self.rt)
        u0 = np.ascontiguousarray(np.asarray(u0).reshape(shape), self.ct)
        u = np.zeros(shape, self.ct)
        fn = self.L._QMnlsPsolve_nls if dim == 1 else self.L._QMnlsPsolve_nls_2d
        fn(_vp(self._r(dt)), _vp(self._r(dx)), self._i(n), self._i(order), self._i(iters),
           _vp(P), _vp(co), _vp(u0), _vp(u))
        return u

    def hamiltonian_1d(self, P, coeffs, u, order, dx):
        """hamiltonian (nls.f90:621) with the band of make_laplacian."""
        n = len(u)
        op = np.ascontiguousarray(self.make_laplacian(n, order, dx).T)
        P = np.ascontiguousarray(P, self.rt)
        co = np.ascontiguousarray(coeffs, self.rt)
        u = np.ascontiguousarray(u, self.ct)
        v = np.zeros(n, self.ct)
        self.L._QMnlsPhamiltonian(_vp(P), _vp(co), _vp(u), _vp(v), _vp(op),
                                  self._i((order - 1) // 2), self._i(n))
        return v

    def chemical_potential(self, dim, dx, n, P, coeffs, u0):
        shape = _shape(dim, n)
        P = np.ascontiguousarray(np.asarray(P).reshape(shape), self.rt)
        co = np.ascontiguousarray(coeffs, self.rt)
        u0 = np.ascontiguousarray(np.asarray(u0).reshape(shape), self.ct)
        if dim == 1:
            mu = np.zeros(1, self.ct)
            self.L._QMnlsPchemical_potential_1d(_vp(self._r(dx)), self._i(n), _vp(P), _vp(co),
                                                _vp(u0), _vp(mu))
            return complex(mu[0])
        mu = np.zeros(1, self.rt)
        self.L._QMnlsPchemical_potential_2d(_vp(self._r(dx)), self._i(n), _vp(P), _vp(co),
                                            _vp(u0), _vp(mu))
        return float(mu[0])
