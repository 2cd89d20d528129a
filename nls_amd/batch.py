"""Batched solves on one GPU: many independent instances of mixed sizes in
shared packed device arrays, advanced together by libnls_amd's batch engine
(``nls_batch_*`` in include/nls_amd.h).

The reference runs one solve per worker process (schedulr/workr.py:39-107);
each member here follows solve_nls / solve_nls_2d (nls.f90:797, :903).
"""
import ctypes as C

import numpy as np

from . import native

CPLX = {"f32": "complex64", "f64": "complex128"}
REAL = {"f32": "float32", "f64": "float64"}


class Batch(object):
    """Members share dim, order, boundary family, reservoir mode and dtype.

    ``members``: sequence of objects with attributes n, dx, dt, coeffs, P,
    psi0 and n0 (e.g. ``configs.Instance``)."""

    def __init__(self, members, dim=2, order=5, bc="periodic", dtype="f32", engine="auto", device=None,
                 align=64, graphs=True):
        import torch
        self.torch = torch
        self.dim, self.order, self.bc, self.dtype = dim, order, bc, dtype
        self.device = device or torch.device("cuda:0")
        self.count = len(members)
        self.sizes = [m.n if dim == 1 else m.n * m.n for m in members]
        offs, o = [], 0
        for sz in self.sizes:
            offs.append(o)
            o += -(-sz // align) * align          # keep members aligned in the packed arrays
        self.offsets = offs
        self.points = o
        arr = (native.Member * self.count)()
        for i, m in enumerate(members):
            arr[i].n, arr[i].offset, arr[i].dt, arr[i].dx = int(m.n), int(offs[i]), float(m.dt), float(m.dx)
            c = np.zeros(23)
            c[:] = np.asarray(m.coeffs, np.float64)[:23]
            arr[i].coeffs[:] = list(c)
        self.dynamic = bool(np.asarray(members[0].coeffs)[10] != 0.0)
        h = C.c_void_p()
        native._check(native.load().nls_batch_create(dim, order, native.BC[bc], native.DTYPE[dtype],
                                                     native.ENGINE[engine], arr, self.count, C.byref(h)))
        self.handle = h
        # step-block graphs (advance calls of >= 128 steps that every member's budget covers)
        native._check(native.load().nls_batch_set_flags(h, 0 if graphs else native.FLAG_EAGER))
        rt = getattr(torch, REAL[dtype])
        ct = getattr(torch, CPLX[dtype])
        npts = max(self.points, int(native.load().nls_batch_points(h)))
        Ph = np.zeros(npts, REAL[dtype])
        psih = np.zeros(npts, CPLX[dtype])
        nh = np.zeros(npts, REAL[dtype])
        for i, m in enumerate(members):
            a, b = offs[i], offs[i] + self.sizes[i]
            Ph[a:b] = np.asarray(m.P).ravel()
            psih[a:b] = np.asarray(m.psi0).ravel()
            if self.dynamic and m.n0 is not None:
                nh[a:b] = np.asarray(m.n0).ravel()
        self.P = torch.from_numpy(Ph).to(self.device)
        self.psi = torch.from_numpy(psih).to(self.device)
        self.nres = torch.from_numpy(nh).to(self.device) if self.dynamic else None
        self._psi0 = self.psi.clone()
        self._n0 = self.nres.clone() if self.dynamic else None
        wsb = int(native.load().nls_batch_workspace_bytes(h))
        self.ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=self.device)
        self.tiles = int(native.load().nls_batch_tiles(h))

    def reset(self):
        """Restore the initial state of every member."""
        self.psi.copy_(self._psi0)
        if self.dynamic:
            self.nres.copy_(self._n0)

    def set_steps(self, steps, stream=None):
        s = stream if stream is not None else self.torch.cuda.current_stream()
        a = np.ascontiguousarray(steps, np.int32)
        assert a.size == self.count
        native._check(native.load().nls_batch_set_steps(self.handle, a.ctypes.data_as(C.c_void_p),
                                                        C.c_void_p(s.cuda_stream)))

    def advance(self, iters, stream=None):
        s = stream if stream is not None else self.torch.cuda.current_stream()
        native._check(native.load().nls_batch_advance(
            self.handle, int(iters), C.c_void_p(self.P.data_ptr()), C.c_void_p(self.psi.data_ptr()),
            C.c_void_p(self.nres.data_ptr()) if self.nres is not None else None,
            C.c_void_p(self.ws.data_ptr()), self.ws.numel(), C.c_void_p(s.cuda_stream)))

    def prepare(self, iters):
        """Capture the step block an advance of `iters` steps would replay
        (nls_batch_prepare); runs no step."""
        native._check(native.load().nls_batch_prepare(
            self.handle, int(iters), C.c_void_p(self.P.data_ptr()), C.c_void_p(self.psi.data_ptr()),
            C.c_void_p(self.nres.data_ptr()) if self.nres is not None else None,
            C.c_void_p(self.ws.data_ptr()), self.ws.numel()))

    def observables(self, stream=None):
        """Per-member (norm, population, peak density, non-finite count),
        accumulated in double precision on the device (nls_batch_observables):
        a (count, 4) float64 array."""
        s = stream if stream is not None else self.torch.cuda.current_stream()
        out = np.zeros((self.count, 4), np.float64)
        native._check(native.load().nls_batch_observables(
            self.handle, C.c_void_p(self.psi.data_ptr()),
            C.c_void_p(self.nres.data_ptr()) if self.nres is not None else None,
            out.ctypes.data_as(C.c_void_p), C.c_void_p(s.cuda_stream)))
        return out

    def energy(self, stream=None):
        """Per-member (energy, kinetic, interaction, reservoir coupling) of the
        conservative part of the right-hand side, accumulated in double
        precision on the device (nls_batch_energy): a (count, 4) float64
        array."""
        s = stream if stream is not None else self.torch.cuda.current_stream()
        out = np.zeros((self.count, 4), np.float64)
        native._check(native.load().nls_batch_energy(
            self.handle, C.c_void_p(self.P.data_ptr()), C.c_void_p(self.psi.data_ptr()),
            C.c_void_p(self.nres.data_ptr()) if self.nres is not None else None,
            out.ctypes.data_as(C.c_void_p), C.c_void_p(s.cuda_stream)))
        return out

    def member(self, i):
        """(psi, n) of member i as host numpy arrays in its grid shape."""
        a, b = self.offsets[i], self.offsets[i] + self.sizes[i]
        n = int(round(self.sizes[i] ** (0.5 if self.dim == 2 else 1.0)))
        shape = (n,) if self.dim == 1 else (n, n)
        psi = self.psi[a:b].cpu().numpy().reshape(shape)
        nres = self.nres[a:b].cpu().numpy().reshape(shape) if self.dynamic else None
        return psi, nres

    def close(self):
        if getattr(self, "handle", None):
            native.load().nls_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
