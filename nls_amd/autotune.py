"""Measured choice of RK4 engine and launch parameters, cached on disk.

Every candidate that can run a given problem is timed on the current device
with synthetic data of the problem's shape (same n, dtype, order, boundary
family and reservoir mode; the solution values do not change the work).  The
fastest candidate is stored in a JSON file keyed by the device name and the
problem class, so each choice is measured once per machine and reused after.

What is tuned:
  * ``plain_engine(desc)``   -- engine of a plain solve (nls_advance): resident
    (instances that fit on chip), fused (2-D) or per-stage, each with its
    step-block graphs;
  * ``stage_tile(dtype)``    -- tile height (4, 8, 16, 32 rows) of the 2-D
    per-stage kernels (nls_tune_stage_tile), applied to the library;
  * ``batch_engine(members, ...)`` -- engine of a batch (nls_batch_create):
    resident, fused or per-stage for that member mix;
  * ``fused_form(members, ...)`` / ``plain_fused_form(desc)`` -- tile shape
    of the fused kernel (nls_tune_fused_form), applied to the library:
    64 x 32 tiles of 1024 threads ("tall"), 64 x 16 tiles of 1024 threads
    ("short"), or, in double precision on periodic grids, 64 x 16 tiles of 512
    threads with one stage buffer ("half").  The halo is not a free
    parameter: one launch per RK4 step needs the 4R rings its four stages
    consume (R the stencil radius).

The cache file is ``$NLS_AMD_TUNE_CACHE`` or ``~/.cache/nls_amd/tune.json``;
``Tuner(path=None)`` keeps the results in memory only.
"""
import json
import os

import numpy as np

from . import native

ENGINES_2D = ("resident", "fused", "stage")
ENGINES_1D = ("resident", "stage")
STAGE_TILES = (4, 8, 16, 32)


def default_cache_path():
    p = os.environ.get("NLS_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "nls_amd", "tune.json")


def _device_name():
    import torch
    return "%s (%s)" % (torch.cuda.get_device_name(0), getattr(torch.cuda.get_device_properties(0), "gcnArchName", ""))


class Tuner(object):
    """``timer(candidate, setup) -> seconds`` may be injected (tests); by
    default candidates run on the GPU and are timed with HIP events."""

    def __init__(self, path="default", steps=64, reps=3, timer=None, device_name=None):
        self.path = default_cache_path() if path == "default" else path
        self.steps, self.reps = int(steps), int(reps)
        self._timer = timer
        self._device = device_name
        self.table = {}
        self.last = {}            # timings of the last measured choice, by candidate
        if self.path and os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    # ---- cache -------------------------------------------------------------
    def device(self):
        if self._device is None:
            self._device = _device_name()
        return self._device

    def _key(self, kind, **fields):
        return "|".join([self.device(), kind] + ["%s=%s" % (k, fields[k]) for k in sorted(fields)])

    def _save(self):
        if not self.path:
            return
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def _choose(self, key, candidates, run):
        """Cached best of candidates; run(candidate) -> seconds per step."""
        hit = self.table.get(key)
        if hit is not None and hit.get("best") in candidates:
            self.last = dict(hit.get("seconds_per_step", {}))
            return hit["best"]
        times = {}
        for c in candidates:
            times[str(c)] = min(run(c) for _ in range(self.reps))
        best = min(candidates, key=lambda c: times[str(c)])
        self.table[key] = {"best": best, "seconds_per_step": times}
        self.last = times
        self._save()
        return best

    # ---- plain solves ------------------------------------------------------
    def plain_candidates(self, desc):
        cands = []
        for e in (ENGINES_2D if desc.dim == 2 else ENGINES_1D):
            if e == "resident" and not _resident_fits(desc):
                continue
            cands.append(e)
        return cands

    def plain_engine(self, desc):
        """Fastest RK4 engine for a plain solve of desc's problem class."""
        if desc.scheme != native.SCHEME["rk4"]:
            return "auto"
        key = self._key("plain", dim=desc.dim, n=desc.n, dtype=desc.dtype, order=desc.order, bc=desc.bc,
                        dyn=int(desc.dynamic))
        return self._choose(key, self.plain_candidates(desc), lambda e: self._time_plain(desc, e))

    def tuned(self, desc):
        """A copy of desc with the tuned engine."""
        d = native.ProblemDesc()
        C_copy(d, desc)
        d.engine = native.ENGINE[self.plain_engine(desc)]
        return d

    def _time_plain(self, desc, engine):
        if self._timer is not None:
            return self._timer(("plain", engine), desc)
        import torch
        d = native.ProblemDesc()
        C_copy(d, desc)
        d.engine = native.ENGINE[engine]
        dev = torch.device("cuda:0")
        rt = torch.float64 if desc.dtype == 1 else torch.float32
        ct = torch.complex128 if desc.dtype == 1 else torch.complex64
        shape = desc.shape
        P = torch.ones(shape, dtype=rt, device=dev)
        psi = torch.full(shape, 0.1, dtype=ct, device=dev)
        nres = torch.zeros(shape, dtype=rt, device=dev) if desc.dynamic else None
        ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device=dev)
        steps = max(2 * 64, self.steps)        # long enough for the step-block graphs
        native.advance(d, steps, P, psi, nres, ws)          # warm-up (and graph capture)
        return _event_seconds(lambda: native.advance(d, steps, P, psi, nres, ws)) / steps

    # ---- per-stage tile height ----------------------------------------------
    def stage_tile(self, dtype, n=1024, order=5, apply=True):
        """Fastest tile height of the 2-D per-stage kernels on an n x n grid;
        applied to the library (nls_tune_stage_tile) when apply."""
        key = self._key("stage_tile", dtype=dtype, n=n, order=order)
        prev = native.stage_tile(dtype) if self._timer is None else None

        def run(rows):
            if self._timer is not None:
                return self._timer(("stage_tile", rows), dtype)
            native.stage_tile(dtype, rows)
            d = native.ProblemDesc.make(2, n, order, 1e-4, 0.05, _coeffs(True), bc="periodic", dtype=dtype,
                                        engine="stage")
            return self._time_plain(d, "stage")

        best = self._choose(key, list(STAGE_TILES), run)
        if self._timer is None:
            native.stage_tile(dtype, best if apply else prev)
        return best

    # ---- fused kernel tile form ---------------------------------------------
    @staticmethod
    def form_candidates(dtype, bc, order):
        if order == 7:
            return ["tall"]                        # one form at order 7
        periodic = bc in ("periodic", 1)
        return (["tall", "short"] + (["half"] if dtype in ("f64", 1) and periodic else []) +
                (["tall512"] if dtype in ("f32", 0) and periodic and order == 5 else []))

    def _form_choice(self, key, dtype, cands, run, apply):
        prev = native.fused_form(dtype) if self._timer is None else None

        def timed(form):
            if self._timer is None:
                native.fused_form(dtype, form)
            return run(form)

        best = self._choose(key, cands, timed)
        if self._timer is None:
            native.fused_form(dtype, best if apply else prev)
        return best

    def fused_form(self, members, dim=2, order=5, bc="periodic", dtype="f32", engine="fused", apply=True):
        """Fastest tile form of the fused kernel for a batch of this member
        mix on `engine` ("fused" or "hybrid"); applied to the library."""
        sizes = {}
        for m in members:
            sizes[int(m.n)] = sizes.get(int(m.n), 0) + 1
        mix = ",".join("%dx%d" % (n, c) for n, c in sorted(sizes.items()))
        dyn = int(np.asarray(members[0].coeffs)[10] != 0.0)
        key = self._key("fused_form", dim=dim, order=order, bc=bc, dtype=dtype, dyn=dyn, mix=mix, engine=engine)

        def run(form):
            if self._timer is not None:
                return self._timer(("fused_form", form), members)
            return self._time_batch(members, dim, order, bc, dtype, engine)

        return self._form_choice(key, dtype, self.form_candidates(dtype, bc, order), run, apply)

    def plain_fused_form(self, desc, apply=True):
        """Fastest tile form of the fused kernel for a plain 2-D solve of
        desc's problem class; applied to the library."""
        dtype = "f64" if desc.dtype == 1 else "f32"
        bc = "periodic" if desc.bc == 1 else "ref"
        key = self._key("plain_fused_form", n=desc.n, dtype=dtype, order=desc.order, bc=bc, dyn=int(desc.dynamic))

        def run(form):
            if self._timer is not None:
                return self._timer(("plain_fused_form", form), desc)
            return self._time_plain(desc, "fused")

        return self._form_choice(key, dtype, self.form_candidates(dtype, bc, desc.order), run, apply)

    # ---- batches -----------------------------------------------------------
    def batch_engine(self, members, dim=2, order=5, bc="periodic", dtype="f32"):
        """Fastest batch engine for this member mix (sizes and count)."""
        sizes = {}
        for m in members:
            sizes[int(m.n)] = sizes.get(int(m.n), 0) + 1
        mix = ",".join("%dx%d" % (n, c) for n, c in sorted(sizes.items()))
        dyn = int(np.asarray(members[0].coeffs)[10] != 0.0)
        key = self._key("batch", dim=dim, order=order, bc=bc, dtype=dtype, dyn=dyn, mix=mix)
        cands = [e for e in (ENGINES_2D if dim == 2 else ENGINES_1D)
                 if e != "resident" or _batch_resident_fits(members, dim, order, dtype)]
        fit = sum(_batch_resident_fits([m], dim, order, dtype) for m in members)
        if 0 < fit < len(members):
            cands.append("hybrid")           # the members that fit resident, the rest on the fused/per-stage engine
        return self._choose(key, cands, lambda e: self._time_batch(members, dim, order, bc, dtype, e))

    def _time_batch(self, members, dim, order, bc, dtype, engine):
        if self._timer is not None:
            return self._timer(("batch", engine), members)
        from .batch import Batch
        b = Batch(members, dim=dim, order=order, bc=bc, dtype=dtype, engine=engine)
        try:
            b.advance(4)
            steps = self.steps
            return _event_seconds(lambda: b.advance(steps)) / steps
        finally:
            b.close()


def _event_seconds(fn):
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e-3


def _coeffs(dynamic):
    from . import configs
    return configs.model_coeffs(dynamic=dynamic)


def C_copy(dst, src):
    import ctypes
    ctypes.memmove(ctypes.addressof(dst), ctypes.addressof(src), ctypes.sizeof(src))


RES_MAXP = 4096


def _resident_fits(desc):
    if desc.dim == 1:
        return desc.n <= RES_MAXP
    return desc.n * desc.n <= RES_MAXP and not (desc.dtype == 1 and desc.order == 7)


def _batch_resident_fits(members, dim, order, dtype):
    for m in members:
        pts = m.n if dim == 1 else m.n * m.n
        if pts > RES_MAXP:
            return False
    return not (dim == 2 and dtype == "f64" and order == 7)
