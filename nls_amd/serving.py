"""Continuous-batching request loop for solves on one GPU.

The reference's only "serving" is a pull queue in which each worker process
runs one solve at a time (schedulr/schedulr.py:78-126 claim/complete,
schedulr/workr.py:39-107).  Here requests share the GPU:

* admission by HBM budget: each request gets a region of a device pool
  (state psi, reservoir n, pumping P) from a first-fit allocator; requests
  that do not fit wait in arrival order, later ones that fit go ahead;
* the state stays on the device for the whole solve; inputs are uploaded
  once at admission -- or, for requests given by parameters, generated in
  place on the device (nls_amd.fields) -- and the result downloaded once at
  completion;
* three pools, each one batch: 1-D requests (n <= 4096) and 2-D requests up
  to 64 x 64 on the resident engine (a workgroup per request runs every step
  of a block in one launch), larger 2-D requests on the fused RK4 engine (one
  launch per step); the HBM budget is split 15% / 15% / 70%;
* every block advances all active requests together, each by
  min(block, remaining) steps (nls_batch_set_steps), so a request that
  finishes mid-block holds its state and its slot is freed at the end of the
  block, when waiting requests are admitted; the block length adapts so one
  block takes about target_block_ms (20 ms), or idle_block_ms (5 ms) while
  the load is light (none waiting, at most idle_max_active in flight), so a
  request waits less for block boundaries;
* each pool runs on its own stream (default priority), so the 1-D lines,
  small grids and large grids of a block run on the device at the same time;
* a member's arithmetic never depends on the other members, so a result is
  bit-identical alone and in any mix.

Latency = time from arrival to the result being on the host.
"""
import bisect
import math
import time

import numpy as np

from . import configs, fields, native

ALIGN = 64          # points; members start on 64-point boundaries


class Request(object):
    """One solve.  Inputs either as host arrays (P, psi0, n0) or as a `spec`
    of parameters generated on the device at admission (nls_amd.fields):
    dict(grid=Grid, pump=Profile, state=Profile, noise=std, seed=int); the
    reservoir then starts at zero."""

    def __init__(self, rid, arrival, dim, n, steps, dt, dx, coeffs, P=None, psi0=None, n0=None, spec=None):
        self.rid, self.arrival, self.dim, self.n, self.steps = rid, arrival, dim, n, steps
        self.dt, self.dx, self.coeffs, self.P, self.psi0, self.n0 = dt, dx, coeffs, P, psi0, n0
        self.spec = spec
        self.remaining = steps
        self.offset = None
        self.admit_time = self.finish_time = None
        self.psi = self.nres = None

    @property
    def points(self):
        return self.n if self.dim == 1 else self.n * self.n

    @property
    def latency(self):
        return None if self.finish_time is None else self.finish_time - self.arrival


class PoolAllocator(object):
    """First-fit allocator over [0, capacity) in ALIGN-point units."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.free = [(0, capacity)]          # sorted (start, length)

    def alloc(self, size):
        size = -(-size // ALIGN) * ALIGN
        for i, (s, l) in enumerate(self.free):
            if l >= size:
                if l == size:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + size, l - size)
                return s
        return None

    def release(self, start, size):
        size = -(-size // ALIGN) * ALIGN
        i = bisect.bisect_left(self.free, (start, 0))
        self.free.insert(i, (start, size))
        merged = []
        for s, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((s, l))
        self.free = merged


class Group(object):
    """Device pool and batch for the requests of one dimensionality."""

    def __init__(self, dim, capacity, dtype, device, engine, order=5, bc="periodic"):
        import torch
        self.torch = torch
        self.dim, self.capacity, self.dtype, self.device = dim, capacity, dtype, device
        self.order, self.bc = order, bc
        self.engine = engine
        rt = torch.float32 if dtype == "f32" else torch.float64
        ct = torch.complex64 if dtype == "f32" else torch.complex128
        self.psi = torch.zeros(capacity, dtype=ct, device=device)
        self.nres = torch.zeros(capacity, dtype=rt, device=device)
        self.P = torch.zeros(capacity, dtype=rt, device=device)
        ws_per_point = (8 if dtype == "f32" else 16) + (4 if dtype == "f32" else 8)
        factor = {"resident": 0, "fused": 1, "stage": 3}[engine]
        self.ws = torch.empty(factor * capacity * ws_per_point + 4096, dtype=torch.uint8, device=device)
        self.alloc = PoolAllocator(capacity)
        self.members = []
        self.handle = None
        self.dirty = True
        # the pool's own stream (default priority): the pools' launches
        # overlap on the device; admission fills and steps are ordered on it
        self.stream = torch.cuda.Stream(device=device)

    @staticmethod
    def bytes_per_point(engine, dtype):
        s = (8 + 4) if dtype == "f32" else (16 + 8)
        r = 4 if dtype == "f32" else 8
        return (s + r) + {"resident": 0, "fused": 1, "stage": 3}[engine] * s

    def admit(self, req):
        off = self.alloc.alloc(req.points)
        if off is None:
            return False
        req.offset = off
        a, b = off, off + req.points
        t = self.torch
        ct = self.psi.dtype
        if req.spec is not None:
            # inputs generated in place on the device: no upload
            sp = req.spec
            fields.fill_profile(self.P[a:b], sp["pump"], req.dim, req.n, sp["grid"])
            fields.fill_state(self.psi[a:b], sp["state"], sp["noise"], sp["seed"], req.dim, req.n, sp["grid"])
            self.nres[a:b].zero_()
        else:
            self.psi[a:b] = t.from_numpy(np.ascontiguousarray(req.psi0.ravel())).to(self.device, ct,
                                                                                    non_blocking=True)
            self.P[a:b] = t.from_numpy(np.ascontiguousarray(req.P.ravel())).to(self.device, self.P.dtype,
                                                                              non_blocking=True)
            n0 = np.zeros(req.points) if req.n0 is None else req.n0.ravel()
            self.nres[a:b] = t.from_numpy(np.ascontiguousarray(n0)).to(self.device, self.nres.dtype,
                                                                        non_blocking=True)
        self.members.append(req)
        self.dirty = True
        return True

    def _rebuild(self):
        import ctypes as C
        if self.handle is not None:
            native.load().nls_batch_destroy(self.handle)
            self.handle = None
        if not self.members:
            return
        arr = (native.Member * len(self.members))()
        for i, r in enumerate(self.members):
            arr[i].n, arr[i].offset, arr[i].dt, arr[i].dx = r.n, r.offset, r.dt, r.dx
            arr[i].coeffs[:] = list(np.asarray(r.coeffs, np.float64)[:23])
        h = C.c_void_p()
        native._check(native.load().nls_batch_create(self.dim, self.order, native.BC[self.bc],
                                                     native.DTYPE[self.dtype], native.ENGINE[self.engine],
                                                     arr, len(self.members), C.byref(h)))
        self.handle = h
        self.dirty = False

    def advance(self, block, stream):
        """Advance every member by min(block, remaining) steps."""
        import ctypes as C
        if not self.members:
            return
        if self.dirty:
            self._rebuild()
        L = native.load()
        budgets = np.array([min(block, r.remaining) for r in self.members], np.int32)
        native._check(L.nls_batch_set_steps(self.handle, budgets.ctypes.data_as(C.c_void_p),
                                            C.c_void_p(stream.cuda_stream)))
        need = int(L.nls_batch_workspace_bytes(self.handle))
        assert need <= self.ws.numel()
        native._check(L.nls_batch_advance(self.handle, int(budgets.max()), C.c_void_p(self.P.data_ptr()),
                                          C.c_void_p(self.psi.data_ptr()), C.c_void_p(self.nres.data_ptr()),
                                          C.c_void_p(self.ws.data_ptr()), self.ws.numel(),
                                          C.c_void_p(stream.cuda_stream)))
        for r, b in zip(self.members, budgets):
            r.remaining -= int(b)

    def retire(self, now_fn):
        """Download and free the members that have finished, with their FP64
        observables (norm, population, peak density, non-finite count,
        energy)."""
        import ctypes as C
        done = [r for r in self.members if r.remaining == 0]
        if not done:
            return []
        obs = np.zeros((len(self.members), 4), np.float64)      # one launch for the whole pool
        native._check(native.load().nls_batch_observables(
            self.handle, C.c_void_p(self.psi.data_ptr()), C.c_void_p(self.nres.data_ptr()),
            obs.ctypes.data_as(C.c_void_p), C.c_void_p(self.stream.cuda_stream)))
        en = np.zeros((len(self.members), 4), np.float64)
        native._check(native.load().nls_batch_energy(
            self.handle, C.c_void_p(self.P.data_ptr()), C.c_void_p(self.psi.data_ptr()),
            C.c_void_p(self.nres.data_ptr()), en.ctypes.data_as(C.c_void_p), C.c_void_p(self.stream.cuda_stream)))
        index = {id(r): i for i, r in enumerate(self.members)}
        for r in done:
            o = obs[index[id(r)]]
            r.observables = {"norm": float(o[0]), "population": float(o[1]), "peak": float(o[2]),
                             "nonfinite": int(o[3]), "energy": float(en[index[id(r)]][0])}
        for r in done:
            a, b = r.offset, r.offset + r.points
            shape = (r.n,) if r.dim == 1 else (r.n, r.n)
            r.psi = self.psi[a:b].cpu().numpy().reshape(shape)
            r.nres = self.nres[a:b].cpu().numpy().reshape(shape)
            r.finish_time = now_fn()
            self.alloc.release(r.offset, r.points)
            self.members.remove(r)
        self.dirty = True
        return done

    def close(self):
        if self.handle is not None:
            native.load().nls_batch_destroy(self.handle)
            self.handle = None


RESIDENT_MAX = {1: 4096, 2: 64}      # largest n a single workgroup holds


class Engine(object):
    """Continuous batching under an HBM budget (bytes).  Three pools: 1-D
    requests and 2-D requests up to 64^2 on the resident engine (a
    workgroup per request, every step of a block in one launch), larger 2-D
    requests on the fused engine (one launch per step)."""

    SHARES = {"line": 0.15, "small": 0.15, "large": 0.70}

    def __init__(self, budget_bytes=8 << 30, dtype="f32", device=None, target_block_ms=20.0,
                 min_block=8, max_block=2000, idle_block_ms=5.0, idle_max_active=32):
        import torch
        self.torch = torch
        self.device = device or torch.device("cuda:0")
        self.dtype = dtype
        spec = {"line": (1, "resident"), "small": (2, "resident"), "large": (2, "fused")}
        self.groups = {}
        for key, (dim, eng) in spec.items():
            cap = int(budget_bytes * self.SHARES[key] // Group.bytes_per_point(eng, dtype)) // ALIGN * ALIGN
            self.groups[key] = Group(dim, cap, dtype, self.device, eng)
        self.target_block_ms, self.min_block, self.max_block = target_block_ms, min_block, max_block
        # shorter blocks while the load is light (no request waits for
        # admission and at most idle_max_active are in flight): a request
        # admitted or finished mid-block waits less for the block's end
        self.idle_block_ms = target_block_ms if idle_block_ms is None else idle_block_ms
        self.idle_max_active = idle_max_active
        self.block = min_block
        self.blocks = 0

    def run(self, requests, clock=None, log=None):
        """Serve `requests` (sorted by arrival, seconds from start).  Returns them
        with admit/finish times filled in."""
        t0 = time.perf_counter()
        now = clock or (lambda: time.perf_counter() - t0)
        pending = list(requests)
        waiting = []
        finished = []
        while pending or waiting or any(g.members for g in self.groups.values()):
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.pop(0))
            still = []
            for r in waiting:                       # arrival order; later ones may pass
                g = self.groups[self.group_of(r)]
                with self.torch.cuda.stream(g.stream):
                    ok = g.admit(r)
                if ok:
                    r.admit_time = now()
                else:
                    still.append(r)
            waiting = still
            active = [g for g in self.groups.values() if g.members]
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.002)))
                continue
            tb = time.perf_counter()
            for g in active:                        # each pool on its own stream, concurrently
                g.advance(self.block, g.stream)
            for g in active:
                g.stream.synchronize()
            for g in active:
                finished += g.retire(now)
            self.blocks += 1
            el = (time.perf_counter() - tb) * 1e3
            per_step = el / max(self.block, 1)
            light = not waiting and sum(len(g.members) for g in active) <= self.idle_max_active
            target = self.idle_block_ms if light else self.target_block_ms
            self.block = int(min(self.max_block, max(self.min_block, target / max(per_step, 1e-4))))
            if log and self.blocks % 50 == 0:
                log("t=%.2fs finished=%d waiting=%d active=%d block=%d" % (
                    now(), len(finished), len(waiting), sum(len(g.members) for g in self.groups.values()),
                    self.block))
        return finished

    @staticmethod
    def group_of(req):
        if req.dim == 1:
            return "line"
        return "small" if req.n <= RESIDENT_MAX[2] else "large"

    def close(self):
        for g in self.groups.values():
            g.close()


# ---------------------------------------------------------------------------
# The serving workload: Poisson arrivals of mixed 1-D / 2-D solves.
# ---------------------------------------------------------------------------

def stable_dt(dim, h, order=5, safety=0.5):
    """safety * 2 sqrt(2) / |lambda_max| of the order-`order` periodic Laplacian."""
    sym = {3: 4.0, 5: 16.0 / 3.0, 7: 272.0 / 45.0}[order]
    return safety * 2.0 * math.sqrt(2.0) / (dim * sym / (h * h))


def make_request(rid, arrival, rng, step_range=(1e3, 1e5), device_inputs=False):
    """A request of the serving mix.  device_inputs: give the pumping and the
    initial state as parameters (Gaussian or single-Gaussian ring pumping,
    0.1 exp(-r^2/4) plus seeded noise of std 1e-3), generated on the device
    at admission, instead of host arrays."""
    if device_inputs:
        return _make_spec_request(rid, arrival, rng, step_range)
    dim = 1 if rng.random() < 0.5 else 2
    if dim == 1:
        n = int(round(math.exp(rng.uniform(math.log(256), math.log(4096))) / 64.0)) * 64
        L = 64.0
        x = -L / 2 + L * np.arange(n) / n
        r = np.abs(x)
        h = L / n
    else:
        n = int(round(math.exp(rng.uniform(math.log(64), math.log(512))) / 64.0)) * 64
        L = 32.0
        X, Y, h = configs.grid_2d(n, L)
        r = np.sqrt(X ** 2 + Y ** 2)
    steps = int(round(math.exp(rng.uniform(math.log(step_range[0]), math.log(step_range[1])))))
    P0, r0, w = rng.uniform(0.5, 5.0), rng.uniform(3.0, 8.0), rng.uniform(0.5, 2.0)
    P = P0 * np.exp(-r ** 2 / (2 * w * w)) if rng.random() < 0.5 else configs.ring(r, P0, r0, w)
    psi0 = 0.1 * np.exp(-r ** 2 / 4.0) + configs.noise(rng, r.shape)
    return Request(rid, arrival, dim, n, steps, stable_dt(dim, h), h, configs.C_MODEL.copy(), P, psi0,
                   np.zeros(r.shape))


def _make_spec_request(rid, arrival, rng, step_range):
    dim = 1 if rng.random() < 0.5 else 2
    if dim == 1:
        n = int(round(math.exp(rng.uniform(math.log(256), math.log(4096))) / 64.0)) * 64
        L = 64.0
    else:
        n = int(round(math.exp(rng.uniform(math.log(64), math.log(512))) / 64.0)) * 64
        L = 32.0
    h = L / n
    steps = int(round(math.exp(rng.uniform(math.log(step_range[0]), math.log(step_range[1])))))
    P0, r0, w = rng.uniform(0.5, 5.0), rng.uniform(3.0, 8.0), rng.uniform(0.5, 2.0)
    pump = (fields.Profile("gaussian", P0, width=w) if rng.random() < 0.5
            else fields.Profile("ring", P0, width=w, radius=r0))
    spec = {"grid": fields.Grid("fraction", -L / 2.0, L), "pump": pump,
            "state": fields.Profile("gaussian", 0.1, width=math.sqrt(2.0)), "noise": 1e-3,
            "seed": int(rng.integers(0, 2 ** 62))}
    return Request(rid, arrival, dim, n, steps, stable_dt(dim, h), h, configs.C_MODEL.copy(), spec=spec)


def poisson_requests(rate, duration, seed=0, step_range=(1e3, 1e5), device_inputs=False):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        out.append(make_request(len(out), t, rng, step_range, device_inputs=device_inputs))
    return out


def summarize(done, wall):
    lat = np.array([r.latency for r in done])
    work = sum(r.points * r.steps for r in done)
    return {"requests": len(done), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "throughput_req_s": len(done) / wall,
            "grid_point_steps_per_s": work / wall}


def sweep(rates=(20, 50, 100, 200, 500), duration=4.0, budget_gb=16, seed=0, step_range=(1e3, 1e5), log=print):
    import torch
    out = []
    for lam in rates:
        eng = Engine(budget_bytes=int(budget_gb * (1 << 30)))
        reqs = poisson_requests(lam, duration, seed=seed, step_range=step_range)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = eng.run(reqs)
        wall = time.perf_counter() - t0
        s = summarize(done, wall)
        s["rate"] = lam
        s["arrival_window_s"] = duration
        s["wall_s"] = wall
        eng.close()
        if log:
            log(s)
        out.append(s)
    return out


if __name__ == "__main__":
    import json
    import sys
    rates = [float(x) for x in sys.argv[1:]] or [20, 50, 100, 200, 500]
    print(json.dumps(sweep(rates)))
