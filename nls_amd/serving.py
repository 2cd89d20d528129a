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
* three pools, each one persistent batch: 1-D requests (n <= 4096) and 2-D
  requests up to 64 x 64 on the resident engine (a workgroup per request
  runs every step of a block in one launch), larger 2-D requests on the
  fused RK4 engine (one launch per step); the HBM budget is split
  15% / 15% / 70%.  A pool's batch is created once and then updated
  (nls_batch_update: members, budgets and tables uploaded asynchronously on
  the pool's stream), never re-created;
* every block advances all active requests together, each by
  min(block, remaining) steps, so a request that finishes mid-block holds
  its state; its FP64 observables, energy and state are copied to pinned
  host memory on the pool's stream right behind the block, and its slot is
  reused by the next admission (ordered after that copy on the same
  stream).  The host keeps up to two blocks enqueued per pool and collects
  a block's results once its end event has passed, so the device does not
  wait for the host between blocks;
* failure detection: every block computes every member's FP64 observables;
  a request whose state holds non-finite points (a dt above the scheme's
  stability limit) is evicted and returned with error = NLS_ENONFINITE,
  while the others continue unchanged (the reference reports nothing,
  nls.f90:725-733);
* the block length adapts so one block takes about target_block_ms (20 ms),
  or idle_block_ms (5 ms) while the load is light (none waiting, at most
  idle_max_active in flight), so a request waits less for block boundaries;
* each pool runs on its own stream (default priority), so the 1-D lines,
  small grids and large grids run on the device at the same time;
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
        self.error = None               # NLS_ENONFINITE when evicted with a non-finite state
        self.observables = None
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


class Block(object):
    """One block issued on a pool: the requests it advanced, their budgets,
    the device-side FP64 observables being copied to pinned host memory, and
    the events that bracket it on the pool's stream."""

    def __init__(self, members, budgets, finishing, start, end, obs, energy, work):
        self.members, self.budgets, self.finishing = members, budgets, finishing
        self.start, self.end, self.obs, self.energy, self.work = start, end, obs, energy, work


class Group(object):
    """Device pool and one persistent batch for the requests of one engine.

    The batch is created at the first admission and then only updated
    (nls_batch_update: members, budgets and tables uploaded asynchronously on
    the pool's stream); blocks are enqueued without waiting for the previous
    one, and their results are collected when their end event has passed."""

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
        self.inflight = []                  # Blocks enqueued and not yet collected
        self.step_ms = None                 # measured time per step of this pool's blocks
        # the pool's own stream (default priority): the pools' launches
        # overlap on the device; admission fills and blocks are ordered on it
        self.stream = torch.cuda.Stream(device=device)

    @staticmethod
    def bytes_per_point(engine, dtype):
        s = (8 + 4) if dtype == "f32" else (16 + 8)
        r = 4 if dtype == "f32" else 8
        return (s + r) + {"resident": 0, "fused": 1, "stage": 3}[engine] * s

    def admit(self, req):
        """Place a request in the pool and enqueue its inputs on the pool's
        stream (generated on the device, or uploaded from pinned memory)."""
        off = self.alloc.alloc(req.points)
        if off is None:
            return False
        req.offset = off
        a, b = off, off + req.points
        t = self.torch
        with t.cuda.stream(self.stream):
            if req.spec is not None:
                sp = req.spec
                fields.fill_profile(self.P[a:b], sp["pump"], req.dim, req.n, sp["grid"], stream=self.stream)
                fields.fill_state(self.psi[a:b], sp["state"], sp["noise"], sp["seed"], req.dim, req.n, sp["grid"],
                                  stream=self.stream)
                self.nres[a:b].zero_()
            else:
                def put(dst, x):
                    h = t.from_numpy(np.ascontiguousarray(np.asarray(x).ravel())).to(dst.dtype).pin_memory()
                    dst.copy_(h, non_blocking=True)
                put(self.psi[a:b], req.psi0)
                put(self.P[a:b], req.P)
                put(self.nres[a:b], np.zeros(req.points) if req.n0 is None else req.n0)
        self.members.append(req)
        self.dirty = True
        return True

    def _table(self):
        arr = (native.Member * len(self.members))()
        for i, r in enumerate(self.members):
            arr[i].n, arr[i].offset, arr[i].dt, arr[i].dx = r.n, r.offset, r.dt, r.dx
            arr[i].coeffs[:] = list(np.asarray(r.coeffs, np.float64)[:23])
        return arr

    def issue(self, block):
        """Enqueue one block: every member advances min(block, remaining)
        steps; FP64 observables of every member (and the energy and result
        download of those that finish) follow on the same stream.  Finished
        members leave the pool at once: later fills of their slots are
        ordered after the downloads on this stream."""
        import ctypes as C
        t = self.torch
        L = native.load()
        st = C.c_void_p(self.stream.cuda_stream)
        members = list(self.members)
        budgets = np.array([min(block, r.remaining) for r in members], np.int32)
        if self.handle is None:
            h = C.c_void_p()
            native._check(L.nls_batch_create(self.dim, self.order, native.BC[self.bc], native.DTYPE[self.dtype],
                                             native.ENGINE[self.engine], self._table(), len(members), C.byref(h)))
            self.handle = h
            native._check(L.nls_batch_set_flags(h, native.FLAG_EAGER))
            self.dirty = True
        if self.dirty:
            native._check(L.nls_batch_update(self.handle, self._table(), len(members),
                                             budgets.ctypes.data_as(C.c_void_p), st))
            self.dirty = False
        else:
            native._check(L.nls_batch_set_steps(self.handle, budgets.ctypes.data_as(C.c_void_p), st))
        assert int(L.nls_batch_workspace_bytes(self.handle)) <= self.ws.numel()
        start, end = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        finishing = [r for r, b in zip(members, budgets) if int(b) == r.remaining]
        with t.cuda.stream(self.stream):
            start.record(self.stream)
            native._check(L.nls_batch_advance(self.handle, int(budgets.max()), C.c_void_p(self.P.data_ptr()),
                                              C.c_void_p(self.psi.data_ptr()), C.c_void_p(self.nres.data_ptr()),
                                              C.c_void_p(self.ws.data_ptr()), self.ws.numel(), st))
            obs_d = t.empty(4 * len(members), dtype=t.float64, device=self.device)
            native._check(L.nls_batch_observables_async(self.handle, C.c_void_p(self.psi.data_ptr()),
                                                        C.c_void_p(self.nres.data_ptr()),
                                                        C.c_void_p(obs_d.data_ptr()), st))
            obs = t.empty(obs_d.shape, dtype=t.float64, pin_memory=True)
            obs.copy_(obs_d, non_blocking=True)
            energy = None
            if finishing:
                en_d = t.empty(4 * len(members), dtype=t.float64, device=self.device)
                native._check(L.nls_batch_energy_async(self.handle, C.c_void_p(self.P.data_ptr()),
                                                       C.c_void_p(self.psi.data_ptr()),
                                                       C.c_void_p(self.nres.data_ptr()),
                                                       C.c_void_p(en_d.data_ptr()), st))
                energy = t.empty(en_d.shape, dtype=t.float64, pin_memory=True)
                energy.copy_(en_d, non_blocking=True)
                for r in finishing:
                    a, b = r.offset, r.offset + r.points
                    r._psi_host = t.empty(r.points, dtype=self.psi.dtype, pin_memory=True)
                    r._nres_host = t.empty(r.points, dtype=self.nres.dtype, pin_memory=True)
                    r._psi_host.copy_(self.psi[a:b], non_blocking=True)
                    r._nres_host.copy_(self.nres[a:b], non_blocking=True)
            end.record(self.stream)
        work = 0
        for r, b in zip(members, budgets):
            r.remaining -= int(b)
            work += r.points * int(b)
        for r in finishing:
            self.alloc.release(r.offset, r.points)
            self.members.remove(r)
            self.dirty = True
        blk = Block(members, budgets, finishing, start, end, obs, energy, work)
        self.inflight.append(blk)
        return blk

    def collect(self, now_fn, wait=False):
        """Results of the blocks whose end event has passed (in issue order):
        finished requests get their state, FP64 observables and energy; a
        request whose state holds non-finite points is evicted with
        error = NLS_ENONFINITE.  Returns (requests done, blocks collected)."""
        done, blocks = [], []
        while self.inflight and (wait or self.inflight[0].end.query()):
            blk = self.inflight.pop(0)
            blk.end.synchronize()
            wait = False
            obs = blk.obs.numpy().reshape(-1, 4)
            en = blk.energy.numpy().reshape(-1, 4) if blk.energy is not None else None
            ms = blk.start.elapsed_time(blk.end)
            steps = int(blk.budgets.max()) if len(blk.budgets) else 0
            if steps:
                per = ms / steps
                self.step_ms = per if self.step_ms is None else 0.5 * (self.step_ms + per)
            fin = set(id(r) for r in blk.finishing)
            for i, r in enumerate(blk.members):
                bad = obs[i, 3] > 0
                if id(r) in fin:
                    r.observables = {"norm": float(obs[i, 0]), "population": float(obs[i, 1]),
                                     "peak": float(obs[i, 2]), "nonfinite": int(obs[i, 3]),
                                     "energy": float(en[i, 0])}
                    shape = (r.n,) if r.dim == 1 else (r.n, r.n)
                    r.psi = r._psi_host.numpy().reshape(shape)
                    r.nres = r._nres_host.numpy().reshape(shape)
                    r._psi_host = r._nres_host = None
                    if bad:
                        r.error = native.NLS_ENONFINITE
                    r.finish_time = now_fn()
                    done.append(r)
                elif bad and r.error is None and r in self.members:
                    # blown up (e.g. dt above the scheme's stability limit):
                    # evicted; the blocks already enqueued only touch its slot
                    r.error = native.NLS_ENONFINITE
                    r.observables = {"norm": float(obs[i, 0]), "population": float(obs[i, 1]),
                                     "peak": float(obs[i, 2]), "nonfinite": int(obs[i, 3]), "energy": None}
                    self.alloc.release(r.offset, r.points)
                    self.members.remove(r)
                    self.dirty = True
                    r.finish_time = now_fn()
                    done.append(r)
            blocks.append(blk)
        return done, blocks

    def close(self):
        if self.handle is not None:
            self.stream.synchronize()
            native.load().nls_batch_destroy(self.handle)
            self.handle = None


RESIDENT_MAX = {1: 4096, 2: 64}      # largest n a single workgroup holds


class Engine(object):
    """Continuous batching under an HBM budget (bytes).  Three pools: 1-D
    requests and 2-D requests up to 64^2 on the resident engine (a
    workgroup per request, every step of a block in one launch), larger 2-D
    requests on the fused engine (one launch per step).  Each pool keeps up
    to `inflight` blocks enqueued, so the device never waits for the host
    between blocks."""

    SHARES = {"line": 0.15, "small": 0.15, "large": 0.70}

    def __init__(self, budget_bytes=8 << 30, dtype="f32", device=None, target_block_ms=20.0,
                 min_block=8, max_block=2000, idle_block_ms=5.0, idle_max_active=32, inflight=2):
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
        self.inflight = inflight
        self.blocks = 0
        self.block_log = []          # (host time the block was collected, grid-point-steps it advanced)

    def _block_len(self, g, light):
        target = self.idle_block_ms if light else self.target_block_ms
        if g.step_ms is None:
            return self.min_block
        return int(min(self.max_block, max(self.min_block, target / max(g.step_ms, 1e-4))))

    def run(self, requests, clock=None, log=None):
        """Serve `requests` (sorted by arrival, seconds from start).  Returns
        them as they finish, with admit/finish times (and, for a request
        evicted because its state went non-finite, error = NLS_ENONFINITE)."""
        t0 = time.perf_counter()
        now = clock or (lambda: time.perf_counter() - t0)
        pending = list(requests)
        for r in pending:
            r.error = None
        pi = 0
        waiting = []
        finished = []
        groups = list(self.groups.values())
        while pi < len(pending) or waiting or any(g.members or g.inflight for g in groups):
            t = now()
            while pi < len(pending) and pending[pi].arrival <= t:
                waiting.append(pending[pi])
                pi += 1
            for g in groups:
                done, blocks = g.collect(now)
                finished += done
                for blk in blocks:
                    self.block_log.append((now(), blk.work))
            still = []
            for r in waiting:                       # arrival order; later ones may pass
                g = self.groups[self.group_of(r)]
                if g.admit(r):
                    r.admit_time = now()
                else:
                    still.append(r)
            waiting = still
            light = not waiting and sum(len(g.members) for g in groups) <= self.idle_max_active
            issued = False
            for g in groups:
                if g.members and len(g.inflight) < self.inflight:
                    g.issue(self._block_len(g, light))
                    self.blocks += 1
                    issued = True
            if issued:
                continue
            if any(g.inflight for g in groups):
                time.sleep(0.0002)                  # every pool has its blocks queued: poll again
            elif pi < len(pending):
                time.sleep(max(0.0, min(pending[pi].arrival - now(), 0.002)))
            if log and self.blocks % 50 == 0:
                log("t=%.2fs finished=%d waiting=%d active=%d" % (
                    now(), len(finished), len(waiting), sum(len(g.members) for g in groups)))
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


def summarize(done, wall, engine=None, window=None, offered=None):
    """Latency percentiles and throughput of a served run.  With the engine's
    block log and the arrival window W: the steady-state rate is the work of
    the blocks collected during [0.2 W, W] (arrivals still coming, the
    first fifth left out as ramp-up) per second, without the drain after the
    last arrival."""
    ok = [r for r in done if r.error is None]
    lat = np.array([r.latency for r in ok])
    work = sum(r.points * r.steps for r in ok)
    out = {"requests": len(done), "failed": len(done) - len(ok), "p50_s": float(np.percentile(lat, 50)),
           "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()), "max_s": float(lat.max()),
           "throughput_req_s": len(ok) / wall, "grid_point_steps_per_s": work / wall}
    if engine is not None and window:
        a, b = 0.2 * window, window
        steady = sum(w for t, w in engine.block_log if a <= t <= b)
        out["steady_grid_point_steps_per_s"] = steady / (b - a)
        out["steady_window_s"] = [a, b]
    if offered is not None and window:
        out["offered_grid_point_steps_per_s"] = offered / window
    return out


def mean_request_work(samples=20000, seed=12345, step_range=(1e3, 1e5)):
    """Mean grid-point-steps of a request of the serving mix (sampled)."""
    rng = np.random.default_rng(seed)
    tot = 0.0
    for _ in range(samples):
        dim = 1 if rng.random() < 0.5 else 2
        lo, hi = (256, 4096) if dim == 1 else (64, 512)
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))) / 64.0)) * 64
        steps = int(round(math.exp(rng.uniform(math.log(step_range[0]), math.log(step_range[1])))))
        tot += (n if dim == 1 else n * n) * steps
    return tot / samples


def window_for(rate, min_requests=240, min_window=3.0):
    """Arrival window (s) at `rate`: about min_requests expected arrivals
    (so a Poisson draw holds at least ~200) and at least min_window seconds,
    so that the steady-state part [0.2 W, W] of an overloaded run is past the
    ramp-up."""
    return max(min_window, min_requests / float(rate))


def serve(rate, window=None, budget_gb=16, seed=0, step_range=(1e3, 1e5), device_inputs=True):
    """One run of the request loop: Poisson arrivals at `rate` over `window`
    seconds (default window_for(rate)), the serving mix; returns the
    summary."""
    import torch
    window = window or window_for(rate)
    eng = Engine(budget_bytes=int(budget_gb * (1 << 30)))
    reqs = poisson_requests(rate, window, seed=seed, step_range=step_range, device_inputs=device_inputs)
    offered = sum(r.points * r.steps for r in reqs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = eng.run(reqs)
    wall = time.perf_counter() - t0
    eng.close()
    s = summarize(done, wall, engine=eng, window=window, offered=offered)
    s.update({"rate_per_s": rate, "arrival_window_s": window, "wall_s": wall, "blocks": eng.blocks})
    return s


def capacity(count=400, budget_gb=16, seed=7, step_range=(1e3, 1e5), device_inputs=True, reference=True):
    """Saturated throughput of the request loop: `count` requests of the
    serving mix all arriving at once, so every pool runs as full as the HBM
    budget lets it.  The rate is the work of the blocks collected until half
    of all the work is done, per second: the time when the loop is loaded,
    before the tail in which only the longest requests remain.  With
    `reference`, beside it the batch engine's rate on the same requests'
    grids (batch_rate)."""
    import torch
    rng = np.random.default_rng(seed)
    reqs = [make_request(i, 0.0, rng, step_range, device_inputs=device_inputs) for i in range(count)]
    total = sum(r.points * r.steps for r in reqs)
    eng = Engine(budget_bytes=int(budget_gb * (1 << 30)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = eng.run(reqs)
    wall = time.perf_counter() - t0
    eng.close()
    acc, t_half = 0, None
    for t, w in eng.block_log:
        acc += w
        if acc >= total / 2:
            t_half = t
            break
    out = {"requests": len(done), "failed": sum(r.error is not None for r in done), "work": total,
           "wall_s": wall, "half_work_s": t_half, "grid_point_steps_per_s": (total / 2) / t_half,
           "grid_point_steps_per_s_incl_tail": total / wall}
    if reference:
        out["batch_engine_same_grids_grid_point_steps_per_s"] = batch_rate(reqs)
    return out


def batch_rate(reqs, steps=256, dtype="f32", warm=True):
    """Grid-point-steps/s of the batch engine on the grids of `reqs`: the
    requests of each pool as one batch on the pool's engine (1-D and 2-D up
    to 64^2 resident, larger 2-D fused), every member advancing `steps`
    steps, the three batches on streams of their own at once -- the request
    loop's pools at full occupancy without admission, budgets, observables or
    result downloads.  The inputs are a Gaussian pumping and state on each
    request's grid (the rate does not depend on them).  This is the ceiling
    the loop's saturated rate is measured against."""
    import torch
    from .batch import Batch

    class _M(object):
        pass

    groups = {}
    for r in reqs:
        key = Engine.group_of(r)
        m = _M()
        m.n, m.dx, m.dt, m.coeffs = r.n, r.dx, r.dt, np.asarray(r.coeffs, np.float64)
        if r.dim == 1:
            x = (np.arange(r.n) - r.n / 2.0) * r.dx
            rr = np.abs(x)
        else:
            x = (np.arange(r.n) - r.n / 2.0) * r.dx
            rr = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2)
        m.P = 2.0 * np.exp(-rr ** 2 / 8.0)
        m.psi0 = (0.1 * np.exp(-rr ** 2 / 4.0)).astype(np.complex128)
        m.n0 = np.zeros(rr.shape)
        groups.setdefault(key, []).append(m)
    engines = {"line": (1, "resident"), "small": (2, "resident"), "large": (2, "fused")}
    batches, streams, work = [], [], 0
    for key, members in groups.items():
        dim, eng = engines[key]
        batches.append(Batch(members, dim=dim, order=5, bc="periodic", dtype=dtype, engine=eng))
        streams.append(torch.cuda.Stream())
        work += sum(m.n if dim == 1 else m.n * m.n for m in members) * steps
    for b in batches:
        b.prepare(steps)
    if warm:
        for b, s in zip(batches, streams):
            b.advance(steps, stream=s)
    torch.cuda.synchronize()
    for b in batches:
        b.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b, s in zip(batches, streams):
        b.advance(steps, stream=s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for b in batches:
        b.close()
    return work / wall


def sweep(rates=(20, 50, 100, 200, 500), budget_gb=16, seed=0, step_range=(1e3, 1e5), log=print):
    out = []
    for lam in rates:
        s = serve(lam, budget_gb=budget_gb, seed=seed, step_range=step_range)
        if log:
            log(s)
        out.append(s)
    return out


if __name__ == "__main__":
    import json
    import sys
    rates = [float(x) for x in sys.argv[1:]] or [20, 50, 100, 200, 500]
    print(json.dumps(sweep(rates)))
