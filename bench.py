"""Benchmark of the time-stepping core on one MI355X.

Workload (sweep512): 512 independent 2-D instances with sides drawn from
{64, 128, 256} (p = 0.5/0.3/0.2) on [-16,16)^2, periodic, order-5 stencil,
ring pumping with per-instance P0 ~ U[0.5,5], r0 ~ U[3,8], w ~ U[0.5,2],
dynamic reservoir, dt = 5e-4, 5000 RK4 steps, seeded noise per instance
(nls_amd.configs.sweep512).  A "step" is one RK4 step of the whole batch.

Reported value: grid-point-steps/s of the batched complex64 solve (FP32
arithmetic) with all inputs resident in HBM.  Beside it: the plain FP64 path
(one instance at a time) on a sample of the same instances, the relative L2
error of the batched FP32 result against that FP64 result after the full step
count, the roofline of the dominant kernel, the CPU baseline, one 2048 x 2048
grid on the plain path (FP64 and FP32, both RK4 engines, split-step Fourier)
and the continuous-batching request loop (p50/p99 latency under Poisson
arrivals).

  python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-point-steps/s batched FP32 vs plain FP64 (relL2≤1e-4); p50/p99 latency"
HBM_PEAK = 8.0e12          # bytes/s, MI355X
FUSED_BYTES_F32_DYN = 28   # algorithmic bytes per grid-point-step of the fused step kernel: 2 S + r
# Vector ALU issue rate of MI355X without packed FP32: 256 CUs x 4 SIMDs, a
# wave64 instruction occupies a 16-lane SIMD for 4 cycles, 2.4 GHz
VALU_LANES_PEAK = 256 * 4 * 16 * 2.4e9      # lane-instructions/s
SWEEP_STEPS = 5000         # the workload's step count (accuracy and the plain FP64 leg use it whatever --steps is)
GRAPH_MIN_STEPS = 128      # a batch call replays 64-step graphs from this many steps on (include/nls_amd.h)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_profile_value(name, key):
    """A `key = value` figure of a profile summary (profiles/<name>): HBM bytes
    per launch from rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, VALU lane-instructions
    per grid-point-step from SQ_INSTS_VALU; None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals[key])
    except (KeyError, ValueError):
        return None


def host_cores():
    """Host cores this process may use (the box's CPU share caps at 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(members, steps):
    """The CPU path timed on this host, on bounded samples:

    * headline: the reference Fortran (oracle/_ref, FP32, BLAS sgbmv, one core)
      running solve_nls_2d on sweep512 members -- or, where it is not built,
      the C restatement on one core;
    * line256_reference: the reference's solve_nls on the 1-D N=256 workload in
      full (2000 steps), one core;
    * oracle_all_cores: the C restatement (FP32, the workload's own periodic
      grid and reservoir equation) on every host core, same members."""
    import ctypes
    import numpy as np
    os.environ["MKL_NUM_THREADS"] = "1"
    from nls_amd import configs
    from oracle import oracle as O
    sample = "sweep512 members 0-%d, %d RK4 steps each" % (len(members) - 1, steps)
    work = sum(m.points for m in members) * steps
    out = None
    try:
        R = O.RefLib(32)
        try:
            ctypes.CDLL("libmkl_rt.so.1").MKL_Set_Num_Threads(1)   # the BLAS the reference links
        except OSError:
            pass
        t0 = time.perf_counter()
        for m in members:
            R.solve(2, m.dt, m.dx, m.n, 5, steps, m.P, m.coeffs, m.psi0)
        t = time.perf_counter() - t0
        out = {"value": work / t, "unit": "grid-point-steps/s", "cores": 1, "kind": "reference",
               "sample": sample + "; reference Fortran solve_nls_2d (FP32, BLAS sgbmv, MKL_NUM_THREADS=1) "
                         "with its own boundaries and algebraic reservoir (it has no periodic grid or "
                         "reservoir equation)"}
        line = configs.line256()
        t0 = time.perf_counter()
        R.solve(1, line.dt, line.dx, line.n, 5, line.steps, line.P, line.coeffs, line.psi0)
        t = time.perf_counter() - t0
        out["line256_reference"] = {
            "value": line.points * line.steps / t, "unit": "grid-point-steps/s", "cores": 1, "seconds": t,
            "sample": "solve_nls, 1-D N=256, %d steps of %g, order 5, in full; the reference's radial grid "
                      "and algebraic reservoir" % (line.steps, line.dt)}
    except (OSError, FileNotFoundError):
        O.set_threads(1)
        t0 = time.perf_counter()
        for m in members:
            O.rk4(2, m.n, 5, 1, m.dt, m.dx, steps, m.P, m.coeffs, m.psi0, m.n0, dtype=np.float32)
        t = time.perf_counter() - t0
        out = {"value": work / t, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
               "sample": sample + "; C restatement, FP32, same boundaries and reservoir"}
    cores = host_cores()
    O.set_threads(cores)
    t0 = time.perf_counter()
    for m in members:
        O.rk4(2, m.n, 5, 1, m.dt, m.dx, steps, m.P, m.coeffs, m.psi0, m.n0, dtype=np.float32)
    t = time.perf_counter() - t0
    out["oracle_all_cores"] = {"value": work / t, "unit": "grid-point-steps/s", "cores": cores, "kind": "port",
                               "sample": sample + "; C restatement, FP32, periodic, reservoir equation, "
                                                  "%d threads" % cores}
    # the C restatement in both precisions, on one thread and on every core:
    # median of 3 runs over a smaller sample of the same members
    few = members[:4]
    few_steps = max(1, steps // 2)
    few_work = sum(m.points for m in few) * few_steps
    table = {}
    for name, dt in (("f32", np.float32), ("f64", np.float64)):
        table[name] = {}
        for threads in (1, cores):
            O.set_threads(threads)
            rates = []
            for _ in range(3):
                t0 = time.perf_counter()
                for m in few:
                    O.rk4(2, m.n, 5, 1, m.dt, m.dx, few_steps, m.P, m.coeffs, m.psi0, m.n0, dtype=dt)
                rates.append(few_work / (time.perf_counter() - t0))
            table[name][str(threads)] = float(np.median(rates))
    out["oracle_threads"] = {"unit": "grid-point-steps/s", "kind": "port", "statistic": "median of 3 runs",
                             "sample": "sweep512 members 0-%d, %d RK4 steps each; C restatement, periodic, "
                                       "reservoir equation" % (len(few) - 1, few_steps),
                             "cores": cores, "rates": table}
    return out


def single_grid_2048(dev, steps=100):
    """One 2048 x 2048 periodic grid (h = 0.0625, order 5) on the plain path:
    FP64 and FP32, dynamic reservoir, each RK4 engine and the split-step
    Fourier scheme, each the median of three timed runs of `steps` steps.
    Rates in grid-point-steps/s; the HBM fraction of each is given on two
    bases: the per-stage (k-storage) scheme's 15 S + 4 r bytes per
    point-step, whose 8 TB/s ceiling is the reference roofline of this path,
    and the compulsory 2 S + r (SSF: 4 c + 3 r)."""
    import torch
    from nls_amd import configs, native
    from nls_amd.autotune import Tuner
    n = 2048
    out = {}
    tuner = Tuner(path=None, steps=32, reps=2)
    for dtype in ("f64", "f32"):
        rows = tuner.stage_tile(dtype, n=n)          # tile height of the per-stage kernels, measured here
        log("2048^2 %s: per-stage tile height %d (%s)" % (dtype, rows, ", ".join(
            "%s rows %.1f us" % (k, v * 1e6) for k, v in tuner.last.items())))
        cs, rs = (16, 8) if dtype == "f64" else (8, 4)
        S = cs + rs
        rt = torch.float64 if dtype == "f64" else torch.float32
        ct = torch.complex128 if dtype == "f64" else torch.complex64
        P = torch.ones((n, n), dtype=rt, device=dev)
        legs = [("rk4_fused", dict(engine="fused"), 15 * S + 4 * rs, 2 * S + rs),
                ("rk4_stage", dict(engine="stage"), 15 * S + 4 * rs, 2 * S + rs),
                ("ssf", dict(scheme="ssf"), None, 4 * cs + 3 * rs)]
        if dtype == "f32":
            # double-precision arithmetic on the single-precision fields
            legs.append(("rk4_fused_f64arith", dict(engine="fused", f64_arith=True), None, 2 * S + rs))
        for name, kw, kbytes, mbytes in legs:
            d = native.ProblemDesc.make(2, n, 5, 1e-4, 0.0625, configs.model_coeffs(dynamic=True), bc="periodic",
                                        dtype=dtype, **kw)
            psi = torch.full((n, n), 0.1, dtype=ct, device=dev)
            nres = torch.zeros((n, n), dtype=rt, device=dev)
            ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device=dev)
            native.advance(d, 4, P, psi, nres, ws)
            torch.cuda.synchronize()
            # median of three timed runs: the first run after the grid's
            # allocation can be several percent slow (clock ramp)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                native.advance(d, steps, P, psi, nres, ws)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            t = sorted(ts)[1]
            rate = n * n * steps / t
            r = {"grid_point_steps_per_s": rate, "us_per_step": t / steps * 1e6,
                 "hbm_frac_compulsory_bytes": rate * mbytes / HBM_PEAK}
            if kbytes:
                r["hbm_frac_kstorage_bytes"] = rate * kbytes / HBM_PEAK
                r["vs_kstorage_ceiling"] = rate / (HBM_PEAK / kbytes)
            if name == "rk4_stage":
                r["tile_rows"] = rows
            out["%s_%s" % (name, dtype)] = r
            log("2048^2 %s %s: %.1f us/step, %.2f G grid-point-steps/s" % (name, dtype, t / steps * 1e6, rate / 1e9))
    return out


XGMI_LINK = 153e9     # bytes/s of one MI355X xGMI link


def slab_shards(dev, steps=20, reps=3):
    """Row-slab decomposition at its shard shape (nls_slab_step): one of 8
    row slabs of a 2048^2 and of a 16384^2 periodic grid (order 5, dynamic
    reservoir), with its 8 ghost rows per side, against 1/8 of the whole-grid
    fused step on this device.  Slab steps are captured in a graph (one
    launch per step, as a device would issue them); the ghost rows a slab
    receives per step and their time over one xGMI link per neighbour are
    estimates (one device here)."""
    import torch
    from nls_amd import configs, native
    out = {}
    h = native.slab_halo_rows(5)
    for n in (2048, 16384):
        for dtype in ("f32", "f64"):
            cs, rs = (16, 8) if dtype == "f64" else (8, 4)
            rt = torch.float64 if dtype == "f64" else torch.float32
            ct = torch.complex128 if dtype == "f64" else torch.complex64
            d = native.ProblemDesc.make(2, n, 5, 1e-4, 0.0625, configs.model_coeffs(dynamic=True), bc="periodic",
                                        dtype=dtype, engine="fused", graphs=False)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            # whole grid
            P = torch.ones((n, n), dtype=rt, device=dev)
            psi = torch.full((n, n), 0.1, dtype=ct, device=dev)
            nres = torch.zeros((n, n), dtype=rt, device=dev)
            ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device=dev)
            native.advance(d, 2, P, psi, nres, ws)
            torch.cuda.synchronize()
            e0.record()
            native.advance(d, steps, P, psi, nres, ws)
            e1.record()
            torch.cuda.synchronize()
            whole_us = e0.elapsed_time(e1) * 1e3 / steps
            del P, psi, nres, ws
            torch.cuda.empty_cache()
            # one slab of n/8 rows
            rows = n // 8
            Ps = torch.ones((rows + 2 * h, n), dtype=rt, device=dev)
            a = torch.full((rows + 2 * h, n), 0.1, dtype=ct, device=dev)
            b = torch.full_like(a, 0.1)
            na, nb = torch.zeros((rows + 2 * h, n), dtype=rt, device=dev), torch.zeros((rows + 2 * h, n), dtype=rt,
                                                                                        device=dev)
            native.slab_step(d, rows, Ps, a, na, b, nb)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                for i in range(steps):
                    src, dst = ((a, na), (b, nb)) if i % 2 == 0 else ((b, nb), (a, na))
                    native.slab_step(d, rows, Ps, src[0], src[1], dst[0], dst[1])
            g.replay()
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                g.replay()
            e1.record()
            torch.cuda.synchronize()
            slab_us = e0.elapsed_time(e1) * 1e3 / (reps * steps)
            del g, Ps, a, b, na, nb
            torch.cuda.empty_cache()
            ghost = h * n * (cs + rs)                 # bytes from each neighbour per step
            key = "%dx%d_%s" % (n, rows, dtype)
            out[key] = {"slab_us_per_step": slab_us, "whole_grid_us_per_step": whole_us,
                        "whole_over_8_us": whole_us / 8, "slab_vs_whole_over_8": slab_us / (whole_us / 8),
                        "ghost_rows_per_side": h, "ghost_bytes_per_neighbour_per_step": ghost,
                        "xgmi_us_per_neighbour_estimate": ghost / XGMI_LINK * 1e6}
            log("slab %s: %.1f us/step vs whole grid %.1f us / 8 = %.1f us (x%.2f); ghost rows %.2f MB per "
                "neighbour per step, ~%.1f us over one xGMI link" % (key, slab_us, whole_us, whole_us / 8,
                                                                     slab_us / (whole_us / 8), ghost / 1e6,
                                                                     ghost / XGMI_LINK * 1e6))
    return out


def config_accuracy(dev):
    """Relative L2 error of the FP32 (complex64) solve against the FP64 solve
    after the full step count of each configuration: line256 (1-D N=256,
    2000 steps), square256 (256^2, 10^4 steps), square2048 (2048^2, 2*10^4
    steps, RK4 and split-step Fourier).  The FP64 GPU path is the reference
    here; the tests pin it to the FP64 oracle (relL2 <= 1e-10).  Beside it,
    the FP64-accumulated norm and energy (nls_observables, nls_energy) of
    both final states."""
    import numpy as np
    import torch
    from nls_amd import configs, native
    out = {}
    cases = [("line256", configs.line256(), "rk4"), ("square256", configs.square256(), "rk4"),
             ("square2048_rk4", configs.square2048(), "rk4"), ("square2048_ssf", configs.square2048(), "ssf")]
    for name, inst, scheme in cases:
        res, obs = {}, {}
        # square2048 RK4 also with double-precision arithmetic on the
        # single-precision fields (NLS_FLAG_F64_ARITH)
        variants = [("f64", "f64", False), ("f32", "f32", False)]
        if name == "square2048_rk4":
            variants.append(("f32_f64arith", "f32", True))
        for key, dtype, wide in variants:
            d = native.ProblemDesc.make(inst.dim, inst.n, 5, inst.dt, inst.dx, inst.coeffs, bc="periodic",
                                        dtype=dtype, scheme=scheme, f64_arith=wide)
            rt = torch.float64 if dtype == "f64" else torch.float32
            ct = torch.complex128 if dtype == "f64" else torch.complex64
            P = torch.as_tensor(np.ascontiguousarray(inst.P), dtype=rt, device=dev)
            psi = torch.as_tensor(np.ascontiguousarray(inst.psi0), dtype=ct, device=dev)
            nres = torch.as_tensor(np.ascontiguousarray(inst.n0), dtype=rt, device=dev)
            ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device=dev)
            native.advance(d, inst.steps, P, psi, nres, ws)
            torch.cuda.synchronize()
            res[key] = psi.cpu().numpy().astype(np.complex128)
            obs[key] = {"norm": native.observables(d, psi, nres)["norm"],
                        "energy": native.energy(d, P, psi, nres)["energy"]}
        err = float(np.linalg.norm(res["f32"] - res["f64"]) / np.linalg.norm(res["f64"]))
        out[name] = {"relL2_f32_vs_f64": err, "steps": inst.steps, "points": inst.points,
                     "fp64_accumulated": {"norm_f64": obs["f64"]["norm"], "norm_f32": obs["f32"]["norm"],
                                          "energy_f64": obs["f64"]["energy"], "energy_f32": obs["f32"]["energy"]}}
        if "f32_f64arith" in res:
            out[name]["relL2_f32_f64arith_vs_f64"] = float(np.linalg.norm(res["f32_f64arith"] - res["f64"]) /
                                                         np.linalg.norm(res["f64"]))
            log("accuracy %s: FP32 fields with FP64 arithmetic vs FP64 relL2 = %.2e" % (
                name, out[name]["relL2_f32_f64arith_vs_f64"]))
        # nls_energy's kinetic term uses the Laplacian the scheme evolves
        out[name]["fp64_accumulated"]["energy_kinetic_operator"] = (
            "spectral" if scheme == "ssf" else "finite-difference order %d" % inst.order)
        log("accuracy %s (%d steps): FP32 vs FP64 relL2 = %.2e; norm %.10g / %.10g, energy %.10g / %.10g "
            "(FP64 / FP32 state, FP64 sums)" % (name, inst.steps, err, obs["f64"]["norm"], obs["f32"]["norm"],
                                               obs["f64"]["energy"], obs["f32"]["energy"]))
    return out


def sweep512_accuracy(members, engine, dev, steps=SWEEP_STEPS):
    """Every sweep512 member after the workload's full step count: the
    batched complex64 solve on `engine` against the batched complex128 solve
    of the same inputs (AUTO engine), relative L2 per member.  The FP64 batch
    stands for the plain FP64 path: the tests hold every batch member bitwise
    equal to the member solved alone and within 1e-10 of the oracle."""
    import numpy as np
    import torch
    from nls_amd.batch import Batch
    states, secs = {}, {}
    offs = sizes = None
    for dtype, eng in (("f32", engine), ("f64", "auto")):
        b = Batch(members, dim=2, order=5, bc="periodic", dtype=dtype, engine=eng, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.advance(steps)
        torch.cuda.synchronize()
        secs[dtype] = time.perf_counter() - t0
        states[dtype] = b.psi.cpu().numpy().astype(np.complex128)
        offs, sizes = b.offsets, b.sizes
        b.close()
    errs = np.empty(len(members))
    for i in range(len(members)):
        a, e = offs[i], offs[i] + sizes[i]
        ref = states["f64"][a:e]
        errs[i] = np.linalg.norm(states["f32"][a:e] - ref) / np.linalg.norm(ref)
    per_size = {}
    for n in sorted({m.n for m in members}):
        sel = [i for i, m in enumerate(members) if m.n == n]
        per_size[str(n)] = {"members": len(sel), "max_relL2": float(errs[sel].max()),
                            "median_relL2": float(np.median(errs[sel]))}
    out = {"steps": steps, "members": len(members), "max_relL2_vs_f64": float(errs.max()), "tolerance": 1e-4,
           "within_tolerance": bool(errs.max() <= 1e-4), "per_size": per_size,
           "f32_batch_s": secs["f32"], "f64_batch_s": secs["f64"],
           "reference": "complex128 batch of the same inputs, %d steps" % steps}
    log("accuracy sweep512 (%d steps, all %d members): max FP32 vs FP64 relL2 = %.2e (%s)" % (
        steps, len(members), errs.max(), ", ".join("%s^2 %.1e" % (k, v["max_relL2"]) for k, v in per_size.items())))
    return out


def plain_f64(members, dev, steps=SWEEP_STEPS):
    """The plain FP64 path (nls_advance, one instance at a time, AUTO engine)
    on one instance of each sweep512 size: warmed up (kernel modules loaded,
    the 64-step graph captured), state reset, then the workload's full step
    count timed; the rate over the workload's size mix weights each size by
    its members."""
    import numpy as np
    import torch
    from nls_amd import native
    sample = {}
    for i, m in enumerate(members):
        sample.setdefault(m.n, i)
    rates = {}
    for n, i in sorted(sample.items()):
        m = members[i]
        d = native.ProblemDesc.make(2, n, 5, m.dt, m.dx, m.coeffs, bc="periodic", dtype="f64")
        P = torch.from_numpy(np.ascontiguousarray(m.P)).to(dev)
        psi0 = torch.from_numpy(np.ascontiguousarray(m.psi0, np.complex128)).to(dev)
        n0 = torch.from_numpy(np.ascontiguousarray(m.n0)).to(dev)
        psi, nres = psi0.clone(), n0.clone()
        ws = torch.empty(max(1, native.workspace_bytes(d)), dtype=torch.uint8, device=dev)
        native.advance(d, GRAPH_MIN_STEPS, P, psi, nres, ws)       # warm-up on the timed buffers
        psi.copy_(psi0)
        nres.copy_(n0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        native.advance(d, steps, P, psi, nres, ws)
        torch.cuda.synchronize()
        t = time.perf_counter() - t1
        rates[n] = n * n * steps / t
        log("plain FP64 %d^2: %.3f G grid-point-steps/s (%d steps, %.1f us/step)" % (n, rates[n] / 1e9, steps,
                                                                                    t / steps * 1e6))
    points = sum(m.points for m in members)
    t_plain = sum(m.points * steps / rates[m.n] for m in members)
    return {"value": points * steps / t_plain, "unit": "grid-point-steps/s", "steps": steps,
            "per_size": {str(k): v for k, v in rates.items()},
            "note": "plain RK4 FP64 path (nls_advance, AUTO engine), one instance at a time; rate per size on one "
                    "instance of each size over the workload's %d steps after a warm-up, weighted by the size "
                    "mix" % steps}


def serving(rates, capacity_ref=None):
    """Continuous batching: Poisson arrivals at each rate over a window of
    at least 240 expected requests and 3 s, the request mix of
    nls_amd.serving.make_request with inputs given by parameters and
    generated on the device at admission; latency is arrival to result on the
    host.  Beside the percentiles: the steady-state rate (work of the blocks
    run while requests keep arriving, the ramp-up and the drain after the
    last arrival left out) and the offered load.  Then the saturated
    capacity: 400 requests arriving at once, the rate until half their work
    is done, beside the batch engine's rate on the same grids; the
    saturation rate is that capacity divided by the mean work of a
    request."""
    from nls_amd import serving as S
    runs = []
    for rate in rates:
        s = S.serve(rate)
        log("serving at %g req/s over %.1fs: %d requests, p50 %.3fs, p99 %.3fs, steady %.1f G grid-point-steps/s "
            "(offered %.1f G)" % (rate, s["arrival_window_s"], s["requests"], s["p50_s"], s["p99_s"],
                                 s.get("steady_grid_point_steps_per_s", 0) / 1e9,
                                 s.get("offered_grid_point_steps_per_s", 0) / 1e9))
        runs.append(s)
    mean_work = S.mean_request_work()
    sat = S.capacity()
    cap = sat["grid_point_steps_per_s"]
    log("serving saturated: %d requests at once, %.1f G grid-point-steps/s until half the work is done (%.1f G "
        "with the tail)" % (sat["requests"], cap / 1e9, sat["grid_point_steps_per_s_incl_tail"] / 1e9))
    out = {"runs": runs, "saturated": sat, "mean_request_grid_point_steps": mean_work,
           "capacity_grid_point_steps_per_s": cap, "saturation_rate_req_s": cap / mean_work}
    same = sat.get("batch_engine_same_grids_grid_point_steps_per_s")
    if same:
        # the batch engine on the saturated run's own grids (every pool full,
        # no admission, budgets, observables or downloads)
        out["capacity_vs_batch_engine_same_grids"] = cap / same
        log("batch engine on the same 400 grids: %.1f G grid-point-steps/s; serving reaches %.0f%% of it" % (
            same / 1e9, 100 * cap / same))
    if capacity_ref:
        out["capacity_vs_batch_engine_sweep512"] = cap / capacity_ref
    log("serving capacity %.1f G grid-point-steps/s = %.0f requests/s of the mix%s" % (
        cap / 1e9, cap / mean_work,
        "" if not capacity_ref else " (%.0f%% of the batch engine on sweep512)" % (100 * cap / capacity_ref)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 comparison, the 2048^2 grid and the row-slab shards")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop latency measurement")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-vs-FP64 accuracy of the line256/square256/square2048 configurations")
    ap.add_argument("--serve-rates", default="20,50,100,200,500",
                    help="Poisson arrival rates (requests/s, comma-separated), one request-loop run each")
    ap.add_argument("--engine", default="tuned", choices=("tuned", "fused", "hybrid", "stage", "resident", "auto"),
                    help="batch engine of the headline workload (tuned: measured on this device)")
    ap.add_argument("--fused-form", default="tuned", choices=("tuned", "auto", "tall", "short", "tall512"),
                    help="tile form of the fused kernel on the headline batch (tuned: measured on this device "
                         "when the engine is tuned; auto: the per-launch rule)")
    ap.add_argument("--grid2048-only", action="store_true",
                    help="only measure the 2048 x 2048 single grid (plain path) and print its JSON")
    ap.add_argument("--slab-only", action="store_true",
                    help="only measure the row-slab shards (one of 8 slabs of 2048^2 and 16384^2) and print its JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every box here has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from nls_amd import configs, native
    from nls_amd.batch import Batch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    K, W = args.steps, args.warmup
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d bytes 5 times" % (src.numel() * 4))
        return

    if args.grid2048_only:
        print(json.dumps({"single_grid_2048": single_grid_2048(dev)}))
        return
    if args.slab_only:
        print(json.dumps({"slab_shards": slab_shards(dev)}))
        return

    t0 = time.perf_counter()
    members = configs.sweep512(steps=K)
    points = sum(m.points for m in members)
    log("sweep512: %d instances, %d points, generated in %.1fs" % (len(members), points, time.perf_counter() - t0))

    # engine of the batch: measured on this device among those that can hold
    # the member mix (nls_amd.autotune), unless given
    tuning = None
    engine = args.engine
    if engine == "tuned":
        from nls_amd.autotune import Tuner
        tuner = Tuner(path=None, steps=32, reps=2)
        engine = tuner.batch_engine(members, dim=2, order=5, bc="periodic", dtype="f32")
        tuning = {k: v * 1e6 for k, v in tuner.last.items()}
        log("tuned batch engine: %s (%s us/step)" % (engine, ", ".join("%s %.1f" % kv for kv in tuning.items())))
        if engine in ("fused", "hybrid"):
            form = tuner.fused_form(members, dim=2, order=5, bc="periodic", dtype="f32", engine=engine)
            tuning = {"engine": tuning, "fused_form": {k: v * 1e6 for k, v in tuner.last.items()}}
            log("tuned fused tile form: %s (%s us/step)" % (form, ", ".join(
                "%s %.1f" % kv for kv in tuning["fused_form"].items())))
    if args.fused_form != "tuned":
        native.fused_form("f32", args.fused_form)
    batch = Batch(members, dim=2, order=5, bc="periodic", dtype="f32", engine=engine, device=dev)
    stream = torch.cuda.current_stream()
    if W > 0:
        batch.advance(W)
    torch.cuda.synchronize()
    batch.reset()
    batch.prepare(K)                 # capture the 64-step block the timed call replays (runs no step)
    torch.cuda.synchronize()
    log("warmup done (%d steps); timing %d steps" % (W, K))

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    batch.advance(K)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    gpu_s = e0.elapsed_time(e1) * 1e-3
    value = points * K / wall
    launch_s = gpu_s / K             # one fused-kernel launch per step; events on its stream
    log("batched FP32: %.3fs wall, %.1f us/step, %.2f G grid-point-steps/s" % (wall, wall / K * 1e6, value / 1e9))

    graphs = K >= GRAPH_MIN_STEPS
    if engine in ("fused", "hybrid"):
        # hybrid: the 64^2 members run on the resident engine on a side stream
        # (state on chip, HBM at the ends of a call); the fused kernel, the
        # only one on this stream, steps the others
        fused_points = points if engine == "fused" else sum(m.points for m in members if m.n * m.n > 4096)
        bytes_per_launch = fused_points * FUSED_BYTES_F32_DYN
        achieved = bytes_per_launch / launch_s
        prof = ("fused_step_f32_sweep512_traffic.txt" if engine == "fused" else
                "fused_step_f32_sweep512_hybrid_traffic.txt")
        traffic = read_profile_value(prof, "traffic_bytes_per_launch")
        valu_per_point = read_profile_value(prof, "valu_lane_insts_per_point_step")
        roofline = {"bound": "valu", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK, "traffic": traffic,
                    "kernel": "k_fused_pairs_2d<periodic,dynamic> (float, order 5)", "kernel_points": fused_points,
                    "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_us": launch_s * 1e6,
                    "note": "achieved/peak/frac: the HBM roofline of the algorithmic bytes (secondary); the "
                            "kernel's limiter is its vector ALU work (valu)"}
        if valu_per_point:
            va = valu_per_point * fused_points / launch_s
            roofline["valu"] = {"lane_insts_per_point_step": valu_per_point, "achieved": va, "peak": VALU_LANES_PEAK,
                                "unit": "lane-instructions/s", "frac": va / VALU_LANES_PEAK,
                                "source": "rocprofv3 SQ_INSTS_VALU x 64 / points per launch (profiles/%s); peak: "
                                          "1024 SIMDs x 16 lanes x 2.4 GHz, no packed FP32" % prof}
    else:
        # per-stage engine: four launches per step, 15 S + 4 r bytes per point-step
        bytes_per_step = points * (15 * 12 + 4 * 4)
        achieved = bytes_per_step / launch_s
        roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "k_batch_stage (4 per step)",
                    "algorithmic_bytes_per_step": bytes_per_step, "avg_step_us": launch_s * 1e6}

    if engine == "resident":
        blocks = "one resident launch per call"
    else:
        part = "fused part" if engine == "hybrid" else "every step"
        blocks = ("%s: 64-step hipGraph replays (%d steps >= %d)" % (part, K, GRAPH_MIN_STEPS) if graphs else
                  "%s: direct launches, no graph replay (%d steps < %d)" % (part, K, GRAPH_MIN_STEPS))
        if engine == "hybrid":
            blocks += "; resident part (64^2 members, own stream): one launch per call"
    out = {"metric": METRIC, "value": value, "unit": "grid-point-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": wall / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "sweep512: 512 2-D instances, sides 64/128/256 (p=0.5/0.3/0.2), periodic, "
                                  "order 5, dynamic reservoir, dt=5e-4, %d RK4 steps, complex64 state" % K,
                      "instances": len(members), "grid_points": points, "engine": engine,
                      "step_blocks": blocks, "graphs_replayed": bool(graphs and engine != "resident"),
                      "fused_form": native.fused_form("f32"),
                      "parallelism": "single GPU"},
           "roofline": roofline}
    if tuning is not None:
        out["engine_tuning_us_per_step"] = tuning
    # FP64 observables of every member after the timed run (one launch)
    obs = batch.observables()
    en = batch.energy()
    out["fp64_observables"] = {"members": int(obs.shape[0]), "nonfinite_members": int((obs[:, 3] > 0).sum()),
                               "norm_min": float(obs[:, 0].min()), "norm_max": float(obs[:, 0].max()),
                               "population_total": float(obs[:, 1].sum()),
                               "energy_min": float(en[:, 0].min()), "energy_max": float(en[:, 0].max())}
    log("FP64 observables: %d members, %d with non-finite points, norm %.4g .. %.4g, energy %.4g .. %.4g" % (
        obs.shape[0], out["fp64_observables"]["nonfinite_members"], obs[:, 0].min(), obs[:, 0].max(),
        en[:, 0].min(), en[:, 0].max()))
    batch.close()
    native.fused_form("f32", "auto")         # the other legs take the per-launch rule

    if not args.no_plain:
        # the headline's qualifier at the workload's full step count, whatever
        # --steps is: every member of the batched FP32 solve against the FP64
        # solve of the same inputs, and the plain FP64 path's rate
        acc = sweep512_accuracy(members, engine, dev)
        out["accuracy"] = acc
        plain = plain_f64(members, dev)
        out["plain_f64"] = plain
        out["speedup_vs_plain_f64"] = value / plain["value"]
        log("batched FP32 / plain FP64 = %.1fx" % out["speedup_vs_plain_f64"])
        out["single_grid_2048"] = single_grid_2048(dev)
        out["slab_shards"] = slab_shards(dev)

    if not args.no_accuracy:
        out["config_accuracy"] = config_accuracy(dev)
        if "accuracy" in out:
            a = out["accuracy"]
            out["config_accuracy"]["sweep512"] = {"relL2_f32_vs_f64": a["max_relL2_vs_f64"], "steps": a["steps"],
                                                  "members": a["members"], "per_size": a["per_size"],
                                                  "note": "largest over all members; see accuracy"}
    if not args.no_serve:
        out["serving"] = serving([float(r) for r in args.serve_rates.split(",") if r], capacity_ref=value)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(members[:16], 200)
        log("cpu baseline: %.3g grid-point-steps/s (%s)" % (out["cpu_baseline"]["value"], out["cpu_baseline"]["kind"]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
