#!/usr/bin/env python3
"""Throughput of the HIP acoustic forward operator on one MI355X.

Default workload (the headline configuration): 3D constant density, float32,
384^3 grid at 12.5 m, space order 8, 40-point cubic damping layer
(alpha 0.001) -> 472^3 extended grid; 6 flat velocity layers 1500..4500 m/s
times (1 + p), p seeded Gaussian-smoothed noise (sigma 8 cells, max |p| 5%);
tf = 1.5 s at the CFL step; one 15 Hz Ricker source at (100, 2393.75,
2393.75) m; 128 x 128 receivers at z = 25 m; saving_stride 0.

A "step" is one complete shot (every time step of the run) with the models,
wavefield slots and shot record resident in HBM; the wavefield is zeroed at
the start of each shot inside the timed region.  value = interior grid-point
updates per second (Gpts/s).

The roofline object is measured live for the dominant kernel (the fused
stencil+source+boundary step kernel) with HIP events around each launch on
the launch stream.  The CPU baseline times the reference's own cpu_openmp
forward (built from its sources into oracle/_ref) or, where that build is
absent, the C restatement in oracle/, on a bounded number of time steps of
the same grid.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
    python bench.py --copy-calibration   # plain device copy of known size
"""
import argparse
import contextlib
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("Gpts/s & % HBM roofline (3D order-8 fp32); "
          "p50/p99 shot latency under load")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ workloads
def headline_3d(dtype=np.float32, seed=0):
    """384^3, so8, 40-cell damping, 6 layers +-5% smooth noise."""
    import simwave_amd as sw
    from scipy.ndimage import gaussian_filter

    n, h = 384, 12.5
    rng = np.random.default_rng(seed)
    noise = gaussian_filter(rng.standard_normal((n, n, n), dtype=np.float32),
                            8.0, mode="reflect", truncate=3.0)
    noise *= np.float32(0.05 / np.abs(noise).max())
    layers = np.linspace(1500.0, 4500.0, 6)
    vel = np.repeat(layers, n // 6).astype(np.float32)[:, None, None] * (1 + noise)
    vel = vel.astype(dtype)
    del noise
    ext = (n - 1) * h
    sm = sw.SpaceModel(bounding_box=(0, ext) * 3, grid_spacing=(h, h, h),
                       velocity_model=vel, space_order=8, dtype=dtype)
    sm.config_boundary(damping_length=500.0,
                       boundary_condition=("null_neumann",) + ("null_dirichlet",) * 5,
                       damping_polynomial_degree=3, damping_alpha=0.001)
    tm = sw.TimeModel(space_model=sm, tf=1.5, saving_stride=0)
    src = sw.Source(sm, coordinates=[(100.0, ext / 2, ext / 2)])
    g = np.linspace(0.0, ext, 128)
    rec = sw.Receiver(sm, coordinates=[(25.0, x, y) for x in g for y in g])
    return dict(space_model=sm, time_model=tm, sources=src, receivers=rec,
                wavelet=sw.RickerWavelet(15.0, tm))


def readme_2d(dtype=np.float32):
    """The reference README example: 512^2 -> 513^2, so4, 512 receivers."""
    import simwave_amd as sw
    vel = np.full((512, 512), 1500.0, dtype=dtype)
    vel[100:] = 2000.0
    sm = sw.SpaceModel(bounding_box=(0, 5120, 0, 5120), grid_spacing=(10, 10),
                       velocity_model=vel, space_order=4, dtype=dtype)
    sm.config_boundary(damping_length=0,
                       boundary_condition=("null_neumann", "null_dirichlet",
                                           "none", "null_dirichlet"))
    tm = sw.TimeModel(space_model=sm, tf=1.0, saving_stride=0)
    src = sw.Source(sm, coordinates=[(2560, 2560)], window_radius=1)
    rec = sw.Receiver(sm, coordinates=[(2560, i) for i in range(0, 5120, 10)],
                      window_radius=1)
    return dict(space_model=sm, time_model=tm, sources=src, receivers=rec,
                wavelet=sw.RickerWavelet(10.0, tm))


def faulted_2d_vd(dtype=np.float32, seed=0):
    """2D variable density, 1000 x 3000 at 5 m, space order 16, 60-cell
    damping; 20 seeded dipping layers (1500-5500 m/s) offset by 3 seeded
    normal faults; density from Gardner's relation rho = 310 v^0.25;
    tf 3.0 s; 25 Hz Ricker at the surface centre; 3000 surface receivers;
    saving_stride 10 (every 10th wavefield returned)."""
    import simwave_amd as sw
    nz, nx, h = 1000, 3000, 5.0
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None] * h
    x = np.arange(nx)[None, :] * h
    dip = np.tan(np.deg2rad(rng.uniform(-8.0, 8.0)))
    depth = z - dip * (x - x.mean())
    for _ in range(3):                      # normal faults: the hanging wall drops
        x0 = rng.uniform(0.2, 0.8) * nx * h
        slope = np.tan(np.deg2rad(rng.uniform(55.0, 75.0)))
        throw = rng.uniform(40.0, 150.0)
        hanging = x > x0 + z / slope
        depth = np.where(hanging, depth - throw, depth)
    edges = np.sort(rng.uniform(0.0, nz * h, 19))
    speeds = np.linspace(1500.0, 5500.0, 20)
    vel = speeds[np.searchsorted(edges, depth)].astype(dtype)
    den = (310.0 * vel.astype(np.float64) ** 0.25).astype(dtype)
    sm = sw.SpaceModel(bounding_box=(0, (nz - 1) * h, 0, (nx - 1) * h),
                       grid_spacing=(h, h), velocity_model=vel, density_model=den,
                       space_order=16, dtype=dtype)
    sm.config_boundary(damping_length=60 * h,
                       boundary_condition=("null_neumann",) + ("null_dirichlet",) * 3,
                       damping_polynomial_degree=3, damping_alpha=0.001)
    tm = sw.TimeModel(space_model=sm, tf=3.0, saving_stride=10)
    src = sw.Source(sm, coordinates=[(2 * h, (nx - 1) * h / 2)])
    rec = sw.Receiver(sm, coordinates=[(2 * h, x) for x in
                                       np.linspace(0.0, (nx - 1) * h, 3000)])
    return dict(space_model=sm, time_model=tm, sources=src, receivers=rec,
                wavelet=sw.RickerWavelet(25.0, tm))


def fwi_3d(dtype=np.float32, seed=0, shot=0):
    """FWI shot: 160 x 256 x 256 at 15 m, space order 8, 30-cell cubic
    damping (228 x 324 x 324 extended), 5 flat layers 1500..4000 m/s times
    (1 + p), p seeded smooth noise (max |p| 5%); tf 1.2 s; source `shot` of
    16 on a 4 x 4 grid at 30 m depth; 100 x 100 receivers at 15 m depth."""
    import simwave_amd as sw
    from scipy.ndimage import gaussian_filter

    nz, nx, h = 160, 256, 15.0
    rng = np.random.default_rng(seed)
    noise = gaussian_filter(rng.standard_normal((nz, nx, nx), dtype=np.float32),
                            6.0, mode="reflect", truncate=3.0)
    noise *= np.float32(0.05 / np.abs(noise).max())
    layers = np.linspace(1500.0, 4000.0, 5)
    vel = np.repeat(layers, nz // 5).astype(np.float32)[:, None, None] * (1 + noise)
    vel = vel.astype(dtype)
    ez, ex = (nz - 1) * h, (nx - 1) * h
    sm = sw.SpaceModel(bounding_box=(0, ez, 0, ex, 0, ex), grid_spacing=(h, h, h),
                       velocity_model=vel, space_order=8, dtype=dtype)
    sm.config_boundary(damping_length=30 * h,
                       boundary_condition=("null_neumann",) + ("null_dirichlet",) * 5,
                       damping_polynomial_degree=3, damping_alpha=0.001)
    tm = sw.TimeModel(space_model=sm, tf=1.2, saving_stride=0)
    sx = ex * (0.5 + (shot % 4)) / 4
    sy = ex * (0.5 + (shot // 4 % 4)) / 4
    src = sw.Source(sm, coordinates=[(30.0, sx, sy)])
    g = np.linspace(0.0, ex, 100)
    rec = sw.Receiver(sm, coordinates=[(15.0, x, y) for x in g for y in g])
    return dict(space_model=sm, time_model=tm, sources=src, receivers=rec,
                wavelet=sw.RickerWavelet(10.0, tm))


def run_fwi(args, desc):
    """Adjoint-state gradient of one FWI shot, device resident: the forward
    (keeping every level when they fit in HBM, as here, else checkpoints),
    then the reverse sweep.  value = the shot's interior updates / gradient
    seconds; the checkpointed (K = 20) gradient is timed beside it."""
    import torch
    import simwave_amd as sw
    from simwave_amd import adjoint
    from simwave_amd.device import DeviceForward

    t0 = time.time()
    solver = sw.Solver(**fwi_3d())
    dev = DeviceForward(solver)
    kw = dev.kwargs
    T, nrec = kw["end_timestep"], kw["num_receivers"]
    res = torch.from_numpy(np.random.default_rng(args.seed).standard_normal(
        (T, nrec)).astype(np.float32)).to(dev.device)
    log("workload %s: grid %s, %d steps, %d receivers, host prep %.1fs"
        % (args.workload, kw["velocity_model"].shape, T, nrec, time.time() - t0))
    every = adjoint.checkpoint_interval(dev)      # every level when they fit in HBM
    log("checkpoint interval %d (%s)" % (every, "all levels in HBM" if every >= T
                                         else "recomputed segments"))
    for _ in range(args.warmup):
        dev.reset()
        adjoint._run(dev, res, every, False)
    torch.cuda.synchronize()
    grad_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        dev.reset()
        adjoint._run(dev, res, every, False)     # synchronises before returning
        grad_ms.append((time.perf_counter() - ts) * 1e3)
        log("gradient %d done (%.1f ms)" % (i, grad_ms[-1]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dev.reset()
    adjoint._run(dev, res, 20, False)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    dev.reset()
    adjoint._run(dev, res, 20, False)
    k20_ms = (time.perf_counter() - ts) * 1e3
    fwd_ms = []
    for _ in range(2):
        dev.reset()
        ts = time.perf_counter()
        dev.run()
        fwd_ms.append((time.perf_counter() - ts) * 1e3)
    upd = dev.updates_per_shot
    out = {
        "metric": METRIC, "value": round(upd / (elapsed / args.steps) / 1e9, 3),
        "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "description": desc,
                   "grid": list(kw["velocity_model"].shape), "time_steps": T,
                   "updates_per_shot": upd,
                   "receivers": nrec, "parallelism": "none"},
        "gradient": {"ms_p50": round(float(np.percentile(grad_ms, 50)), 2),
                     "checkpoint_every": every,
                     "ms_with_checkpoints_every_20": round(k20_ms, 2),
                     "forward_ms": round(min(fwd_ms), 2),
                     "gradient_over_forward": round(float(np.percentile(grad_ms, 50))
                                                    / min(fwd_ms), 3)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_host_workload(args, builder, desc):
    """Times Solver.forward() (the reference-compatible host entry: models
    and every returned snapshot cross PCIe inside the call)."""
    import torch
    import simwave_amd as sw
    t0 = time.time()
    parts = builder()
    solver = sw.Solver(**parts)
    kw = solver.forward_arguments()
    log("workload %s: grid %s, %d steps, %d snapshots, host prep %.1fs"
        % (args.workload, kw["velocity_model"].shape, kw["end_timestep"],
           kw["num_snapshots"], time.time() - t0))
    r = kw["space_order"] // 2
    upd = int(np.prod([n - 2 * r for n in kw["velocity_model"].shape])) * kw["end_timestep"]
    for _ in range(args.warmup):
        with contextlib.redirect_stdout(sys.stderr):   # its timing line
            solver.forward()
    times = []
    for i in range(args.steps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        with contextlib.redirect_stdout(sys.stderr):
            u, rec = solver.forward()
        times.append(time.perf_counter() - ts)
        log("shot %d: %.3fs" % (i, times[-1]))
    el = float(np.median(times))
    out = {
        "metric": METRIC, "value": round(upd / el / 1e9, 3), "unit": "Gpts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc,
                   "grid": list(kw["velocity_model"].shape),
                   "time_steps": int(kw["end_timestep"]),
                   "snapshots": int(kw["num_snapshots"]), "updates_per_shot": upd,
                   "path": "host entry (Solver.forward), PCIe-inclusive"},
        "shot_latency_ms": {"p50": round(float(np.percentile(times, 50)) * 1e3, 2),
                            "p99": round(float(np.percentile(times, 99)) * 1e3, 2),
                            "shots": len(times), "load": "one shot at a time"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def serving_requests(count, seed=0, n2=(256, 1024), n3=(64, 160), steps=(100, 300)):
    """Seeded stream of shot requests for the serving workload: 2D or 3D
    (50/50), 2D grids n2[0]^2..n2[1]^2 and 3D grids n3[0]^3..n3[1]^3 at
    10 m, space order uniform in {2,4,8,12,16,20}, constant or variable
    density 50/50 (Gardner), 3-8 random flat layers of 1500-5000 m/s, a
    5-40 Hz Ricker at the shallow centre, 64-4096 surface receivers, 20-cell
    damping, float32, tf set for steps[0]..steps[1] time steps."""
    import simwave_amd as sw
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        dim = 2 if rng.random() < 0.5 else 3
        n = int(rng.integers(*(n2 if dim == 2 else n3), endpoint=True))
        so = int(rng.choice([2, 4, 8, 12, 16, 20]))
        dens = bool(rng.random() < 0.5)
        layers = np.sort(rng.uniform(1500.0, 5000.0, int(rng.integers(3, 8, endpoint=True))))
        f0 = float(rng.uniform(5.0, 40.0))
        nrec = int(rng.integers(64, 4096, endpoint=True))
        nsteps = int(rng.integers(*steps, endpoint=True))
        out.append(dict(dim=dim, n=n, so=so, dens=dens, layers=layers, f0=f0,
                        nrec=nrec, nsteps=nsteps))

    def build(req):
        dim, n, h = req["dim"], req["n"], 10.0
        shape = (n,) * dim
        nl = len(req["layers"])
        vel = np.repeat(req["layers"], -(-n // nl))[:n].astype(np.float32)
        vel = np.broadcast_to(vel.reshape((n,) + (1,) * (dim - 1)), shape).copy()
        den = (310.0 * vel.astype(np.float64) ** 0.25).astype(np.float32) if req["dens"] else None
        ext = (n - 1) * h
        sm = sw.SpaceModel(bounding_box=(0, ext) * dim, grid_spacing=(h,) * dim,
                           velocity_model=vel, density_model=den,
                           space_order=req["so"], dtype=np.float32)
        sm.config_boundary(damping_length=20 * h,
                           boundary_condition=("null_neumann",) + ("null_dirichlet",) * (2 * dim - 1))
        dt = sw.TimeModel(space_model=sm, tf=1.0).dt
        tm = sw.TimeModel(space_model=sm, tf=float(dt) * (req["nsteps"] - 1))
        src = sw.Source(sm, coordinates=[(2 * h,) + (ext / 2,) * (dim - 1)])
        if dim == 2:
            xs = np.linspace(0.0, ext, req["nrec"])
            rc = [(h, x) for x in xs]
        else:
            m = int(np.ceil(np.sqrt(req["nrec"])))
            g = np.linspace(0.0, ext, m)
            rc = [(h, x, y) for x in g for y in g][:req["nrec"]]
        rec = sw.Receiver(sm, coordinates=rc)
        return sw.Solver(space_model=sm, time_model=tm, sources=src, receivers=rec,
                         wavelet=sw.RickerWavelet(req["f0"], tm))

    return out, build


def graph_replay_effect(reps=3):
    """A launch-bound shot (2D 128^2, order 2, 960 steps) run through a plan
    in 48-step chunks, with direct launches and with graph replay."""
    import torch
    import simwave_amd as sw
    from simwave_amd.device import DeviceForward, FLAG_GRAPH
    vel = np.full((128, 128), 2000.0, dtype=np.float32)
    sm = sw.SpaceModel(bounding_box=(0, 1270, 0, 1270), grid_spacing=(10, 10),
                       velocity_model=vel, space_order=2, dtype=np.float32)
    sm.config_boundary(damping_length=100.0,
                       boundary_condition=("null_neumann",) + ("null_dirichlet",) * 3)
    dt = sw.TimeModel(space_model=sm, tf=1.0).dt
    tm = sw.TimeModel(space_model=sm, tf=float(dt) * 959)
    solver = sw.Solver(space_model=sm, time_model=tm,
                       sources=sw.Source(sm, coordinates=[(20.0, 635.0)]),
                       receivers=sw.Receiver(sm, coordinates=[(10.0, x) for x in
                                                              np.linspace(0, 1270, 128)]),
                       wavelet=sw.RickerWavelet(15.0, tm))
    stream = torch.cuda.Stream()
    out = {"shot": "2D 128^2 so2, %d steps, 48-step chunks" % tm.timesteps}
    for name, flags in (("direct_ms", 0), ("graph_ms", FLAG_GRAPH)):
        best = None
        for _ in range(reps):
            dev = DeviceForward(solver)
            dev.plan(stream, flags=flags)
            stream.synchronize()
            t0 = time.perf_counter()
            left = dev.steps
            while left > 0:
                left = dev.run_steps(48, stream)
            stream.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            dev.release_plan()
            best = ms if best is None else min(best, ms)
        out[name] = round(best, 2)
    return out


def run_serving(args):
    """Serving workload: one JSON line with the aggregate Gpts/s of the
    server under Poisson arrivals, p50/p99 latency and the plain path (the
    same shots one at a time) beside it."""
    import torch
    from simwave_amd.device import DeviceForward
    from simwave_amd.serving import ShotServer

    reqs, build = serving_requests(args.serve_requests, seed=args.seed)
    t0 = time.time()
    solvers = [build(r) for r in reqs]
    log("serving: %d requests built in %.1fs" % (len(solvers), time.time() - t0))
    warm = [build(r) for r in serving_requests(2, seed=args.seed + 1000,
                                               n2=(128, 128), n3=(32, 32),
                                               steps=(20, 20))[0]]
    log("serving: warm-up (%d small shots)" % len(warm))
    with ShotServer(hbm_budget=args.hbm_budget * 1e9, max_active=args.max_active,
                    chunk_steps=args.chunk_steps, plan_flags=args.plan_flags) as srv:
        for f in [srv.submit(s) for s in warm]:
            f.result(timeout=120)
    log("serving: warm-up done")
    rng = np.random.default_rng(args.seed + 1)
    gaps = rng.exponential(1.0 / args.rate, size=len(solvers))
    torch.cuda.synchronize()
    with ShotServer(hbm_budget=args.hbm_budget * 1e9, max_active=args.max_active,
                    chunk_steps=args.chunk_steps, plan_flags=args.plan_flags) as srv:
        start = time.perf_counter()
        futs = []
        due = start
        for s, g in zip(solvers, gaps):
            due += g
            delay = due - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(s, arrival=due))
        for i, f in enumerate(futs):
            f.result(timeout=150)
            log("serving: request %d/%d done" % (i + 1, len(futs)))
    summ = srv.stats.summary()
    log("serving: %s" % summ)
    # plain path: the drop-in Solver.forward(), one shot at a time
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for i, s in enumerate(solvers):
        with contextlib.redirect_stdout(sys.stderr):   # its timing line
            s.forward()
        if (i + 1) % 8 == 0:
            log("plain path (Solver.forward): %d/%d shots" % (i + 1, len(solvers)))
    plain_s = time.perf_counter() - tp
    plain = summ["updates"] / plain_s / 1e9
    log("plain path: %d shots in %.2fs, %.2f Gpts/s" % (len(solvers), plain_s, plain))
    # device-resident path: DeviceForward one shot at a time
    torch.cuda.synchronize()
    tp = time.perf_counter()
    upd = 0
    for i, s in enumerate(solvers):
        dev = DeviceForward(s)
        dev.run()
        dev.results()
        upd += dev.updates_per_shot
        del dev
        if (i + 1) % 8 == 0:
            log("device path: %d/%d shots" % (i + 1, len(solvers)))
    torch.cuda.synchronize()
    dev_s = time.perf_counter() - tp
    devp = upd / dev_s / 1e9
    log("device path: %d shots in %.2fs, %.2f Gpts/s" % (len(solvers), dev_s, devp))
    graph = graph_replay_effect()
    log("graph replay on a small shot: %s" % graph)
    out = {
        "metric": METRIC, "value": round(summ["gpts"], 3), "unit": "Gpts/s",
        "n_gpus": 1, "steps": len(solvers), "warmup": len(warm),
        "ms_per_step": round(summ["makespan_s"] / len(solvers) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serving", "requests": len(solvers), "seed": args.seed,
                   "arrival_rate_hz": args.rate, "hbm_budget_gb": args.hbm_budget,
                   "max_active": args.max_active, "chunk_steps": args.chunk_steps,
                   "plan_flags": args.plan_flags,
                   "description": serving_requests.__doc__.split("\n")[0]},
        "shot_latency_ms": {"p50": round(summ["latency_p50_s"] * 1e3, 2),
                            "p99": round(summ["latency_p99_s"] * 1e3, 2),
                            "load": "Poisson arrivals at %g/s" % args.rate},
        "serving": {"updates": summ["updates"], "makespan_s": round(summ["makespan_s"], 3),
                    "queue_p50_ms": round(summ["queue_p50_s"] * 1e3, 2),
                    "max_active_seen": summ["max_active"],
                    "prep_mean_ms": round(summ["prep_mean_s"] * 1e3, 2),
                    "prep_build_mean_ms": round(summ["prep_build_mean_s"] * 1e3, 2),
                    "prep_plan_mean_ms": round(summ["prep_plan_mean_s"] * 1e3, 2),
                    "device_mean_ms": round(summ["device_mean_s"] * 1e3, 2),
                    "fetch_mean_ms": round(summ["fetch_mean_s"] * 1e3, 2),
                    "plain_path": "Solver.forward(), one shot at a time",
                    "plain_path_gpts": round(plain, 3),
                    "speedup_vs_plain": round(summ["gpts"] / plain, 3),
                    "device_path_gpts": round(devp, 3),
                    "speedup_vs_device_path": round(summ["gpts"] / devp, 3),
                    "graph_replay": graph},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


WORKLOADS = {
    "3d_cd_so8_384": (headline_3d,
                      "3D constant density so8 384^3 (472^3 extended), "
                      "12.5 m, 40-pt cubic damping, tf 1.5 s, 128x128 receivers"),
    "2d_cd_so4_readme": (readme_2d,
                         "2D constant density so4 512^2 (README example)"),
    "serving": (None, "mixed shot stream through the continuous-batching server"),
    "fwi_3d_so8_256": (fwi_3d, "adjoint-state velocity gradient of one FWI shot: 3D so8 "
                               "160x256x256 (228x324x324 extended), 15 m, tf 1.2 s, "
                               "100x100 receivers"),
    "2d_vd_so16_stride10": (faulted_2d_vd,
                            "2D variable density so16 1000x3000 (1136x3136 extended), "
                            "5 m, 60-pt damping, faulted layers, tf 3.0 s, 3000 "
                            "receivers, saving_stride 10 (host entry)"),
}


# ---------------------------------------------------------------- baselines
def _cpu_library(variant):
    ref = os.path.join(HERE, "oracle", "_ref", "libsw_ref_%s_omp.so" % variant)
    if os.path.exists(ref):
        return ref, "reference"
    port = os.path.join(HERE, "oracle", "_build", "libsw_oracle_%s_omp.so" % variant)
    if os.path.exists(port):
        return port, "port"
    return None, None


def time_cpu(kwargs, variant, target_s=15.0):
    from simwave_amd.middleware import Middleware
    path, kind = _cpu_library(variant)
    if path is None:
        return None
    lib = ctypes.CDLL(path)
    fn = lib.forward
    fn.restype = ctypes.c_double
    r = kwargs["space_order"] // 2
    pts = 1
    for n in kwargs["velocity_model"].shape:
        pts *= n - 2 * r

    def run(nsteps):
        kw = dict(kwargs)
        kw["u_full"] = np.zeros((3,) + kwargs["velocity_model"].shape,
                                dtype=kwargs["velocity_model"].dtype)
        kw["shot_record"] = np.zeros_like(kwargs["shot_record"])
        kw["saving_stride"] = 0
        kw["num_snapshots"] = 3
        kw["begin_timestep"] = 1
        kw["end_timestep"] = nsteps
        argtypes, args = Middleware.ordered(kw)
        fn.argtypes = argtypes
        return fn(*args)

    probe = 2
    t_probe = run(probe)
    nsteps = int(max(probe, min(kwargs["end_timestep"],
                                target_s / max(t_probe / probe, 1e-6))))
    secs = run(nsteps)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(pts * nsteps / secs / 1e9, 4), "unit": "Gpts/s",
            "cores": threads, "kind": kind,
            "sample": "%d of %d time steps (cpu_openmp forward, %d OpenMP threads)"
                      % (nsteps, kwargs["end_timestep"], threads),
            "seconds": round(secs, 3)}


def read_traffic(kernel, workload, dtype):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary (float32
    runs; no summary covers float64)."""
    if dtype != np.float32:
        return None
    path = os.path.join(HERE, "profiles", "%s_%s_traffic.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain streaming kernel of known traffic: torch's elementwise
    y = -x over `nbytes` (reads nbytes, writes nbytes), `reps` times.  It
    scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"kernel": "elementwise neg (float32)", "bytes_read": nbytes,
                      "bytes_written": nbytes, "dispatches": reps}), flush=True)


# --------------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="3d_cd_so8_384", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="f32", choices=("f32", "f64"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--time-steps", type=int, default=None,
                    help="truncate each shot to this many time steps (profiling)")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--rate", type=float, default=400.0, help="serving arrivals per second")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="serving budget, GB")
    ap.add_argument("--max-active", type=int, default=16)
    ap.add_argument("--chunk-steps", type=int, default=64)
    ap.add_argument("--plan-flags", type=int, default=None,
                    help="serving: plan flags (default: the server's, graph replay)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (counter scale)")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shots shard across GPUs with no exchange; run one process "
              "per GPU).", file=sys.stderr)
        sys.exit(2)

    import torch
    import simwave_amd as sw
    from simwave_amd.device import DeviceForward, FLAG_TIME_STEPS, FLAG_OVERLAP_RECEIVERS

    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "serving":
        run_serving(args)
        return
    if args.workload == "fwi_3d_so8_256":
        run_fwi(args, WORKLOADS[args.workload][1])
        return
    if args.workload == "2d_vd_so16_stride10":
        builder, desc = WORKLOADS[args.workload]
        run_host_workload(args, lambda: builder(
            dtype=np.float32 if args.dtype == "f32" else np.float64), desc)
        return
    dtype = np.float32 if args.dtype == "f32" else np.float64
    builder, desc = WORKLOADS[args.workload]
    t0 = time.time()
    parts = builder(dtype=dtype)
    solver = sw.Solver(**parts)
    dev = DeviceForward(solver, end_timestep=args.time_steps)
    t_host = time.time() - t0
    kw = dev.kwargs
    steps_per_shot = kw["end_timestep"] - kw["begin_timestep"] + 1
    updates = dev.updates_per_shot
    log("workload %s: grid %s, %d time steps, %.3e updates/shot, host prep %.1fs"
        % (args.workload, kw["velocity_model"].shape, steps_per_shot, updates, t_host))

    for _ in range(args.warmup):
        dev.reset()
        dev.run()
    torch.cuda.synchronize()
    shot_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        dev.reset()
        dev.run()            # returns after the shot's stream has drained
        shot_ms.append((time.perf_counter() - ts) * 1e3)
        log("shot %d done (%.1f ms)" % (i, shot_ms[-1]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    gpts = updates / (elapsed / args.steps) / 1e9
    fused = dev.timings()[6] == 1.0

    roof = None
    if not args.no_roofline:
        dev.reset()
        dev.run(flags=FLAG_TIME_STEPS | FLAG_OVERLAP_RECEIVERS)
        tm = dev.timings()
        avg = tm[4] / max(tm[5], 1)
        r = kw["space_order"] // 2
        pts = updates // steps_per_shot
        bpp = (20 if dtype == np.float32 else 40) + (
            (4 if dtype == np.float32 else 8) if kw.get("density_model") is not None else 0)
        achieved = pts * bpp / avg / 1e9
        ndim = kw["velocity_model"].ndim
        if not fused:
            kname = "k_stencil_generic"
        elif ndim == 3 and tm[7] == 1.0:
            kname = "k_step3d_vec"
        elif tm[7] == 2.0:
            kname = "k_step%dd_z" % ndim
        else:
            kname = "k_step%dd" % ndim
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(kname, args.workload, dtype),
                "kernel": kname, "avg_launch_us": round(avg * 1e6, 2),
                "launches": int(tm[5]), "algorithmic_bytes_per_launch": pts * bpp}
        log("step kernel: %d launches, avg %.1f us, %.1f GB/s algorithmic"
            % (tm[5], avg * 1e6, achieved))

    cpu = None
    if not args.no_cpu_baseline:
        from simwave_amd.middleware import Middleware
        cpu = time_cpu(Middleware.normalize(**solver.forward_arguments()), dev.variant,
                       args.cpu_seconds)
        if cpu:
            log("cpu baseline: %s" % cpu)

    out = {
        "metric": METRIC,
        "value": round(gpts, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc,
                   "grid": list(kw["velocity_model"].shape),
                   "time_steps": steps_per_shot, "updates_per_shot": updates,
                   "space_order": kw["space_order"],
                   "receivers": kw["num_receivers"], "parallelism": "none",
                   "kernel_path": "fused" if fused else "generic"},
        "shot_latency_ms": {"p50": round(float(np.percentile(shot_ms, 50)), 2),
                            "p99": round(float(np.percentile(shot_ms, 99)), 2),
                            "shots": len(shot_ms), "load": "one shot at a time"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
