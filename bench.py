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

Other workloads (--workload): the README example (2d_cd_so4_readme), the
faulted 2D variable-density run with snapshots (2d_vd_so16_stride10, host
entry, PCIe-inclusive), the FWI gradient iteration over 16 shots
(fwi_3d_so8_256) and the mixed shot stream through the server (serving).
The experiments are built by simwave_amd.workloads.

The roofline object is measured live for the dominant kernel with HIP events
around each launch on the launch stream.  The CPU baseline times the
reference's own cpu_openmp forward (built from its sources into oracle/_ref)
or, where that build is absent, the C restatement in oracle/, on a bounded
number of time steps of the same grid, in a child process with
OMP_PROC_BIND=close and OMP_PLACES=cores (set there only, so that the
placement binds no thread of this process).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
    python bench.py --copy-calibration   # plain device copy of known size
"""
import argparse
import contextlib
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("Gpts/s & % HBM roofline (3D order-8 fp32); "
          "p50/p99 shot latency under load")
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "3d_cd_so8_384": "3D constant density so8 384^3 (472^3 extended), 12.5 m, 40-pt cubic "
                     "damping, tf 1.5 s, 128x128 receivers",
    "2d_cd_so4_readme": "2D constant density so4 512^2 resampled to 513^2 (README example), "
                        "tf 1.0 s, 512 receivers",
    "2d_vd_so16_stride10": "2D variable density so16 1000x3000 (1136x3136 extended), 5 m, "
                           "60-pt damping, faulted layers, tf 3.0 s, 3000 receivers, "
                           "saving_stride 10 (host entry)",
    "fwi_3d_so8_256": "FWI gradient iteration: L2-misfit velocity gradient summed over 16 "
                      "shots, 3D so8 160x256x256 (228x324x324 extended), 15 m, tf 1.2 s, "
                      "100x100 receivers, checkpoints every 20 steps",
    "serving": "mixed shot stream through the continuous-batching server",
    "3d_vd_so8_256": "3D variable density so8 256^3, 12.5 m, 10-pt damping, tf 0.3 s, "
                     "64x64 receivers (profiling grid of the variable-density kernel)",
    "3d_vd_so16_256": "3D variable density so16 256^3, 12.5 m, 10-pt damping, tf 0.3 s, "
                      "64x64 receivers (profiling grid of the variable-density kernel)",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _dtype(args):
    return np.float32 if args.dtype == "f32" else np.float64


def _variant(kw):
    vel = kw["velocity_model"]
    dens = "vd" if kw.get("density_model") is not None else "cd"
    return "%dd_%s_%s" % (vel.ndim, dens, "f32" if vel.dtype == np.float32 else "f64")


# ---------------------------------------------------------------- baselines
def host_info():
    """Host cores of the CPU baseline: logical CPUs, those this process may
    run on, the CPU model and the OpenMP placement."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def _cpu_library(variant):
    ref = os.path.join(HERE, "oracle", "_ref", "libsw_ref_%s_omp.so" % variant)
    if os.path.exists(ref):
        return ref, "reference"
    port = os.path.join(HERE, "oracle", "_build", "libsw_oracle_%s_omp.so" % variant)
    if os.path.exists(port):
        return port, "port"
    return None, None


def time_cpu(kwargs, variant, target_s=15.0, what="forward"):
    """The cpu_openmp forward on the same grid over a bounded number of time
    steps (about target_s seconds), with saving_stride 0."""
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
    info = host_info()
    threads = int(info["omp_num_threads"] or info["usable_cpus"] or 1)
    out = {"value": round(pts * nsteps / secs / 1e9, 4), "unit": "Gpts/s",
           "cores": threads, "kind": kind,
           "sample": "%d of %d time steps of the %s (cpu_openmp forward, %d OpenMP threads)"
                     % (nsteps, kwargs["end_timestep"], what, threads),
           "seconds": round(secs, 3)}
    out.update(info)
    return out


def cpu_workload(workload, dtype, seed):
    """(forward keyword arguments, variant, description) of the run the CPU
    baseline times for `workload`."""
    import simwave_amd as sw
    from simwave_amd import workloads
    from simwave_amd.middleware import Middleware
    what = "forward"
    if workload == "fwi_3d_so8_256":
        true, init = workloads.fwi_models(seed=seed)
        parts = workloads.fwi_shot(init, 0, dtype=dtype, dt=workloads.fwi_dt(true))
        what = "forward of shot 0 (the reference has no adjoint)"
    elif workload == "2d_vd_so16_stride10":
        parts = workloads.faulted_2d_vd(dtype)
    elif workload == "2d_cd_so4_readme":
        parts = workloads.readme_2d(dtype)
    elif workload.startswith("3d_vd_so"):
        parts = workloads.vd_3d(int(workload.split("_")[2][2:]), dtype)
    else:
        parts = workloads.headline_3d(dtype)
    kw = Middleware.normalize(**sw.Solver(**parts).forward_arguments())
    return kw, _variant(kw), what


def cpu_baseline(args):
    """The CPU baseline of args.workload, timed in a child process whose
    OpenMP threads are placed one per core (OMP_PROC_BIND=close,
    OMP_PLACES=cores); None when no CPU build is present."""
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-only", "--workload",
           args.workload, "--dtype", args.dtype, "--cpu-seconds", str(args.cpu_seconds),
           "--seed", str(args.seed)]
    res = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        log("cpu baseline failed:\n" + res.stderr[-2000:])
        return None
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    out = json.loads(lines[-1]) if lines else None
    log("cpu baseline: %s" % out)
    return out


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


def step_roofline(dev, workload, dtype):
    """Roofline of the step kernel of a device-resident shot: every step
    kernel bracketed by HIP events on its stream (FLAG_TIME_STEPS); the
    algorithmic bytes are 20 (float32) / 40 (float64) per interior update,
    plus the density (4 / 8) for variable density."""
    from simwave_amd.device import FLAG_TIME_STEPS, FLAG_OVERLAP_RECEIVERS
    kw = dev.kwargs
    dev.reset()
    dev.run(flags=FLAG_TIME_STEPS | FLAG_OVERLAP_RECEIVERS)
    tm = dev.timings()
    avg = tm[4] / max(tm[5], 1)
    steps = kw["end_timestep"] - kw["begin_timestep"] + 1
    pts = dev.updates_per_shot // steps
    esz = np.dtype(dtype).itemsize
    bpp = 5 * esz + (esz if kw.get("density_model") is not None else 0)
    achieved = pts * bpp / avg / 1e9
    ndim = kw["velocity_model"].ndim
    if tm[6] != 1.0:
        kname = "k_stencil_generic"
    elif ndim == 3 and tm[7] == 1.0:
        kname = "k_step3d_vec"
    elif tm[7] == 2.0:
        kname = "k_step%dd_z" % ndim
    else:
        kname = "k_step%dd" % ndim
    log("step kernel %s: %d launches, avg %.2f us, %.1f GB/s algorithmic"
        % (kname, tm[5], avg * 1e6, achieved))
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(kname, workload, dtype),
            "kernel": kname, "avg_launch_us": round(avg * 1e6, 2),
            "launches": int(tm[5]), "algorithmic_bytes_per_launch": pts * bpp,
            "bytes_per_update": bpp}


def _record(args, value, ms, config, **extra):
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gpts/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": args.dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


# ------------------------------------------------------- device-resident shot
def run_device_workload(args, parts):
    """Whole shots with everything resident in HBM (DeviceForward): the
    headline and the README workloads."""
    import torch
    import simwave_amd as sw
    from simwave_amd.device import DeviceForward

    dtype = _dtype(args)
    t0 = time.time()
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
    fused = dev.timings()[6] == 1.0
    roof = None if args.no_roofline else step_roofline(dev, args.workload, dtype)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args)
    config = {"workload": args.workload, "description": WORKLOADS[args.workload],
              "grid": list(kw["velocity_model"].shape), "time_steps": steps_per_shot,
              "updates_per_shot": updates, "space_order": kw["space_order"],
              "receivers": kw["num_receivers"], "parallelism": "none",
              "kernel_path": "fused" if fused else "generic"}
    return _record(args, updates / (elapsed / args.steps) / 1e9,
                   elapsed / args.steps * 1e3, config,
                   shot_latency_ms={"p50": round(float(np.percentile(shot_ms, 50)), 2),
                                    "p99": round(float(np.percentile(shot_ms, 99)), 2),
                                    "shots": len(shot_ms), "load": "one shot at a time"},
                   roofline=roof, cpu_baseline=cpu)


# ------------------------------------------------- host entry with snapshots
def run_host_workload(args, parts):
    """Solver.forward() (the reference-compatible host entry: models and
    every returned snapshot cross PCIe inside the call).  The roofline is
    the step kernel's on the same shot resident in HBM."""
    import torch
    import simwave_amd as sw
    from simwave_amd.device import DeviceForward

    dtype = _dtype(args)
    t0 = time.time()
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
            solver.forward()
        times.append(time.perf_counter() - ts)
        log("shot %d: %.3fs" % (i, times[-1]))
    el = float(np.median(times))
    roof = None
    if not args.no_roofline:
        dev = DeviceForward(solver)
        roof = step_roofline(dev, args.workload, dtype)
        del dev
        torch.cuda.empty_cache()
    cpu = None if args.no_cpu_baseline else cpu_baseline(args)
    config = {"workload": args.workload, "description": WORKLOADS[args.workload],
              "grid": list(kw["velocity_model"].shape),
              "time_steps": int(kw["end_timestep"]),
              "snapshots": int(kw["num_snapshots"]), "updates_per_shot": upd,
              "path": "host entry (Solver.forward), PCIe-inclusive"}
    return _record(args, upd / el / 1e9, el * 1e3, config,
                   shot_latency_ms={"p50": round(float(np.percentile(times, 50)) * 1e3, 2),
                                    "p99": round(float(np.percentile(times, 99)) * 1e3, 2),
                                    "shots": len(times), "load": "one shot at a time"},
                   roofline=roof, cpu_baseline=cpu)


# ------------------------------------------------------------ FWI iteration
def run_fwi(args):
    """One FWI iteration per step: the L2-misfit gradient summed over the 16
    shots of the experiment (4 x 4 sources) on the smoothed starting model,
    against data observed on the true model (+10% sphere), each shot's
    forward, residual and reverse sweep with checkpoints every 20 steps.
    value = the 16 shots' forward updates / iteration seconds.  Outside the
    timed region the same gradient in float64 gives the float32 error.  The
    roofline is the reverse update kernel's (k_adj_update_z, 13 grid
    accesses per extended point: W, lambda^{n+1}, e, v, lambda^n, three
    forward levels and the gradient sum read; lambda^n, lambda^{n+1},
    lambda^{n-1} and the gradient sum written)."""
    import torch
    from simwave_amd import fwi, workloads

    dtype = _dtype(args)
    F64 = np.float64
    t0 = time.time()
    true, init = workloads.fwi_models(seed=args.seed)
    dt = workloads.fwi_dt(true)
    shots = list(range(workloads.FWI_SHOTS))
    obs = fwi.observed_data([workloads.fwi_shot(true, s, dtype=F64, dt=dt) for s in shots])
    log("fwi: observed data of %d shots (float64, true model) in %.1fs"
        % (len(shots), time.time() - t0))
    t0 = time.time()
    devs = fwi.modelled_shots([workloads.fwi_shot(init, s, dtype=dtype, dt=dt)
                               for s in shots])
    kw = devs[0].kwargs
    T, nrec = kw["end_timestep"], kw["num_receivers"]
    upd = sum(d.updates_per_shot for d in devs)
    log("fwi: %d modelled shots, grid %s, %d steps, %d receivers, built in %.1fs"
        % (len(devs), kw["velocity_model"].shape, T, nrec, time.time() - t0))
    K = 20
    for _ in range(args.warmup):
        fwi.misfit_gradient(devs, obs, checkpoint_every=K)
    torch.cuda.synchronize()
    it_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        misfit, grad = fwi.misfit_gradient(devs, obs, checkpoint_every=K)
        it_ms.append((time.perf_counter() - ts) * 1e3)
        log("iteration %d: misfit %.6g, %.1f ms" % (i, misfit, it_ms[-1]))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # the same iteration with its reverse update kernels timed, and the
    # forward alone (shot 0)
    stats = {}
    fwi.misfit_gradient(devs, obs, checkpoint_every=K, stats=stats)
    devs[0].reset()
    tf0 = time.perf_counter()
    devs[0].run()
    fwd_ms = (time.perf_counter() - tf0) * 1e3
    roof = None
    if stats.get("update_launches"):
        avg = stats["update_s"] / stats["update_launches"]
        npts = int(np.prod(kw["velocity_model"].shape))
        nbytes = 13 * np.dtype(dtype).itemsize * npts
        achieved = nbytes / avg / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic("k_adj_update_z", args.workload, dtype),
                "kernel": "k_adj_update_z", "avg_launch_us": round(avg * 1e6, 2),
                "launches": int(stats["update_launches"]),
                "algorithmic_bytes_per_launch": int(nbytes)}
        log("reverse update: %d launches, avg %.1f us, %.1f GB/s algorithmic"
            % (stats["update_launches"], avg * 1e6, achieved))
    err = None
    if dtype != F64 and not args.no_f64_check:
        g32 = grad.clone()
        del devs
        torch.cuda.empty_cache()
        t0 = time.time()
        devs64 = fwi.modelled_shots([workloads.fwi_shot(init, s, dtype=F64, dt=dt)
                                     for s in shots])
        m64, g64 = fwi.misfit_gradient(devs64, obs, checkpoint_every=K)
        err = fwi.relative_l2(g32, g64)
        log("float64 gradient in %.1fs: relative L2 error of float32 %.3g (misfit %.9g "
            "vs %.9g)" % (time.time() - t0, err, misfit, m64))
        del devs64
    cpu = None if args.no_cpu_baseline else cpu_baseline(args)
    config = {"workload": args.workload, "description": WORKLOADS[args.workload],
              "grid": list(kw["velocity_model"].shape), "time_steps": T,
              "shots": len(shots), "updates_per_iteration": upd, "receivers": nrec,
              "checkpoint_every": K, "parallelism": "none"}
    return _record(args, upd / (elapsed / args.steps) / 1e9, elapsed / args.steps * 1e3,
                   config,
                   fwi={"iteration_ms_p50": round(float(np.percentile(it_ms, 50)), 1),
                        "gradient_ms_per_shot": round(float(np.percentile(it_ms, 50))
                                                      / len(shots), 2),
                        "forward_ms_per_shot": round(fwd_ms, 2),
                        "misfit": misfit,
                        "gradient_rel_l2_vs_f64": err},
                   roofline=roof, cpu_baseline=cpu)


# ----------------------------------------------------------------- serving
def graph_replay_effect(reps=3):
    """A launch-bound shot (2D 128^2, order 2, 960 steps) run through a plan
    in 48-step chunks, with direct launches and with graph replay."""
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
    from simwave_amd import hipmem
    stream = hipmem.Stream(hipmem.default_library())
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
    server under Poisson arrivals, p50/p99 latency, and the plain path and
    the device-resident path (the same shots one at a time) beside it."""
    import torch
    from simwave_amd import workloads
    from simwave_amd.device import DeviceForward
    from simwave_amd.serving import ShotServer

    n2 = tuple(int(x) for x in args.serve_2d.split(","))
    n3 = tuple(int(x) for x in args.serve_3d.split(","))
    reqs, build = workloads.serving_requests(args.serve_requests, seed=args.seed, n2=n2,
                                             n3=n3, tf_scale=args.tf_scale)
    warm = workloads.serving_requests(4, seed=args.seed + 1000, n2=(128, 128),
                                      n3=(32, 32), steps=(20, 20))
    log("serving: warm-up (%d small shots)" % len(warm[0]))
    with ShotServer(hbm_budget=args.hbm_budget * 1e9, max_active=args.max_active,
                    chunk_steps=args.chunk_steps, plan_flags=args.plan_flags) as srv:
        for f in [srv.submit(warm[1](r)) for r in warm[0]]:
            f.result(timeout=120)
    log("serving: warm-up done")
    rng = np.random.default_rng(args.seed + 1)
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs))
    torch.cuda.synchronize()
    with ShotServer(hbm_budget=args.hbm_budget * 1e9, max_active=args.max_active,
                    chunk_steps=args.chunk_steps, plan_flags=args.plan_flags,
                    build_ahead=args.build_ahead) as srv:
        start = time.perf_counter()
        futs = []
        due = start
        for r, g in zip(reqs, gaps):
            due += g
            delay = due - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(lambda r=r: build(r), arrival=due, keep_result=False))
        for i, f in enumerate(futs):
            f.result(timeout=900)
            if (i + 1) % 50 == 0 or i + 1 == len(futs):
                log("serving: request %d/%d done" % (i + 1, len(futs)))
    summ = srv.stats.summary()
    log("serving: %s" % summ)
    # the paths one shot at a time, on a prefix of the same stream
    sample = reqs[:min(len(reqs), args.plain_sample)]
    torch.cuda.synchronize()
    tp = time.perf_counter()
    upd = 0
    for i, r in enumerate(sample):
        s = build(r)
        with contextlib.redirect_stdout(sys.stderr):   # its timing line
            s.forward()
        upd += _updates(s)
        if (i + 1) % 25 == 0:
            log("plain path (Solver.forward): %d/%d shots" % (i + 1, len(sample)))
    plain_s = time.perf_counter() - tp
    plain = upd / plain_s / 1e9
    log("plain path: %d shots in %.2fs, %.2f Gpts/s" % (len(sample), plain_s, plain))
    torch.cuda.synchronize()
    tp = time.perf_counter()
    upd = 0
    for i, r in enumerate(sample):
        dev = DeviceForward(build(r))
        dev.run()
        dev.results()
        upd += dev.updates_per_shot
        del dev
        if (i + 1) % 25 == 0:
            log("device path: %d/%d shots" % (i + 1, len(sample)))
    torch.cuda.synchronize()
    dev_s = time.perf_counter() - tp
    devp = upd / dev_s / 1e9
    log("device path: %d shots in %.2fs, %.2f Gpts/s" % (len(sample), dev_s, devp))
    graph = graph_replay_effect()
    log("graph replay on a small shot: %s" % graph)
    args.steps, args.warmup = len(reqs), len(warm[0])
    config = {"workload": "serving", "requests": len(reqs), "seed": args.seed,
              "arrival_rate_hz": args.rate, "hbm_budget_gb": args.hbm_budget,
              "max_active": args.max_active, "chunk_steps": args.chunk_steps,
              "build_ahead": args.build_ahead,
              "plan_flags": args.plan_flags, "grids_2d": list(n2), "grids_3d": list(n3),
              "tf_scale": args.tf_scale,
              "description": workloads.serving_requests.__doc__.split("\n")[0]}
    return _record(
        args, summ["gpts"], summ["makespan_s"] / len(reqs) * 1e3, config,
        shot_latency_ms={"p50": round(summ["latency_p50_s"] * 1e3, 2),
                         "p99": round(summ["latency_p99_s"] * 1e3, 2),
                         "load": "Poisson arrivals at %g/s" % args.rate},
        serving={"updates": summ["updates"], "makespan_s": round(summ["makespan_s"], 3),
                 "queue_p50_ms": round(summ["queue_p50_s"] * 1e3, 2),
                 "max_active_seen": summ["max_active"],
                 "prep_mean_ms": round(summ["prep_mean_s"] * 1e3, 2),
                 "device_mean_ms": round(summ["device_mean_s"] * 1e3, 2),
                 "fetch_mean_ms": round(summ["fetch_mean_s"] * 1e3, 2),
                 "batched": summ.get("batched"),
                 "one_at_a_time_sample": len(sample),
                 "plain_path": "Solver.forward(), one shot at a time",
                 "plain_path_gpts": round(plain, 3),
                 "speedup_vs_plain": round(summ["gpts"] / plain, 3),
                 "device_path_gpts": round(devp, 3),
                 "speedup_vs_device_path": round(summ["gpts"] / devp, 3),
                 "graph_replay": graph},
        roofline=None, cpu_baseline=None)


def _updates(solver):
    sm, tm = solver.space_model, solver.time_model
    r = sm.space_order // 2
    return int(np.prod([n - 2 * r for n in sm.extended_shape])) * int(tm.timesteps)


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
    ap.add_argument("--no-f64-check", action="store_true",
                    help="fwi: skip the float64 gradient beside the timed float32 one")
    ap.add_argument("--time-steps", type=int, default=None,
                    help="truncate each shot to this many time steps (profiling)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-2d", default="256,2048", help="serving: 2D grid edge range")
    ap.add_argument("--serve-3d", default="128,384", help="serving: 3D grid edge range")
    ap.add_argument("--tf-scale", type=float, default=0.1,
                    help="serving: duration as a fraction of the slowest layer's "
                         "crossing time of the grid's depth")
    ap.add_argument("--plain-sample", type=int, default=200,
                    help="serving: requests of the stream run one at a time for the "
                         "plain and device-resident paths")
    ap.add_argument("--rate", type=float, default=30.0,
                    help="serving arrivals per second (the default stream of 2000 "
                         "requests is sustainable at it: about 85 Gpts/s offered)")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="serving budget, GB")
    ap.add_argument("--max-active", type=int, default=32)
    ap.add_argument("--chunk-steps", type=int, default=48)
    ap.add_argument("--build-ahead", type=int, default=4,
                    help="serving: requests built ahead of admission on a thread pool")
    ap.add_argument("--plan-flags", type=int, default=None,
                    help="serving: plan flags (default: the server's, graph replay)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (counter scale)")
    ap.add_argument("--cpu-only", action="store_true",
                    help="print only the CPU baseline of the workload (no GPU)")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration()
        return
    if args.cpu_only:
        info = host_info()      # before the OpenMP runtime binds this thread
        kw, variant, what = cpu_workload(args.workload, _dtype(args), args.seed)
        out = time_cpu(kw, variant, args.cpu_seconds, what)
        if out:
            out.update(info)
        print(json.dumps(out), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(shots shard across GPUs with no exchange; run one process "
              "per GPU).", file=sys.stderr)
        sys.exit(2)

    import torch
    from simwave_amd import workloads
    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    dtype = _dtype(args)
    if args.workload == "serving":
        out = run_serving(args)
    elif args.workload == "fwi_3d_so8_256":
        out = run_fwi(args)
    elif args.workload == "2d_vd_so16_stride10":
        out = run_host_workload(args, workloads.faulted_2d_vd(dtype))
    elif args.workload == "2d_cd_so4_readme":
        out = run_device_workload(args, workloads.readme_2d(dtype))
    elif args.workload.startswith("3d_vd_so"):
        out = run_device_workload(args, workloads.vd_3d(int(args.workload.split("_")[2][2:]),
                                                        dtype))
    else:
        out = run_device_workload(args, workloads.headline_3d(dtype))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
