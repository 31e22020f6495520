"""The benchmark experiments and the reference's own accelerator tests on
the HIP path (needs an MI355X).

* The reference's golden wavefields (tests/golden/*.npz, copied as data
  from the reference's tests/reference_solution) on the HIP library under
  the reference's accelerator criterion (tests/test_gpu_solution.py:21-139
  of the reference, there atol 1e-4; here 1e-5), and bit identity with the
  digests of the reference C's outputs on its golden configurations
  (tests/golden/ref_outputs.json, written by make_fixtures.py).
* The analytic 2D response and the convergence rates in time and space of
  the reference's tests/test_convergence.py:52-139 and
  tests/test_gpu_convergence.py, computed by the HIP library.
* The README, headline 3D (float32 at its full 1254 steps, float64 over
  400) and faulted 2D variable-density (all 6361 steps, every one of its
  637 snapshots) experiments at full size, bit for bit against the
  oracle's OpenMP build on the same inputs.
* The FWI gradient experiment at its full size (160 x 256 x 256, so8,
  228 x 324 x 324 extended, 736 steps, 10 000 receivers): float32 against
  float64 gradients of the summed L2 misfit, a dot-product test, central
  finite differences on 10 cells and checkpointing every 20 steps against
  every level kept.
* A serving mix drawn from the full size ranges of the mixed shot stream,
  served concurrently, each result bit for bit what Solver.forward()
  returns for it alone.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np
import pytest

import cases
import harness
import simwave_amd as sw
from simwave_amd import adjoint, fwi, workloads
from simwave_amd.device import DeviceForward
from simwave_amd.serving import ShotServer

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)
import make_fixtures  # noqa: E402

pytestmark = pytest.mark.gpu

F32, F64 = np.float32, np.float64


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8)).hexdigest()


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- golden
@pytest.mark.parametrize("so", [2, 8])
def test_reference_golden_wavefields_on_hip(gpu, so):
    u, _ = harness.run(cases.golden_2d(so), "hip")
    with np.load(os.path.join(GOLDEN, "wavefield-2d-so-%d.npz" % so),
                 allow_pickle=False) as z:
        ref = z["u"]
    assert u.shape == ref.shape
    err = harness.rel_l2(u, ref)
    print("golden so%d: max |diff| %.3g rel L2 %.3g" % (so, np.abs(u - ref).max(), err))
    assert np.allclose(u, ref, atol=1e-05)
    assert err < 1e-5


REF = _load("ref_outputs.json")


@pytest.mark.parametrize("name", sorted(REF))
def test_reference_digests_on_hip(gpu, name):
    """The HIP library returns the reference C's bits on the reference's
    golden configurations (2D and 3D) and the kernel-level cases."""
    print("reference digest", name, flush=True)
    u, rec = harness.run(make_fixtures.setup_of(name), "hip")
    want = REF[name]
    assert list(u.shape) == want["u"]["shape"]
    assert list(rec.shape) == want["rec"]["shape"]
    assert _sha(u) == want["u"]["sha256"], "u differs from the reference C"
    assert _sha(rec) == want["rec"]["sha256"], "shot record differs"


# ------------------------------------------------ analytic accuracy, rates
def _hankel_error(space_order=8, spacing=0.5, dt=0.1):
    setup = cases.hankel_accuracy_case(space_order=space_order, spacing=spacing, dt=dt)
    _, rec = harness.run(setup, "hip")
    num = rec.ravel() / spacing ** 2
    exact = cases.hankel_reference(setup)
    return float(np.linalg.norm(num - exact) / np.sqrt(num.size))


def test_hankel_accuracy_on_hip(gpu):
    """Reference bound 1e-4 (tests/test_convergence.py:75-87)."""
    err = _hankel_error()
    print("hankel error %.3g" % err)
    assert err < 1e-4


def test_convergence_in_time_on_hip(gpu):
    """Slope of log error against log dt above 1.7 over dt 0.1 .. 0.025
    (tests/test_convergence.py:90-112)."""
    dts = [0.1, 0.075, 0.04, 0.025]
    errs = [_hankel_error(dt=dt) for dt in dts]
    rate = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    print("time errors", errs, "rate %.3f" % rate)
    assert rate > 1.7


def test_convergence_in_space_on_hip(gpu):
    """Slopes of log error against log spacing above 1.7, 4, 6, 7.7, 8.7 for
    space orders 2 .. 10 over spacings 2, 2.5, 4 at dt 0.025
    (tests/test_convergence.py:115-139)."""
    spacings = [2.0, 2.5, 4.0]
    for order, floor in zip([2, 4, 6, 8, 10], [1.7, 4, 6, 7.7, 8.7]):
        errs = [_hankel_error(space_order=order, spacing=h, dt=0.025) for h in spacings]
        rate = np.polyfit(np.log(spacings), np.log(errs), 1)[0]
        print("order %d errors %s rate %.3f" % (order, errs, rate), flush=True)
        assert rate > floor, (order, rate)


# ------------------------------------------- benchmark experiments, full size
EXPERIMENTS = {
    # README example: 513^2 so4, 328 steps
    "readme_2d": lambda: workloads.readme_2d(F32),
    # headline: 472^3 so8, all 1254 steps
    "headline_3d_f32": lambda: workloads.headline_3d(F32),
    # headline in float64, cut to tf 0.48 s (400 steps) to bound the oracle's time
    "headline_3d_f64_400": lambda: workloads.headline_3d(F64, tf=0.48),
    # faulted variable density: 1136 x 3136 so16, all 6361 steps, 637 snapshots
    "faulted_2d_vd_f32": lambda: workloads.faulted_2d_vd(F32),
}


@pytest.mark.parametrize("name", sorted(EXPERIMENTS))
def test_benchmark_experiment_matches_oracle(gpu, name):
    """Full-size runs through Solver.forward() (the drop-in path) against
    the oracle's OpenMP build on the same inputs: every returned snapshot
    and the shot record bit for bit (the stated bars, 1e-5 float32 and
    1e-12 float64 relative L2, are reported and asserted too)."""
    print("experiment", name, flush=True)
    t0 = time.time()
    parts = EXPERIMENTS[name]()
    solver = sw.Solver(**parts)
    t1 = time.time()
    u, rec = solver.forward()
    t2 = time.time()
    print("%s: host build %.1f s, HIP forward %.2f s, u %s" % (name, t1 - t0, t2 - t1, u.shape),
          flush=True)
    oracle = sw.Solver(compiler=harness.LibraryCompiler("oracle", omp=True), **parts)
    oracle._middleware.verbose = False
    u_o, rec_o = oracle.forward()
    print("%s: oracle (OpenMP, %s threads) %.1f s" % (name, os.environ.get("OMP_NUM_THREADS"),
                                                      time.time() - t2), flush=True)
    tol = 1e-12 if u.dtype == np.float64 else 1e-5
    eu, er = harness.rel_l2(u, u_o), harness.rel_l2(rec, rec_o)
    msg = "%s: rel_l2 u=%.3g rec=%.3g" % (name, eu, er)
    print(msg, flush=True)
    assert u.shape == u_o.shape and rec.shape == rec_o.shape
    assert np.abs(rec_o).max() > 0
    assert eu <= tol and er <= tol, msg
    assert np.array_equal(rec, rec_o), "shot record not bit-identical: " + msg
    assert np.array_equal(u, u_o), "wavefield not bit-identical: " + msg


# --------------------------------------------------- FWI gradient, full size
FWI_SHOTS = (0, 5, 10, 15)     # the diagonal of the 4 x 4 source grid


@pytest.fixture(scope="module")
def fwi_setup(gpu):
    true, init = workloads.fwi_models()
    dt = workloads.fwi_dt(true)
    return true, init, dt


def _shot(vel, shot, dtype, dt):
    return workloads.fwi_shot(vel, shot, dtype=dtype, dt=dt)


def test_fwi_gradient_float32_matches_float64(fwi_setup):
    """Summed L2-misfit gradient over 4 shots on the smoothed model, data
    observed on the true model: float32 within 1e-4 relative L2 of
    float64; checkpoints every 20 steps equal to every level kept."""
    true, init, dt = fwi_setup
    grads, misfits = {}, {}
    for dtype in (F64, F32):
        t0 = time.time()
        obs = fwi.observed_data([_shot(true, s, F64, dt) for s in FWI_SHOTS])
        shots = fwi.modelled_shots([_shot(init, s, dtype, dt) for s in FWI_SHOTS])
        m, g = fwi.misfit_gradient(shots, obs, checkpoint_every=20)
        grads[dtype], misfits[dtype] = g, m
        print("%s: misfit %.9g, %.1f s" % (np.dtype(dtype).name, m, time.time() - t0),
              flush=True)
        if dtype is F32:
            m_all, g_all = fwi.misfit_gradient(shots, obs,
                                               checkpoint_every=shots[0].steps + 1)
            err_k = fwi.relative_l2(g_all, g)
            print("f32 every level vs K=20: %.3g" % err_k, flush=True)
            assert err_k < 1e-13
        del shots, obs
    err = fwi.relative_l2(grads[F32], grads[F64])
    print("f32 vs f64 gradient rel L2 %.3g, misfit rel diff %.3g"
          % (err, abs(misfits[F32] - misfits[F64]) / misfits[F64]), flush=True)
    assert float(grads[F64].abs().max()) > 0
    assert err < 1e-4


def test_fwi_wavelet_dot_product(fwi_setup):
    """<F s, r> = <s, F^T r> at full size (float64), F the map from wavelet
    samples to the shot record."""
    _, init, dt = fwi_setup
    parts = _shot(init, 5, F64, dt)
    rng = np.random.default_rng(11)
    T = parts["time_model"].timesteps
    s = rng.standard_normal((T, 1))
    parts["wavelet"] = sw.MultiWavelet(s, parts["time_model"])
    solver = sw.Solver(**parts)
    rr = rng.standard_normal((T, solver.receivers.count))
    out = adjoint.gradient(solver, rr, checkpoint_every=20, wavelet_gradient=True)
    lhs = float(np.sum(out["shot_record"] * rr))
    rhs = float(np.sum(s.reshape(out["grad_wavelet"].shape) * out["grad_wavelet"]))
    print("dot product %.17g %.17g" % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


def test_fwi_gradient_finite_differences(fwi_setup):
    """Central differences of the float64 misfit of one shot on 10 random
    cells where the gradient is at least 1% of its maximum."""
    import torch
    true, init, dt = fwi_setup
    shot = 5
    obs = fwi.observed_data([_shot(true, shot, F64, dt)])[0]
    solver = sw.Solver(**_shot(init, shot, F64, dt))
    dev = DeviceForward(solver, buffers="torch")
    vel = dev.tensors["velocity_model"].clone()
    m0, g = adjoint.shot_misfit_gradient(dev, obs, checkpoint_every=20)
    g = g.cpu().numpy()
    gmax = np.abs(g).max()
    cells = fwi.random_interior_cells(g, 10, seed=12)
    for idx in cells:
        h = 1e-4 * float(vel[idx])
        js = []
        for sign in (1, -1):
            vp = vel.clone()
            vp[idx] += sign * h
            d = DeviceForward(solver, velocity=vp)
            d.run()
            res = d.tensors["shot_record"] - obs
            js.append(0.5 * float((res * res).sum()))
            del d
        fd = (js[0] - js[1]) / (2 * h)
        print("cell %s: fd %.9g grad %.9g" % (idx, fd, g[idx]), flush=True)
        assert abs(fd - g[idx]) <= 1e-4 * abs(g[idx]) + 1e-6 * gmax, (idx, fd, g[idx])
    torch.cuda.synchronize()


# ------------------------------------------------------------- serving mix
def test_serving_mix_from_the_stream_distribution(gpu):
    """20 requests drawn from the stream's full size ranges (2D 256^2 ..
    2048^2, 3D 128^3 .. 384^3, orders 2-20, constant or variable density,
    64-4096 receivers), cut to 20-50 time steps, arriving at 50/s and
    served with up to 8 in flight: each result is bit for bit the
    Solver.forward() result of the same request alone."""
    reqs, build = workloads.serving_requests(20, seed=3, steps=(20, 50))
    alone = []
    for r in reqs:
        print("alone", {k: r[k] for k in ("dim", "n", "so", "dens", "nrec", "nsteps")},
              flush=True)
        alone.append(build(r).forward())
    rng = np.random.default_rng(4)
    gaps = rng.exponential(1 / 50.0, len(reqs))
    with ShotServer(hbm_budget=200e9, max_active=8, chunk_steps=48) as srv:
        futs = []
        for r, gap in zip(reqs, gaps):
            time.sleep(gap)
            futs.append(srv.submit(lambda r=r: build(r)))
        results = [f.result(timeout=600) for f in futs]
    for r, (u_a, r_a), (u, rec) in zip(reqs, alone, results):
        assert u.shape == u_a.shape and rec.shape == r_a.shape, r
        assert np.array_equal(u, u_a), r
        assert np.array_equal(rec, r_a), r
    s = srv.stats.summary()
    print("serving summary", s, flush=True)
    assert s["requests"] == len(reqs) and s["max_active"] > 1
