"""DeviceForward builds the wavefield slots, shot record, extended models
and damping mask on the device; they must equal the arrays the frontend
builds on the host (Solver.forward_arguments()) bit for bit."""
import numpy as np
import pytest

import cases
import simwave_amd as sw
from simwave_amd import hipmem
from simwave_amd.device import DeviceForward
from simwave_amd.middleware import Middleware

pytestmark = pytest.mark.gpu

CASES = {
    "2d_cd": dict(dim=2, n=40, space_order=4),
    "2d_vd_f64": dict(dim=2, n=36, space_order=8, density=True, dtype=np.float64),
    "3d_cd": dict(dim=3, shape=(20, 26, 31), space_order=8, damping=30.0),
    "3d_vd": dict(dim=3, n=22, space_order=4, density=True, damping=50.0),
    "3d_nodamp": dict(dim=3, n=18, space_order=2, damping=0.0),
    "2d_stride": dict(dim=2, n=30, stride=3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_arrays_equal_host_arrays(name):
    print("device arrays", name, flush=True)
    solver = sw.Solver(**cases.make_case(**CASES[name])())
    host = Middleware.normalize(**solver.forward_arguments())
    dev = DeviceForward(solver)
    assert dev.buffers == "library"
    for key, t in dev.tensors.items():
        ref = host[key]
        got = t.numpy()
        assert got.shape == ref.shape and got.dtype == ref.dtype, key
        assert np.array_equal(got, ref), key


@pytest.mark.parametrize("name", ["3d_cd", "2d_vd_f64"])
def test_shot_built_on_one_stream_runs_on_another(name):
    """The shot's uploads and gathers are queued on the stream it is built
    on, which the constructor synchronises; a plan, its chunks and the
    results fetched on another of the library's streams see them and give
    the same arrays as the shot's own stream."""
    print("cross-stream shot", name, flush=True)
    solver = sw.Solver(**cases.make_case(**CASES[name])())
    ref = DeviceForward(solver)
    ref.run()
    u_ref, rec_ref = ref.results()
    lib = hipmem.default_library()
    build, other = hipmem.Stream(lib), hipmem.Stream(lib)
    dev = DeviceForward(solver, stream=build)
    dev.plan(other)
    left = dev.steps
    while left > 0:
        left = dev.run_steps(7, other)
    u, rec = dev.results(other)
    dev.release_plan()
    assert np.array_equal(u, u_ref) and np.array_equal(rec, rec_ref)


@pytest.mark.parametrize("name,env", [("3d_cd", {"SIMWAVE_HIP_TILE": "10"}),
                                      ("3d_cd", {"SIMWAVE_HIP_ZMARCH": "1"}),
                                      ("2d_cd", {}), ("2d_vd_f64", {}), ("3d_vd", {})])
def test_timed_steps_keep_results(name, env, monkeypatch):
    """Step timing (flag 1) launches the step kernels with runtime-set
    start and end events; the results are the untimed run's bits, the
    launch count is the step count and every launch took some time."""
    print("timed steps", name, env, flush=True)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    solver = sw.Solver(**cases.make_case(**CASES[name])())
    dev = DeviceForward(solver)
    dev.run(flags=0)
    u0, r0 = dev.results()
    for flags in (1, 1 | 4):
        dev.reset()
        dev.run(flags=flags)
        t = dev.timings()
        u, r = dev.results()
        assert np.array_equal(u, u0) and np.array_equal(r, r0), flags
        assert t[5] == dev.steps and t[4] > 0.0, (flags, t)
        assert t[4] / t[5] < 0.05, t   # seconds per launch of a tiny grid
