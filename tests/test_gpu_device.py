"""DeviceForward builds the wavefield slots, shot record, extended models
and damping mask on the device; they must equal the arrays the frontend
builds on the host (Solver.forward_arguments()) bit for bit."""
import numpy as np
import pytest

import cases
import simwave_amd as sw
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
    for key, t in dev.tensors.items():
        ref = host[key]
        got = t.cpu().numpy()
        assert got.shape == ref.shape and got.dtype == ref.dtype, key
        assert np.array_equal(got, ref), key


@pytest.mark.parametrize("name", ["3d_cd", "2d_vd_f64"])
def test_shot_built_on_one_stream_runs_on_another(name):
    """The shot's uploads and gathers are queued on the stream it is built
    on; a plan, its chunks and the results fetched on another stream see
    them (the shot's ready event), and give the same arrays as one stream."""
    import torch
    print("cross-stream shot", name, flush=True)
    solver = sw.Solver(**cases.make_case(**CASES[name])())
    ref = DeviceForward(solver)
    ref.run()
    u_ref, rec_ref = ref.results()
    build, other = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(build):
        dev = DeviceForward(solver)
    dev.plan(other)
    left = dev.steps
    while left > 0:
        left = dev.run_steps(7, other)
    u, rec = dev.results(other)
    dev.release_plan()
    assert np.array_equal(u, u_ref) and np.array_equal(rec, rec_ref)
