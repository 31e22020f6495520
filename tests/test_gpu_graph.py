"""Persistent plans with graph replay (flag 8): a stride-0 shot's 48-step
chunks replay one captured hipGraph, the step number coming from a device
counter.  Every case must equal the direct launches bit for bit, for chunk
sizes that align with the graph, straddle it, and do not fit it."""
import numpy as np
import pytest
import torch

import cases
import simwave_amd as sw
from simwave_amd.device import DeviceForward, FLAG_GRAPH, FLAG_GENERIC

pytestmark = pytest.mark.gpu

CASES = {
    "2d_cd": dict(dim=2, n=60, space_order=4, tf=0.3),
    "2d_vd_multi": dict(dim=2, n=50, space_order=8, density=True, multi=True,
                        src=[(200.0, 210.0), (230.0, 180.0)], tf=0.3),
    "3d_cd_tiled": dict(dim=3, n=30, space_order=8, tf=0.2),
    "3d_vd": dict(dim=3, n=26, space_order=4, density=True, tf=0.2),
    "2d_src_halo": dict(dim=2, damping=0.0, src=[(0.0, 200.0), (5.0, 470.0)], multi=True,
                        tf=0.3),
    "3d_f64": dict(dim=3, n=24, space_order=4, dtype=np.float64, tf=0.2),
}


def run(solver, chunk, flags):
    dev = DeviceForward(solver)
    stream = torch.cuda.current_stream()
    dev.plan(stream, flags=flags)
    left = dev.steps
    while left > 0:
        left = dev.run_steps(chunk, stream)
    out = dev.results(stream)
    dev.release_plan()
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replay_equals_direct_launches(name):
    print("graph replay", name, flush=True)
    solver = sw.Solver(**cases.make_case(**CASES[name])())
    assert solver.time_model.timesteps > 100
    u0, r0 = run(solver, 10 ** 6, 0)
    for chunk in (48, 96, 50, 7, 10 ** 6):
        u, r = run(solver, chunk, FLAG_GRAPH)
        assert np.array_equal(u, u0) and np.array_equal(r, r0), (name, chunk)


def test_graph_replay_generic_path():
    solver = sw.Solver(**cases.make_case(dim=2, n=40, tf=0.3)())
    u0, r0 = run(solver, 10 ** 6, FLAG_GENERIC)
    u, r = run(solver, 48, FLAG_GENERIC | FLAG_GRAPH)
    assert np.array_equal(u, u0) and np.array_equal(r, r0)


def test_graph_flag_ignored_with_snapshots():
    """stride > 0 slots do not repeat every 3 steps: direct launches."""
    solver = sw.Solver(**cases.make_case(dim=2, n=40, stride=3, tf=0.2)())
    u0, r0 = run(solver, 10 ** 6, 0)
    u, r = run(solver, 48, FLAG_GRAPH)
    assert np.array_equal(u, u0) and np.array_equal(r, r0)
