"""Batched multi-shot runs (simwave_hip_plans_run_<variant>) on the GPU.

Shots of several sizes, orders and densities advance together in batched
launches (one step kernel and one receiver kernel per time step for all
shots of a radius), joining the running set at different times and
leaving it as they finish; each result is bit for bit the result of the
same shot run alone.  Shots that cannot be batched (tiled 3D kernel,
saving_stride > 0, the generic path) advance through the same call on
their own."""
import numpy as np
import pytest

import cases
import simwave_amd as sw
from simwave_amd import hipmem
from simwave_amd.device import DeviceForward

pytestmark = pytest.mark.gpu

F64 = np.float64

# (case parameters, batchable)
MIX = [
    (dict(dim=2, n=64, space_order=4, tf=0.2), True),
    (dict(dim=2, n=90, space_order=4, tf=0.15), True),
    (dict(dim=2, n=70, space_order=8, tf=0.12), True),
    (dict(dim=2, n=80, space_order=20, tf=0.1), True),
    (dict(dim=2, n=48, space_order=16, density=True, tf=0.1), True),
    (dict(dim=2, n=60, space_order=4, density=True, tf=0.12), True),
    (dict(dim=2, n=56, space_order=4, dtype=F64, tf=0.1), True),
    (dict(dim=2, n=40, stride=2, tf=0.08), False),
    (dict(dim=3, n=28, space_order=4, tf=0.08), True),
    (dict(dim=3, n=30, space_order=8, tf=0.07), True),
    (dict(dim=3, n=26, space_order=2, tf=0.09), True),
    (dict(dim=3, n=26, space_order=8, density=True, tf=0.06), True),
    (dict(dim=3, shape=(24, 28, 30), space_order=4, density=True, tf=0.06), False),
    (dict(dim=3, shape=(26, 30, 140), space_order=12, tf=0.05), False),
]


def _alone(params):
    dev = DeviceForward(sw.Solver(**cases.make_case(**params)()))
    dev.run(flags=0)
    return dev.results()


def test_batched_runs_equal_shots_alone(gpu):
    alone = []
    for params, _ in MIX:
        print("alone", params, flush=True)
        alone.append(_alone(params))
    stream = hipmem.Stream(hipmem.default_library())
    shots = []
    for params, want in MIX:
        dev = DeviceForward(sw.Solver(**cases.make_case(**params)()))
        dev.plan(stream, flags=0)
        assert dev.batchable == want, params
        shots.append(dev)
    # half the shots start two rounds late; rounds of 7 steps per call
    late = set(range(1, len(shots), 2))
    left = {i: s.steps for i, s in enumerate(shots)}
    rnd = 0
    while any(v > 0 for v in left.values()):
        running = [i for i in left if left[i] > 0 and (rnd >= 2 or i not in late)]
        by_variant = {}
        for i in running:
            by_variant.setdefault(shots[i].variant, []).append(i)
        for idx in by_variant.values():
            out = DeviceForward.run_batched([shots[i] for i in idx], 7, stream)
            for i, n in zip(idx, out):
                assert n == max(0, left[i] - 7)
                left[i] = n
        rnd += 1
    assert rnd > 3
    for (params, _), s, (u_a, r_a) in zip(MIX, shots, alone):
        u, r = s.results(stream)
        s.release_plan()
        assert u.shape == u_a.shape and r.shape == r_a.shape, params
        assert np.array_equal(u, u_a), params
        assert np.array_equal(r, r_a), params


def test_batched_run_rejects_mixed_libraries(gpu):
    stream = hipmem.Stream(hipmem.default_library())
    a = DeviceForward(sw.Solver(**cases.make_case(dim=2, n=40)()))
    b = DeviceForward(sw.Solver(**cases.make_case(dim=3, n=24, tf=0.05)()))
    a.plan(stream)
    b.plan(stream)
    with pytest.raises(ValueError):
        DeviceForward.run_batched([a, b], 5, stream)
    a.release_plan()
    b.release_plan()
