"""The shot server on the GPU: a mix of shots of different dimension, order,
density, precision and snapshot stride, admitted against a budget that
forces queueing and run concurrently, returns for every shot exactly what
Solver.forward() returns for it alone (bit for bit)."""
import numpy as np
import pytest

import cases
import harness
import simwave_amd as sw
from simwave_amd.serving import ShotServer, shot_bytes

pytestmark = pytest.mark.gpu

MIX = [
    dict(dim=2, n=64, space_order=4),
    dict(dim=3, n=28, space_order=8, tf=0.08),
    dict(dim=2, n=48, space_order=16, density=True),
    dict(dim=3, n=24, space_order=2, density=True, tf=0.06),
    dict(dim=2, n=80, space_order=20, tf=0.1),
    dict(dim=3, shape=(26, 30, 270), space_order=12, tf=0.05),
    dict(dim=2, n=56, space_order=8, dtype=np.float64),
    dict(dim=2, n=40, stride=3, tf=0.08),
    dict(dim=3, n=26, space_order=6, stride=2, tf=0.05),
]


def test_results_identical_alone_and_in_the_mix(gpu):
    alone = []
    for p in MIX:
        print("alone", p, flush=True)
        alone.append(harness.run(cases.make_case(**p), "hip"))
    solvers = [sw.Solver(**cases.make_case(**p)()) for p in MIX]
    probe = shot_bytes(solvers[1])
    budget = 3.5 * max(shot_bytes(s) for s in solvers)
    print("budget %.3g MB (3D so8 shot %.3g MB)" % (budget / 1e6, probe / 1e6), flush=True)
    with ShotServer(hbm_budget=budget, max_active=4, chunk_steps=7) as srv:
        futs = [srv.submit(s) for s in solvers]
        results = [f.result(timeout=300) for f in futs]
    for p, (u_a, r_a), (u, r) in zip(MIX, alone, results):
        assert u.shape == u_a.shape and r.shape == r_a.shape, p
        assert np.array_equal(u, u_a), p
        assert np.array_equal(r, r_a), p
    s = srv.stats.summary()
    print("serving summary", s, flush=True)
    assert s["requests"] == len(MIX)
    assert s["max_active"] > 1


def test_server_result_matches_oracle(gpu):
    p = dict(dim=3, n=26, space_order=8, density=True, tf=0.06)
    setup = cases.make_case(**p)
    with ShotServer(hbm_budget=8e9, max_active=2) as srv:
        u, r = srv.submit(sw.Solver(**setup())).result(timeout=300)
    u_o, r_o = harness.run(setup, "oracle")
    assert harness.rel_l2(u, u_o) <= 1e-5 and harness.rel_l2(r, r_o) <= 1e-5
    assert np.array_equal(u, u_o) and np.array_equal(r, r_o)
