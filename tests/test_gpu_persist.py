"""Whole-run 2D launches (k_step2d_persist): all steps of a small stride-0
2D shot in one cooperative launch with grid barriers between steps, taken
by the device entry (DeviceForward.run).  Every case is compared bit for
bit with the oracle on the same inputs, and with the step-by-step launches
(SIMWAVE_HIP_PERSIST=0); the timings report which path ran."""
import numpy as np
import pytest

import cases
import harness
import simwave_amd as sw
from simwave_amd import workloads
from simwave_amd.device import DeviceForward

pytestmark = pytest.mark.gpu

F32, F64 = np.float32, np.float64

CASES = {
    "cd_so4": dict(dim=2, n=90, space_order=4, tf=0.2),
    "cd_so2_mixed_bc": dict(dim=2, shape=(70, 96), space_order=2, tf=0.2,
                            bc=("null_neumann", "null_dirichlet", "none", "null_neumann")),
    "cd_so20": dict(dim=2, n=80, space_order=20, tf=0.1),
    "vd_so8": dict(dim=2, n=64, space_order=8, density=True, tf=0.12),
    "vd_so16": dict(dim=2, n=60, space_order=16, density=True, tf=0.1),
    "cd_so4_f64": dict(dim=2, n=56, space_order=4, dtype=F64, tf=0.1),
    "cd_multi_src": dict(dim=2, n=72, space_order=8, multi=True,
                         src=[(200.0, 300.0), (350.0, 120.0), (210.0, 310.0)], tf=0.12),
    "cd_tiny": dict(dim=2, n=24, space_order=4, tf=0.05),
}


def _device(parts, persist, flags=0):
    import os
    old = os.environ.get("SIMWAVE_HIP_PERSIST")
    if persist is None:
        os.environ.pop("SIMWAVE_HIP_PERSIST", None)
    else:
        os.environ["SIMWAVE_HIP_PERSIST"] = persist
    try:
        dev = DeviceForward(sw.Solver(**parts))
        dev.run(flags=flags)
        t = dev.timings()
        u, rec = dev.results()
    finally:
        if old is None:
            os.environ.pop("SIMWAVE_HIP_PERSIST", None)
        else:
            os.environ["SIMWAVE_HIP_PERSIST"] = old
    return u, rec, t, dev.steps


@pytest.mark.parametrize("name", sorted(CASES))
def test_whole_run_matches_oracle(name):
    print("whole-run 2D", name, flush=True)
    setup = cases.make_case(**CASES[name])
    u_o, r_o = harness.run(setup, "oracle")
    u, r, t, _ = _device(setup(), "1")
    assert t[7] == 3.0, "the whole-run launch did not run: %s" % t
    assert u.shape == u_o.shape and r.shape == r_o.shape
    assert np.array_equal(u, u_o) and np.array_equal(r, r_o)
    u1, r1, t1, _ = _device(setup(), "0")
    assert t1[7] != 3.0
    assert np.array_equal(u1, u) and np.array_equal(r1, r)


def test_whole_run_timing():
    """Flag 1 on a whole run: one event pair around the launch, the step
    count as the launch count."""
    u0, r0, _, _ = _device(cases.make_case(**CASES["cd_so4"])(), "1")
    u, r, t, steps = _device(cases.make_case(**CASES["cd_so4"])(), "1", flags=1)
    assert t[7] == 3.0 and t[5] == steps and t[4] > 0.0, t
    assert np.array_equal(u, u0) and np.array_equal(r, r0)


def test_readme_experiment_whole_run():
    """The README example (513^2 so4, 328 steps) takes the whole-run launch
    by default and matches the oracle bit for bit."""
    parts = workloads.readme_2d(F32)
    u, r, t, _ = _device(parts, None, flags=0)
    assert t[7] == 3.0, t
    oracle = sw.Solver(compiler=harness.LibraryCompiler("oracle"), **workloads.readme_2d(F32))
    oracle._middleware.verbose = False
    u_o, r_o = oracle.forward()
    assert np.array_equal(u, u_o) and np.array_equal(r, r_o)


def test_failed_roll_call_falls_back(monkeypatch):
    """A roll call that cannot complete (forced from the host: the kernel
    waits for one block more than it launched, with a short bounded wait)
    leaves every slot untouched; the device entry then runs the steps
    launch by launch and the results are the oracle's."""
    monkeypatch.setenv("SIMWAVE_HIP_PERSIST_TEST_FALLBACK", "1")
    setup = cases.make_case(**CASES["cd_so4"])
    u_o, r_o = harness.run(setup, "oracle")
    u, r, t, _ = _device(setup(), "1")
    assert t[7] != 3.0, t
    assert np.array_equal(u, u_o) and np.array_equal(r, r_o)
