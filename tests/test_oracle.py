"""Pins the oracle (oracle/sw_oracle.c, the CPU restatement of the
reference's forward) before it is trusted as the GPU parity checker.

* the reference's own golden 2D wavefields (tests/golden/*.npz, copied from
  the reference's tests/reference_solution) under the reference's own
  criterion, atol 1e-5 (tests/test_solution.py:136-143 of the reference);
* bit identity with the reference C compiled from its sources: recorded
  digests (tests/golden/ref_outputs.json) and, where oracle/_ref is built, a
  live comparison on further cases;
* the invariants the reference's tests assert: saving_stride does not
  change the last snapshot (test_u_saving.py), OpenMP equals sequential
  (test_parallel_solution.py), unit-density variable density equals
  constant density (test_solution.py), and the analytic 2D Hankel response
  (test_convergence.py:52-87);
* a planted mistake in the oracle's inputs (a dropped stencil term) fails
  the golden check, so the check has teeth.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import cases
import harness

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

pytestmark = pytest.mark.skipif(not harness.have("oracle", omp=True),
                                reason="oracle not built (make -C oracle)")


def _ref_outputs():
    with open(os.path.join(GOLDEN, "ref_outputs.json")) as f:
        return json.load(f)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _golden_u(so):
    with np.load(os.path.join(GOLDEN, "wavefield-2d-so-%d.npz" % so),
                 allow_pickle=False) as z:
        return z["u"]


@pytest.mark.parametrize("so, density", [(2, False), (8, False), (2, True)])
def test_reference_golden_wavefields(so, density):
    u, _ = harness.run(cases.golden_2d(so, density=density), "oracle", omp=True)
    ref = _golden_u(so)
    assert u.shape == ref.shape
    err = harness.rel_l2(u, ref)
    print("so%d density=%s: max |diff| %.3g, rel L2 %.3g"
          % (so, density, np.abs(u - ref).max(), err))
    assert np.allclose(u, ref, atol=1e-05)
    assert err < 1e-5


def test_planted_mistake_fails_golden():
    """Dropping the outermost stencil term (a plausible indexing slip)
    moves the so-8 wavefield far outside the golden tolerance."""
    import simwave_amd as sw

    setup = cases.golden_2d(8)
    parts = setup()
    solver = sw.Solver(compiler=harness.LibraryCompiler("oracle", True), **parts)
    kw = solver.forward_arguments()
    c2 = kw["second_order_fd_coefficients"].copy()
    c2[-1] = 0
    kw["second_order_fd_coefficients"] = c2
    u, _ = solver._middleware.exec("forward", **kw)
    u = parts["time_model"].remove_time_halo_region(u)
    u = parts["space_model"].remove_halo_region(u)
    assert not np.allclose(u, _golden_u(8), atol=1e-05)


REF = _ref_outputs()


@pytest.mark.parametrize("name", sorted(REF))
def test_bit_identical_to_reference_c(name):
    """Digests recorded from the reference's C forward built from source."""
    import sys
    sys.path.insert(0, GOLDEN)
    import make_fixtures
    u, rec = harness.run(make_fixtures.setup_of(name), "oracle", omp=True)
    want = REF[name]
    assert list(u.shape) == want["u"]["shape"]
    assert list(rec.shape) == want["rec"]["shape"]
    if "rec_values" in want:
        np.testing.assert_array_equal(rec.astype(np.float64).ravel(),
                                      np.array(want["rec_values"]))
    assert _sha(u) == want["u"]["sha256"], "u differs from the reference C"
    assert _sha(rec) == want["rec"]["sha256"], "shot record differs"


LIVE = {
    "2d_vd_so16": dict(dim=2, space_order=16, density=True, n=40),
    "2d_bc_neumann": dict(dim=2, bc=("null_neumann",) * 4, damping=0.0, tf=0.3),
    "2d_stride4": dict(dim=2, stride=4, tf=0.08),
    "2d_src_halo": dict(dim=2, damping=0.0, src=[(0.0, 200.0), (5.0, 470.0)],
                        multi=True),
    "3d_vd_so8": dict(dim=3, n=28, space_order=8, density=True, tf=0.08),
    "3d_cd_so20": dict(dim=3, n=28, space_order=20, tf=0.08),
    "3d_stride3": dict(dim=3, n=24, stride=3, tf=0.05),
    "3d_vd_f64_so4": dict(dim=3, n=26, space_order=4, density=True,
                          dtype=np.float64, tf=0.08),
}


@pytest.mark.skipif(not harness.have("ref", omp=False),
                    reason="reference C not built here (make -C oracle ref)")
@pytest.mark.parametrize("name", sorted(LIVE))
def test_bit_identical_to_reference_c_live(name):
    setup = cases.make_case(**LIVE[name])
    u_o, r_o = harness.run(setup, "oracle")
    u_r, r_r = harness.run(setup, "ref")
    np.testing.assert_array_equal(u_o, u_r)
    np.testing.assert_array_equal(r_o, r_r)


@pytest.mark.parametrize("density", [False, True])
@pytest.mark.parametrize("stride", [1, 2, 5])
def test_saving_stride_invariance_2d(stride, density):
    base, _ = harness.run(cases.saving_2d(0, density), "oracle", omp=True)
    u, _ = harness.run(cases.saving_2d(stride, density), "oracle", omp=True)
    np.testing.assert_array_equal(base[-1], u[-1])


@pytest.mark.parametrize("density", [False, True])
@pytest.mark.parametrize("stride", [1, 2, 3, 4])
def test_saving_stride_invariance_3d(stride, density):
    base, _ = harness.run(cases.saving_3d(0, density), "oracle", omp=True)
    u, _ = harness.run(cases.saving_3d(stride, density), "oracle", omp=True)
    np.testing.assert_array_equal(base[-1], u[-1])


@pytest.mark.parametrize("params", [dict(dim=2, space_order=8, density=True),
                                    dict(dim=3, n=26, space_order=4, tf=0.08)])
def test_openmp_matches_sequential(params):
    setup = cases.make_case(**params)
    u_s, r_s = harness.run(setup, "oracle", omp=False)
    u_p, r_p = harness.run(setup, "oracle", omp=True)
    np.testing.assert_array_equal(u_s, u_p)
    np.testing.assert_array_equal(r_s, r_p)


def test_hankel_accuracy():
    """Receiver trace vs the analytic Green's function response, the
    reference's bound 1e-4 (tests/test_convergence.py:75-87)."""
    setup = cases.hankel_accuracy_case()
    _, rec = harness.run(setup, "oracle", omp=True)
    num = rec.ravel() / 0.5 ** 2
    exact = cases.hankel_reference(setup)
    err = np.linalg.norm(num - exact) / np.sqrt(num.size)
    print("hankel error %.3g" % err)
    assert err < 1e-4
