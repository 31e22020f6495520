"""Pins the float64 adjoint oracle (oracle/adjoint_ref.py) before it is
trusted as the checker of the HIP adjoint: its forward against the C
oracle in float64, its velocity gradient against central finite
differences of its own misfit, its wavelet adjoint by the dot-product
test <F s, r> = <s, F^T r>, and its cross-correlation image as the
derivative of the data under a scaling of every wavefield level."""
import os
import sys

import numpy as np
import pytest

import cases
import harness

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import adjoint_ref  # noqa: E402

CASES = cases.ADJOINT_CASES


def shot(name, **over):
    parts, _, kw = cases.adjoint_shot(name, **over)
    return parts, kw


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_forward_matches_c_oracle(name):
    parts, kw = shot(name)
    ref = adjoint_ref.Shot(kw)
    _, rec = ref.forward()
    _, rec_c = harness.run(lambda: parts, "oracle")
    assert rec.shape == rec_c.shape
    assert harness.rel_l2(rec, rec_c) < 1e-12


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_gradient_matches_finite_differences(name):
    parts, kw = shot(name)
    ref = adjoint_ref.Shot(kw)
    _, observed = ref.forward()
    # perturb the model so the misfit is not at its minimum
    rng = np.random.default_rng(1)
    vel = kw["velocity_model"] * (1 + 0.02 * rng.standard_normal(kw["velocity_model"].shape))
    base = adjoint_ref.Shot(adjoint_ref.with_velocity(kw, vel))
    _, g = base.misfit_gradient(observed)
    gmax = np.abs(g).max()
    assert gmax > 0
    # cells carrying signal (|g| >= 1% of the largest) and random cells
    strong = np.argwhere(np.abs(g) >= 0.01 * gmax)
    r = base.r
    inner = np.argwhere(np.ones([n - 2 * r for n in vel.shape], bool)) + r
    picks = list(strong[rng.choice(len(strong), min(4, len(strong)), replace=False)])
    picks += list(inner[rng.choice(len(inner), 3, replace=False)])
    for idx in map(tuple, picks):
        h = 1e-4 * vel[idx]
        vp, vm = vel.copy(), vel.copy()
        vp[idx] += h
        vm[idx] -= h
        jp, _ = adjoint_ref.Shot(adjoint_ref.with_velocity(kw, vp)).misfit_gradient(observed)
        jm, _ = adjoint_ref.Shot(adjoint_ref.with_velocity(kw, vm)).misfit_gradient(observed)
        fd = (jp - jm) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-5 * abs(g[idx]) + 1e-7 * gmax, (idx, fd, g[idx], gmax)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_wavelet_dot_product(name):
    parts, kw = shot(name, multi=True)
    ref = adjoint_ref.Shot(kw)
    rng = np.random.default_rng(2)
    s = rng.standard_normal(np.asarray(kw["wavelet"]).shape)
    u, d = ref.forward(wavelet=s)
    rr = rng.standard_normal(d.shape)
    _, gw = ref.adjoint(u, rr)
    lhs = float(np.sum(d * rr))
    rhs = float(np.sum(s.reshape(gw.shape) * gw))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_image_is_the_level_scaling_derivative(name):
    """sum_p m_p I_p = d/de <r, d(e)> where every new level is multiplied by
    (1 + e m) before its boundary phases: the image correlates each level
    with its complete adjoint."""
    parts, kw = shot(name)
    ref = adjoint_ref.Shot(kw)
    u, d = ref.forward()
    rng = np.random.default_rng(4)
    rr = rng.standard_normal(d.shape)
    img, _ = ref.adjoint(u, rr, image=True)
    m = rng.uniform(0.5, 1.5, size=ref.shape)
    eps = 1e-6
    jp = np.sum(rr * ref.forward(scale=1 + eps * m)[1])
    jm = np.sum(rr * ref.forward(scale=1 - eps * m)[1])
    fd = (jp - jm) / (2 * eps)
    lhs = float(np.sum(m * img))
    assert abs(fd - lhs) <= 1e-7 * abs(lhs), (fd, lhs)
