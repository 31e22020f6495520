"""Central finite-difference weights and the CFL time step.

Reference: simwave/kernel/frontend/fd.py.  The reference obtains the weights
from the third-party ``findiff`` package (``findiff.coefficients(deriv, acc)``,
center scheme, fd.py:50; ``findiff>=0.8.9`` in requirements.txt), which is not
installed here.  This module computes the same centred stencil weights exactly,
as rationals, by solving the Taylor (Vandermonde) moment conditions on the
offsets -p..p with p = (deriv + acc - 1) // 2, then rounds them to float64.
For deriv in {1, 2} and even acc this is the (acc+1)-point centred stencil.
"""
from fractions import Fraction
from functools import lru_cache

import numpy as np

__all__ = ["coefficients", "half_coefficients", "calculate_dt"]


@lru_cache(maxsize=None)
def _centred_weights(derivative_order, space_order):
    """Exact centred weights as a tuple of Fractions for offsets -p..p
    (cached: the exact solve is the frontend's costliest step)."""
    if derivative_order < 1:
        raise ValueError("derivative order must be >= 1")
    if space_order < 1:
        raise ValueError("accuracy order must be >= 1")
    p = (derivative_order + space_order - 1) // 2
    offsets = list(range(-p, p + 1))
    npts = len(offsets)
    # moment m: sum_k w_k * o_k**m = m! * delta(m, deriv)
    rows = [[Fraction(o) ** m for o in offsets] for m in range(npts)]
    rhs = [Fraction(0)] * npts
    fact = 1
    for i in range(1, derivative_order + 1):
        fact *= i
    rhs[derivative_order] = Fraction(fact)
    # Gauss-Jordan elimination over the rationals
    a = [row[:] + [rhs[i]] for i, row in enumerate(rows)]
    for col in range(npts):
        piv = next(r for r in range(col, npts) if a[r][col] != 0)
        a[col], a[piv] = a[piv], a[col]
        inv = 1 / a[col][col]
        a[col] = [v * inv for v in a[col]]
        for r in range(npts):
            if r != col and a[r][col] != 0:
                f = a[r][col]
                a[r] = [vr - f * vc for vr, vc in zip(a[r], a[col])]
    return tuple(a[i][npts] for i in range(npts))


def coefficients(derivative_order, space_order):
    """Full centred FD weights (float64), offsets -p..p.

    Reference: fd.py:31-52.
    """
    return np.array([float(w) for w in
                     _centred_weights(derivative_order, space_order)],
                    dtype=np.float64)


def half_coefficients(derivative_order, space_order):
    """Centre and right-side weights [c0, c1, ..., cp] (float64).

    Reference: fd.py:5-28.
    """
    coeffs = coefficients(derivative_order, space_order)
    return coeffs[len(coeffs) // 2:]


def calculate_dt(dimension, space_order, grid_spacing, velocity_model):
    """CFL time step sqrt(4 / (dim * sum|c2|)) * min(h) / max(v).

    Reference: fd.py:55-95 (Lines & Slawinski 1999 acoustic bound).
    """
    weight_sum = dimension * np.sum(np.abs(coefficients(2, space_order)))
    factor = np.sqrt(4 / weight_sum)
    return factor * np.min(grid_spacing) / np.max(velocity_model)
