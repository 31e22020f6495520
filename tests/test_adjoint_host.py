"""Host-side checks of the adjoint wrapper (no device needed): the cases
the library does not cover are rejected before any device work."""
import numpy as np
import pytest

import cases
from simwave_amd import adjoint


def test_variable_density_is_rejected():
    parts = cases.make_case(dim=2, n=24, density=True, dtype=np.float64, tf=0.02)()
    import simwave_amd as sw
    solver = sw.Solver(**parts)
    with pytest.raises(NotImplementedError, match="constant density"):
        adjoint.gradient(solver, np.zeros((1, 1)))


def test_snapshot_stride_is_rejected():
    parts = cases.make_case(dim=2, n=24, stride=2, dtype=np.float64, tf=0.02)()
    import simwave_amd as sw
    solver = sw.Solver(**parts)
    with pytest.raises(ValueError, match="saving_stride 0"):
        adjoint.misfit_gradient(solver, np.zeros((1, 1)))
    with pytest.raises(ValueError, match="saving_stride 0"):
        adjoint.image(solver, np.zeros((1, 1)))
