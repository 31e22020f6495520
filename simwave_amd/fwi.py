"""Full-waveform-inversion gradient over many shots (constant density).

The L2 misfit of a set of shots, J(v) = sum_s 1/2 ||d_s(v) - o_s||^2, and
its gradient with respect to the extended velocity model v.  Each shot is
one call of the library's adjoint entry in residual mode
(``simwave_hip_gradient_<variant>``, flag 2): the forward, the residual
d_s - o_s formed on the device, and the reverse sweep with checkpoints.
The shots share one device copy of v, so an inversion updates it in place
between iterations; the per-shot gradients are summed on the device.
"""
import numpy as np

from . import adjoint
from .device import DeviceForward
from .solver import Solver

__all__ = ["observed_data", "modelled_shots", "misfit_gradient"]


def observed_data(setups, device="cuda"):
    """Shot records of the experiments `setups` (Solver keyword
    dictionaries), one device tensor [timesteps, receivers] each."""
    out = []
    for parts in setups:
        dev = DeviceForward(Solver(**parts), device=device, buffers="torch")
        dev.run(flags=0)
        out.append(dev.tensors["shot_record"].clone())
        del dev
    return out


def modelled_shots(setups, device="cuda"):
    """Device-resident shots of the experiments `setups`, all reading the
    velocity tensor of the first (``shots[0].tensors["velocity_model"]``)."""
    shots = []
    vel = None
    for parts in setups:
        dev = DeviceForward(Solver(**parts), device=device, velocity=vel, buffers="torch")
        adjoint._check(dev.solver)
        vel = dev.tensors["velocity_model"]
        shots.append(dev)
    return shots


def misfit_gradient(shots, observed, checkpoint_every=20, stats=None):
    """(J, dJ/dv): the summed misfit (float) and gradient (float64 device
    tensor of the extended shape) of `shots` against `observed`.  With a
    `stats` dictionary the reverse update kernels are timed (HIP events)
    and their summed seconds and launches added to stats["update_s"] and
    stats["update_launches"]."""
    import torch
    total = None
    misfit = 0.0
    for dev, obs in zip(shots, observed):
        m, g = adjoint.shot_misfit_gradient(dev, obs, checkpoint_every,
                                            time_update=stats is not None)
        if stats is not None:
            t = dev.timings()
            stats["update_s"] = stats.get("update_s", 0.0) + t[4]
            stats["update_launches"] = stats.get("update_launches", 0) + int(t[5])
        misfit += m
        total = g if total is None else total.add_(g)
    torch.cuda.synchronize()
    return misfit, total


def relative_l2(a, b):
    """||a - b|| / ||b|| of two device or host arrays, in float64."""
    import torch
    a = torch.as_tensor(a).double()
    b = torch.as_tensor(b).double().to(a.device)
    nb = float(torch.linalg.vector_norm(b))
    return float(torch.linalg.vector_norm(a - b)) / nb if nb else float(
        torch.linalg.vector_norm(a))


def random_interior_cells(grad, count, seed=0, frac=0.01):
    """`count` random cells where |grad| >= frac * max |grad| (numpy index
    tuples), for finite-difference checks."""
    g = np.abs(np.asarray(grad))
    strong = np.argwhere(g >= frac * g.max())
    rng = np.random.default_rng(seed)
    picks = strong[rng.choice(len(strong), min(count, len(strong)), replace=False)]
    return [tuple(int(i) for i in p) for p in picks]
