"""Velocity gradient of the shot record by the adjoint-state method.

``gradient`` runs the library's adjoint entry
(``simwave_hip_gradient_<variant>``, include/simwave_hip.h): the forward of
a constant-density shot, then the transpose of every time step, backwards.
When every time level fits in HBM the forward keeps them all; otherwise it
saves checkpoints every ``checkpoint_every`` steps and each segment is
recomputed from its two saved levels.  Given the adjoint of
the shot record (for the L2 misfit 1/2 ||d - d_obs||^2 that is the
residual d - d_obs) it returns the gradient with respect to the extended
velocity model and with respect to the wavelet samples.

``ShotRecord`` wraps it as a ``torch.autograd.Function``: forward returns
the shot record of an extended velocity tensor, backward the
vector-Jacobian product with respect to that tensor.  PyTorch is used
only for device memory and the autograd plumbing.
"""
import ctypes
import functools

import numpy as np

from .device import DeviceForward
from .middleware import _last_error

__all__ = ["gradient", "misfit_gradient", "shot_misfit_gradient", "image", "ShotRecord",
           "checkpoint_interval"]

IMAGE = 1     # adjoint entry flag: cross-correlation image instead of gradient
RESIDUAL = 2  # adjoint entry flag: rec_adjoint is observed data, the source the residual
OVERLAP_RECEIVERS = 4   # forward receiver gather on a second stream (large grids)
TIME_UPDATE = 8         # adjoint entry flag: HIP-event timing of every reverse update kernel


def _check(solver):
    if solver.space_model.density_model is not None:
        raise NotImplementedError("the adjoint covers constant density only")
    if solver.time_model.saving_stride != 0:
        raise ValueError("the adjoint needs saving_stride 0")


def checkpoint_interval(dev, checkpoint_every=None):
    """The interval the entry gets: `checkpoint_every` if given; otherwise
    every time level kept in HBM (no recomputation) when they fit in 80% of
    the free device memory, else checkpoints every 20 steps."""
    if checkpoint_every is not None:
        return int(checkpoint_every)
    import torch
    T = dev.kwargs["end_timestep"]
    vel = dev.tensors["velocity_model"]
    need = (T + 2) * vel.numel() * vel.element_size()
    # memory an earlier run freed into the library's pool counts as used in
    # the query below until the pool hands it back
    with torch.cuda.device(dev.device):
        if dev._lib.simwave_hip_trim_pool() != 0:
            raise RuntimeError("pool trim failed: {}".format(_last_error(dev._lib)))
    free = torch.cuda.mem_get_info(dev.device)[0]
    return T if need <= 0.8 * free else 20


def _run(dev, rec_adjoint, checkpoint_every, want_wavelet, stream=None, image=False,
         overlap_receivers=True, residual=False, time_update=False):
    import torch
    kw = dev.kwargs
    checkpoint_every = checkpoint_interval(dev, checkpoint_every)
    lib = dev._lib
    fn = getattr(lib, "simwave_hip_gradient_" + dev.variant)
    fn.restype = ctypes.c_double
    fn.argtypes = list(dev._argtypes) + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint]
    vel = dev.tensors["velocity_model"]
    grad = torch.zeros(vel.shape, dtype=torch.float64, device=dev.device)
    rec = torch.as_tensor(rec_adjoint, device=dev.device).to(vel.dtype).contiguous()
    if tuple(rec.shape) != tuple(dev.tensors["shot_record"].shape):
        raise ValueError("shot-record adjoint must have shape %s"
                         % (tuple(dev.tensors["shot_record"].shape),))
    gw = None
    if want_wavelet:
        gw = torch.zeros((kw["end_timestep"], kw["num_sources"]), dtype=vel.dtype,
                         device=dev.device)
    if stream is None:
        stream = torch.cuda.current_stream(dev.device)
    # the zeroed outputs above are queued on the current stream, the shot's
    # uploads on the stream it was built on
    stream.wait_stream(torch.cuda.current_stream(dev.device))
    secs = fn(*dev._args, ctypes.c_void_p(rec.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
              ctypes.c_void_p(gw.data_ptr() if gw is not None else 0),
              ctypes.c_size_t(int(checkpoint_every)), ctypes.c_void_p(stream.cuda_stream),
              ctypes.c_uint((IMAGE if image else 0) | (RESIDUAL if residual else 0)
                            | (OVERLAP_RECEIVERS if overlap_receivers else 0)
                            | (TIME_UPDATE if time_update else 0)))
    if secs < 0:
        raise RuntimeError("adjoint failed: {}".format(_last_error(lib)))
    return grad, gw


def gradient(solver, rec_adjoint, checkpoint_every=None, wavelet_gradient=False,
             velocity=None, device="cuda"):
    """Shot record and adjoint gradients of one shot.

    Parameters
    ----------
    solver : Solver
        Constant density, saving_stride 0.
    rec_adjoint : array [timesteps, receivers]
        Adjoint of the shot record (the residual for the L2 misfit).
    checkpoint_every : int, optional
        Forward levels saved every that many steps and recomputed per
        segment; >= timesteps keeps every level.  Default: every level when
        they fit in HBM (see checkpoint_interval), else 20.
    velocity : torch tensor, optional
        Extended velocity on the device (default: the solver's model).

    Returns
    -------
    dict with ``shot_record`` (numpy), ``grad_velocity`` (float64 numpy,
    extended shape) and, with wavelet_gradient, ``grad_wavelet``
    ([timesteps, sources]).
    """
    _check(solver)
    dev = DeviceForward(solver, device=device, velocity=velocity, buffers="torch")
    grad, gw = _run(dev, rec_adjoint, checkpoint_every, wavelet_gradient)
    out = {"shot_record": dev.tensors["shot_record"].cpu().numpy(),
           "grad_velocity": grad.cpu().numpy()}
    if wavelet_gradient:
        out["grad_wavelet"] = gw.cpu().numpy()
    return out


def image(solver, rec_adjoint, checkpoint_every=None, velocity=None, device="cuda"):
    """Zero-lag cross-correlation image sum_k u^k lambda^k of the forward
    wavefield u and the adjoint wavefield lambda driven by `rec_adjoint`
    (reverse-time migration of the data `rec_adjoint`), float64 over the
    extended grid, zero outside the interior."""
    _check(solver)
    dev = DeviceForward(solver, device=device, velocity=velocity, buffers="torch")
    img, _ = _run(dev, rec_adjoint, checkpoint_every, False, image=True)
    return img.cpu().numpy()


def misfit_gradient(solver, observed, checkpoint_every=None, velocity=None, device="cuda"):
    """1/2 ||d - observed||^2 and its gradient with respect to the extended
    velocity (float64, extended shape)."""
    _check(solver)
    dev = DeviceForward(solver, device=device, velocity=velocity, buffers="torch")
    misfit, grad = shot_misfit_gradient(dev, observed, checkpoint_every)
    return misfit, grad.cpu().numpy()


def shot_misfit_gradient(dev, observed, checkpoint_every=None, stream=None,
                         time_update=False):
    """One call of the adjoint entry in residual mode on a device-resident
    shot: the forward, the residual d - observed on the device, the reverse
    sweep.  Returns (misfit 1/2 ||d - observed||^2 as a float, gradient
    with respect to the extended velocity as a float64 device tensor).
    With time_update the reverse update kernels are timed with HIP events
    (``dev.timings()[4:6]``: summed seconds, launches)."""
    import torch
    rec = dev.tensors["shot_record"]
    obs = torch.as_tensor(observed, device=dev.device).to(rec.dtype).contiguous()
    dev.reset()
    grad, _ = _run(dev, obs, checkpoint_every, False, stream=stream, residual=True,
                   time_update=time_update)
    res = rec.double() - obs.double()
    return 0.5 * float((res * res).sum()), grad


@functools.lru_cache(maxsize=None)
def _autograd():
    """The autograd Function class (built once, on first use, so that
    importing this module does not import torch)."""
    import torch

    class ShotRecord(torch.autograd.Function):
        """d = ShotRecord.apply(velocity_ext, solver, checkpoint_every):
        the shot record of `solver` with the extended velocity tensor
        `velocity_ext` (device, model dtype); differentiable in it."""

        @staticmethod
        def forward(ctx, velocity_ext, solver, checkpoint_every=None):
            _check(solver)
            vel = velocity_ext.detach().contiguous()
            dev = DeviceForward(solver, device=vel.device, velocity=vel, buffers="torch")
            dev.run()
            ctx.solver = solver
            ctx.checkpoint_every = checkpoint_every
            ctx.save_for_backward(vel)
            return dev.tensors["shot_record"].clone()

        @staticmethod
        def backward(ctx, grad_rec):
            (vel,) = ctx.saved_tensors
            dev = DeviceForward(ctx.solver, device=vel.device, velocity=vel, buffers="torch")
            grad, _ = _run(dev, grad_rec.detach(), ctx.checkpoint_every, False)
            return grad.to(vel.dtype), None, None

    return ShotRecord


class _Lazy:
    def __getattr__(self, name):
        return getattr(_autograd(), name)

    def apply(self, *args):
        return _autograd().apply(*args)


ShotRecord = _Lazy()
