"""Device-resident forward runs.

``DeviceForward`` keeps one experiment's wavefield slots, models and shot
record in HBM and calls the library's device entry
(``simwave_hip_forward_device_<variant>``, declared in include/simwave_hip.h)
on a HIP stream, so repeated shots pay no host<->device traffic.  Its plan
methods use the persistent-plan entries (``simwave_hip_plan_*``) to run a
shot in chunks of time steps on any stream.  Device memory, streams and
pinned staging come from PyTorch (plumbing only); every computation on the
device, the edge extension of the models included, is in the HIP library.
"""
import ctypes
import threading

import numpy as np

from .middleware import KEY_ORDER, Middleware

__all__ = ["DeviceForward", "FLAG_TIME_STEPS", "FLAG_GENERIC", "FLAG_OVERLAP_RECEIVERS",
           "FLAG_GRAPH"]

FLAG_TIME_STEPS = 1   # accumulate the step kernel's time with HIP events
FLAG_GENERIC = 2      # use the generic kernels instead of the fused one
FLAG_OVERLAP_RECEIVERS = 4   # receiver gather of large grids on a second stream
FLAG_GRAPH = 8               # plans of stride-0 shots: replay a captured 48-step hipGraph

_TORCH_DTYPE = {"float32": None, "float64": None}

# arguments that live in device memory for the device entry
_DEVICE_KEYS = ("u_full", "velocity_model", "density_model", "damping_mask",
                "shot_record")


def _variant(kwargs):
    vel = kwargs["velocity_model"]
    dens = "vd" if kwargs.get("density_model") is not None else "cd"
    prec = {"float32": "f32", "float64": "f64"}[str(vel.dtype)]
    return "%dd_%s_%s" % (vel.ndim, dens, prec)


_STAGE = threading.local()


def _staging(torch, nbytes):
    """This thread's pinned staging buffer (uint8) of at least `nbytes`,
    grown in powers of two: the D2H copies of results() run at pinned
    speed, and a long-running server holds one buffer per worker thread
    instead of one per result kept by its callers."""
    buf = getattr(_STAGE, "buf", None)
    if buf is None or buf.numel() < nbytes:
        _STAGE.buf = None
        size = 1 << max(20, (int(nbytes) - 1).bit_length())
        buf = torch.empty(size, dtype=torch.uint8, pin_memory=True)
        _STAGE.buf = buf
    return buf


class DeviceForward:
    """One experiment resident on a GPU.

    Parameters
    ----------
    solver : simwave_amd.Solver
        Supplies the frontend arrays (same inputs as Solver.forward).
    device : str or torch.device, optional
    """

    def __init__(self, solver, device="cuda", end_timestep=None, velocity=None):
        """velocity: optional device tensor of the extended shape used in
        place of the solver's extended velocity model."""
        import torch
        _TORCH_DTYPE.update(float32=torch.float32, float64=torch.float64)

        self._torch = torch
        self.solver = solver
        # the big arrays (slots, record, extended models, damping mask) are
        # built on the device; the host keeps placeholders of their shape
        kwargs = Middleware.normalize(**solver._arguments(light=True))
        if end_timestep is not None:
            kwargs["end_timestep"] = min(int(end_timestep), kwargs["end_timestep"])
        self.variant = _variant(kwargs)
        lib = solver._middleware.library(
            dimension=kwargs["velocity_model"].ndim,
            density=kwargs.get("density_model"),
            dtype=kwargs["velocity_model"].dtype)
        self._lib = lib
        self._fn = getattr(lib, "simwave_hip_forward_device_" + self.variant)
        self.device = torch.device(device)
        self.tensors = {}
        sm = solver.space_model
        for key in _DEVICE_KEYS:
            host = kwargs.get(key)
            if host is None:
                continue
            if key == "velocity_model" and velocity is not None:
                if tuple(velocity.shape) != host.shape or \
                        velocity.dtype != _TORCH_DTYPE[str(host.dtype)]:
                    raise ValueError("velocity tensor must have shape %s and the model dtype"
                                     % (host.shape,))
                self.tensors[key] = velocity.contiguous()
            elif key in ("u_full", "shot_record"):
                # the Solver's fresh slots and record are zeros
                self.tensors[key] = torch.zeros(host.shape, device=self.device,
                                                dtype=_TORCH_DTYPE[str(host.dtype)])
            elif key == "damping_mask":
                small, index = sm.damping_mask_factors()
                self.tensors[key] = self._gather(small, index)
            else:
                model = sm.velocity_model if key == "velocity_model" else sm.density_model
                self.tensors[key] = self._gather(model, sm.extension_index())
            if tuple(self.tensors[key].shape) != host.shape:
                raise RuntimeError("device %s has shape %s, expected %s"
                                   % (key, tuple(self.tensors[key].shape), host.shape))
        self.kwargs = kwargs
        argtypes, args = Middleware.ordered(kwargs)
        keys = [k for k in KEY_ORDER if kwargs.get(k) is not None]
        for pos, key in enumerate(keys):
            if key in self.tensors:
                argtypes[pos] = ctypes.c_void_p
                args[pos] = ctypes.c_void_p(self.tensors[key].data_ptr())
        self._fn.restype = ctypes.c_double
        self._fn.argtypes = argtypes + [ctypes.c_void_p, ctypes.c_uint]
        self._args = args
        self._plan_create = getattr(lib, "simwave_hip_plan_create_" + self.variant)
        self._plan_create.restype = ctypes.c_void_p
        self._plan_create.argtypes = argtypes + [ctypes.c_void_p, ctypes.c_uint]
        self._plan_run = getattr(lib, "simwave_hip_plan_run_" + self.variant)
        self._plan_run.restype = ctypes.c_longlong
        self._plan_run.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        self._plan_destroy = getattr(lib, "simwave_hip_plan_destroy_" + self.variant)
        self._plan_destroy.restype = ctypes.c_int
        self._plan_destroy.argtypes = [ctypes.c_void_p]
        self._plans_run = getattr(lib, "simwave_hip_plans_run_" + self.variant)
        self._plans_run.restype = ctypes.c_int
        self._plans_run.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                    ctypes.c_size_t, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_longlong)]
        self._plan_batchable = getattr(lib, "simwave_hip_plan_batchable_" + self.variant)
        self._plan_batchable.restype = ctypes.c_int
        self._plan_batchable.argtypes = [ctypes.c_void_p]
        self._plan = None
        self._argtypes = argtypes
        # the uploads and gathers above are queued on the current stream;
        # work on any other stream waits for them (device side)
        self._ready = torch.cuda.Event()
        self._ready.record(torch.cuda.current_stream(self.device))

    def _upload(self, array):
        """Host array -> device tensor (a plain copy; staging through
        freshly pinned memory made 16 concurrent shots slower)."""
        return self._torch.from_numpy(np.ascontiguousarray(array)).to(self.device)

    def _gather(self, small, index):
        """small[np.ix_(*index)] built on the device: one upload of `small`
        and of the index arrays, then the library's gather kernel
        (``simwave_hip_gather_<variant>``) on the current stream (copies
        only, so the values are the host's, bit for bit)."""
        torch = self._torch
        small = np.ascontiguousarray(small)
        src = self._upload(small)
        idx = [self._upload(np.asarray(i, dtype=np.int64)) for i in index]
        shape = tuple(len(i) for i in index)
        out = torch.empty(shape, dtype=src.dtype, device=self.device)
        fn = getattr(self._lib, "simwave_hip_gather_" + self.variant)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                       ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        pad = (1,) * (3 - len(shape))
        n = (ctypes.c_longlong * 3)(*(shape + pad))
        m = (ctypes.c_longlong * 3)(*(tuple(small.shape) + pad))
        ptr = [t.data_ptr() for t in idx] + [None] * (3 - len(idx))
        stream = torch.cuda.current_stream(self.device)
        rc = fn(out.data_ptr(), src.data_ptr(), len(shape), n, m, ptr[0], ptr[1], ptr[2],
                ctypes.c_void_p(stream.cuda_stream))
        if rc != 0:
            from .middleware import _last_error
            raise RuntimeError("device gather failed: {}".format(_last_error(self._lib)))
        return out

    @property
    def device_bytes(self):
        """HBM this shot holds while planned: slots, models, shot record and
        the library's per-point arrays (q, damping term, (dt^2/s)/den and
        the density copy)."""
        total = sum(t.numel() * t.element_size() for t in self.tensors.values())
        vel = self.tensors["velocity_model"]
        return total + 4 * vel.numel() * vel.element_size()

    @property
    def steps(self):
        return self.kwargs["end_timestep"] - self.kwargs["begin_timestep"] + 1

    def plan(self, stream=None, flags=0):
        """Prepares the shot for chunked runs (``run_steps``) on the device."""
        torch = self._torch
        if self._plan is not None:
            raise RuntimeError("plan already created")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        stream.wait_event(self._ready)
        h = self._plan_create(*self._args, ctypes.c_void_p(stream.cuda_stream),
                              ctypes.c_uint(flags))
        if not h:
            from .middleware import _last_error
            raise RuntimeError("plan creation failed: {}".format(_last_error(self._lib)))
        self._plan = h
        return self

    def run_steps(self, steps, stream=None):
        """Enqueues the next `steps` time steps on `stream` (no wait);
        returns the steps left."""
        torch = self._torch
        if self._plan is None:
            raise RuntimeError("no plan (call plan() first)")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        left = self._plan_run(ctypes.c_void_p(self._plan), ctypes.c_size_t(int(steps)),
                              ctypes.c_void_p(stream.cuda_stream))
        if left < 0:
            from .middleware import _last_error
            raise RuntimeError("plan run failed: {}".format(_last_error(self._lib)))
        return left

    @property
    def batchable(self):
        """Whether the planned shot can advance in batched launches with
        other shots of its library (``run_batched``): a stride-0 run on the
        fused z-marching kernels with no source in the halo."""
        return bool(self._plan) and self._plan_batchable(ctypes.c_void_p(self._plan)) == 1

    @staticmethod
    def run_batched(shots, steps, stream):
        """Enqueues up to `steps` further steps of every planned shot in
        `shots` (all of one library) on `stream`, batched where they can be
        (``simwave_hip_plans_run_<variant>``); returns the steps each has
        left."""
        if not shots:
            return []
        lib = shots[0]
        for s in shots:
            if s._plan is None:
                raise RuntimeError("no plan (call plan() first)")
            if s._lib is not lib._lib or s.variant != lib.variant:
                raise ValueError("batched shots must share one library")
        n = len(shots)
        plans = (ctypes.c_void_p * n)(*[s._plan for s in shots])
        left = (ctypes.c_longlong * n)()
        if lib._plans_run(plans, n, ctypes.c_size_t(int(steps)),
                          ctypes.c_void_p(stream.cuda_stream), left) != 0:
            from .middleware import _last_error
            raise RuntimeError("batched run failed: {}".format(_last_error(lib._lib)))
        return list(left)

    def release_plan(self):
        """Frees the plan's device memory once its enqueued work is done."""
        if self._plan is not None:
            h, self._plan = self._plan, None
            if self._plan_destroy(ctypes.c_void_p(h)) != 0:
                from .middleware import _last_error
                raise RuntimeError("plan destroy failed: {}".format(_last_error(self._lib)))

    def __del__(self):
        try:
            self.release_plan()
        except Exception:
            pass

    @property
    def updates_per_shot(self):
        """Interior grid-point updates of one shot (points x time steps)."""
        kw = self.kwargs
        r = kw["space_order"] // 2
        pts = 1
        for n in kw["velocity_model"].shape:
            pts *= n - 2 * r
        return pts * (kw["end_timestep"] - kw["begin_timestep"] + 1)

    def reset(self):
        """Zeroes the wavefield slots (a fresh shot)."""
        self.tensors["u_full"].zero_()

    def run(self, stream=None, flags=FLAG_OVERLAP_RECEIVERS):
        """Runs one shot on `stream` (default: torch's current stream)."""
        torch = self._torch
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        stream.wait_event(self._ready)
        secs = self._fn(*self._args, ctypes.c_void_p(stream.cuda_stream),
                        ctypes.c_uint(flags))
        if secs < 0:
            from .middleware import _last_error
            raise RuntimeError("device forward failed: {}".format(
                _last_error(self._lib)))
        return secs

    def timings(self):
        """Phase timings of the last call on this thread:
        [total s, prepare s, loop s, download s, step-kernel s, step launches,
         fused(1)/generic(0), step kernel: 3D tiled(1)/z-marching(2)/other(0)]."""
        out = (ctypes.c_double * 8)()
        fn = self._lib.simwave_hip_last_timings
        fn.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        fn.restype = ctypes.c_int
        fn(out, 8)
        return list(out)

    def results(self, stream=None):
        """(u_full, shot_record) as the Solver returns them.  The returned
        snapshots (one slot for saving_stride 0) are cut out of the slots on
        the device, so only they cross PCIe, into this thread's pinned
        staging buffer, on `stream` (default: torch's current stream), which
        is synchronised; the arrays returned are ordinary (pageable) copies,
        so results a caller keeps hold no page-locked memory."""
        torch = self._torch
        sm, tm = self.solver.space_model, self.solver.time_model
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        stream.wait_event(self._ready)
        end = self.kwargs["end_timestep"]
        if tm.saving_stride == 0:
            # the last level a full run returns is level T in slot T % 3; a
            # run cut at end_timestep E returns its level E (slot E % 3)
            slots = self.tensors["u_full"][end % 3:end % 3 + 1]
        elif end == tm.timesteps:
            slots = tm.remove_time_halo_region(self.tensors["u_full"])
        else:
            raise ValueError("results() of a saving_stride run needs every time step "
                             "(end_timestep %d < %d)" % (end, tm.timesteps))
        with torch.cuda.stream(stream):
            u = sm.remove_halo_region(slots)
            u = u.contiguous()
            rec = self.tensors["shot_record"]
            nu = u.numel() * u.element_size()
            nr = rec.numel() * rec.element_size()
            buf = _staging(torch, nu + nr)
            u_h = buf[:nu].view(u.dtype).view(u.shape)
            r_h = buf[nu:nu + nr].view(rec.dtype).view(rec.shape)
            u_h.copy_(u, non_blocking=True)
            r_h.copy_(rec, non_blocking=True)
        stream.synchronize()
        return u_h.numpy().copy(), r_h.numpy().copy()
