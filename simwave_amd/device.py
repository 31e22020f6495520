"""Device-resident forward runs.

``DeviceForward`` keeps one experiment's wavefield slots, models and shot
record in HBM and calls the library's device entry
(``simwave_hip_forward_device_<variant>``, declared in include/simwave_hip.h)
on a HIP stream, so repeated shots pay no host<->device traffic.  Its plan
methods use the persistent-plan entries (``simwave_hip_plan_*``) to run a
shot in chunks of time steps on any stream.  Device memory, pinned staging
and streams come from the same library (``simwave_amd.hipmem``), and every
computation on the device, the edge extension of the models included, is
in it.  The gradient side (``simwave_amd.adjoint``, ``simwave_amd.fwi``),
which hands its arrays to torch.autograd, asks for torch tensors instead
(``buffers="torch"``).
"""
import ctypes

import numpy as np

from . import hipmem
from .middleware import KEY_ORDER, Middleware

__all__ = ["DeviceForward", "FLAG_TIME_STEPS", "FLAG_GENERIC", "FLAG_OVERLAP_RECEIVERS",
           "FLAG_GRAPH"]

FLAG_TIME_STEPS = 1   # accumulate the step kernel's time with HIP events
FLAG_GENERIC = 2      # use the generic kernels instead of the fused one
FLAG_OVERLAP_RECEIVERS = 4   # receiver gather of large grids on a second stream
FLAG_GRAPH = 8               # plans of stride-0 shots: replay a captured 48-step hipGraph

# arguments that live in device memory for the device entry
_DEVICE_KEYS = ("u_full", "velocity_model", "density_model", "damping_mask",
                "shot_record")


def _variant(kwargs):
    vel = kwargs["velocity_model"]
    dens = "vd" if kwargs.get("density_model") is not None else "cd"
    prec = {"float32": "f32", "float64": "f64"}[str(vel.dtype)]
    return "%dd_%s_%s" % (vel.ndim, dens, prec)


def _is_torch_tensor(x):
    return type(x).__module__.split(".")[0] == "torch"


class DeviceForward:
    """One experiment resident on a GPU.

    Parameters
    ----------
    solver : simwave_amd.Solver
        Supplies the frontend arrays (same inputs as Solver.forward).
    device : str or torch.device, optional
        The device of torch buffers; library buffers live on the current
        HIP device.
    end_timestep : int, optional
        Cut the run at this time step.
    velocity : device array, optional
        Used in place of the solver's extended velocity model (its shape
        and dtype); a torch tensor selects torch buffers.
    buffers : {"library", "torch"}, optional
        Who allocates the device arrays: the HIP library (default) or torch
        (default when `velocity` is a torch tensor).  With torch buffers the
        default stream of every method is torch's current stream, with
        library buffers a stream of the shot's own.
    stream : optional
        Stream of the construction's copies and gathers (default: the
        default stream above); it is synchronised before the constructor
        returns, so later work on any stream sees the arrays.  Library
        buffers are stream-ordered on it: they return to the library's pool
        after the work enqueued there when the shot is collected, so work
        enqueued on other streams must be complete by then (``results``
        and ``synchronize`` wait for their stream).
    """

    def __init__(self, solver, device="cuda", end_timestep=None, velocity=None,
                 buffers=None, stream=None):
        if buffers is None:
            buffers = "torch" if velocity is not None and _is_torch_tensor(velocity) \
                else "library"
        if buffers not in ("library", "torch"):
            raise ValueError("buffers must be 'library' or 'torch'")
        self.buffers = buffers
        self._torch = None
        self.device = device
        if buffers == "torch":
            import torch
            self._torch = torch
            self.device = torch.device(device)
        self._own_stream = None
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
        st = self._stream(stream)
        keep = []           # host arrays of copies still in flight
        self.tensors = {}
        sm = solver.space_model
        for key in _DEVICE_KEYS:
            host = kwargs.get(key)
            if host is None:
                continue
            if key == "velocity_model" and velocity is not None:
                if tuple(velocity.shape) != host.shape or \
                        np.dtype(str(velocity.dtype).replace("torch.", "")) != host.dtype:
                    raise ValueError("velocity array must have shape %s and the model dtype"
                                     % (host.shape,))
                if _is_torch_tensor(velocity):
                    velocity = velocity.contiguous()
                self.tensors[key] = velocity
            elif key in ("u_full", "shot_record"):
                # the Solver's fresh slots and record are zeros
                self.tensors[key] = self._zeros(host.shape, host.dtype, st)
            elif key == "damping_mask":
                small, index = sm.damping_mask_factors()
                self.tensors[key] = self._gather(small, index, st, keep)
            else:
                model = sm.velocity_model if key == "velocity_model" else sm.density_model
                self.tensors[key] = self._gather(model, sm.extension_index(), st, keep)
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
        # the copies and gathers above are done before any other stream
        # uses the arrays (a host wait, no cross-stream device waits)
        self._sync(st)
        del keep

    def _stream(self, stream):
        """`stream` as given, or this shot's default stream (module docs)."""
        if stream is not None:
            return stream
        if self._torch is not None:
            return self._torch.cuda.current_stream(self.device)
        if self._own_stream is None:
            self._own_stream = hipmem.Stream(self._lib)
        return self._own_stream

    def _sync(self, stream):
        hipmem._call(self._lib, "simwave_hip_stream_synchronize",
                     ctypes.c_void_p(hipmem.stream_handle(stream)))

    def _empty(self, shape, dtype, stream):
        if self._torch is not None:
            torch = self._torch
            tdt = {"float32": torch.float32, "float64": torch.float64,
                   "int64": torch.int64}[np.dtype(dtype).name]
            return torch.empty(tuple(shape), dtype=tdt, device=self.device)
        return hipmem.DeviceArray(self._lib, shape, dtype, stream)

    def _zeros(self, shape, dtype, stream):
        out = self._empty(shape, dtype, stream)
        nbytes = out.numel() * out.element_size()
        hipmem._call(self._lib, "simwave_hip_zero", ctypes.c_void_p(out.data_ptr()),
                     ctypes.c_ulonglong(nbytes),
                     ctypes.c_void_p(hipmem.stream_handle(stream)))
        return out

    def _upload(self, array, stream, keep):
        """Host array -> device array, a stream-ordered copy (the host array
        is kept alive in `keep` until the constructor's final wait)."""
        a = np.ascontiguousarray(array)
        keep.append(a)
        out = self._empty(a.shape, a.dtype, stream)
        hipmem._call(self._lib, "simwave_hip_copy", ctypes.c_void_p(out.data_ptr()),
                     ctypes.c_void_p(a.ctypes.data), ctypes.c_ulonglong(a.nbytes),
                     ctypes.c_void_p(hipmem.stream_handle(stream)))
        return out

    def _gather(self, small, index, stream, keep):
        """small[np.ix_(*index)] built on the device: one upload of `small`
        and of the index arrays, then the library's gather kernel
        (``simwave_hip_gather_<variant>``) on `stream` (copies only, so the
        values are the host's, bit for bit)."""
        small = np.ascontiguousarray(small)
        src = self._upload(small, stream, keep)
        idx = [self._upload(np.asarray(i, dtype=np.int64), stream, keep) for i in index]
        keep.extend([src] + idx)
        shape = tuple(len(i) for i in index)
        out = self._empty(shape, small.dtype, stream)
        fn = getattr(self._lib, "simwave_hip_gather_" + self.variant)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                       ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        pad = (1,) * (3 - len(shape))
        n = (ctypes.c_longlong * 3)(*(shape + pad))
        m = (ctypes.c_longlong * 3)(*(tuple(small.shape) + pad))
        ptr = [t.data_ptr() for t in idx] + [None] * (3 - len(idx))
        rc = fn(out.data_ptr(), src.data_ptr(), len(shape), n, m, ptr[0], ptr[1], ptr[2],
                ctypes.c_void_p(hipmem.stream_handle(stream)))
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
        if self._plan is not None:
            raise RuntimeError("plan already created")
        stream = self._stream(stream)
        h = self._plan_create(*self._args, ctypes.c_void_p(hipmem.stream_handle(stream)),
                              ctypes.c_uint(flags))
        if not h:
            from .middleware import _last_error
            raise RuntimeError("plan creation failed: {}".format(_last_error(self._lib)))
        self._plan = h
        return self

    def run_steps(self, steps, stream=None):
        """Enqueues the next `steps` time steps on `stream` (no wait);
        returns the steps left."""
        if self._plan is None:
            raise RuntimeError("no plan (call plan() first)")
        stream = self._stream(stream)
        left = self._plan_run(ctypes.c_void_p(self._plan), ctypes.c_size_t(int(steps)),
                              ctypes.c_void_p(hipmem.stream_handle(stream)))
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
                          ctypes.c_void_p(hipmem.stream_handle(stream)), left) != 0:
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

    def reset(self, stream=None):
        """Zeroes the wavefield slots (a fresh shot) on `stream`."""
        u = self.tensors["u_full"]
        hipmem._call(self._lib, "simwave_hip_zero", ctypes.c_void_p(u.data_ptr()),
                     ctypes.c_ulonglong(u.numel() * u.element_size()),
                     ctypes.c_void_p(hipmem.stream_handle(self._stream(stream))))

    def run(self, stream=None, flags=FLAG_OVERLAP_RECEIVERS):
        """Runs one shot on `stream` (default: the shot's default stream)."""
        stream = self._stream(stream)
        secs = self._fn(*self._args, ctypes.c_void_p(hipmem.stream_handle(stream)),
                        ctypes.c_uint(flags))
        if secs < 0:
            from .middleware import _last_error
            raise RuntimeError("device forward failed: {}".format(
                _last_error(self._lib)))
        return secs

    def synchronize(self, stream=None):
        """Waits until `stream` (default: the shot's default stream) has run
        everything enqueued on it."""
        self._sync(self._stream(stream))

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
        """(u_full, shot_record) as the Solver returns them.  The slots of
        the returned snapshots (one slot for saving_stride 0, a contiguous
        range otherwise) and the record are copied into a pinned staging
        buffer of the library's pool on `stream` (default: the shot's
        default stream),
        which is synchronised; the halo is cut on the host into ordinary
        (pageable) arrays, so results a caller keeps hold no page-locked
        memory."""
        sm, tm = self.solver.space_model, self.solver.time_model
        stream = self._stream(stream)
        end = self.kwargs["end_timestep"]
        u = self.tensors["u_full"]
        nslots = u.shape[0]
        if tm.saving_stride == 0:
            # the last level a full run returns is level T in slot T % 3; a
            # run cut at end_timestep E returns its level E (slot E % 3)
            first, count = end % 3, 1
        elif end == tm.timesteps:
            first, count = 1, nslots - 2       # remove_time_halo_region
        else:
            raise ValueError("results() of a saving_stride run needs every time step "
                             "(end_timestep %d < %d)" % (end, tm.timesteps))
        dtype = np.dtype(self.kwargs["u_full"].dtype)
        slot_bytes = int(np.prod(u.shape[1:], dtype=np.int64)) * dtype.itemsize
        rec = self.tensors["shot_record"]
        nu = count * slot_bytes
        nr = rec.numel() * rec.element_size()
        h = ctypes.c_void_p(hipmem.stream_handle(stream))
        with hipmem.staging(self._lib, nu + nr) as buf:
            hipmem._call(self._lib, "simwave_hip_copy", ctypes.c_void_p(buf.ctypes.data),
                         ctypes.c_void_p(u.data_ptr() + first * slot_bytes),
                         ctypes.c_ulonglong(nu), h)
            hipmem._call(self._lib, "simwave_hip_copy", ctypes.c_void_p(buf.ctypes.data + nu),
                         ctypes.c_void_p(rec.data_ptr()), ctypes.c_ulonglong(nr), h)
            self._sync(stream)
            slots = buf[:nu].view(dtype).reshape((count,) + tuple(u.shape[1:]))
            u_h = np.array(sm.remove_halo_region(slots), order="C", copy=True)
            r_h = buf[nu:nu + nr].view(dtype).reshape(tuple(rec.shape)).copy()
        return u_h, r_h
