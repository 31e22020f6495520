"""Device memory, pinned host memory and streams owned by the HIP library.

Thin ctypes wrappers over the library's memory, copy, zero, stream and
event entry points (``simwave_hip_device_alloc`` ... ``simwave_hip_event_*``,
declared in include/simwave_hip.h), so that the device-resident forward
path and the serving loop need no other GPU runtime.  Every object keeps
the library handle that made it and frees its resource when collected.
"""
import ctypes
import threading

import numpy as np

__all__ = ["DeviceArray", "Stream", "Event", "stream_handle", "staging", "default_library"]

_VP = ctypes.c_void_p
_ULL = ctypes.c_ulonglong
_SIGS = {
    "simwave_hip_device_alloc": [_ULL, _VP, ctypes.POINTER(_VP)],
    "simwave_hip_device_free": [_VP, _VP],
    "simwave_hip_host_alloc": [_ULL, ctypes.POINTER(_VP)],
    "simwave_hip_host_free": [_VP],
    "simwave_hip_copy": [_VP, _VP, _ULL, _VP],
    "simwave_hip_zero": [_VP, _ULL, _VP],
    "simwave_hip_stream_create": [ctypes.POINTER(_VP)],
    "simwave_hip_stream_destroy": [_VP],
    "simwave_hip_stream_synchronize": [_VP],
    "simwave_hip_event_create": [ctypes.POINTER(_VP)],
    "simwave_hip_event_record": [_VP, _VP],
    "simwave_hip_event_synchronize": [_VP],
    "simwave_hip_event_destroy": [_VP],
}
_BOUND = set()
_BIND_LOCK = threading.Lock()


def _fn(lib, name):
    key = (id(lib), name)
    fn = getattr(lib, name)
    if key not in _BOUND:
        with _BIND_LOCK:
            fn.argtypes = _SIGS[name]
            fn.restype = ctypes.c_int
            _BOUND.add(key)
    return fn


def _call(lib, name, *args):
    if _fn(lib, name)(*args) != 0:
        from .middleware import _last_error
        raise RuntimeError("%s failed: %s" % (name, _last_error(lib)))


_DEFAULT_LIB = []


def default_library():
    """A loaded variant library (2D constant density, float32) for memory,
    streams and events that belong to no particular shot; every variant
    library exports the same entry points for them."""
    if not _DEFAULT_LIB:
        from . import build
        path = build.library_path(2, "cd", "f32")
        with _BIND_LOCK:
            if not _DEFAULT_LIB:
                _DEFAULT_LIB.append(ctypes.CDLL(path))
    return _DEFAULT_LIB[0]


def stream_handle(stream):
    """The hipStream_t of `stream`: a Stream of this module, any object with
    a ``cuda_stream`` attribute (a torch stream), an integer handle, or None
    (the null stream)."""
    if stream is None:
        return None
    if isinstance(stream, Stream):
        return stream.handle
    h = getattr(stream, "cuda_stream", None)
    if h is not None:
        return int(h)
    return int(stream)


class Stream:
    """A non-blocking HIP stream of the library (default priority)."""

    def __init__(self, lib):
        self._lib = lib
        h = _VP()
        _call(lib, "simwave_hip_stream_create", ctypes.byref(h))
        self.handle = h.value

    def synchronize(self):
        _call(self._lib, "simwave_hip_stream_synchronize", _VP(self.handle))

    def record_event(self):
        """An Event recorded on this stream now."""
        ev = Event(self._lib)
        ev.record(self)
        return ev

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h:
            try:
                _fn(self._lib, "simwave_hip_stream_destroy")(_VP(h))
            except Exception:
                pass


class Event:
    """A HIP event without timing, for host-side waits."""

    def __init__(self, lib):
        self._lib = lib
        h = _VP()
        _call(lib, "simwave_hip_event_create", ctypes.byref(h))
        self.handle = h.value

    def record(self, stream):
        _call(self._lib, "simwave_hip_event_record", _VP(self.handle),
              _VP(stream_handle(stream)))

    def synchronize(self):
        _call(self._lib, "simwave_hip_event_synchronize", _VP(self.handle))

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h:
            try:
                _fn(self._lib, "simwave_hip_event_destroy")(_VP(h))
            except Exception:
                pass


class DeviceArray:
    """A C-contiguous array in device memory allocated by the library.

    Exposes what the device path reads of an array (``shape``, ``dtype``,
    ``data_ptr()``, ``numel()``, ``element_size()``) and stream-ordered
    copies to and from host arrays.  The memory is stream-ordered on
    `stream` (the library's default pool): usable by work enqueued there
    after the allocation (other streams once that point is reached), and
    returned to the pool after the work enqueued there before the array is
    collected; work on other streams that reads it must be complete by
    then."""

    def __init__(self, lib, shape, dtype, stream=None):
        self._lib = lib
        self.shape = tuple(int(n) for n in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._stream = stream
        h = _VP()
        _call(lib, "simwave_hip_device_alloc", _ULL(self.nbytes), _VP(stream_handle(stream)),
              ctypes.byref(h))
        self._ptr = h.value

    @classmethod
    def from_host(cls, lib, array, stream):
        """A device copy of the host array `array`, made on `stream` (which
        is synchronised, so the host array may go at once)."""
        a = np.ascontiguousarray(array)
        out = cls(lib, a.shape, a.dtype, stream)
        out.copy_from_host(a, stream)
        return out

    def data_ptr(self):
        return self._ptr

    def numel(self):
        return self.nbytes // self.dtype.itemsize

    def element_size(self):
        return self.dtype.itemsize

    @property
    def ndim(self):
        return len(self.shape)

    def zero_(self, stream):
        _call(self._lib, "simwave_hip_zero", _VP(self._ptr), _ULL(self.nbytes),
              _VP(stream_handle(stream)))
        return self

    def copy_from_host(self, array, stream):
        """Copies the host array in on `stream` and waits for the copy (the
        source, perhaps a temporary, is read by the copy engine)."""
        a = np.ascontiguousarray(array, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("host array has %d bytes, device array %d" % (a.nbytes, self.nbytes))
        _call(self._lib, "simwave_hip_copy", _VP(self._ptr), _VP(a.ctypes.data),
              _ULL(self.nbytes), _VP(stream_handle(stream)))
        _call(self._lib, "simwave_hip_stream_synchronize", _VP(stream_handle(stream)))

    def numpy(self, stream=None):
        """A host copy (synchronises `stream`)."""
        out = np.empty(self.shape, dtype=self.dtype)
        _call(self._lib, "simwave_hip_copy", _VP(out.ctypes.data), _VP(self._ptr),
              _ULL(self.nbytes), _VP(stream_handle(stream)))
        _call(self._lib, "simwave_hip_stream_synchronize", _VP(stream_handle(stream)))
        return out

    def __del__(self):
        p, self._ptr = getattr(self, "_ptr", None), None
        if p:
            try:
                _fn(self._lib, "simwave_hip_device_free")(_VP(p),
                                                         _VP(stream_handle(self._stream)))
            except Exception:
                pass


class _Pinned:
    """Page-locked host memory of the library, viewed as a numpy uint8 array."""

    def __init__(self, lib, nbytes):
        self._lib = lib
        h = _VP()
        _call(lib, "simwave_hip_host_alloc", _ULL(nbytes), ctypes.byref(h))
        self._ptr = h.value
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self._ptr))

    def __del__(self):
        p, self._ptr = getattr(self, "_ptr", None), None
        if p:
            self.array = None
            try:
                _fn(self._lib, "simwave_hip_host_free")(_VP(p))
            except Exception:
                pass


_POOL = []                  # idle staging buffers (_Pinned)
_POOL_LOCK = threading.Lock()
_POOL_KEEP = 4 << 30        # idle page-locked bytes kept for reuse


class staging:
    """``with staging(lib, nbytes) as buf:`` a pinned staging buffer (a
    numpy uint8 array over library-allocated page-locked memory) of at least
    `nbytes`, taken from a process-wide pool and returned to it on exit, so
    device-to-host copies run at pinned speed without page-locking memory
    per call or per thread.  Sizes are powers of two; up to 4 GiB of idle
    buffers are kept."""

    def __init__(self, lib, nbytes):
        self._lib = lib
        self._need = max(1, int(nbytes))
        self._buf = None

    def __enter__(self):
        with _POOL_LOCK:
            fits = [b for b in _POOL if b.nbytes >= self._need]
            if fits:
                self._buf = min(fits, key=lambda b: b.nbytes)
                _POOL.remove(self._buf)
        if self._buf is None:
            size = 1 << max(20, (self._need - 1).bit_length())
            self._buf = _Pinned(self._lib, size)
        return self._buf.array

    def __exit__(self, *exc):
        buf, self._buf = self._buf, None
        with _POOL_LOCK:
            if sum(b.nbytes for b in _POOL) + buf.nbytes <= _POOL_KEEP:
                _POOL.append(buf)
                buf = None
        del buf          # freed outside the lock when not kept
        return False
