"""ctypes marshalling between the Python frontend and a backend library.

Reference: simwave/kernel/backend/middleware.py.  The argument table, the
key order and the numpy/ctypes type mapping are the reference's
(middleware.py:160-285): they define the C ABI of ``forward`` that every
backend library exports.  A negative return value is this build's error
signal; the message comes from the library's ``simwave_hip_last_error``.
"""
import ctypes

import numpy as np
from numpy.ctypeslib import ndpointer

from .compiler import Compiler

__all__ = ["Middleware", "KEY_ORDER"]

# C parameter order of forward (reference middleware.py:166-202); keys whose
# value is None are not passed.
KEY_ORDER = (
    "u_full", "velocity_model", "density_model", "damping_mask", "wavelet",
    "wavelet_size", "wavelet_count", "second_order_fd_coefficients",
    "first_order_fd_coefficients", "boundary_condition",
    "src_points_interval", "src_points_interval_size", "src_points_values",
    "src_points_values_size", "src_points_values_offset",
    "rec_points_interval", "rec_points_interval_size", "rec_points_values",
    "rec_points_values_size", "rec_points_values_offset", "shot_record",
    "num_sources", "num_receivers", "nz", "nx", "ny", "dz", "dx", "dy",
    "saving_stride", "dt", "begin_timestep", "end_timestep", "space_order",
    "num_snapshots",
)

_BC_CODES = {"none": 0, "null_dirichlet": 1, "null_neumann": 2}

_CTYPE = {
    "int": ctypes.c_size_t,
    "float": ctypes.c_float,
    "float32": ctypes.c_float,
    "float64": ctypes.c_double,
    "np(uint64)": ndpointer(ctypes.c_size_t, flags="C_CONTIGUOUS"),
    "np(float32)": ndpointer(ctypes.c_float, flags="C_CONTIGUOUS"),
    "np(float64)": ndpointer(ctypes.c_double, flags="C_CONTIGUOUS"),
}

_LOADED = {}


class Middleware:
    """Calls a backend library's ``forward`` with the reference arguments."""

    def __init__(self, compiler=None, verbose=True):
        self._compiler = Compiler() if compiler is None else compiler
        self.verbose = verbose

    @property
    def compiler(self):
        return self._compiler

    def library(self, dimension, density, dtype):
        """Loads the backend library for a variant (middleware.py:26-51)."""
        precision = {"float32": "-DFLOAT", "float64": "-DDOUBLE"}[str(dtype)]
        dens = "constant_density" if density is None else "variable_density"
        path = self.compiler.compile(dimension=dimension, density=dens,
                                     float_precision=precision,
                                     operator="forward")
        lib = _LOADED.get(path)
        if lib is None:
            lib = ctypes.cdll.LoadLibrary(path)
            _LOADED[path] = lib
        return lib

    @staticmethod
    def normalize(**kwargs):
        """Reference exec() preprocessing (middleware.py:70-101): boundary
        names to codes, density arguments dropped for constant density, grid
        shape and spacing unpacked into nz/nx[/ny] and dz/dx[/dy]."""
        kwargs["boundary_condition"] = Middleware._convert_boundary_condition(
            kwargs.get("boundary_condition"))
        if kwargs.get("density_model") is None:
            kwargs.pop("density_model", None)
            kwargs.pop("first_order_fd_coefficients", None)
        shape = kwargs.get("velocity_model").shape
        for name, n in zip(("nz", "nx", "ny"), shape):
            kwargs[name] = n
        spacing = kwargs.pop("grid_spacing")
        for name, h in zip(("dz", "dx", "dy"), spacing):
            kwargs[name] = h
        return kwargs

    @staticmethod
    def ordered(kwargs):
        """(argtypes, args) of forward in C parameter order."""
        argtypes, args = [], []
        for key in KEY_ORDER:
            value = kwargs.get(key)
            if value is None:
                continue
            argtypes.append(Middleware._ctype_of(value))
            args.append(value)
        return argtypes, args

    def exec(self, operator, **kwargs):
        """Normalises the keyword arguments and runs `operator`."""
        kwargs = self.normalize(**kwargs)
        if operator == "forward":
            return self._exec_forward(**kwargs)
        raise ValueError("operator {} not available".format(operator))

    def _exec_forward(self, **kwargs):
        vel = kwargs.get("velocity_model")
        lib = self.library(dimension=len(vel.shape),
                           density=kwargs.get("density_model"),
                           dtype=vel.dtype)
        argtypes, args = self.ordered(kwargs)
        fn = lib.forward
        fn.restype = ctypes.c_double
        fn.argtypes = argtypes
        seconds = fn(*args)
        if seconds < 0:
            raise RuntimeError("forward failed: {}".format(
                _last_error(lib) or "unknown error"))
        if self.verbose:
            print("Run forward in %f seconds." % seconds)
        return kwargs.get("u_full"), kwargs.get("shot_record")

    @staticmethod
    def _ctype_of(value):
        if isinstance(value, np.ndarray):
            return _CTYPE["np({})".format(value.dtype)]
        return _CTYPE[type(value).__name__]

    @staticmethod
    def _convert_boundary_condition(boundary_condition):
        return np.uint([_BC_CODES[b] for b in boundary_condition])


def _last_error(lib):
    try:
        fn = lib.simwave_hip_last_error
    except AttributeError:
        return None
    fn.restype = ctypes.c_char_p
    fn.argtypes = []
    msg = fn()
    return msg.decode() if msg else None
