"""simwave_amd: simwave's acoustic forward operator on MI355X (gfx950, HIP).

Drop-in for simwave's plain forward path: the same frontend classes
(simwave/__init__.py:1-41) in front of hand-written HIP kernels reached
through a C-ABI library loaded with ctypes.
"""
from .compiler import Compiler
from .model import SpaceModel, TimeModel
from .source import Source, Receiver, Wavelet, RickerWavelet, MultiWavelet
from .solver import Solver
from .middleware import Middleware

__all__ = [
    "Compiler",
    "SpaceModel",
    "TimeModel",
    "Source",
    "Receiver",
    "Wavelet",
    "RickerWavelet",
    "MultiWavelet",
    "Solver",
    "Middleware",
]
