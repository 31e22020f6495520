"""Sources, receivers and wavelets.

Reference: simwave/kernel/frontend/source.py (Source :5-164, Receiver = Source
:168, Wavelet :171-209, RickerWavelet :212-257, MultiWavelet :260-302).
"""
import math

import numpy as np

from . import kws

__all__ = ["Source", "Receiver", "Wavelet", "RickerWavelet", "MultiWavelet"]


class Source:
    """A set of point sources (or receivers) at physical coordinates.

    Parameters
    ----------
    space_model : SpaceModel
    coordinates : sequence or ndarray
        One (z, x[, y]) tuple per position, in metres.
    window_radius : int, optional
        Kaiser window half-width in grid points, 1..10 (default 4).
    """

    def __init__(self, space_model, coordinates, window_radius=4):
        self._space_model = space_model
        self._window_radius = window_radius
        if not isinstance(coordinates, (list, tuple, np.ndarray)):
            raise ValueError("Source/Receiver coordinates must be "
                             "represented as lists, tuples or ndarrays.")
        coords = np.asarray(coordinates, dtype=space_model.dtype)
        if coords.ndim == 1:
            coords = coords.reshape((1,) + coords.shape)
        elif coords.ndim != 2 or coords.shape[1] not in (2, 3):
            raise ValueError("Invalid source/receiver coordinates format.")
        self._coordinates = coords

    @property
    def space_model(self):
        return self._space_model

    @property
    def coordinates(self):
        return self._coordinates

    @property
    def window_radius(self):
        return self._window_radius

    @property
    def count(self):
        return self.coordinates.shape[0]

    @property
    def grid_positions(self):
        """(coord - min) / h per axis (source.py:62-103).  Coordinates, box
        and spacing are all held in the model dtype, so the subtraction and
        the division are operations of that dtype, elementwise here."""
        sm = self.space_model
        box, h = sm.bounding_box, sm.grid_spacing
        coords = np.asarray(self.coordinates, dtype=sm.dtype)
        if coords.size == 0:
            return np.zeros((0, sm.dimension), dtype=sm.dtype)
        dim = coords.shape[1]
        if dim not in (2, 3):
            raise Exception("Dimension %d not supported." % dim)
        lo = np.array([box[2 * a] for a in range(dim)], dtype=sm.dtype)
        hi = np.array([box[2 * a + 1] for a in range(dim)], dtype=sm.dtype)
        bad = ((coords < lo) | (coords > hi)).any(axis=1)
        if bad.any():
            raise Exception("Coordinates %s out of bounds."
                            % (tuple(self.coordinates[int(np.argmax(bad))]),))
        step = np.array([h[a] for a in range(dim)], dtype=sm.dtype)
        return ((coords - lo) / step).astype(sm.dtype)

    @property
    def adjusted_grid_positions(self):
        """Grid positions shifted by the leading damping layer and halo
        (source.py:105-121), added in the model dtype."""
        sm = self.space_model
        origin = np.array([sm.nbl[2 * a] + sm.halo_size[2 * a]
                           for a in range(sm.dimension)], dtype=sm.dtype)
        return (self.grid_positions + origin).astype(sm.dtype)

    @property
    def interpolated_points_and_values(self):
        """(intervals uint64, weights in dtype, offsets uint64).

        Per position: [begin, end] per axis (inclusive), the per-axis
        windowed-sinc weights (float32-rounded), and offsets[i] = start of
        position i's weights (source.py:123-159).
        """
        sm = self.space_model
        positions = self.adjusted_grid_positions
        if positions.size == 0:
            return (np.array([], dtype=np.uint),
                    np.array([], dtype=sm.dtype),
                    np.asarray([0], dtype=np.uint))
        points, values = kws.points_and_values(
            sm.extended_shape, positions, self.window_radius)
        spans = (points[1::2].astype(np.int64) - points[0::2].astype(np.int64) + 1)
        sizes = spans.reshape(self.count, sm.dimension).sum(axis=1)
        offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint)
        return points, values.astype(sm.dtype), offsets


# a receiver is a source used for sampling
Receiver = Source


class Wavelet:
    """A source time function given by a callable and its keyword arguments."""

    def __init__(self, function, **kwargs):
        self._function = function
        self._kwargs = kwargs

    @property
    def function(self):
        return self._function

    @property
    def kwargs(self):
        return self._kwargs

    @property
    def values(self):
        return self.function(**self.kwargs)

    @property
    def num_sources(self):
        return 1

    @property
    def timesteps(self):
        return len(self.values)


class RickerWavelet(Wavelet):
    """Ricker wavelet A (1 - 2 r^2) exp(-r^2), r = pi f (t - 1/f).

    Sampled at TimeModel.time_values (linspace(t0, tf, timesteps)).
    Reference: source.py:212-257.
    """

    def __init__(self, peak_frequency, time_model, amplitude=1):
        self._peak_frequency = peak_frequency
        self._time_model = time_model
        self._amplitude = amplitude
        super().__init__(self._function, peak_frequency=peak_frequency,
                         time_model=time_model, amplitude=amplitude)

    @property
    def peak_frequency(self):
        return self._peak_frequency

    @property
    def time_model(self):
        return self._time_model

    @property
    def amplitude(self):
        return self._amplitude

    def _function(self, peak_frequency, time_model, amplitude):
        # operand order and types follow the reference expression, so the
        # samples round identically in the model dtype
        delay = 1 / peak_frequency
        arg = math.pi * peak_frequency * (time_model.time_values - delay)
        return amplitude * (1 - 2.0 * arg ** 2) * np.exp(-arg ** 2)


class MultiWavelet(Wavelet):
    """One wavelet per source: values[timesteps][sources] (source.py:260-302)."""

    def __init__(self, values, time_model):
        self._values = values
        self._time_model = time_model
        if self.timesteps != self.time_model.timesteps:
            raise ValueError("Wavelet must have {} timesteps.".format(
                self.time_model.timesteps))

    @property
    def values(self):
        return np.ascontiguousarray(self._values, dtype=self.dtype)

    @property
    def num_sources(self):
        return self.values.shape[1]

    @property
    def timesteps(self):
        return self.values.shape[0]

    @property
    def time_model(self):
        return self._time_model

    @property
    def dtype(self):
        return self.time_model.dtype
