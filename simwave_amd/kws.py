"""Kaiser-windowed sinc interpolation weights (Hicks 2002).

Reference: simwave/kernel/frontend/kws.py.  The reference evaluates the window
over every point of an axis and then scans the result point by point in
Python (kws.py:121-124), which dominates host time for thousands of
receivers.  Here the same elementwise expressions are evaluated only on the
few points around each position, for all positions at once; every weight is
computed by the same numpy/scipy operations on the same operand types, so the
values are identical.
"""

import numpy as np
from scipy.special import i0

__all__ = ["get_kaiser_half_width", "kaiser_windowing_sinc",
           "get_kws_valid_points", "get_source_points", "points_and_values"]

# Optimal Kaiser b per window half-width (Hicks 2002, table 2),
# reference kws.py:28-39.
_KAISER_B = {1: 1.24, 2: 2.94, 3: 4.53, 4: 6.31, 5: 7.91,
             6: 9.42, 7: 10.95, 8: 12.53, 9: 14.09, 10: 14.18}


def get_kaiser_half_width(half_width):
    """Kaiser b parameter for a window half-width in 1..10 (kws.py:6-41)."""
    if half_width not in _KAISER_B:
        raise Exception(
            "Kaiser windowing half-width {} not supported".format(half_width))
    return _KAISER_B[half_width]


def _kws(x, half_width):
    """Windowed sinc at float32 distances x (NaN outside the window).

    The scalar normaliser i0(b) takes the dtype of x, as scalars did under
    NumPy's value-based casting when the reference's fixtures were made
    (tests/test_source.py:111-164 of the reference pins these float32 values
    bit for bit; a float64 normaliser changes the last bit).
    """
    # NumPy's error state is per thread (serving builds shots on several
    # threads at once; the warnings module's filters are process-wide)
    with np.errstate(invalid="ignore"):
        root = np.sqrt(1 - (x / half_width) ** 2)
    b = get_kaiser_half_width(half_width)
    kaiser = i0(b * root) / x.dtype.type(i0(b))
    return kaiser * np.sinc(x)


def kaiser_windowing_sinc(num_points, source_point, half_width):
    """Weights on every point of an axis, NaN outside the window.

    Reference: kws.py:44-94.
    """
    x = np.linspace(start=0, stop=num_points - 1, num=num_points,
                    dtype=np.float32)
    return _kws(x - np.float32(source_point), half_width)


def get_kws_valid_points(kaiser_windowed_array):
    """First/last valid index and the valid values as float32 (kws.py:97-135)."""
    valid = np.flatnonzero(~np.isnan(kaiser_windowed_array))
    if valid.size == 0:
        raise Exception(
            "There is no valid point in the source/receiver location")
    values = np.asarray(kaiser_windowed_array[valid], dtype=np.float32)
    return int(valid[0]), int(valid[-1]), values


def get_source_points(grid_shape, source_location, half_width):
    """Intervals and values of one position (kws.py:138-183)."""
    if len(grid_shape) != len(source_location):
        raise Exception(
            "Grid and source/receiver location must have the same dimension.")
    points, values = points_and_values(
        grid_shape, np.asarray([source_location]), half_width)
    return points.astype(np.uint), values


def points_and_values(grid_shape, positions, half_width):
    """Vectorised intervals and weights for many positions.

    positions: (n, dim) array in grid units (the adjusted positions).
    Returns (points uint64 [n*2*dim], values float32 [sum of box sizes]) laid
    out like the reference: per position, per axis, [begin, end] and the
    axis weights (source.py:140-159).
    """
    positions = np.asarray(positions)
    if positions.ndim != 2 or positions.shape[1] != len(grid_shape):
        raise Exception(
            "Grid and source/receiver location must have the same dimension.")
    get_kaiser_half_width(half_width)
    count, dim = positions.shape
    if count == 0:
        return np.zeros(0, dtype=np.uint), np.zeros(0, dtype=np.float32)
    span = 2 * half_width + 4
    starts = np.empty((count, dim), dtype=np.int64)
    begins = np.empty((count, dim), dtype=np.int64)
    ends = np.empty((count, dim), dtype=np.int64)
    per_axis = []
    for axis in range(dim):
        n = int(grid_shape[axis])
        # every quantity below is a function of the coordinate alone:
        # evaluate it once per distinct coordinate (receiver lines and grids
        # repeat them) and spread the rows back
        pos, back = np.unique(positions[:, axis], return_inverse=True)
        back = back.reshape(-1)
        lo = np.floor(pos.astype(np.float64)).astype(np.int64) - half_width - 1
        lo = np.clip(lo, 0, max(n - 1, 0))
        idx = lo[:, None] + np.arange(span)[None, :]
        inside = idx < n
        grid = np.where(inside, idx, 0).astype(np.float32)
        # float32 grid minus the position taken as float32 (the reference's
        # float32 linspace minus a position scalar under value-based casting)
        x = grid - pos.astype(np.float32)[:, None]
        w = _kws(x, half_width)
        w = np.where(inside, w, np.nan)
        valid = ~np.isnan(w)
        has = valid.any(axis=1)
        if not has.all():
            raise Exception(
                "There is no valid point in the source/receiver location")
        first = valid.argmax(axis=1)
        last = span - 1 - valid[:, ::-1].argmax(axis=1)
        lo, w, first, last = lo[back], w[back], first[back], last[back]
        starts[:, axis] = lo
        begins[:, axis] = lo + first
        ends[:, axis] = lo + last
        per_axis.append((w, first, last))
    points = np.empty((count, 2 * dim), dtype=np.uint)
    points[:, 0::2] = begins
    points[:, 1::2] = ends
    # per position, per axis, the weights first..last: a (count, dim, span)
    # gather in row-major order
    k = np.arange(span)[None, :]
    w_all = np.stack([w for w, _, _ in per_axis], axis=1).astype(np.float32)
    keep = np.stack([(k >= first[:, None]) & (k <= last[:, None])
                     for _, first, last in per_axis], axis=1)
    values = w_all[keep]
    return points.reshape(-1), values
