"""Spatial and temporal models of a simulation.

Reference: simwave/kernel/frontend/model.py (SpaceModel :6-515,
TimeModel :518-643).  Same public attributes, the same validation errors and
the same array values; the extended arrays are built with numpy's own padding
routines exactly as the reference builds them.
"""
import numpy as np
from scipy.interpolate import RegularGridInterpolator

from . import fd

__all__ = ["SpaceModel", "TimeModel"]

_BOUNDARY_NAMES = ("none", "null_dirichlet", "null_neumann")



def _pad_edge(a, widths):
    """np.pad(a, widths, mode='edge'): a copy with every face replicated
    outwards, axis by axis (the corners follow, as in numpy)."""
    a = np.asarray(a)
    out = np.empty(tuple(n + w0 + w1 for n, (w0, w1) in zip(a.shape, widths)),
                   dtype=a.dtype)
    core = tuple(slice(w0, w0 + n) for n, (w0, _) in zip(a.shape, widths))
    out[core] = a
    for ax, (w0, w1) in enumerate(widths):
        lo, hi = w0, w0 + a.shape[ax]
        idx = [slice(None)] * a.ndim
        if w0:
            idx[ax] = slice(0, lo)
            src = [slice(None)] * a.ndim
            src[ax] = slice(lo, lo + 1)
            out[tuple(idx)] = out[tuple(src)]
        if w1:
            idx[ax] = slice(hi, hi + w1)
            src = [slice(None)] * a.ndim
            src[ax] = slice(hi - 1, hi)
            out[tuple(idx)] = out[tuple(src)]
    return out


def _pad_zero(a, widths):
    """np.pad(a, widths) (constant zero)."""
    out = np.zeros(tuple(n + w0 + w1 for n, (w0, w1) in zip(a.shape, widths)),
                   dtype=a.dtype)
    out[tuple(slice(w0, w0 + n) for n, (w0, _) in zip(a.shape, widths))] = a
    return out

class SpaceModel:
    """Grid, velocity/density models, damping layer and halo of a run.

    Parameters
    ----------
    bounding_box : tuple of float
        (z_min, z_max, x_min, x_max[, y_min, y_max]) in metres.
    grid_spacing : tuple of float
        Spacing (dz, dx[, dy]) in metres.
    velocity_model : ndarray
        P-wave velocity (m/s), 2D or 3D, spanning the bounding box.
    density_model : ndarray, optional
        Density; selects the variable-density operator.
    space_order : int, optional
        Even spatial order 2..20 (default 2).
    dtype : numpy float type, optional
        np.float32 (default) or np.float64.
    """

    def __init__(self, bounding_box, grid_spacing, velocity_model,
                 density_model=None, space_order=2, dtype=np.float32):
        self._dtype = dtype
        self._bounding_box = tuple(dtype(v) for v in bounding_box)
        self._grid_spacing = tuple(dtype(v) for v in grid_spacing)
        self._space_order = space_order
        # reference model.py:41-51
        if space_order % 2 != 0:
            raise ValueError(
                "Odd space order {} not supported".format(space_order))
        if not 2 <= space_order <= 20:
            raise ValueError("Space order limited from 2 to 20.")
        self._dimension = len(velocity_model.shape)
        self._velocity_model = self.interpolate(velocity_model)
        self._density_model = (None if density_model is None
                               else self.interpolate(density_model))

    # ------------------------------------------------------------ attributes
    @property
    def bounding_box(self):
        return self._bounding_box

    @property
    def grid_spacing(self):
        return self._grid_spacing

    @property
    def velocity_model(self):
        return self._velocity_model

    @property
    def density_model(self):
        return self._density_model

    @property
    def space_order(self):
        return self._space_order

    @property
    def dimension(self):
        return self._dimension

    @property
    def dtype(self):
        return self._dtype

    @property
    def shape(self):
        """Points per axis: int((max - min) / h) + 1 (model.py:93-112)."""
        box = self.bounding_box
        return tuple(int((box[2 * a + 1] - box[2 * a]) / self.grid_spacing[a]) + 1
                     for a in range(self.dimension))

    @property
    def extended_shape(self):
        """shape + damping layers + halo on both sides (model.py:114-133)."""
        nbl, halo = self.nbl, self.halo_size
        return tuple(self.shape[a] + nbl[2 * a] + nbl[2 * a + 1]
                     + halo[2 * a] + halo[2 * a + 1]
                     for a in range(self.dimension))

    @property
    def grid(self):
        return np.zeros(self.shape, dtype=self.dtype)

    @property
    def nbl(self):
        """Damping-layer points per face, (z_b, z_a, x_b, x_a[, y_b, y_a])."""
        return getattr(self, "_nbl", (0,) * self.dimension * 2)

    @property
    def damping_length(self):
        return getattr(self, "_damping_length", (0.0,) * self.dimension * 2)

    @property
    def boundary_condition(self):
        return getattr(self, "_boundary_condition",
                       ("none",) * self.dimension * 2)

    @property
    def damping_polynomial_degree(self):
        return getattr(self, "_damping_polynomial_degree", 3)

    @property
    def damping_alpha(self):
        return getattr(self, "_damping_alpha", 0.001)

    @property
    def halo_size(self):
        """Halo of space_order // 2 points on every face (model.py:263-269)."""
        return (self.space_order // 2,) * self.dimension * 2

    def fd_coefficients(self, derivative_order):
        """Centre + right-side FD weights in the model dtype (model.py:192-208)."""
        return self.dtype(fd.half_coefficients(derivative_order,
                                               self.space_order))

    # ---------------------------------------------------------- resampling
    def interpolate(self, data):
        """Linear resampling of `data` onto the bounding-box grid.

        Reference: model.py:210-261 (scipy RegularGridInterpolator over
        linspace node coordinates).  When the data already has the grid's
        shape the nodes coincide and the interpolant returns the data
        unchanged, so that case skips the evaluation.
        """
        shape = self.shape
        box = self.bounding_box
        if tuple(data.shape) == tuple(shape):
            return self.dtype(np.asarray(data))
        src_axes = tuple(np.linspace(box[2 * a], box[2 * a + 1], data.shape[a])
                         for a in range(self.dimension))
        dst_axes = [np.linspace(box[2 * a], box[2 * a + 1], shape[a])
                    for a in range(self.dimension)]
        interpolant = RegularGridInterpolator(src_axes, data)
        out = np.empty(shape, dtype=self.dtype)
        # evaluate one z plane at a time (pointwise identical, bounded memory)
        rest = list(np.meshgrid(*dst_axes[1:], indexing="ij"))
        for iz, z in enumerate(dst_axes[0]):
            pts = np.stack([np.full(rest[0].shape, z)] + rest, axis=-1)
            out[iz] = self.dtype(interpolant(pts))
        return out

    # ------------------------------------------------------------- boundary
    def config_boundary(self, damping_length=0.0, boundary_condition="none",
                        damping_polynomial_degree=3, damping_alpha=0.001):
        """Damping layers (metres) and face conditions (model.py:271-349)."""
        self._damping_polynomial_degree = damping_polynomial_degree
        self._damping_alpha = damping_alpha
        faces = self.dimension * 2
        if isinstance(damping_length, (float, int)):
            self._damping_length = (self.dtype(damping_length),) * faces
        else:
            self._damping_length = tuple(self.dtype(v) for v in damping_length)
        if isinstance(boundary_condition, str):
            self._boundary_condition = (boundary_condition,) * faces
        else:
            self._boundary_condition = boundary_condition
        for bc in self.boundary_condition:
            if bc not in _BOUNDARY_NAMES:
                raise ValueError('Boundary condition {} not available.'.format(
                    self.boundary_condition))
        self._nbl = tuple(int(self._damping_length[f] / self.grid_spacing[f // 2])
                          for f in range(faces))

    @property
    def nbl_pad_width(self):
        nbl = self.nbl
        return tuple((nbl[2 * a], nbl[2 * a + 1]) for a in range(self.dimension))

    @property
    def halo_pad_width(self):
        halo = self.halo_size
        return tuple((halo[2 * a], halo[2 * a + 1])
                     for a in range(self.dimension))

    @property
    def damping_mask(self):
        """alpha * d**degree in the damping layers, zero elsewhere.

        d is numpy's linear ramp from 0 at the inner edge to nbl at the outer
        edge, padded axis by axis; the halo is zero (model.py:377-406).
        """
        return _pad_zero(self._damping_values(self.shape), self.halo_pad_width)

    def _damping_values(self, core):
        """The damping layer around a zero interior of shape `core`."""
        mask = np.pad(np.zeros(core, dtype=self.dtype),
                      pad_width=self.nbl_pad_width, mode="linear_ramp",
                      end_values=self.nbl_pad_width)
        deg = self.damping_polynomial_degree
        if isinstance(deg, (int, np.integer)) and deg >= 3:
            # the same power, evaluated only where the ramp is non-zero
            # (0**deg == 0); the interior is most of the grid
            mask = np.power(mask, deg, out=np.zeros_like(mask), where=mask != 0)
        else:
            mask = mask ** deg
        return mask * self.damping_alpha

    def damping_mask_factors(self):
        """(small, index) with damping_mask == small[np.ix_(*index)].

        np.pad's linear ramps depend only on the edge values and the pad
        widths, never on the interior size, so the mask around a one-cell
        interior holds every value of the full mask; `index` maps each
        extended index to its cell of `small` (for building the mask where
        the data lives, e.g. on a device, without the full host array)."""
        small = _pad_zero(self._damping_values((1,) * self.dimension),
                          self.halo_pad_width)
        nbl, halo = self.nbl, self.halo_size
        index = []
        for a in range(self.dimension):
            lo = halo[2 * a] + nbl[2 * a]
            n = self.shape[a]
            e = np.arange(self.extended_shape[a], dtype=np.int64)
            index.append(np.where(e < lo, e, np.where(e < lo + n, lo, e - n + 1)))
        return small, index

    def extension_index(self):
        """Per axis, the model index each extended index copies under the
        edge padding of extended_velocity_model: clip(e - lo, 0, n - 1)."""
        nbl, halo = self.nbl, self.halo_size
        return [np.clip(np.arange(self.extended_shape[a], dtype=np.int64)
                        - (halo[2 * a] + nbl[2 * a]), 0, self.shape[a] - 1)
                for a in range(self.dimension)]

    @property
    def extended_grid(self):
        return np.pad(np.pad(self.grid, pad_width=self.nbl_pad_width),
                      pad_width=self.halo_pad_width)

    def _edge_extend(self, model):
        """np.pad(np.pad(model, nbl, 'edge'), halo, 'edge') as one copy:
        both are edge replication, so the result replicates the model's
        edges over nbl + halo."""
        widths = tuple((a[0] + b[0], a[1] + b[1])
                       for a, b in zip(self.nbl_pad_width, self.halo_pad_width))
        return _pad_edge(model, widths)

    @property
    def extended_velocity_model(self):
        """Velocity edge-padded by the damping layers and the halo."""
        return self._edge_extend(self.velocity_model)

    @property
    def extended_density_model(self):
        """Density edge-padded like the velocity, or None."""
        if self.density_model is None:
            return None
        return self._edge_extend(self.density_model)

    def remove_halo_region(self, u):
        """Drops the halo of every snapshot; keeps the damping layers."""
        h = self.halo_size[0]
        if self.dimension not in (2, 3):
            raise Exception("Wavefield dimension not supported.")
        return u[(slice(None),) + (slice(h, -h),) * self.dimension]

    def remove_nbl(self, u):
        """Drops the damping layers of every snapshot (model.py:494-515)."""
        nbl = self.nbl
        if self.dimension not in (2, 3):
            raise Exception("Wavefield dimension not supported.")
        return u[(slice(None),) + tuple(slice(nbl[2 * a], -nbl[2 * a + 1])
                                        for a in range(self.dimension))]


class TimeModel:
    """Time axis of a simulation (model.py:518-643).

    Parameters
    ----------
    space_model : SpaceModel
    tf : float
        End time (s).
    dt : float, optional
        Time step; must not exceed the CFL step.
    t0 : float, optional
        Start time (s), default 0.
    saving_stride : int, optional
        Snapshot stride; 0 keeps only the last wavefield.
    """

    def __init__(self, space_model, tf, dt=None, t0=0.0, saving_stride=0):
        self._space_model = space_model
        self._tf = space_model.dtype(tf)
        self._t0 = space_model.dtype(t0)
        self._saving_stride = saving_stride
        self._dt = fd.calculate_dt(
            dimension=space_model.dimension,
            space_order=space_model.space_order,
            grid_spacing=space_model.grid_spacing,
            velocity_model=space_model.velocity_model)
        if dt is not None:
            self.dt = dt
        if not 0 <= self.saving_stride <= self.timesteps:
            raise Exception(
                "Saving jumps can not be less than zero or "
                "greater than the number of timesteps.")

    @property
    def space_model(self):
        return self._space_model

    @property
    def tf(self):
        return self._tf

    @property
    def t0(self):
        return self._t0

    @property
    def saving_stride(self):
        return self._saving_stride

    @property
    def dtype(self):
        return self.space_model.dtype

    @property
    def dt(self):
        return self.dtype(self._dt)

    @dt.setter
    def dt(self, value):
        if value < 0:
            raise ValueError("Time step cannot be negative.")
        if value > self.dt:
            raise ValueError("Time step value violates CFL condition.")
        self._dt = value

    @property
    def timesteps(self):
        """ceil((tf - t0 + dt) / dt) in the model dtype, raised until
        timesteps % stride == 1 for 1 < stride <= timesteps (model.py:599-609).
        """
        steps = int(np.ceil((self.tf - self.t0 + self.dt) / self.dt))
        stride = self.saving_stride
        if 1 < stride <= steps:
            steps += (1 - steps) % stride
        return steps

    @property
    def time_indexes(self):
        return np.linspace(0, self.timesteps - 1, self.timesteps, dtype=np.uint)

    @property
    def time_values(self):
        return np.linspace(self.t0, self.tf, self.timesteps, dtype=self.dtype)

    def remove_time_halo_region(self, u):
        """Keeps the returned snapshots (model.py:623-643): the last level
        (slot timesteps % 3) for stride 0, otherwise slots 1..-2."""
        if self.saving_stride == 0:
            last = self.timesteps % 3
            return u[last:last + 1]
        return u[1:-1]
