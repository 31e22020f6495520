"""Acoustic forward solver (reference: simwave/kernel/frontend/solver.py)."""
import numpy as np

from .middleware import Middleware

__all__ = ["Solver"]


class Solver:
    """Runs the forward operator of a configured experiment.

    Parameters
    ----------
    space_model : SpaceModel
    time_model : TimeModel
    sources : Source
    receivers : Receiver
    wavelet : Wavelet
    compiler : Compiler, optional
        Backend selection; default ``Compiler()`` (the HIP kernels).
    """

    def __init__(self, space_model, time_model, sources, receivers, wavelet,
                 compiler=None):
        self._space_model = space_model
        self._time_model = time_model
        self._sources = sources
        self._receivers = receivers
        self._wavelet = wavelet
        self._compiler = compiler
        self._middleware = Middleware(compiler=compiler)

    @property
    def space_model(self):
        return self._space_model

    @property
    def time_model(self):
        return self._time_model

    @property
    def sources(self):
        return self._sources

    @property
    def receivers(self):
        return self._receivers

    @property
    def wavelet(self):
        return self._wavelet

    @property
    def compiler(self):
        return self._compiler

    @property
    def snapshot_indexes(self):
        """Time indexes kept in u_full (solver.py:67-83)."""
        tm = self.time_model
        if tm.saving_stride == 0:
            return [tm.time_indexes[-1]]
        return list(range(tm.time_indexes[0], tm.timesteps, tm.saving_stride))

    @property
    def num_snapshots(self):
        return len(self.snapshot_indexes)

    @property
    def shot_record(self):
        """Zeroed receiver record [timesteps][receivers]."""
        return np.zeros((self.time_model.timesteps, self.receivers.count),
                        dtype=self.space_model.dtype)

    @property
    def u_full(self):
        """Zeroed wavefield slots: snapshots + 2 time-halo slots."""
        shape = (self.num_snapshots + 2,) + self.space_model.extended_shape
        return np.zeros(shape, dtype=self.space_model.dtype)

    def forward_arguments(self):
        """Keyword arguments of the backend's forward, before marshalling
        (the argument set of solver.py:131-163)."""
        return self._arguments(light=False)

    def _arguments(self, light):
        """forward_arguments(); with `light`, the wavefield slots, shot
        record and extended models are zero-stride placeholders of the right
        shape and dtype (no memory), for callers that build those arrays on
        a device (simwave_amd.device.DeviceForward)."""
        sm, tm = self.space_model, self.time_model
        src_points, src_values, src_offsets = \
            self.sources.interpolated_points_and_values
        rec_points, rec_values, rec_offsets = \
            self.receivers.interpolated_points_and_values
        if light:
            def hollow(shape):
                return np.broadcast_to(np.zeros((), dtype=sm.dtype), shape)
            ext = sm.extended_shape
            u_full = hollow((self.num_snapshots + 2,) + ext)
            vel, damp = hollow(ext), hollow(ext)
            den = None if sm.density_model is None else hollow(ext)
            record = hollow((tm.timesteps, self.receivers.count))
        else:
            u_full = self.u_full
            vel, den = sm.extended_velocity_model, sm.extended_density_model
            damp = sm.damping_mask
            record = self.shot_record
        return dict(
            u_full=u_full,
            velocity_model=vel,
            density_model=den,
            damping_mask=damp,
            wavelet=self.wavelet.values,
            wavelet_size=self.wavelet.timesteps,
            wavelet_count=self.wavelet.num_sources,
            second_order_fd_coefficients=sm.fd_coefficients(2),
            first_order_fd_coefficients=sm.fd_coefficients(1),
            boundary_condition=sm.boundary_condition,
            src_points_interval=src_points,
            src_points_interval_size=len(src_points),
            src_points_values=src_values,
            src_points_values_offset=src_offsets,
            src_points_values_size=len(src_values),
            rec_points_interval=rec_points,
            rec_points_interval_size=len(rec_points),
            rec_points_values=rec_values,
            rec_points_values_offset=rec_offsets,
            rec_points_values_size=len(rec_values),
            shot_record=record,
            num_sources=self.sources.count,
            num_receivers=self.receivers.count,
            grid_spacing=sm.grid_spacing,
            saving_stride=tm.saving_stride,
            dt=tm.dt,
            begin_timestep=1,
            end_timestep=tm.timesteps,
            space_order=sm.space_order,
            num_snapshots=u_full.shape[0],
        )

    def forward(self):
        """Runs the propagation; returns (u_full, shot_record).

        u_full keeps the damping layers but not the halo; for saving_stride
        0 it holds only the last wavefield (solver.py:112-169).
        """
        u_full, recv = self._middleware.exec(operator="forward",
                                             **self.forward_arguments())
        u_full = self.time_model.remove_time_halo_region(u_full)
        u_full = self.space_model.remove_halo_region(u_full)
        return u_full, recv
