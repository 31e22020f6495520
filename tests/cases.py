"""Experiment builders shared by the CPU and GPU parity tests.

Each builder returns a zero-argument callable producing the keyword
arguments of simwave_amd.Solver (space_model, time_model, sources,
receivers, wavelet), so the same experiment can be run on several backends.
"""
import numpy as np

import simwave_amd as sw

BC_NAMES = ("none", "null_dirichlet", "null_neumann")


def layered(shape, vmin=1500.0, vmax=3000.0, dtype=np.float32, seed=0,
            noise=0.0):
    """Velocity with flat layers along z, optionally perturbed."""
    nz = shape[0]
    vel = np.empty(shape, dtype=np.float64)
    layers = np.linspace(vmin, vmax, 4)
    for i in range(nz):
        vel[i] = layers[min(3, 4 * i // nz)]
    if noise:
        rng = np.random.default_rng(seed)
        vel *= 1.0 + noise * rng.uniform(-1, 1, size=shape)
    return vel.astype(dtype)


def make_case(dim=2, n=48, h=10.0, space_order=4, density=False,
              dtype=np.float32, stride=0, bc=None, damping=40.0, tf=0.12,
              src=None, rec=None, window=4, multi=False, seed=0, dt=None,
              shape=None):
    """Builds an experiment; returns a callable for harness.run()."""

    def setup():
        shp = tuple(shape) if shape is not None else (n,) * dim
        box = []
        for a in range(dim):
            box += [0.0, (shp[a] - 1) * h]
        vel = layered(shp, dtype=dtype, seed=seed, noise=0.05)
        den = None
        if density:
            rng = np.random.default_rng(seed + 1)
            den = (310.0 * vel.astype(np.float64) ** 0.25
                   * (1 + 0.05 * rng.uniform(-1, 1, size=shp))).astype(dtype)
        sm = sw.SpaceModel(bounding_box=tuple(box), grid_spacing=(h,) * dim,
                           velocity_model=vel, density_model=den,
                           space_order=space_order, dtype=dtype)
        faces = bc if bc is not None else ("null_neumann", "null_dirichlet") * dim
        sm.config_boundary(damping_length=damping, boundary_condition=faces,
                           damping_polynomial_degree=3, damping_alpha=0.001)
        tm = sw.TimeModel(space_model=sm, tf=tf, saving_stride=stride, dt=dt)
        ext = [(shp[a] - 1) * h for a in range(dim)]
        if src is None:
            s_coords = [tuple(0.5 * e + 0.37 * h for e in ext)]
        else:
            s_coords = src
        if rec is None:
            r_coords = [(0.3 * ext[0], x * ext[1] / 7.0) + tuple(
                0.45 * e for e in ext[2:]) for x in range(8)]
        else:
            r_coords = rec
        source = sw.Source(sm, coordinates=s_coords, window_radius=window)
        receiver = sw.Receiver(sm, coordinates=r_coords, window_radius=window)
        if multi:
            rng = np.random.default_rng(seed + 7)
            base = sw.RickerWavelet(15.0, tm).values
            vals = np.stack([base * (1 + 0.5 * rng.uniform(-1, 1))
                             for _ in range(source.count)], axis=1)
            vals[::5, 0] = 0  # exercise the zero-amplitude skip
            wav = sw.MultiWavelet(vals, tm)
        else:
            wav = sw.RickerWavelet(15.0, tm)
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=wav)

    return setup


def readme_2d():
    """The reference README example (2D, so4, 512^2 resampled to 513^2)."""

    def setup():
        vel = np.zeros(shape=(512, 512), dtype=np.float32)
        vel[:] = 1500.0
        vel[100:] = 2000.0
        sm = sw.SpaceModel(bounding_box=(0, 5120, 0, 5120),
                           grid_spacing=(10, 10), velocity_model=vel,
                           space_order=4, dtype=np.float32)
        sm.config_boundary(damping_length=0,
                           boundary_condition=("null_neumann", "null_dirichlet",
                                               "none", "null_dirichlet"),
                           damping_polynomial_degree=3, damping_alpha=0.001)
        tm = sw.TimeModel(space_model=sm, tf=1.0, saving_stride=0)
        source = sw.Source(sm, coordinates=[(2560, 2560)], window_radius=1)
        receiver = sw.Receiver(sm, coordinates=[(2560, i) for i in range(0, 5112, 10)],
                               window_radius=1)
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=sw.RickerWavelet(10.0, tm))

    return setup


def golden_2d(space_order, density=False):
    """Configuration of the reference's golden 2D wavefields
    (tests/reference_solution/generator.py:6-50 of the reference); with
    density=True the unit-density variable-density run its tests compare
    against the same wavefield (tests/test_solution.py:68-81)."""

    def setup():
        shape = (500, 500)
        vel = np.zeros(shape=shape, dtype=np.float32)
        vel[:] = 1500.0
        vel[shape[0] // 2:] = 2000.0
        den = np.ones(shape, dtype=np.float32) if density else None
        sm = sw.SpaceModel(bounding_box=(0, 5000, 0, 5000), grid_spacing=(10, 10),
                           velocity_model=vel, density_model=den,
                           space_order=space_order, dtype=np.float32)
        sm.config_boundary(damping_length=100,
                           boundary_condition=("null_neumann", "null_dirichlet",
                                               "none", "null_dirichlet"),
                           damping_polynomial_degree=3, damping_alpha=0.001)
        tm = sw.TimeModel(space_model=sm, tf=1.0, saving_stride=0)
        source = sw.Source(sm, coordinates=[(2500, 2500)], window_radius=1)
        receiver = sw.Receiver(sm, coordinates=[(2500, 2500)], window_radius=1)
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=sw.RickerWavelet(10.0, tm))

    return setup


def golden_3d(space_order, density=False):
    """The reference's 3D golden configuration (generator.py:20-32); its
    .npy files are not shipped, so parity uses the oracle."""

    def setup():
        shape = (100, 100, 100)
        vel = np.zeros(shape=shape, dtype=np.float32)
        vel[:] = 1500.0
        vel[shape[0] // 2:] = 2000.0
        den = None
        if density:
            den = np.ones(shape, dtype=np.float32)
        sm = sw.SpaceModel(bounding_box=(0, 1000, 0, 1000, 0, 1000),
                           grid_spacing=(10, 10, 10), velocity_model=vel,
                           density_model=den, space_order=space_order,
                           dtype=np.float32)
        sm.config_boundary(damping_length=50,
                           boundary_condition=("null_neumann", "null_dirichlet",
                                               "none", "null_dirichlet",
                                               "null_neumann", "null_dirichlet"),
                           damping_polynomial_degree=3, damping_alpha=0.001)
        tm = sw.TimeModel(space_model=sm, tf=0.4, saving_stride=0)
        source = sw.Source(sm, coordinates=[(500, 495, 505)], window_radius=1)
        receiver = sw.Receiver(sm, coordinates=[(500, 495, 505)], window_radius=1)
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=sw.RickerWavelet(15.0, tm))

    return setup


def saving_2d(saving_stride, density=False):
    """The reference's snapshot-stride experiment in 2D
    (tests/test_u_saving.py:7-71 of the reference)."""

    def setup():
        vel = np.full((512, 512), 1500.0, dtype=np.float32)
        vel[250:] = 3000.0
        den = np.full((512, 512), 5.0, dtype=np.float32) if density else None
        sm = sw.SpaceModel(bounding_box=(0, 5120, 0, 5120), grid_spacing=(10, 10),
                           velocity_model=vel, density_model=den, space_order=4,
                           dtype=np.float32)
        sm.config_boundary(damping_length=0,
                           boundary_condition=("null_neumann", "null_dirichlet",
                                               "none", "null_dirichlet"),
                           damping_polynomial_degree=3, damping_alpha=0.001)
        tm = sw.TimeModel(space_model=sm, tf=1.0, saving_stride=saving_stride)
        source = sw.Source(sm, coordinates=[(2560, 2560)], window_radius=4)
        receiver = sw.Receiver(sm, coordinates=[(2560, i) for i in range(0, 5120, 10)],
                               window_radius=4)
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=sw.RickerWavelet(10.0, tm))

    return setup


def saving_3d(saving_stride, density=False):
    """The reference's snapshot-stride experiment in 3D
    (tests/test_u_saving.py:74-143 of the reference)."""

    def setup():
        vel = np.full((100, 100, 100), 1500.0, dtype=np.float32)
        den = np.full((100, 100, 100), 5.0, dtype=np.float32) if density else None
        sm = sw.SpaceModel(bounding_box=(0, 1000, 0, 1000, 0, 1000),
                           grid_spacing=(10, 10, 10), velocity_model=vel,
                           density_model=den, space_order=4, dtype=np.float32)
        sm.config_boundary(damping_length=0,
                           boundary_condition=("null_neumann", "null_dirichlet",
                                               "null_dirichlet", "null_dirichlet",
                                               "null_dirichlet", "null_dirichlet"),
                           damping_polynomial_degree=3, damping_alpha=0.001)
        tm = sw.TimeModel(space_model=sm, tf=0.4, saving_stride=saving_stride)
        source = sw.Source(sm, coordinates=[(500, 500, 500)], window_radius=4)
        receiver = sw.Receiver(sm, coordinates=[(500, 500, i) for i in range(0, 1000, 10)],
                               window_radius=4)
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=sw.RickerWavelet(15.0, tm))

    return setup


def hankel_accuracy_case(space_order=8, spacing=0.5, dt=0.1, tf=150.075):
    """The reference's analytic-accuracy experiment: 2D homogeneous 1.5 km/s,
    0.09 kHz Ricker at (200, 200), one receiver at (260, 260)
    (tests/test_convergence.py:9-87 of the reference)."""

    def setup():
        sm = sw.SpaceModel(bounding_box=(-40, 440, -40, 440),
                           grid_spacing=(spacing, spacing),
                           velocity_model=1.5 * np.ones((100, 100), dtype=np.float64),
                           space_order=space_order)
        tm = sw.TimeModel(space_model=sm, t0=0, tf=tf)
        tm.dt = dt
        source = sw.Source(sm, coordinates=[(200, 200)])
        receiver = sw.Receiver(sm, coordinates=[(260, 260)])
        return dict(space_model=sm, time_model=tm, sources=source,
                    receivers=receiver, wavelet=sw.RickerWavelet(0.09, tm))

    return setup


def hankel_reference(setup, freq=0.09, src=(200.0, 200.0), rec=(260.0, 260.0)):
    """Analytic 2D Green's function response to the Ricker wavelet (the
    zeroth-order Hankel function of the second kind, evaluated in the
    frequency domain), as in tests/test_convergence.py:27-49 of the
    reference.  Returns the trace over the run's time steps."""
    from scipy.special import hankel2
    parts = setup()
    sm, tm0 = parts["space_model"], parts["time_model"]
    tm = sw.TimeModel(space_model=sm, t0=0, tf=3000)
    tm.dt = tm0.dt
    wav = sw.RickerWavelet(freq, tm).values
    nt = tm.timesteps
    nf = int(nt / 2 + 1)
    df = 1 / tm.tf
    f = df * np.arange(nf)
    qf = np.fft.fft(wav)[:nf]
    r = np.hypot(rec[0] - src[0], rec[1] - src[1])
    k = 2 * np.pi * f / np.unique(sm.velocity_model)
    g = np.zeros(nf, dtype=complex)
    g[1:-1] = hankel2(0, k[1:-1] * r)
    trace = np.real(np.fft.ifft(-1j * np.pi * g * qf, nt)) / (2 * np.pi)
    return trace[:tm0.timesteps]


# Small constant-density shots for the adjoint (stride 0, source boxes in
# the interior): damping, every boundary code, several sources, 3D.
ADJOINT_CASES = {
    "2d_damped": dict(dim=2, n=30, space_order=4, tf=0.08, damping=40.0),
    "2d_neumann_dirichlet": dict(dim=2, n=28, space_order=8, tf=0.07, damping=30.0,
                                 bc=("null_neumann", "null_dirichlet", "none",
                                     "null_neumann")),
    "2d_multi": dict(dim=2, n=30, space_order=2, tf=0.06, multi=True,
                     src=[(120.0, 130.0), (160.0, 100.0)], window=2),
    "3d_mixed": dict(dim=3, n=18, space_order=4, tf=0.04, damping=30.0,
                     bc=("null_neumann", "null_dirichlet", "null_dirichlet",
                         "null_neumann", "none", "null_neumann")),
    "3d_so8": dict(dim=3, shape=(22, 16, 19), space_order=8, tf=0.04, damping=20.0),
}


def adjoint_shot(name, dtype=np.float64, **over):
    """(setup parts, Solver, normalised forward keyword arguments)."""
    from simwave_amd.middleware import Middleware
    params = dict(ADJOINT_CASES[name], dtype=dtype)
    params.update(over)
    parts = make_case(**params)()
    solver = sw.Solver(**parts)
    kw = Middleware.normalize(**solver.forward_arguments())
    return parts, solver, kw
