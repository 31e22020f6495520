"""Frontend known answers (CPU).

Expected values are the ones the reference's own tests assert
(tests/test_space_model.py, tests/test_source.py, tests/test_time_model.py,
tests/test_compiler.py of HPCSys-Lab/simwave), plus closed forms where the
reference relies on a third-party library (findiff, scipy.special.i0).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import simwave_amd as sw
from simwave_amd import fd, kws


# ------------------------------------------------------------ SpaceModel
@pytest.mark.parametrize("dimension", [2, 3])
def test_space_model_properties(dimension):
    vel = 1500 * np.ones(shape=(101,) * dimension)
    den = 15 * np.ones(shape=(101,) * dimension)
    bbox = (0, 1000) * dimension
    spacing = (10,) * dimension
    sm = sw.SpaceModel(bounding_box=bbox, grid_spacing=spacing,
                       velocity_model=vel, density_model=den, space_order=4,
                       dtype=np.float32)
    assert sm.bounding_box == bbox
    assert sm.grid_spacing == spacing
    assert np.array_equal(sm.velocity_model, vel)
    assert np.array_equal(sm.density_model, den)
    assert sm.space_order == 4 and sm.dimension == dimension
    assert sm.dtype == np.float32
    assert sm.damping_length == (0.0,) * dimension * 2
    assert sm.boundary_condition == ("none",) * dimension * 2
    assert sm.damping_polynomial_degree == 3
    assert sm.damping_alpha == 0.001


@pytest.mark.parametrize("dimension, bbox, spacing, shape", [
    (2, (0, 1000, 0, 1000), (10, 10), (101, 101)),
    (2, (100., 1000, 0, 1250.5), (10, 10), (91, 126)),
    (2, (-100, 1000, 500, 1000), (20, 10), (56, 51)),
    (3, (0, 1000, 0, 1000, 0, 1000.0), (5, 10, 20), (201, 101, 51)),
    (3, (10, 100, -10, 100, 0.0, 100.0), (5, 5, 5), (19, 23, 21)),
])
def test_shape(dimension, bbox, spacing, shape):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=bbox, grid_spacing=spacing, velocity_model=vel)
    assert sm.shape == shape


@pytest.mark.parametrize("dimension, damping_length, space_order, extended_shape", [
    (2, 100, 2, (123, 123)),
    (2, 100, 4, (125, 125)),
    (2, 150, 16, (147, 147)),
    (2, (150, 100, 150, 100), 8, (134, 134)),
    (2, (150, 100, 0, 100), 8, (134, 119)),
    (3, 100, 2, (123, 123, 123)),
    (3, 100, 4, (125, 125, 125)),
    (3, 150, 16, (147, 147, 147)),
    (3, (150, 100, 150, 100, 150, 100), 8, (134, 134, 134)),
    (3, (150, 100, 0, 100, 0, 150), 8, (134, 119, 124)),
])
def test_extended_shape(dimension, damping_length, space_order, extended_shape):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 1000) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel,
                       space_order=space_order)
    sm.config_boundary(damping_length=damping_length)
    assert sm.extended_shape == extended_shape
    # every extended array has that shape
    assert sm.damping_mask.shape == extended_shape
    assert sm.extended_grid.shape == extended_shape
    assert sm.extended_velocity_model.shape == extended_shape


@pytest.mark.parametrize("dimension, damping_length, nbl", [
    (2, None, (0, 0, 0, 0)),
    (3, None, (0, 0, 0, 0, 0, 0)),
    (2, 120, (12, 12, 12, 12)),
    (3, 90, (9, 9, 9, 9, 9, 9)),
    (2, (50, 60, 75, 80), (5, 6, 7, 8)),
    (3, (0, 10, 8, 50, 20, 30), (0, 1, 0, 5, 2, 3)),
])
def test_nbl(dimension, damping_length, nbl):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 1000) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel)
    if damping_length is not None:
        sm.config_boundary(damping_length=damping_length)
    assert sm.nbl == nbl


@pytest.mark.parametrize("space_order, coeffs", [
    (2, [-2., 1.]),
    (4, [-2.5, 1.3333334, -0.08333334]),
    (8, [-2.8472223e+00, 1.6e+00, -2.0e-01, 2.5396826e-02, -1.7857143e-03]),
])
def test_fd_coefficients_known_answers(space_order, coeffs):
    vel = 1500 * np.ones(shape=(50, 50))
    sm = sw.SpaceModel(bounding_box=(0, 500, 0, 500), grid_spacing=(10, 10),
                       velocity_model=vel, space_order=space_order)
    assert np.allclose(sm.fd_coefficients(2), np.asarray(coeffs, dtype=sm.dtype))


@pytest.mark.parametrize("space_order", range(2, 21, 2))
def test_fd_closed_forms(space_order):
    """Centred weights vs the closed forms (m = space_order/2):
    c2[k] = 2(-1)^(k+1) (m!)^2 / (k^2 (m-k)! (m+k)!), c2[0] = -2 sum c2[k];
    c1[k] = (-1)^(k+1) (m!)^2 / (k (m-k)! (m+k)!), c1[0] = 0."""
    m = space_order // 2
    f = math.factorial
    c2 = [Fraction(2 * (-1) ** (k + 1) * f(m) ** 2, k * k * f(m - k) * f(m + k))
          for k in range(1, m + 1)]
    c2 = [-2 * sum(c2)] + c2
    c1 = [Fraction(0)] + [Fraction((-1) ** (k + 1) * f(m) ** 2, k * f(m - k) * f(m + k))
                          for k in range(1, m + 1)]
    assert list(fd._centred_weights(2, space_order)[m:]) == c2
    assert list(fd._centred_weights(1, space_order)[m:]) == c1
    assert np.array_equal(fd.half_coefficients(2, space_order),
                          np.array([float(c) for c in c2]))
    # antisymmetry / symmetry of the full stencils
    full1 = fd.coefficients(1, space_order)
    full2 = fd.coefficients(2, space_order)
    assert np.array_equal(full1, -full1[::-1])
    assert np.array_equal(full2, full2[::-1])


@pytest.mark.parametrize("dimension, space_order, halo", [
    (2, 2, (1,) * 4), (2, 4, (2,) * 4), (2, 20, (10,) * 4),
    (3, 2, (1,) * 6), (3, 8, (4,) * 6), (3, 10, (5,) * 6),
])
def test_halo_size(dimension, space_order, halo):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 500) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel,
                       space_order=space_order)
    assert sm.halo_size == halo


@pytest.mark.parametrize("dimension, damping_length, boundary_condition, degree, alpha", [
    (2, 500, "none", 2, 0.1),
    (2, (50, 50, 40, 40), "null_neumann", 4, 0.001),
    (3, 100, "null_dirichlet", 4, 0.001),
    (2, 75, ("none", "null_neumann", "none", "null_dirichlet"), 4, 0.001),
    (3, (5, 5, 5, 5, 6, 6), "null_dirichlet", 1, 0.001),
])
def test_config_boundary(dimension, damping_length, boundary_condition, degree, alpha):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 500) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel)
    sm.config_boundary(damping_length=damping_length,
                       boundary_condition=boundary_condition,
                       damping_polynomial_degree=degree, damping_alpha=alpha)
    if isinstance(damping_length, (float, int)):
        assert sm.damping_length == (damping_length,) * dimension * 2
    else:
        assert sm.damping_length == damping_length
    if isinstance(boundary_condition, str):
        assert sm.boundary_condition == (boundary_condition,) * dimension * 2
    else:
        assert sm.boundary_condition == boundary_condition
    assert sm.damping_alpha == alpha
    assert sm.damping_polynomial_degree == degree


def test_invalid_inputs():
    vel = 1500 * np.ones((20, 20))
    with pytest.raises(ValueError):
        sw.SpaceModel((0, 190, 0, 190), (10, 10), vel, space_order=3)
    with pytest.raises(ValueError):
        sw.SpaceModel((0, 190, 0, 190), (10, 10), vel, space_order=22)
    sm = sw.SpaceModel((0, 190, 0, 190), (10, 10), vel)
    with pytest.raises(ValueError):
        sm.config_boundary(boundary_condition="periodic")
    tm = sw.TimeModel(sm, tf=0.1)
    with pytest.raises(ValueError):
        tm.dt = -1.0
    with pytest.raises(ValueError):
        tm.dt = 1.0
    # out-of-box coordinates raise when positions are computed (source.py:75-77)
    with pytest.raises(Exception):
        sw.Source(sm, coordinates=[(500.0, 10.0)]).grid_positions


def test_damping_mask_closed_form():
    """Damping = alpha * d**degree with d the distance (1..nbl) from the
    inner domain along one axis; zero inside and in the halo."""
    vel = 1500 * np.ones((30, 40))
    sm = sw.SpaceModel((0, 290, 0, 390), (10, 10), vel, space_order=4)
    sm.config_boundary(damping_length=(50, 30, 0, 20), damping_polynomial_degree=3,
                       damping_alpha=0.001)
    mask = sm.damping_mask.astype(np.float64)
    h = 2
    nz0, nz1, nx0, nx1 = sm.nbl
    inner = mask[h + nz0:-h - nz1, h + nx0:-h - nx1]
    assert np.all(inner == 0)
    assert np.all(mask[:h] == 0) and np.all(mask[-h:] == 0)
    assert np.all(mask[:, :h] == 0) and np.all(mask[:, -h:] == 0)
    # a column through the inner x range ramps 1..nbl towards the top face
    col = mask[h:h + nz0, h + nx0 + 3]
    expected = 0.001 * np.arange(nz0, 0, -1, dtype=np.float64) ** 3
    assert np.allclose(col, expected, rtol=1e-6)


def test_interpolation_identity_and_linear():
    """Same-shape data is returned unchanged; resampled data matches scipy's
    linear RegularGridInterpolator on the full node grid."""
    rng = np.random.default_rng(0)
    data = rng.uniform(1500, 3000, size=(21, 31)).astype(np.float32)
    sm = sw.SpaceModel((0, 200, 0, 300), (10, 10), data)
    assert np.array_equal(sm.velocity_model, data)
    sm2 = sw.SpaceModel((0, 200, 0, 300), (10, 7.5), data)
    from scipy.interpolate import RegularGridInterpolator
    z = np.linspace(0, 200, 21)
    x = np.linspace(0, 300, 31)
    it = RegularGridInterpolator((z, x), data)
    zz = np.linspace(0, 200, sm2.shape[0])
    xx = np.linspace(0, 300, sm2.shape[1])
    X, Z = np.meshgrid(xx, zz)
    assert np.array_equal(sm2.velocity_model, np.float32(it((Z, X))))


# ------------------------------------------------------------ TimeModel
def test_time_model_properties():
    vel = 1500 * np.ones(shape=(10, 10))
    sm = sw.SpaceModel(bounding_box=(0, 100, 0, 100), grid_spacing=(10, 10),
                       velocity_model=vel)
    tm = sw.TimeModel(space_model=sm, tf=1.0, t0=0.0, saving_stride=1)
    tm.dt = 0.001
    assert tm.space_model == sm
    assert tm.tf == 1.0 and tm.t0 == 0.0 and tm.saving_stride == 1
    assert tm.dt == sm.dtype(0.001)


@pytest.mark.parametrize("dt, tf, t0, saving_stride, timesteps", [
    (0.001, 1.0, 0.0, 0, 1001),
    (0.001, 1.0, 0.5, 0, 501),
    (0.002, 2.0, 0.0, 0, 1001),
    (0.001, 1.0, 0.0, 1, 1001),
    (0.001, 1.0, 0.5, 1, 501),
    (0.002, 2.0, 0.0, 1, 1001),
    (0.001, 1.0, 0.0, 2, 1001),
    (0.001, 1.0, 0.5, 6, 505),
    (0.002, 2.0, 0.0, 3, 1003),
])
def test_timesteps(dt, tf, t0, saving_stride, timesteps):
    vel = 1500 * np.ones(shape=(10, 10))
    sm = sw.SpaceModel(bounding_box=(0, 100, 0, 100), grid_spacing=(10, 10),
                       velocity_model=vel)
    tm = sw.TimeModel(space_model=sm, tf=tf, t0=t0, saving_stride=saving_stride)
    tm.dt = dt
    assert tm.timesteps == timesteps


def test_cfl_dt_closed_form():
    """dt = sqrt(4 / (dim * sum|c2|)) * min(h) / max(v) (Lines et al. 1999)."""
    vel = np.full((20, 30), 1500.0, dtype=np.float32)
    vel[10:] = 3000.0
    sm = sw.SpaceModel((0, 190, 0, 290), (10, 10), vel, space_order=4)
    tm = sw.TimeModel(sm, tf=1.0)
    # so4: |c2| sums to 1/12 + 4/3 + 5/2 + 4/3 + 1/12 = 16/3
    expected = math.sqrt(4 / (2 * 16 / 3)) * 10 / 3000
    assert abs(float(tm._dt) - expected) < 1e-15


# ---------------------------------------------------------------- Source
@pytest.mark.parametrize("dimension, coords", [
    (2, [(0, 25)]), (2, [(0, 25), [250, 250]]), (3, [(0.5, 50.8, 500)]),
])
def test_source_properties(dimension, coords):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 500) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel)
    src = sw.Source(sm, coordinates=coords, window_radius=8)
    assert src.space_model == sm and src.window_radius == 8
    assert src.count == len(coords)
    assert np.array_equal(src.coordinates, sm.dtype(coords))


@pytest.mark.parametrize("dimension, bbox, spacing, coords, grid_positions", [
    (2, (0, 5120, 0, 5120), (10, 10), (0, 512), (0, 51.2)),
    (2, (0, 5120, 0, 5120), (10, 5), (120, 500), (12, 100)),
    (3, (0, 500, 20, 500, 0, 200), (20, 20, 20), (10, 20, 100), (0.5, 0.0, 5.0)),
])
def test_grid_positions(dimension, bbox, spacing, coords, grid_positions):
    vel = np.full((50,) * dimension, 1500.0, dtype=np.float32)
    sm = sw.SpaceModel(bounding_box=bbox, grid_spacing=spacing, velocity_model=vel)
    src = sw.Source(sm, coordinates=coords, window_radius=4)
    assert np.array_equal(src.grid_positions,
                          np.asarray([grid_positions], dtype=sm.dtype))


@pytest.mark.parametrize("dimension, damping_length, space_order, coords, expected", [
    (2, 500, 2, (0, 0), [(51, 51)]),
    (2, 0, 2, (0, 0), [(1, 1)]),
    (2, 0, 4, (250, 100), [(27, 12)]),
    (2, 50, 16, (255, 255), [(38.5, 38.5)]),
    (3, (50, 0, 50, 0, 0, 0), 16, (255, 255, 0), [(38.5, 38.5, 8)]),
    (3, 2, 4, (250, 100, 100), [(27, 12, 12)]),
])
def test_adjusted_grid_positions(dimension, damping_length, space_order, coords, expected):
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 500) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel,
                       space_order=space_order)
    sm.config_boundary(damping_length=damping_length)
    src = sw.Source(sm, coordinates=coords)
    assert np.array_equal(src.adjusted_grid_positions, sm.dtype(expected))


@pytest.mark.parametrize("dimension, window_radius, damping_length, space_order, coords, points, values", [
    (2, 1, 50, 2, (250, 250), [30, 32, 30, 32],
     [-1.9556131e-08, 1.0000001e+00, -1.9556131e-08,
      -1.9556131e-08, 1.0000001e+00, -1.9556131e-08]),
    (3, 4, 0, 4, (255, 250, 100), [24, 31, 23, 31, 8, 16],
     [-5.1983586e-03, 3.6357623e-02, -1.3933791e-01, 6.0838646e-01,
      6.0838646e-01, -1.3933791e-01, 3.6357623e-02, -5.1983586e-03,
      3.1170111e-10, -3.7239498e-10, 1.2889303e-08, -2.3163784e-08,
      9.9999994e-01, -2.3163784e-08, 1.2889303e-08, -3.7239498e-10,
      3.1170111e-10, 3.1170111e-10, -3.7239498e-10, 1.2889303e-08,
      -2.3163784e-08, 9.9999994e-01, -2.3163784e-08, 1.2889303e-08,
      -3.7239498e-10, 3.1170111e-10]),
])
def test_interpolated_points_and_values(dimension, window_radius, damping_length,
                                        space_order, coords, points, values):
    """Exact float32 windowed-sinc weights of the reference's fixtures."""
    vel = 1500 * np.ones(shape=(50,) * dimension)
    sm = sw.SpaceModel(bounding_box=(0, 500) * dimension,
                       grid_spacing=(10,) * dimension, velocity_model=vel,
                       space_order=space_order)
    sm.config_boundary(damping_length=damping_length)
    src = sw.Source(sm, coordinates=coords, window_radius=window_radius)
    p, v, off = src.interpolated_points_and_values
    assert np.array_equal(p, np.asarray(points))
    assert np.array_equal(v, sm.dtype(values))
    assert list(off) == [0, len(values)]


@pytest.mark.parametrize("half_width", range(1, 11))
def test_kws_vectorised_matches_full_axis(half_width):
    """The windowed evaluation equals the full-axis evaluation + NaN scan
    (kws.py:44-135 of the reference) for many off-grid positions."""
    rng = np.random.default_rng(half_width)
    n = 60
    for pos in np.concatenate([rng.uniform(0, n - 1, 40), [0.0, n - 1.0, 7.0, 7.5]]):
        pos = np.float32(pos)
        full = kws.kaiser_windowing_sinc(n, pos, half_width)
        b, e, vals = kws.get_kws_valid_points(full)
        pts, v2 = kws.points_and_values((n,), np.array([[pos]]), half_width)
        assert (int(pts[0]), int(pts[1])) == (b, e)
        assert np.array_equal(v2, vals)


def test_kws_window_closed_form():
    """Kaiser window times sinc (Hicks 2002): weight 1 at an on-grid
    position; elsewhere i0(b sqrt(1-(x/hw)^2)) / i0(b) * sin(pi x)/(pi x)."""
    from scipy.special import i0
    w = kws.kaiser_windowing_sinc(21, np.float32(10.0), 4)
    assert w[10] == np.float32(1.0) or abs(w[10] - 1) < 1e-6
    x = np.float32(3.3)
    w = kws.kaiser_windowing_sinc(21, np.float32(10.0) - x, 4)[10]
    b = 6.31
    expect = i0(b * math.sqrt(1 - (3.3 / 4) ** 2)) / i0(b) * math.sin(math.pi * 3.3) / (math.pi * 3.3)
    assert abs(float(w) - expect) < 1e-5


# --------------------------------------------------------------- Wavelets
def test_ricker_closed_form():
    vel = 1500 * np.ones((20, 20))
    sm = sw.SpaceModel((0, 190, 0, 190), (10, 10), vel, dtype=np.float64)
    tm = sw.TimeModel(sm, tf=0.3)
    f = 15.0
    wav = sw.RickerWavelet(f, tm).values
    t = np.linspace(0.0, 0.3, tm.timesteps)
    r = np.pi * f * (t - 1 / f)
    assert np.allclose(wav, (1 - 2 * r ** 2) * np.exp(-r ** 2), atol=1e-14)
    assert wav.dtype == np.float64


def test_multiwavelet():
    vel = 1500 * np.ones((20, 20))
    sm = sw.SpaceModel((0, 190, 0, 190), (10, 10), vel)
    tm = sw.TimeModel(sm, tf=0.1)
    vals = np.ones((tm.timesteps, 3))
    mw = sw.MultiWavelet(vals, tm)
    assert mw.num_sources == 3 and mw.timesteps == tm.timesteps
    assert mw.values.dtype == np.float32
    with pytest.raises(ValueError):
        sw.MultiWavelet(np.ones((tm.timesteps + 1, 3)), tm)


# --------------------------------------------------------------- Compiler
@pytest.mark.parametrize("cc, language, cflags, expected", [
    ("gcc", "c", None, "-O3 -fPIC -Wall -std=c99 -shared"),
    ("gcc", "c", "-O3 -fPIC", "-O3 -fPIC -shared"),
    ("icc", "c", "-O3 -shared", "-O3 -shared"),
    ("icc", "c", "-O3", "-O3 -shared"),
    ("clang", "c", None, "-O3 -fPIC -Wall -std=c99 -shared"),
    ("gcc", "cpu_openmp", "-O3 -fPIC", "-O3 -fPIC -shared -fopenmp"),
    ("pgcc", "cpu_openmp", "-shared", "-shared -mp"),
    ("icc", "cpu_openmp", "-O3", "-O3 -shared -openmp"),
    ("clang", "cpu_openmp", "-O3", "-O3 -shared -fopenmp"),
    ("foo", "cpu_openmp", "-O3", "-O3 -shared"),
])
def test_compiler_cflags(cc, language, cflags, expected):
    assert sw.Compiler(cc=cc, language=language, cflags=cflags).cflags == expected


@pytest.mark.parametrize("cc, omp_flag", [
    ("gcc", "-fopenmp"), ("icc", "-openmp"), ("pgcc", "-mp"),
    ("clang", "-fopenmp"), ("foo", None),
])
def test_compiler_openmp_flag(cc, omp_flag):
    assert sw.Compiler(cc=cc).get_openmp_flag() == omp_flag


def test_compiler_languages():
    for lang in ("hip", "c", "cpu_openmp"):
        assert sw.Compiler(language=lang).language == lang
    assert sw.Compiler().language == "hip"
    for bad in ("cpu_mpi", "cuda", "gpu_openacc"):
        with pytest.raises(ValueError):
            sw.Compiler(language=bad)
    with pytest.raises(ValueError):
        sw.Compiler(language="c").compile(2, "constant_density", "-DFLOAT", "forward")


def test_compiler_resolves_hip_library():
    from simwave_amd import build
    path = sw.Compiler().compile(3, "constant_density", "-DFLOAT", "forward")
    assert path == build.library_path(3, "cd", "f32")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_vectorised_positions_match_scalar_expressions(dtype):
    """grid_positions / adjusted_grid_positions, evaluated on whole arrays,
    equal the reference's per-coordinate scalar expressions
    (source.py:62-121 of the reference: dtype scalars throughout)."""
    rng = np.random.default_rng(3)
    sm = sw.SpaceModel(bounding_box=(0, 4787.5, -13.3, 4000.7, 0.1, 999.9),
                       grid_spacing=(12.5, 7.3, 3.3),
                       velocity_model=np.full((20, 20, 20), 1500.0, dtype=dtype),
                       space_order=8, dtype=dtype)
    sm.config_boundary(damping_length=(100.0, 50.0, 70.0, 10.0, 33.0, 0.0))
    coords = np.stack([rng.uniform(0, 4787.5, 2000), rng.uniform(-13.3, 4000.7, 2000),
                       rng.uniform(0.1, 999.9, 2000)], 1)
    src = sw.Source(sm, coordinates=coords)
    box, h = sm.bounding_box, sm.grid_spacing
    origin = [sm.nbl[2 * a] + sm.halo_size[2 * a] for a in range(3)]
    ref = np.asarray([tuple((c[a] - box[2 * a]) / h[a] for a in range(3))
                      for c in src.coordinates], dtype=dtype)
    ref_adj = np.asarray([tuple(p[a] + origin[a] for a in range(3)) for p in ref],
                         dtype=dtype)
    assert np.array_equal(src.grid_positions, ref)
    assert np.array_equal(src.adjusted_grid_positions, ref_adj)


def test_points_and_values_layout_matches_per_position_calls():
    """The vectorised gather lays weights out per position, per axis, as
    concatenating single-position results does (source.py:140-159)."""
    rng = np.random.default_rng(5)
    shape = (61, 47, 73)
    pos = np.stack([rng.uniform(5, n - 6, 40) for n in shape], 1).astype(np.float32)
    pts, vals = kws.points_and_values(shape, pos, 4)
    one = [kws.points_and_values(shape, pos[i:i + 1], 4) for i in range(len(pos))]
    assert np.array_equal(pts, np.concatenate([p for p, _ in one]))
    assert np.array_equal(vals, np.concatenate([v for _, v in one]))


def test_padding_helpers_equal_numpy_pad():
    from simwave_amd.model import _pad_edge, _pad_zero
    rng = np.random.default_rng(0)
    for shape in [(5, 7), (3, 4, 6), (1, 1), (2, 3, 1)]:
        a = rng.standard_normal(shape).astype(np.float32)
        for w in [((0, 0),) * len(shape), ((1, 2),) * len(shape),
                  tuple((i, len(shape) - i + 1) for i in range(len(shape)))]:
            assert np.array_equal(_pad_edge(a, w), np.pad(a, w, mode="edge"))
            assert np.array_equal(_pad_zero(a, w), np.pad(a, w))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("deg,alpha", [(3, 0.001), (2, 0.002), (1.5, 0.01), (4, 0.0005)])
def test_damping_mask_factors_rebuild_the_mask(dim, dtype, deg, alpha):
    """damping_mask == small[ix(index)] bit for bit, for uneven layers."""
    rng = np.random.default_rng(dim * 7 + int(deg * 10))
    for _ in range(3):
        shape = tuple(int(rng.integers(5, 23)) for _ in range(dim))
        h = 10.0
        box = []
        for n in shape:
            box += [0.0, (n - 1) * h]
        sm = sw.SpaceModel(bounding_box=tuple(box), grid_spacing=(h,) * dim,
                           velocity_model=np.full(shape, 2000.0, dtype=dtype),
                           space_order=int(rng.choice([2, 4, 8, 16])), dtype=dtype)
        lengths = tuple(float(rng.integers(0, 9)) * h for _ in range(2 * dim))
        sm.config_boundary(damping_length=lengths,
                           boundary_condition=("null_neumann",) * (2 * dim),
                           damping_polynomial_degree=deg, damping_alpha=alpha)
        small, index = sm.damping_mask_factors()
        full = sm.damping_mask
        rebuilt = small[np.ix_(*index)]
        assert rebuilt.dtype == full.dtype
        assert np.array_equal(rebuilt, full)


@pytest.mark.parametrize("dim", [2, 3])
def test_extension_index_rebuilds_the_extended_models(dim):
    rng = np.random.default_rng(dim)
    shape = (11, 14, 9)[:dim]
    h = 5.0
    box = []
    for n in shape:
        box += [0.0, (n - 1) * h]
    vel = rng.uniform(1500, 3000, size=shape).astype(np.float32)
    den = rng.uniform(1000, 2000, size=shape).astype(np.float32)
    sm = sw.SpaceModel(bounding_box=tuple(box), grid_spacing=(h,) * dim,
                       velocity_model=vel, density_model=den, space_order=6,
                       dtype=np.float32)
    sm.config_boundary(damping_length=tuple(float(w) * h for w in (3, 0, 5, 2, 1, 4)[:2 * dim]),
                       boundary_condition=("null_dirichlet",) * (2 * dim))
    index = sm.extension_index()
    assert np.array_equal(sm.velocity_model[np.ix_(*index)], sm.extended_velocity_model)
    assert np.array_equal(sm.density_model[np.ix_(*index)], sm.extended_density_model)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_points_and_values_with_repeated_coordinates(dtype):
    """Receiver grids repeat coordinates per axis (evaluated once per
    distinct value): same intervals and weights as per-position calls."""
    rng = np.random.default_rng(11)
    shape = (40, 57, 66)
    xs = rng.uniform(5, 50, 7)
    ys = rng.uniform(5, 60, 6)
    pos = np.array([(3.25, x, y) for x in xs for y in ys] + [(3.25, xs[0], 9.0)] * 3,
                   dtype=dtype)
    pos = pos[rng.permutation(len(pos))]
    pts, vals = kws.points_and_values(shape, pos, 4)
    one = [kws.points_and_values(shape, pos[i:i + 1], 4) for i in range(len(pos))]
    assert np.array_equal(pts, np.concatenate([p for p, _ in one]))
    assert np.array_equal(vals, np.concatenate([v for _, v in one]))
