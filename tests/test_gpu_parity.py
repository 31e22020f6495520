"""HIP forward vs the oracle on the same inputs (needs an MI355X).

Each case runs the full simwave_amd frontend twice, once on the CPU oracle
library and once on the HIP library, and compares every returned snapshot
and the shot record.  The kernels reproduce the reference C's evaluation
order, so the bar is bit equality; the relative-L2 bound of the contract
(1e-5 fp32, 1e-12 fp64) is reported in the failure message.
"""
import os

import numpy as np
import pytest

import cases
import harness

pytestmark = pytest.mark.gpu

F32, F64 = np.float32, np.float64

CASES = {
    # 2D constant density, every supported radius
    **{"2d_cd_so%d" % so: dict(dim=2, space_order=so) for so in (2, 4, 6, 8, 12, 16, 20)},
    "2d_vd_so4": dict(dim=2, space_order=4, density=True),
    "2d_vd_so16": dict(dim=2, space_order=16, density=True, n=40),
    "2d_cd_f64_so8": dict(dim=2, space_order=8, dtype=F64),
    "2d_vd_f64_so8": dict(dim=2, space_order=8, density=True, dtype=F64),
    "2d_vd_f64_so16": dict(dim=2, space_order=16, density=True, dtype=F64, n=40),
    # boundary codes on every face
    "2d_bc_none": dict(dim=2, bc=("none",) * 4, damping=0.0, tf=0.3),
    "2d_bc_dirichlet": dict(dim=2, bc=("null_dirichlet",) * 4, damping=0.0, tf=0.3),
    "2d_bc_neumann": dict(dim=2, bc=("null_neumann",) * 4, damping=0.0, tf=0.3),
    "2d_bc_mixed": dict(dim=2, bc=("none", "null_neumann", "null_dirichlet",
                                   "null_neumann"), damping=0.0, tf=0.3),
    "2d_bc_damped": dict(dim=2, bc=("null_neumann", "null_dirichlet", "none",
                                    "null_neumann"), damping=30.0, tf=0.4),
    # snapshot strides
    **{"2d_stride%d" % s: dict(dim=2, stride=s, tf=0.08) for s in (1, 2, 3, 4, 5)},
    # sources: at the surface without damping (boxes reach the halo),
    # several overlapping with per-source wavelets, more than the fused cap
    "2d_src_halo": dict(dim=2, damping=0.0, src=[(0.0, 200.0), (5.0, 470.0)],
                        multi=True),
    "2d_src_overlap": dict(dim=2, src=[(200.0, 200.0), (215.0, 230.0), (230.0, 215.0)],
                           multi=True, window=6),
    "2d_src_many": dict(dim=2, src=[(100.0 + 3 * i, 50.0 + 5 * i) for i in range(70)],
                        window=2),
    "2d_tiny_so20": dict(dim=2, space_order=20, n=8, damping=0.0, window=2,
                         src=[(35.0, 35.0)], rec=[(20.0, 40.0)]),
    # 3D
    **{"3d_cd_so%d" % so: dict(dim=3, n=28, space_order=so, tf=0.08)
       for so in (2, 4, 8, 12, 20)},
    "3d_vd_so4": dict(dim=3, n=28, space_order=4, density=True, tf=0.08),
    "3d_vd_so8": dict(dim=3, n=28, space_order=8, density=True, tf=0.08),
    "3d_vd_quirk": dict(dim=3, shape=(26, 30, 24), space_order=4, density=True,
                        tf=0.08),
    "3d_vd_bc": dict(dim=3, n=24, density=True, damping=0.0,
                     bc=("null_neumann", "null_dirichlet", "null_neumann", "none",
                         "null_dirichlet", "null_neumann"),
                     src=[(0.0, 100.0, 120.0), (8.0, 5.0, 200.0)], multi=True, tf=0.06),
    "3d_vd_so16_wide": dict(dim=3, shape=(30, 90, 90), space_order=16, density=True,
                            damping=30.0, tf=0.03),
    "3d_cd_f64_so8": dict(dim=3, n=26, space_order=8, dtype=F64, tf=0.08),
    "3d_vd_f64_so4": dict(dim=3, n=26, space_order=4, density=True, dtype=F64,
                          tf=0.08),
    # variable density from space order 12: density derivative sums made
    # once per run (the z-marching kernels)
    "2d_vd_so12": dict(dim=2, space_order=12, density=True, n=44),
    "2d_vd_so20_bc": dict(dim=2, space_order=20, density=True, n=52, damping=0.0,
                          bc=("null_neumann", "null_dirichlet", "null_neumann", "none")),
    "3d_vd_so12": dict(dim=3, n=34, space_order=12, density=True, tf=0.05),
    "3d_vd_so20_bc": dict(dim=3, n=40, space_order=20, density=True, damping=0.0,
                          bc=("null_dirichlet", "null_neumann", "null_neumann", "none",
                              "null_dirichlet", "null_neumann"), tf=0.04),
    "3d_vd_f64_so12": dict(dim=3, n=30, space_order=12, density=True, dtype=F64,
                           tf=0.05),
    "3d_bc_neumann": dict(dim=3, n=26, bc=("null_neumann",) * 6, damping=0.0,
                          tf=0.15),
    "3d_bc_dirichlet": dict(dim=3, n=26, bc=("null_dirichlet",) * 6, damping=0.0,
                            tf=0.15),
    "3d_bc_mixed": dict(dim=3, n=26, bc=("null_dirichlet", "null_neumann", "none",
                                         "null_neumann", "null_neumann",
                                         "null_dirichlet"), damping=0.0, tf=0.15),
    "3d_stride2": dict(dim=3, n=24, stride=2, tf=0.05),
    "3d_stride3": dict(dim=3, n=24, stride=3, tf=0.05),
    "3d_src_halo": dict(dim=3, n=24, damping=0.0,
                        src=[(0.0, 100.0, 120.0), (8.0, 5.0, 200.0)], multi=True,
                        tf=0.06),
    "3d_rect": dict(dim=3, shape=(30, 22, 40), space_order=6, tf=0.06),
    # several y tiles, interior widths that are not whole vectors
    "3d_wide_so8": dict(dim=3, shape=(26, 30, 301), space_order=8, damping=30.0,
                        tf=0.05),
    "3d_wide_f64_so8": dict(dim=3, shape=(22, 27, 263), space_order=8, damping=30.0,
                            dtype=F64, tf=0.04),
    "3d_wide_bc": dict(dim=3, shape=(25, 29, 275), space_order=8, damping=0.0,
                       bc=("null_neumann", "null_dirichlet", "null_neumann",
                           "null_neumann", "null_neumann", "null_neumann"),
                       src=[(0.0, 100.0, 2700.0), (60.0, 140.0, 1300.0)], multi=True,
                       tf=0.05),
    "3d_wide_so6": dict(dim=3, shape=(24, 26, 270), space_order=6, damping=30.0,
                        tf=0.05),
    # odd interior widths at radius 6 (lanes of 2 floats straddling the
    # row end), float32 and float64
    "3d_wide_so12": dict(dim=3, shape=(26, 28, 263), space_order=12, damping=30.0,
                         tf=0.04),
    "3d_wide_f64_so12": dict(dim=3, shape=(24, 26, 241), space_order=12, damping=30.0,
                             dtype=F64, tf=0.03),
    # deep grids: with 8-plane z chunks most tiles touch no boundary face and
    # no source box (the tiled kernel's branch-free march)
    "3d_deep_so8": dict(dim=3, shape=(60, 30, 301), space_order=8, damping=30.0,
                        tf=0.04),
    "3d_deep_so8_whole": dict(dim=3, shape=(58, 40, 264), space_order=8, damping=20.0,
                              tf=0.04),
    "3d_deep_none": dict(dim=3, shape=(56, 34, 200), space_order=8, damping=0.0,
                         bc=("none",) * 6, src=[(300.0, 150.0, 1000.0)], tf=0.05),
    "3d_deep_f64": dict(dim=3, shape=(50, 28, 140), space_order=8, damping=30.0,
                        dtype=F64, tf=0.03),
    "3d_deep_so4": dict(dim=3, shape=(52, 30, 260), space_order=4, damping=30.0,
                        tf=0.04),
    # Dirichlet on every face: the tiled kernel's clean tiles zero the face
    # planes by selects, corners and edges included
    "3d_deep_dirichlet": dict(dim=3, shape=(58, 30, 301), space_order=8, damping=30.0,
                              bc=("null_dirichlet",) * 6, tf=0.04),
    # no boundary condition and 12 damping points: clean tiles march over
    # planes that hold damping (the damped update and its operand loads)
    "3d_deep_damped": dict(dim=3, shape=(60, 30, 301), space_order=8, damping=120.0,
                           bc=("none",) * 6, tf=0.04),
    # spacings whose squares B 2^t (B odd) have t <= 0 (12.5^2 = 625/4,
    # 15^2 = 225): the float32 tiled kernel forms the Laplacian quotients
    # as float64 products with RN(1/h^2)
    "3d_deep_h12p5": dict(dim=3, shape=(60, 30, 301), h=12.5, space_order=8,
                          damping=37.5, tf=0.05),
    "3d_wide_h15_bc": dict(dim=3, shape=(25, 29, 275), h=15.0, space_order=8, damping=0.0,
                           bc=("null_neumann", "null_dirichlet", "null_neumann",
                               "null_neumann", "null_neumann", "null_neumann"),
                           src=[(0.0, 150.0, 4050.0), (90.0, 210.0, 1950.0)], multi=True,
                           tf=0.06),
}


def _compare(name, params, generic):
    setup = cases.make_case(**params)
    if generic:
        os.environ["SIMWAVE_HIP_GENERIC"] = "1"
    try:
        u_h, r_h = harness.run(setup, "hip")
    finally:
        os.environ.pop("SIMWAVE_HIP_GENERIC", None)
    u_o, r_o = harness.run(setup, "oracle")
    tol = 1e-12 if params.get("dtype", F32) == F64 else 1e-5
    eu, er = harness.rel_l2(u_h, u_o), harness.rel_l2(r_h, r_o)
    msg = "%s generic=%s: rel_l2 u=%.3g rec=%.3g" % (name, generic, eu, er)
    print(msg, flush=True)
    assert u_h.shape == u_o.shape and r_h.shape == r_o.shape
    assert eu <= tol and er <= tol, msg
    assert np.array_equal(u_h, u_o), "not bit-identical: " + msg
    assert np.array_equal(r_h, r_o), "not bit-identical: " + msg


@pytest.mark.parametrize("name", sorted(CASES))
def test_parity(gpu, name):
    print("case", name, flush=True)
    _compare(name, CASES[name], generic=False)


@pytest.mark.parametrize("name", ["2d_cd_so8", "2d_vd_so4", "2d_bc_mixed",
                                  "2d_stride2", "3d_cd_so8", "3d_vd_so4",
                                  "3d_bc_mixed", "3d_src_halo"])
def test_parity_generic_path(gpu, name):
    print("case", name, "(generic kernels)", flush=True)
    _compare(name, CASES[name], generic=True)


# Tile configurations of the 3D tiled kernel (SIMWAVE_HIP_TILE): radius 4
# carries 0..23 (17..23 with 8-byte lanes), other radii 0, 4, 8, 17 and 18;
# SIMWAVE_HIP_ORDER picks the tile order.
R4_TILES = list(range(24))
OTHER_TILES = [0, 4, 8, 17, 18]
TILED_R4 = ["3d_cd_so8", "3d_wide_so8", "3d_wide_f64_so8", "3d_wide_bc", "3d_wide_h15_bc"]
TILED_OTHER = ["3d_cd_so2", "3d_cd_so20", "3d_bc_mixed", "3d_src_halo", "3d_rect",
               "3d_wide_so6", "3d_wide_so12", "3d_wide_f64_so12"]


def _tiled(name, tile, order):
    print("case", name, "tile", tile, "order", order, flush=True)
    os.environ["SIMWAVE_HIP_TILE"] = str(tile)
    os.environ["SIMWAVE_HIP_ORDER"] = str(order)
    try:
        _compare(name, CASES[name], generic=False)
    finally:
        os.environ.pop("SIMWAVE_HIP_TILE", None)
        os.environ.pop("SIMWAVE_HIP_ORDER", None)


DEEP_R4 = ["3d_deep_so8", "3d_deep_so8_whole", "3d_deep_none", "3d_deep_f64", "3d_deep_h12p5",
           "3d_deep_dirichlet", "3d_deep_damped"]


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("tile", R4_TILES)
@pytest.mark.parametrize("name", DEEP_R4)
def test_parity_clean_tiles_r4(gpu, name, tile, order):
    """Tiles away from every face and source (8-plane z chunks), every
    radius-4 configuration."""
    os.environ["SIMWAVE_HIP_ZLEN"] = "8"
    try:
        _tiled(name, tile, order)
    finally:
        os.environ.pop("SIMWAVE_HIP_ZLEN", None)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("tile", OTHER_TILES)
def test_parity_clean_tiles_r2(gpu, tile, order):
    os.environ["SIMWAVE_HIP_ZLEN"] = "6"
    try:
        _tiled("3d_deep_so4", tile, order)
    finally:
        os.environ.pop("SIMWAVE_HIP_ZLEN", None)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("tile", R4_TILES)
@pytest.mark.parametrize("name", TILED_R4)
def test_parity_tiled_configs_r4(gpu, name, tile, order):
    """Every tile configuration of the radius-4 tiled 3D kernel."""
    _tiled(name, tile, order)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("tile", OTHER_TILES)
@pytest.mark.parametrize("name", TILED_OTHER)
def test_parity_tiled_configs(gpu, name, tile, order):
    """The tile configurations the other radii carry."""
    _tiled(name, tile, order)


HOST_ENTRY = {
    "2d_stride0": dict(dim=2, n=40, stride=0, tf=0.06, damping=0.0,
                       bc=("null_neumann", "null_dirichlet", "none", "null_neumann"),
                       src=[(0.0, 200.0)]),
    "2d_stride3": dict(dim=2, n=40, stride=3, tf=0.06),
    "2d_stride4": dict(dim=2, n=36, stride=4, tf=0.06, density=True),
    "3d_stride1": dict(dim=3, n=22, stride=1, tf=0.03, space_order=8),
    "3d_stride2": dict(dim=3, n=24, stride=2, tf=0.04, damping=0.0,
                       bc=("null_neumann",) * 6, src=[(0.0, 100.0, 120.0)]),
}


@pytest.mark.parametrize("name", sorted(HOST_ENTRY))
def test_host_entry_keeps_initial_halo_values(gpu, name):
    """The host entry uploads only the halo frame of slots first written by
    a step; with random initial contents in every slot (the reference
    contract allows any), the result must still equal the oracle's."""
    import ctypes
    import simwave_amd as sw
    from simwave_amd.middleware import Middleware

    print("case", name, "(host entry, random initial slots)", flush=True)
    parts = cases.make_case(**HOST_ENTRY[name])()
    solver = sw.Solver(**parts)
    kw = Middleware.normalize(**solver.forward_arguments())
    rng = np.random.default_rng(11)
    kw["u_full"] = rng.standard_normal(kw["u_full"].shape).astype(kw["u_full"].dtype)
    outs = []
    for kind in ("hip", "oracle"):
        k2 = dict(kw)
        k2["u_full"] = kw["u_full"].copy()
        k2["shot_record"] = np.zeros_like(kw["shot_record"])
        path = harness.LibraryCompiler(kind).compile(
            kw["velocity_model"].ndim,
            "variable_density" if kw.get("density_model") is not None else "constant_density",
            "-DFLOAT" if kw["velocity_model"].dtype == np.float32 else "-DDOUBLE", "forward")
        lib = ctypes.CDLL(path)
        argtypes, args = Middleware.ordered(k2)
        lib.forward.argtypes = argtypes
        lib.forward.restype = ctypes.c_double
        assert lib.forward(*args) >= 0
        outs.append((k2["u_full"], k2["shot_record"]))
    (u_h, r_h), (u_o, r_o) = outs
    assert np.array_equal(u_h, u_o)
    assert np.array_equal(r_h, r_o)


# The one-point-per-thread z-marching kernel (k_step3d_z): forced on for
# representative cases, and taken by itself where planes are too large for
# the tiled kernel's 32-bit buffer offsets.
ZMARCH = ["3d_cd_so2", "3d_cd_so8", "3d_cd_so20", "3d_cd_f64_so8", "3d_bc_neumann",
          "3d_bc_mixed", "3d_src_halo", "3d_wide_bc", "3d_deep_so8",
          "3d_vd_so4", "3d_vd_so8", "3d_vd_f64_so4", "3d_vd_bc", "3d_vd_so16_wide",
          "3d_vd_f64_so12", "2d_vd_f64_so16", "3d_vd_so20_bc",
          "2d_cd_so2", "2d_cd_so8", "2d_cd_so20", "2d_vd_so4", "2d_vd_so16", "2d_cd_f64_so8",
          "2d_vd_f64_so8", "2d_bc_mixed", "2d_bc_neumann", "2d_src_halo", "2d_src_overlap",
          "2d_stride3"]


@pytest.mark.parametrize("name", ZMARCH)
def test_parity_zmarch_kernel(gpu, name):
    print("case", name, "(z-marching kernel)", flush=True)
    os.environ["SIMWAVE_HIP_ZMARCH"] = "1"
    try:
        _compare(name, CASES[name], generic=False)
    finally:
        os.environ.pop("SIMWAVE_HIP_ZMARCH", None)


def _step_kernel_of_last_run(dim, dens, prec):
    """The library's record of its last call on this thread: 1 tiled 3D
    kernel, 2 z-marching kernel, 0 other (k_step3d, k_step2d)."""
    import ctypes
    import simwave_amd as sw
    lib = ctypes.CDLL(sw.build.library_path(dim, dens, prec))
    out = (ctypes.c_double * 8)()
    lib.simwave_hip_last_timings.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.simwave_hip_last_timings.restype = ctypes.c_int
    assert lib.simwave_hip_last_timings(out, 8) == 8
    return int(out[7])


# (case, tile forced, expected 3D step kernel): constant-density float32
# grids of radius <= 3, and small ones of radius 4, take the z-marching
# kernel; radius > 4 and float64 the tiled one; a forced tile configuration
# always the tiled one.  Variable density and 2D: z-marching for float32
# and for float64 up to radius 4, the LDS kernels (k_step3d, k_step2d)
# above.
KERNEL_CHOICE = [("3d_cd_so8", None, 2), ("3d_cd_so2", None, 2), ("3d_wide_so6", None, 2),
                 ("3d_cd_so20", None, 1),
                 ("3d_cd_f64_so8", None, 1), ("3d_cd_so8", "10", 1),
                 ("3d_vd_so8", None, 2), ("3d_vd_so16_wide", None, 2),
                 ("3d_vd_f64_so4", None, 2), ("3d_vd_f64_so12", None, 0),
                 ("2d_cd_so8", None, 2), ("2d_vd_so16", None, 2), ("2d_cd_f64_so8", None, 2),
                 ("2d_vd_f64_so16", None, 0)]


@pytest.mark.parametrize("name,tile,expect", KERNEL_CHOICE)
def test_default_step_kernel_choice(gpu, name, tile, expect):
    print("case", name, "tile", tile, flush=True)
    if tile is not None:
        os.environ["SIMWAVE_HIP_TILE"] = tile
    try:
        _compare(name, CASES[name], generic=False)
        prec = "f64" if CASES[name].get("dtype") == F64 else "f32"
        dens = "vd" if CASES[name].get("density") else "cd"
        assert _step_kernel_of_last_run(CASES[name]["dim"], dens, prec) == expect
    finally:
        os.environ.pop("SIMWAVE_HIP_TILE", None)


# The LDS kernels k_step3d and k_step2d (SIMWAVE_HIP_UNTILED): the float64
# defaults above radius 4 (3D variable density, 2D), forced here for the
# other cases too.
UNTILED = ["3d_vd_so4", "3d_vd_so8", "3d_vd_bc", "3d_vd_so16_wide", "3d_cd_so8",
           "3d_bc_mixed", "3d_src_halo", "2d_cd_so2", "2d_cd_so8", "2d_cd_so20",
           "2d_vd_so4", "2d_vd_so16", "2d_bc_mixed", "2d_src_halo", "2d_stride3",
           "2d_cd_f64_so8", "3d_vd_f64_so4"]


@pytest.mark.parametrize("name", UNTILED)
def test_parity_lds_tiled_kernel(gpu, name):
    print("case", name, "(k_step3d)", flush=True)
    os.environ["SIMWAVE_HIP_UNTILED"] = "1"
    try:
        _compare(name, CASES[name], generic=False)
    finally:
        os.environ.pop("SIMWAVE_HIP_UNTILED", None)


# 2D: small grids default to one-row blocks; longer z marches stay covered.
@pytest.mark.parametrize("zlen", ["3", "17"])
@pytest.mark.parametrize("name", ["2d_cd_so8", "2d_vd_so16", "2d_bc_mixed", "2d_src_halo"])
def test_parity_2d_z_marches(gpu, name, zlen):
    print("case", name, "zlen", zlen, flush=True)
    os.environ["SIMWAVE_HIP_ZLEN"] = zlen
    try:
        _compare(name, CASES[name], generic=False)
    finally:
        os.environ.pop("SIMWAVE_HIP_ZLEN", None)


def test_parity_planes_beyond_tiled_offsets(gpu):
    """1920 x 1920 float64 planes (29.5 MB): the tiled kernel's offsets do
    not fit, the z-marching kernel runs."""
    params = dict(dim=3, shape=(14, 1912, 1912), space_order=8, dtype=F64, damping=0.0,
                  tf=0.004)
    print("case big planes", params["shape"], flush=True)
    _compare("3d_big_planes", params, generic=False)
