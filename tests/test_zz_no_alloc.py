"""The allocation-free device entry (simwave_hip_forward_device_noalloc_<v>)
against the allocating one (simwave_hip_forward_device_<v>) through the C
ABI, on the same device arrays: bit for bit, on two calls in a row, the
second with a smaller grid (the static scratch array keeps the first
call's contents), for 2D/3D, constant/variable density, float32/float64,
the z-marching, tiled and generic kernels and a saving_stride run.
Grids beyond the array's size are refused with an error."""
import ctypes

import numpy as np
import pytest

import cases
import simwave_amd as sw
from simwave_amd import hipmem
from simwave_amd.device import DeviceForward
from simwave_amd.middleware import _last_error

pytestmark = pytest.mark.gpu

F64 = np.float64

# (first call, second call with a smaller grid)
PAIRS = {
    "3d_cd_so8": (dict(dim=3, n=40, space_order=8, tf=0.06),
                  dict(dim=3, n=28, space_order=8, tf=0.05)),
    "3d_cd_so12_tiled": (dict(dim=3, shape=(34, 30, 140), space_order=12, tf=0.05),
                         dict(dim=3, n=30, space_order=12, tf=0.04)),
    "3d_vd_f64": (dict(dim=3, n=30, space_order=4, density=True, dtype=F64, tf=0.05),
                  dict(dim=3, n=24, space_order=4, density=True, dtype=F64, tf=0.04)),
    "2d_vd_so16": (dict(dim=2, n=90, space_order=16, density=True),
                   dict(dim=2, n=50, space_order=16, density=True)),
    "2d_cd_stride3": (dict(dim=2, n=70, stride=3, tf=0.1),
                      dict(dim=2, n=40, stride=3, tf=0.08)),
}


def _call(dev, noalloc, flags=0):
    name = ("simwave_hip_forward_device_noalloc_" if noalloc
            else "simwave_hip_forward_device_") + dev.variant
    fn = getattr(dev._lib, name)
    fn.restype = ctypes.c_double
    fn.argtypes = list(dev._argtypes) + [ctypes.c_void_p, ctypes.c_uint]
    stream = dev._stream(None)
    dev.reset(stream)
    dev.tensors["shot_record"].zero_(stream)
    secs = fn(*dev._args, ctypes.c_void_p(hipmem.stream_handle(stream)), ctypes.c_uint(flags))
    if secs < 0:
        raise RuntimeError(_last_error(dev._lib))
    dev.synchronize(stream)
    return dev.tensors["u_full"].numpy(), dev.tensors["shot_record"].numpy()


@pytest.mark.parametrize("generic", [0, 2], ids=["fused", "generic"])
@pytest.mark.parametrize("name", sorted(PAIRS))
def test_noalloc_matches_device_entry(gpu, name, generic):
    for params in PAIRS[name]:
        print("noalloc", name, params, "flags", generic, flush=True)
        dev = DeviceForward(sw.Solver(**cases.make_case(**params)()))
        u_a, r_a = _call(dev, noalloc=False, flags=generic)
        u_n, r_n = _call(dev, noalloc=True, flags=generic)
        assert np.abs(u_a).max() > 0
        assert np.array_equal(u_n, u_a), name
        assert np.array_equal(r_n, r_a), name


def test_noalloc_refuses_large_grids_and_flags(gpu):
    dev = DeviceForward(sw.Solver(**cases.make_case(dim=3, n=180, space_order=4,
                                                    tf=0.01)()))
    limit = getattr(dev._lib, "simwave_hip_scratch_bytes_" + dev.variant)
    limit.restype = ctypes.c_size_t
    assert limit() == 64 << 20
    with pytest.raises(RuntimeError, match="static scratch"):
        _call(dev, noalloc=True)
    small = DeviceForward(sw.Solver(**cases.make_case(dim=2, n=40)()))
    with pytest.raises(RuntimeError, match="allocation-free entry takes only"):
        _call(small, noalloc=True, flags=4)
    # the library still works after the refusals
    u_a, _ = _call(small, noalloc=False)
    u_n, _ = _call(small, noalloc=True)
    assert np.array_equal(u_n, u_a)
