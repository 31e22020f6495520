"""The C ABI: every entry point include/simwave_hip.h declares is exported
by the library of its variant, the reference FFI symbol `forward` is
exported by every library, and a call fails loudly (negative return and a
message, no exit) where it cannot run.  No kernel is launched here."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import simwave_amd as sw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "simwave_hip.h")

pytestmark = pytest.mark.skipif(
    not all(os.path.exists(sw.build.library_path(*v)) for v in sw.build.VARIANTS),
    reason="HIP libraries not built (python -m simwave_amd.build)")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(simwave_hip_\w+)\s*\(", text)))


def test_header_declares_every_variant():
    names = declared()
    for v in sw.build.VARIANTS:
        vn = sw.build.variant_name(*v)
        for base in ("simwave_hip_forward_", "simwave_hip_forward_device_",
                     "simwave_hip_fixed_div_", "simwave_hip_plan_create_",
                     "simwave_hip_plan_run_", "simwave_hip_plan_destroy_",
                     "simwave_hip_gradient_", "simwave_hip_f64_quotient_"):
            assert base + vn in names
    for common in ("simwave_hip_last_error", "simwave_hip_last_timings",
                   "simwave_hip_variant", "simwave_hip_trim_pool",
                   "simwave_hip_overlap_acquire", "simwave_hip_overlap_release",
                   "simwave_hip_overlap_holders", "simwave_hip_device_alloc",
                   "simwave_hip_device_free", "simwave_hip_host_alloc",
                   "simwave_hip_host_free", "simwave_hip_copy", "simwave_hip_zero",
                   "simwave_hip_stream_create", "simwave_hip_stream_destroy",
                   "simwave_hip_stream_synchronize", "simwave_hip_event_create",
                   "simwave_hip_event_record", "simwave_hip_event_synchronize",
                   "simwave_hip_event_destroy"):
        assert common in names


def test_header_compiles_as_c():
    subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", HEADER],
                   check=True)


@pytest.mark.parametrize("variant", sw.build.VARIANTS,
                         ids=[sw.build.variant_name(*v) for v in sw.build.VARIANTS])
def test_library_exports(variant):
    vn = sw.build.variant_name(*variant)
    lib = ctypes.CDLL(sw.build.library_path(*variant))
    assert hasattr(lib, "forward")
    for name in declared():
        if re.search(r"_[23]d_(cd|vd)_f(32|64)$", name):
            assert hasattr(lib, name) == name.endswith(vn), name
        else:
            assert hasattr(lib, name), name
    lib.simwave_hip_variant.restype = ctypes.c_char_p
    assert lib.simwave_hip_variant().decode() == vn


def test_fixed_div_helper():
    lib = ctypes.CDLL(sw.build.library_path(3, "cd", "f32"))
    fn = lib.simwave_hip_fixed_div_3d_cd_f32
    fn.argtypes = [ctypes.c_float] + [ctypes.POINTER(ctypes.c_float)] * 3
    fn.restype = ctypes.c_int
    y, lo, hi = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    assert fn(156.25, ctypes.byref(y), ctypes.byref(lo), ctypes.byref(hi)) == 1
    assert abs(y.value * 156.25 - 1.0) < 1e-6 and lo.value < 1e-20 < 1e20 < hi.value
    assert fn(-1.0, ctypes.byref(y), ctypes.byref(lo), ctypes.byref(hi)) == 0


_FAIL_SCRIPT = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
import cases
import simwave_amd as sw
from simwave_amd.middleware import Middleware
parts = cases.make_case(dim=2, n=24, tf=0.02)()
solver = sw.Solver(**parts)
kw = Middleware.normalize(**solver.forward_arguments())
kw["space_order"] = %(so)d
lib = ctypes.CDLL(sw.build.library_path(2, "cd", "f32"))
argtypes, args = Middleware.ordered(kw)
lib.forward.argtypes = argtypes
lib.forward.restype = ctypes.c_double
rc = lib.forward(*args)
lib.simwave_hip_last_error.restype = ctypes.c_char_p
print("RC", rc)
print("ERR", lib.simwave_hip_last_error().decode())
"""


def _run_failing(so):
    code = _FAIL_SCRIPT % dict(root=ROOT, tests=os.path.join(ROOT, "tests"), so=so)
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"      # no device, also on a GPU host
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300, env=env)
    assert res.returncode == 0, res.stderr[-2000:]
    out = dict(line.split(" ", 1) for line in res.stdout.splitlines()
               if line.startswith(("RC", "ERR")))
    return float(out["RC"]), out.get("ERR", "")


def test_forward_fails_loudly_without_device():
    rc, err = _run_failing(4)
    assert rc < 0
    assert "device" in err.lower()


def test_forward_rejects_bad_space_order():
    """Validation comes first: an odd order is rejected with its own message
    whether or not a device is present."""
    rc, err = _run_failing(3)
    assert rc < 0
    assert "space_order 3" in err


_LEASE_SCRIPT = r"""
import ctypes, sys
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
import cases
import simwave_amd as sw
from simwave_amd.middleware import Middleware
a = ctypes.CDLL(sw.build.library_path(3, "cd", "f32"))
b = ctypes.CDLL(sw.build.library_path(2, "vd", "f64"))
# one lease per process, whichever variant library takes it
assert a.simwave_hip_overlap_acquire() == 1
assert b.simwave_hip_overlap_holders() == 1
assert b.simwave_hip_overlap_acquire() == 0
parts = cases.make_case(dim=2, n=24, tf=0.02)()
kw = Middleware.normalize(**sw.Solver(**parts).forward_arguments())
lib = ctypes.CDLL(sw.build.library_path(2, "cd", "f32"))
argtypes, args = Middleware.ordered(kw)
lib.simwave_hip_last_error.restype = ctypes.c_char_p
for name in ("simwave_hip_forward_device_2d_cd_f32", "simwave_hip_plan_create_2d_cd_f32"):
    fn = getattr(lib, name)
    fn.argtypes = argtypes + [ctypes.c_void_p, ctypes.c_uint]
    fn.restype = ctypes.c_double if "forward" in name else ctypes.c_void_p
    r = fn(*args, None, 4)
    print("CALL", name, r if r else 0, lib.simwave_hip_last_error().decode())
    assert b.simwave_hip_overlap_holders() == 1
b.simwave_hip_overlap_release()
assert a.simwave_hip_overlap_holders() == 0
print("DONE")
"""


def test_receiver_overlap_lease_is_process_wide():
    """Flag 4 (receiver gather on a second stream) is refused while another
    run or plan of the process holds it, across variant libraries, and a
    refused call leaves the holder's lease in place."""
    code = _LEASE_SCRIPT % dict(root=ROOT, tests=os.path.join(ROOT, "tests"))
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300, env=env)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    calls = [line for line in res.stdout.splitlines() if line.startswith("CALL")]
    assert len(calls) == 2
    for line in calls:
        assert "receiver overlap (flag 4) is already in use" in line, line
    assert "DONE" in res.stdout


def test_memory_entry_points_check_their_arguments():
    """The memory, copy and zero entry points refuse bad arguments with an
    error return and a message, before any device work."""
    lib = ctypes.CDLL(sw.build.library_path(3, "cd", "f32"))
    lib.simwave_hip_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    lib.simwave_hip_zero.argtypes = [vp, ctypes.c_ulonglong, vp]
    lib.simwave_hip_copy.argtypes = [vp, vp, ctypes.c_ulonglong, vp]
    lib.simwave_hip_device_alloc.argtypes = [ctypes.c_ulonglong, vp, vp]
    lib.simwave_hip_device_free.argtypes = [vp, vp]
    lib.simwave_hip_host_alloc.argtypes = [ctypes.c_ulonglong, vp]
    assert lib.simwave_hip_zero(None, 6, None) == -1
    assert "multiple of 4" in lib.simwave_hip_last_error().decode()
    assert lib.simwave_hip_zero(None, 8, None) == -1
    assert "NULL pointer" in lib.simwave_hip_last_error().decode()
    assert lib.simwave_hip_copy(None, None, 8, None) == -1
    assert "NULL pointer" in lib.simwave_hip_last_error().decode()
    assert lib.simwave_hip_device_alloc(16, None, None) == -1
    assert "NULL result pointer" in lib.simwave_hip_last_error().decode()
    assert lib.simwave_hip_host_alloc(16, None) == -1
    assert "NULL result pointer" in lib.simwave_hip_last_error().decode()
    # zero-sized work and NULL frees are no-ops
    assert lib.simwave_hip_zero(None, 0, None) == 0
    assert lib.simwave_hip_copy(None, None, 0, None) == 0
    assert lib.simwave_hip_device_free(None, None) == 0
    assert lib.simwave_hip_host_free(None) == 0
