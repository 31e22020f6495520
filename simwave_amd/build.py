"""Builds the HIP forward libraries for gfx950 in-tree.

One shared library per reference variant (dimension x density x precision),
each exporting the reference FFI symbol ``forward``
(simwave/kernel/backend/middleware.py:149-154) plus suffixed and device
entry points declared in include/simwave_hip.h.

    python -m simwave_amd.build          # builds simwave_amd/lib/*.so
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")

ARCH = "gfx950"

# -ffp-contract=off keeps the reference C's separate multiply and add
# roundings; packed FP32 ops are disabled for correctness on this pool.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

VARIANTS = [(dim, dens, prec)
            for dim in (2, 3) for dens in ("cd", "vd") for prec in ("f32", "f64")]


def variant_name(dim, dens, prec):
    return "%dd_%s_%s" % (dim, dens, prec)


def library_path(dim, dens, prec):
    return os.path.join(LIB_DIR,
                        "libsimwave_hip_%s.so" % variant_name(dim, dens, prec))


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the HIP libraries cannot be built")
    return exe


COMMON = "libsimwave_hip_common.so"


def common_library_path():
    return os.path.join(LIB_DIR, COMMON)


def _compile_common(verbose=False):
    """Process-wide state shared by the variant libraries (sw_common.cpp):
    each variant links it, so one copy is mapped per process."""
    out = common_library_path()
    tmp = out + ".tmp%d" % os.getpid()
    cxx = shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("g++ not found; the common library cannot be built")
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           "-Wl,-soname," + COMMON, os.path.join(CSRC, "sw_common.cpp"), "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError("common library build failed:\n%s\n%s"
                           % (" ".join(cmd), res.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def _compile(variant, verbose=False):
    dim, dens, prec = variant
    out = library_path(dim, dens, prec)
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [hipcc(), "--offload-arch=" + ARCH] + HIPCC_FLAGS + [
        "-DSW_DIM=%d" % dim,
        "-DSW_VD=%d" % (1 if dens == "vd" else 0),
    ]
    if prec == "f64":
        cmd.append("-DSW_DOUBLE")
    cmd += [os.path.join(CSRC, "sw_forward.hip"), "-o", tmp,
            "-L" + LIB_DIR, "-l:" + COMMON, "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    errors = "\n".join(line for line in res.stderr.splitlines()
                       if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError("hipcc failed for %s:\n%s\n%s"
                           % (variant_name(*variant), " ".join(cmd), errors))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build(verbose=True, jobs=None):
    """Compiles the common library and all eight variants (always; cached
    objects are not reused)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    _compile_common(verbose)
    jobs = jobs or min(len(VARIANTS), os.cpu_count() or 1, 8)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        outs = list(pool.map(lambda v: _compile(v, verbose), VARIANTS))
    return outs


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
