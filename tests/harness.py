"""Test harness: runs the simwave_amd frontend against any backend library.

The oracle libraries (oracle/_build, the CPU restatement) and, where present,
the reference's own C compiled from source (oracle/_ref) are plugged in
through the Compiler interface, exactly as a custom kernel would be; the
product default is the HIP library.
"""
import os

import numpy as np

import simwave_amd as sw

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_BUILD = os.path.join(REPO, "oracle", "_build")
ORACLE_REF = os.path.join(REPO, "oracle", "_ref")


def variant(dimension, density, float_precision):
    dens = {"constant_density": "cd", "variable_density": "vd"}[density]
    prec = {"-DFLOAT": "f32", "-DDOUBLE": "f64"}[float_precision]
    return "%dd_%s_%s" % (dimension, dens, prec)


class LibraryCompiler:
    """Compiler stand-in that resolves a prebuilt library per variant.

    kind: 'oracle' (CPU restatement), 'ref' (reference C from source),
    'hip' (the product library).
    """

    def __init__(self, kind="oracle", omp=False):
        self.kind = kind
        self.omp = omp

    def path(self, v):
        suffix = "_omp" if self.omp else ""
        if self.kind == "oracle":
            return os.path.join(ORACLE_BUILD, "libsw_oracle_%s%s.so" % (v, suffix))
        if self.kind == "ref":
            return os.path.join(ORACLE_REF, "libsw_ref_%s%s.so" % (v, suffix))
        raise ValueError(self.kind)

    def compile(self, dimension, density, float_precision, operator):
        if self.kind == "hip":
            return sw.Compiler(language="hip").compile(
                dimension, density, float_precision, operator)
        p = self.path(variant(dimension, density, float_precision))
        if not os.path.exists(p):
            raise FileNotFoundError(p)
        return p


def have(kind, omp=False):
    c = LibraryCompiler(kind, omp)
    if kind == "hip":
        return all(os.path.exists(sw.build.library_path(*v))
                   for v in sw.build.VARIANTS)
    return all(os.path.exists(c.path(sw.build.variant_name(*v)))
               for v in sw.build.VARIANTS)


def run(setup, kind, omp=False, quiet=True):
    """Builds the experiment with `setup()` and runs it on a backend.

    setup() -> dict(space_model, time_model, sources, receivers, wavelet).
    Returns (u_full, shot_record).
    """
    parts = setup()
    solver = sw.Solver(compiler=LibraryCompiler(kind, omp), **parts)
    solver._middleware.verbose = not quiet
    return solver.forward()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)
