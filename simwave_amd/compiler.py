"""Backend library resolver.

Reference: simwave/kernel/backend/compiler.py.  The reference compiles its C
kernels at run time with the selected compiler and caches the shared object.
Here the default language ``'hip'`` resolves a prebuilt gfx950 library (one
per dimension x density x precision, built by ``simwave_amd.build``), so a
GPU host does no compilation.  The reference's plugin hook stays: with
``cfile=`` any C source exporting ``forward`` with the reference signature is
compiled by ``cc`` exactly as the reference does it (compiler.py:115-229).
No CPU kernel ships with this package, so the CPU languages require ``cfile``.
"""
import os
import subprocess
from hashlib import sha1

from . import build as _build

__all__ = ["Compiler"]

_LANGUAGES = ("hip", "c", "cpu_openmp")
_UNSUPPORTED = ("gpu_openmp", "gpu_openacc", "cuda")


class Compiler:
    """Selects the backend library that ``Solver`` calls.

    Parameters
    ----------
    cc : str, optional
        C compiler for ``cfile`` plugins (default gcc).
    language : str, optional
        'hip' (default; prebuilt gfx950 kernels), or 'c' / 'cpu_openmp' with
        a ``cfile`` plugin.
    cflags : str, optional
        Flags for ``cfile`` plugins; default '-O3 -fPIC -Wall -std=c99 -shared'.
    cfile : str, optional
        Path to a C kernel exporting ``forward``.
    """

    def __init__(self, cc="gcc", language="hip", cflags=None, cfile=None):
        self.cc = cc
        self.language = language
        self.cflags = cflags
        self.cfile = cfile

    @property
    def cc(self):
        return self._cc

    @cc.setter
    def cc(self, value):
        if not isinstance(value, str):
            raise TypeError("Compiler.cc attribute must be str.")
        self._cc = value

    @property
    def language(self):
        return self._language

    @language.setter
    def language(self, value):
        if not isinstance(value, str):
            raise TypeError("Compiler.language attribute must be str.")
        if value in _UNSUPPORTED:
            raise ValueError(
                "Compiler.language {} is not available on MI355X; use 'hip'."
                .format(value))
        if value not in _LANGUAGES:
            raise ValueError(
                "Compiler.language {} not implemented.".format(value))
        self._language = value

    @property
    def cflags(self):
        return self._cflags

    @cflags.setter
    def cflags(self, value):
        # composition rules of compiler.py:45-68
        if value is None:
            value = "-O3 -fPIC -Wall -std=c99 -shared"
        if not isinstance(value, str):
            raise TypeError("Compiler.cflags attribute must be str.")
        if "-shared" not in value:
            value += " -shared"
        if self.language == "cpu_openmp":
            flag = self.get_openmp_flag()
            if flag is None:
                print("WARNING: make sure OpenMP flag is provided in cflags.")
            elif flag not in value:
                value += " {}".format(flag)
        self._cflags = value

    @property
    def cfile(self):
        return self._cfile

    @cfile.setter
    def cfile(self, value):
        self._cfile = value

    def get_openmp_flag(self):
        """OpenMP flag of the C compiler, or None (compiler.py:96-113)."""
        return {"gcc": "-fopenmp", "icc": "-openmp", "pgcc": "-mp",
                "clang": "-fopenmp"}.get(self.cc)

    def compile(self, dimension, density, float_precision, operator):
        """Path of the shared object implementing `operator` for a variant.

        density: 'constant_density' | 'variable_density';
        float_precision: '-DFLOAT' | '-DDOUBLE'.
        """
        if operator != "forward":
            raise ValueError("operator {} not available".format(operator))
        if self.cfile is not None:
            return self._compile_plugin(dimension, density, float_precision,
                                        operator)
        if self.language != "hip":
            raise ValueError(
                "language '{}' needs cfile=: simwave_amd ships GPU kernels "
                "only".format(self.language))
        dens = {"constant_density": "cd", "variable_density": "vd"}[density]
        prec = {"-DFLOAT": "f32", "-DDOUBLE": "f64"}[float_precision]
        path = _build.library_path(int(dimension), dens, prec)
        if not os.path.exists(path):
            raise RuntimeError(
                "HIP library {} is not built; run `python -m simwave_amd.build`"
                .format(path))
        return path

    def _compile_plugin(self, dimension, density, float_precision, operator):
        with open(self.cfile, "r", encoding="utf-8") as f:
            text = f.read()
        lang = {"cpu_openmp": " -DCPU_OPENMP"}.get(self.language, "")
        key = "wave {} {} {} {}d {} {} {} {} \n{}".format(
            self.language, self.cc, self.cflags, dimension, operator, density,
            float_precision, lang, text)
        out_dir = os.path.join(os.getcwd(), "tmp")
        path = os.path.join(out_dir, sha1(key.encode()).hexdigest() + ".so")
        if os.path.exists(path):
            print("Shared object already compiled in:", path)
            return path
        args = [self.cc, self.cfile]
        args += [f for f in self.cflags.split(" ") if f.strip()]
        args.append(float_precision.strip())
        if lang.strip():
            args.append(lang.strip())
        args += ["-o", path]
        print("Compilation command:", " ".join(args))
        os.makedirs(out_dir, exist_ok=True)
        if subprocess.run(args).returncode != 0:
            raise Exception("Compilation failed")
        return path
