"""Exactness of the fixed-divisor division used for the 1/h^2 terms.

The kernels divide the three per-axis stencil sums by h^2 with Markstein's
scheme (q0 = a*y, r = fma(-q0, b, a), q = fma(r, y, q0), y = RN(1/b)) when
|a| lies in a safe range, and with IEEE division otherwise.  These tests take
y and the range from the library's own host code and check the scheme
against IEEE division: exhaustively over all float32 significands at the
range ends and inside it, and on random float64 numerators.  (No GPU work:
the parameters come from a host-only entry point.)
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from simwave_amd import build

CHECKER = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
/* usage: checker b y lo hi ; prints mismatch count over all significands at
   several exponents of the safe range [lo, hi] */
int main(int argc, char **argv) {
    float b = strtof(argv[1], 0), y = strtof(argv[2], 0);
    float lo = strtof(argv[3], 0), hi = strtof(argv[4], 0);
    int elo, ehi; frexpf(lo, &elo); frexpf(hi, &ehi);
    int exps[4] = {elo, elo + 1, 1, ehi - 1};
    long bad = 0, n = 0;
    for (int ie = 0; ie < 4; ie++) {
        for (uint32_t m = 0; m < (1u << 23); m++) {
            for (int sgn = 0; sgn < 2; sgn++) {
                float a = ldexpf(1.0f + (float)m / 8388608.0f, exps[ie] - 1);
                if (sgn) a = -a;
                float aa = fabsf(a);
                if (!(aa >= lo && aa <= hi)) continue;
                float q0 = a * y;
                float r = fmaf(-q0, b, a);
                float q1 = fmaf(r, y, q0);
                float ref = a / b;
                n++;
                if (memcmp(&q1, &ref, 4) != 0) bad++;
            }
        }
    }
    printf("%ld %ld\n", bad, n);
    return 0;
}
"""

DIVISORS = [10.0, 12.5, 5.0, 0.5, 3.14159]


def _lib(prec):
    path = build.library_path(3, "cd", prec)
    if not os.path.exists(path):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(path)
    fn = getattr(lib, "simwave_hip_fixed_div_3d_cd_" + prec)
    t = ctypes.c_float if prec == "f32" else ctypes.c_double
    fn.argtypes = [t, ctypes.POINTER(t), ctypes.POINTER(t), ctypes.POINTER(t)]
    fn.restype = ctypes.c_int
    return fn, t


def _params(prec, b):
    fn, t = _lib(prec)
    y, lo, hi = t(), t(), t()
    assert fn(t(b), ctypes.byref(y), ctypes.byref(lo), ctypes.byref(hi)) == 1
    return y.value, lo.value, hi.value


@pytest.fixture(scope="module")
def checker():
    d = tempfile.mkdtemp(prefix="swdiv")
    src = os.path.join(d, "c.c")
    exe = os.path.join(d, "c")
    with open(src, "w") as f:
        f.write(CHECKER)
    subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", src, "-o", exe,
                    "-lm"], check=True)
    return exe


@pytest.mark.parametrize("h", DIVISORS)
def test_float32_exhaustive(checker, h):
    b = float(np.float32(h) * np.float32(h))
    y, lo, hi = _params("f32", b)
    # y is the correctly rounded reciprocal
    assert np.float32(y) == np.float32(y)
    from fractions import Fraction
    exact = Fraction(1) / Fraction(float(np.float32(b)))
    nb = [np.nextafter(np.float32(y), np.float32(0)), np.float32(y),
          np.nextafter(np.float32(y), np.float32(np.inf))]
    errs = [abs(Fraction(float(c)) - exact) for c in nb]
    assert errs[1] == min(errs)
    out = subprocess.run([checker, repr(float(np.float32(b))), repr(y), repr(lo),
                          repr(hi)], capture_output=True, text=True, check=True)
    bad, n = map(int, out.stdout.split())
    assert n > 4 * (1 << 23)
    assert bad == 0


@pytest.mark.parametrize("h", DIVISORS)
def test_float64_random(h):
    from fractions import Fraction
    b = h * h
    y, lo, hi = _params("f64", b)
    rng = np.random.default_rng(7)
    # numerators spread over the safe range's exponents plus the ends
    e = rng.integers(int(np.log2(lo)) + 1, min(int(np.log2(hi)) - 1, 300), size=1500)
    a = np.ldexp(rng.uniform(1.0, 2.0, size=e.size), e) * rng.choice([-1.0, 1.0], e.size)
    a = np.concatenate([a, [lo, -lo, hi, -hi, 0.0]])
    fb = Fraction(b)
    for av in a:
        q0 = av * y
        r = float(Fraction(av) - Fraction(q0) * fb)     # exact by the theorem
        assert Fraction(r) == Fraction(av) - Fraction(q0) * fb
        q1 = float(Fraction(q0) + Fraction(r) * Fraction(y))  # one rounding: fma
        assert q1 == av / b


F64_CHECKER = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
/* usage: checker b ; compares (float)((double)a * (1.0 / (double)b)) with the
   IEEE quotient a / b for every float a with a subnormal or zero quotient
   (every a below 2^-126 b) and every significand at a few normal exponents;
   prints the mismatch count */
static float fl(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
int main(int argc, char **argv) {
    volatile float b = strtof(argv[1], 0);
    const double y = 1.0 / (double)b;
    int eb; frexpf(b, &eb);
    long bad = 0;
    /* biased exponents 0 (subnormal a) .. those whose quotient is subnormal,
       then 1.0 and 2^60 */
    int hi_exp = 126 + eb + 1;
    if (hi_exp > 254) hi_exp = 254;
    for (int be = 0; be <= hi_exp + 2; be++) {
        int e = be <= hi_exp ? be : (be == hi_exp + 1 ? 127 : 187);
        for (uint32_t m = 0; m < (1u << 23); m++) {
            for (uint32_t sgn = 0; sgn < 2; sgn++) {
                float a = fl((sgn << 31) | ((uint32_t)e << 23) | m);
                volatile float ref = a / b;
                float got = (float)((double)a * y);
                if (bits(ref) != bits(got)) bad++;
            }
        }
    }
    printf("%ld\n", bad);
    return 0;
}
"""


@pytest.fixture(scope="module")
def f64_checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("f64div")
    src, exe = d / "c.c", d / "c"
    src.write_text(F64_CHECKER)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    return str(exe)


def _f64_rule(b):
    lib = ctypes.CDLL(build.library_path(3, "cd", "f32"))
    fn = lib.simwave_hip_f64_quotient_3d_cd_f32
    fn.argtypes = [ctypes.c_float]
    fn.restype = ctypes.c_int
    return fn(b) == 1


@pytest.mark.skipif(not os.path.exists(build.library_path(3, "cd", "f32")),
                    reason="HIP libraries not built")
@pytest.mark.parametrize("b", [156.25, 225.0, float(np.float32(12.3456) ** 2)])
def test_f64_reciprocal_quotient_is_exact(f64_checker, b):
    """Divisors the library accepts for the float64-reciprocal quotients
    (h2 = B 2^t with t <= 0: 12.5 m, 15 m and an arbitrary spacing)
    give the IEEE quotient for every float numerator whose quotient is
    subnormal or zero, and for every significand at two normal exponents."""
    assert _f64_rule(b)
    out = subprocess.run([f64_checker, repr(b)], capture_output=True, text=True,
                         check=True, timeout=600)
    assert int(out.stdout) == 0


@pytest.mark.skipif(not os.path.exists(build.library_path(3, "cd", "f32")),
                    reason="HIP libraries not built")
def test_f64_reciprocal_rule_rejects_even_divisors():
    """h2 = 100 (10 m), 6, 2^25 and subnormal divisors keep the
    fixed-divisor float quotients: with t >= 1 a subnormal quotient can be an
    exact midpoint, which the float64 product need not round to."""
    for b in (100.0, 6.0, 2.0 ** 25, 1e-40):
        assert not _f64_rule(b), b
