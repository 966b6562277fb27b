"""How far the engine's float transcendentals are from the float libm.

Every float transcendental of the engine and the oracle is evaluated in
double and rounded once to float (`m_*` in tracklooper_amd/csrc/lst_math.h,
`f*` in oracle/orc_common.h).  The reference calls the float overloads of the
C++ math library on the CPU (alpaka::math::*, e.g. SDL/Hit.h:125-137,
SDL/MiniDoublet.h:352-401, SDL/Segment.h:264-356), i.e. glibc's float
functions, which are not all correctly rounded.  This test measures, on random
inputs over the ranges the cuts use, how often (float)f((double)x) and glibc's
f(x) are different floats and by how many units in the last place, and pins
both: at most one unit (two for sinh), on a fraction of the inputs recorded
per function (glibc 2.35).  That last-place difference is the deviation the
engine accepts; the GPU and the oracle agree with each other bit for bit.
"""
import ctypes
import ctypes.util

import numpy as np
import pytest

_LIBM = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")

# name, arity, input range (uniform), bound on the differing fraction (measured with glibc 2.35: sin 0.017,
# cos 0.011, exp 0.0008, log 0.0002, tan 0.040, asin 0.072, acos 0.077, atan 0.023, atan2 0.163,
# acosh 0.052, sinh 0.265), bound on the difference in units in the last place
CASES = [
    ("sin", 1, (-4.0, 4.0), 0.03, 1),
    ("cos", 1, (-4.0, 4.0), 0.02, 1),
    ("exp", 1, (-10.0, 10.0), 0.002, 1),
    ("log", 1, (1e-3, 1e3), 0.001, 1),
    ("tan", 1, (-1.5, 1.5), 0.06, 1),
    ("asin", 1, (-1.0, 1.0), 0.10, 1),
    ("acos", 1, (-1.0, 1.0), 0.10, 1),
    ("atan", 1, (-50.0, 50.0), 0.04, 1),
    ("atan2", 2, (-120.0, 120.0), 0.22, 1),
    ("acosh", 1, (1.0, 200.0), 0.08, 1),
    ("sinh", 1, (-5.0, 5.0), 0.35, 2),
]
N = 20000


def _fns(name, arity):
    f32 = getattr(_LIBM, name + "f")
    f32.restype, f32.argtypes = ctypes.c_float, [ctypes.c_float] * arity
    f64 = getattr(_LIBM, name)
    f64.restype, f64.argtypes = ctypes.c_double, [ctypes.c_double] * arity
    return f32, f64


def _ulps(a, b):
    ia, ib = np.array([a, b], np.float32).view(np.int32)
    return abs(int(ia) - int(ib))


def compare(name, arity, lo, hi, n=N, seed=7):
    """(fraction of inputs where the two floats differ, largest difference in units in the last place)"""
    f32, f64 = _fns(name, arity)
    rng = np.random.default_rng(seed)
    xs = rng.uniform(lo, hi, size=(n, arity)).astype(np.float32)
    diff, worst = 0, 0
    for row in xs:
        a = [float(v) for v in row]
        ref = np.float32(f32(*a))
        ours = np.float32(f64(*a))  # the double result rounded once to float
        if not (ref == ours or (np.isnan(ref) and np.isnan(ours))):
            diff += 1
            worst = max(worst, _ulps(ref, ours))
    return diff / n, worst


@pytest.mark.parametrize("name,arity,rng,bound,ulps", CASES, ids=[c[0] for c in CASES])
def test_double_then_round_is_within_last_place_of_float_libm(name, arity, rng, bound, ulps):
    frac, worst = compare(name, arity, *rng)
    assert worst <= ulps, f"{name}: {worst} units in the last place from glibc's {name}f"
    assert frac <= bound, f"{name}: {frac:.5f} of inputs differ from glibc's {name}f"


def test_a_planted_single_rounding_error_is_seen():
    # rounding the double result through a truncation instead of round-to-nearest must show up as
    # differences: the comparison is able to see a last-place change
    f32, f64 = _fns("atan2", 2)
    rng = np.random.default_rng(3)
    diff = 0
    for y, x in rng.uniform(-100, 100, size=(2000, 2)).astype(np.float32):
        d = f64(float(y), float(x))
        trunc = np.float32(d)
        if abs(float(trunc)) > abs(d):  # round toward zero instead
            trunc = np.nextafter(trunc, np.float32(0.0))
        diff += int(trunc != np.float32(f32(float(y), float(x))))
    assert diff > 100
