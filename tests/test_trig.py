"""The refinement's float atan2, cos and sin (fast/decode.c:740-744) must
be correctly rounded: the reference's goldens were made with such a libm
(std41 and std52 fail with glibc 2.35's atan2f). tests/golden/trig_cases.npz
(make_trig_golden.py) holds inputs next to float rounding boundaries, found
by search, where rounding a double result to float goes wrong, and seeded
random inputs, with their correctly rounded results from 200-bit mpmath.
Bit-exact: the oracle's quad-precision functions (CPU) and the device's
(Ziv's rounding test with a double-double fallback), in its normal mode and
with the fallback forced for every input."""
import ctypes as C

import numpy as np
import pytest

from atg_testlib import GOLDEN, ORACLE_SO

CASES = np.load(GOLDEN / "trig_cases.npz")
SETS = ("hard", "random")


def oracle_fns():
    L = C.CDLL(str(ORACLE_SO))
    for name, nargs in (("cr_atan2f", 2), ("cr_cosf", 1), ("cr_sinf", 1)):
        f = getattr(L, name)
        f.restype = C.c_float
        f.argtypes = [C.c_float] * nargs
    return L


@pytest.mark.parametrize("s", SETS)
def test_oracle_cr_functions(s):
    L = oracle_fns()
    y, x, th = CASES[s + "_y"], CASES[s + "_x"], CASES[s + "_th"]
    assert np.array_equal(np.array([L.cr_atan2f(a, b) for a, b in zip(y, x)], np.float32), CASES[s + "_atan2"])
    assert np.array_equal(np.array([L.cr_cosf(v) for v in th], np.float32), CASES[s + "_cos"])
    assert np.array_equal(np.array([L.cr_sinf(v) for v in th], np.float32), CASES[s + "_sin"])


def test_hard_cases_are_hard():
    """Rounding glibc's double atan2 to float misses on some of the hard
    inputs: they are the cases that tell correct rounding apart."""
    y, x = CASES["hard_y"], CASES["hard_x"]
    naive = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    assert (naive != CASES["hard_atan2"]).sum() >= 1
    assert len(y) >= 20 and len(CASES["hard_th"]) >= 6


def test_fixture_against_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 200

    def exact_is_nearest(v, f):
        f = np.float32(f)
        up, dn = np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))
        d = abs(mp.mpf(float(f)) - v)
        return d <= abs(mp.mpf(float(up)) - v) and d <= abs(mp.mpf(float(dn)) - v)
    for i in range(0, len(CASES["hard_y"])):
        v = mp.atan2(mp.mpf(float(CASES["hard_y"][i])), mp.mpf(float(CASES["hard_x"][i])))
        assert exact_is_nearest(v, CASES["hard_atan2"][i])
    for i in range(0, len(CASES["random_th"]), 97):
        t = mp.mpf(float(CASES["random_th"][i]))
        assert exact_is_nearest(mp.cos(t), CASES["random_cos"][i])
        assert exact_is_nearest(mp.sin(t), CASES["random_sin"][i])


@pytest.mark.gpu
@pytest.mark.parametrize("s", SETS)
@pytest.mark.parametrize("slow", [False, True])
def test_gpu_cr_functions(ctx, s, slow):
    y, x, th = CASES[s + "_y"], CASES[s + "_x"], CASES[s + "_th"]
    n = max(len(y), len(th))
    at, co, si = ctx.trig(np.resize(y, n), np.resize(x, n), np.resize(th, n), slow=slow)
    assert np.array_equal(at[:len(y)], CASES[s + "_atan2"])
    assert np.array_equal(co[:len(th)], CASES[s + "_cos"])
    assert np.array_equal(si[:len(th)], CASES[s + "_sin"])
