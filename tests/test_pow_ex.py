"""The kernels' integer-exponent power (x^(1-n_exp) of power_law!,
SphericalStokes_diapir.jl:262) is the correctly rounded x^-m: checked with
exact rational arithmetic (Python's Fraction -> float conversion rounds
correctly) on the host build of the kernel code, and against the host
libm pow that the CPU oracle calls (glibc's pow is within 0.52 ulp: it
misses the correctly rounded value in about 1 case of 1 400, where the two
then differ by 1 ulp).  numpy's power is a different implementation (about
5 % of its results are 1 ulp off) and is not used here."""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

from emu import emu
from helpers import max_ulp


_LIBM = ctypes.CDLL("libm.so.6")
_LIBM.pow.restype = ctypes.c_double
_LIBM.pow.argtypes = [ctypes.c_double, ctypes.c_double]


def _libm_pow(x, ex):
    return np.array([_LIBM.pow(float(v), ex) for v in x])


def _exact_pow_neg(x, m):
    return float(Fraction(1) / Fraction(x) ** m)


@pytest.mark.parametrize("n_exp", [2.0, 3.0, 5.0, 9.0])
def test_pow_ex_correctly_rounded(n_exp):
    ex = 1.0 - n_exp
    m = int(round(-ex))
    rng = np.random.default_rng(int(n_exp))
    x = np.exp(rng.uniform(-12.0, 12.0, 3000))       # tII/tau_C over many decades
    got = emu.pow_ex(x, ex)
    exact = np.array([_exact_pow_neg(v, m) for v in x])
    assert np.array_equal(got, exact)
    ref = _libm_pow(x, ex)                              # libm pow, as the oracle
    assert max_ulp(got, ref) <= 1
    assert np.count_nonzero(got != ref) <= 10


def test_pow_ex_special_values():
    got = emu.pow_ex(np.array([0.0, 1.0, np.inf, np.nan, 1e200, 1e-200]), -2.0)
    assert got[0] == np.inf and got[1] == 1.0 and got[2] == 0.0 and math.isnan(got[3])
    assert got[4] == 0.0 and got[5] == np.inf
    assert np.array_equal(emu.pow_ex(np.array([0.0, 2.5]), 0.0), [1.0, 1.0])
    # non-integer exponents go to pow
    x = np.array([0.3, 2.0, 7.5])
    assert np.array_equal(emu.pow_ex(x, -1.5), _libm_pow(x, -1.5))
