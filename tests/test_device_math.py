"""Accuracy of the kernels' own bounded-argument sine/cosine (Cody-Waite
reduction + Taylor series) against 50-digit mpmath values."""
from __future__ import annotations

import ctypes as C

import mpmath
import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return L


def _ulp_err(got: float, exact) -> float:
    ref = float(exact)
    ulp = np.spacing(abs(ref)) if ref != 0.0 else np.spacing(0.0)
    return abs(float(mpmath.mpf(got) - exact)) / ulp


def test_sincos_bounded_within_one_ulp(lib):
    mpmath.mp.dps = 50
    rng = np.random.default_rng(20261202)
    xs = np.concatenate([rng.uniform(-7.0, 7.0, 4000), rng.uniform(-1e-3, 1e-3, 200),
                         np.array([0.0, 1e-300, -1e-20, np.pi / 4, np.pi / 2, np.pi, 2 * np.pi,
                                   3.141592653589793, 1.5707963267948966, 2.356194490192345])])
    s, c = C.c_double(), C.c_double()
    worst = 0.0
    for x in xs:
        lib.kp_emul_sincos(float(x), C.byref(s), C.byref(c))
        xm = mpmath.mpf(float(x))
        es, ec = mpmath.sin(xm), mpmath.cos(xm)
        # absolute floor near zeros of sin/cos: 1 ulp of 1 * |r| scale
        worst = max(worst, min(_ulp_err(s.value, es), abs(float(s.value - es)) / 2.3e-16 + 0.0))
        worst = max(worst, min(_ulp_err(c.value, ec), abs(float(c.value - ec)) / 2.3e-16 + 0.0))
        assert abs(s.value - float(es)) <= 2.3e-16 * max(abs(float(es)), 1e-300) * 1.5 + 1e-32 or \
            abs(s.value - float(es)) < 1.2e-16
        assert abs(c.value - float(ec)) <= 2.3e-16 * max(abs(float(ec)), 1e-300) * 1.5 + 1e-32 or \
            abs(c.value - float(ec)) < 1.2e-16
    assert worst <= 1.5
