"""Pins of the oracle's synchrotron fits and GRMHD cell reconstruction that
the reference's tests hold no numeric checkpoint for, checked against what
the mathematics fixes:

* the Dexter (2016) thermal emissivities (thermal_synchrotron.hpp:253-309,
  529-563) against the exact ultrarelativistic Maxwell-Juettner integrals of
  the single-electron synchrotron functions (computed here with scipy), and
  the large-X saddle point that fixes the fits' exponent (3/2) 2^(1/3) and
  amplitudes pi sqrt(2/3) and pi 2^(1/6) / sqrt(3);
* the thermal Faraday rotation and conversion fits (thermal_synchrotron.hpp:
  461-506) against the exact high-frequency relativistic forms (Bessel-
  function ratios from scipy) and the cold-plasma textbook limits, and the
  published closed forms of the large-x constants of the conversion fit;
* the kappa rotativity fits at every anchor kappa = 3.5, 4, 4.5, 5
  (nonthermal_synchrotron.hpp:393-489) against the cold, high-frequency
  Faraday rotation;
* the GRMHD four-vectors and derived scalars (grmhd_grid.hpp:980-1014,
  iharm_loader.cpp:420-505) against a known physical fluid velocity pushed
  into primitives with the iharm normal-observer convention, the magnetic
  four-vector's definition, and the R_high/R_low energy partition.

Each check fails on a plausible slip: a wrong constant digit in an exponent
or amplitude, a dropped term, a swapped Bessel ratio, a wrong power of the
frequency, a sign of the primitive velocity or of the shift.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, grmhd_synth as gs

sp = pytest.importorskip("scipy.special")
integrate = pytest.importorskip("scipy.integrate")

# constants of the reference's thermal fits (thermal_synchrotron.hpp)
EE = 4.803204712570263e-10
ME = 9.1093837015e-28
CL = 2.99792458e10


def _thermal(fit, nu, ne, thetae, b, theta, max_pol=0.99):
    c = oracle.Coeffs()
    oracle.lib().kpo_thermal_coefficients(fit, nu, ne, thetae, b, theta, math.sin(theta), math.cos(theta),
                                          max_pol, 1e-4, 1.0, 1.0, 1.0, C.byref(c))
    return c


# ------------------------------------------------ exact synchrotron integrals

def _tail_k(order, x):
    """int_x^inf K_order(t) dt = e^-x int_{ln x} K_order(e^s) e^(x - e^s) e^s ds
    (log variable for the t -> 0 singularity, e^x scaling against underflow)."""
    f = lambda s: sp.kve(order, math.exp(s)) * math.exp(s) * math.exp(x - math.exp(s))
    top = math.log(x + 100.0)
    q = lambda a, b: integrate.quad(f, a, b, limit=200, epsabs=0.0, epsrel=1e-11)[0]
    v = q(math.log(x), 0.0) + q(0.0, top) if x < 1.0 else q(math.log(x), top)
    return math.exp(-x) * v


def _F(x):  # total-power synchrotron function x int_x^inf K_5/3
    return x * _tail_k(5.0 / 3.0, x)


def _G(x):  # linear-polarisation function x K_2/3(x)
    return x * sp.kv(2.0 / 3.0, x)


def _H(x):  # circular-polarisation function x K_1/3(x) + int_x^inf K_1/3
    return x * sp.kv(1.0 / 3.0, x) + _tail_k(1.0 / 3.0, x)


def _thermal_integral(X, kern, power):
    """(1/X) int_0^inf z^power e^-z kern(X / z^2) dz: the ultrarelativistic
    thermal emissivity with gamma = z Theta_e and X = nu / nu_c(Theta_e),
    nu_c = 3 e B sin(theta) Theta_e^2 / (4 pi m c) (Rybicki & Lightman 6.17c,
    integrated over the Maxwell-Juettner distribution n gamma^2 e^(-gamma/Theta)/(2 Theta^3))."""
    z0 = (2.0 * X) ** (1.0 / 3.0)
    f = lambda z: z ** power * math.exp(-z) * kern(X / (z * z)) if z > 0.0 else 0.0
    a = integrate.quad(f, 0.0, z0, limit=400, epsabs=0.0, epsrel=1e-10)[0]
    b = integrate.quad(f, z0, np.inf, limit=400, epsabs=0.0, epsrel=1e-10)[0]
    return (a + b) / X


def _dexter_state(X, thetae=20.0, theta=1.1, b=10.0):
    """(nu, prefactor_I, prefactor_V) with nu / nu_s = X for the fit's
    nu_s = 3 e B sin(theta) Theta_e^2 / (4 pi m c) + 1."""
    nus = 3.0 * EE * b * math.sin(theta) / (4.0 * math.pi * ME * CL) * thetae * thetae + 1.0
    return X * nus


def _dexter_reduced(X, thetae=20.0, theta=1.1, b=10.0, ne=1.0):
    """The oracle's Dexter jI, jQ, jV divided by their prefactors: the fit
    functions I(X), I_Q(X), I_V(X)."""
    nu = _dexter_state(X, thetae, theta, b)
    c = _thermal(abi.EMISSION_THERMAL_DEXTER, nu, ne, thetae, b, theta, max_pol=1.0)
    pref = ne * EE * EE * nu / (2.0 * math.sqrt(3.0) * CL * thetae * thetae)
    pref_v = 2.0 * ne * EE * EE * nu / (math.tan(theta) * 3.0 * math.sqrt(3.0) * CL * thetae ** 3)
    nu2 = nu * nu
    return c.jI * nu2 / pref, c.jQ * nu2 / pref, c.jV * nu2 / pref_v


@pytest.mark.parametrize("X", [1e-3, 1e-2, 0.1, 1.0, 3.0, 10.0, 100.0, 1e3, 1e4, 1e5])
def test_dexter_fit_against_exact_thermal_integrals(X):
    """Dexter's three fits agree with the exact ultrarelativistic integrals
    to their fitting accuracy (measured: below 0.9 % for X in [1e-3, 1e5]);
    a dropped term or a wrong leading digit of a coefficient moves one of
    them by several per cent somewhere in this range."""
    i_fit, q_fit, v_fit = _dexter_reduced(X)
    i_ex = _thermal_integral(X, _F, 2)
    q_ex = _thermal_integral(X, _G, 2)
    v_ex = _thermal_integral(X, _H, 1)
    assert i_fit == pytest.approx(i_ex, rel=1.2e-2, abs=0.0)
    assert q_fit == pytest.approx(q_ex, rel=1.2e-2, abs=0.0)
    assert v_fit == pytest.approx(v_ex, rel=1.2e-2, abs=0.0)


def _large_x_fit(reduced, power):
    """Least-squares fit of log(reduced(u^3) u^power) over u = X^(1/3) in
    [80, 220] to log a - d u + c1/u + c2/u^2 + c3/u^3; returns (a, d)."""
    us = np.linspace(80.0, 220.0, 15)
    ys = np.array([math.log(reduced(u ** 3) * u ** power) for u in us])
    A = np.stack([np.ones_like(us), -us, 1.0 / us, 1.0 / us ** 2, 1.0 / us ** 3], axis=1)
    coef, *_ = np.linalg.lstsq(A, ys, rcond=None)
    return math.exp(coef[0]), coef[1]


def test_dexter_large_x_saddle_point():
    """At large X the thermal integrals are dominated by the saddle at
    z = (2X)^(1/3): I, I_Q -> pi sqrt(2/3) e^(-d X^(1/3)) and
    I_V -> pi 2^(1/6)/sqrt(3) X^(-1/3) e^(-d X^(1/3)), d = (3/2) 2^(1/3).
    The fits' published constants (2.5651, 1.8899, 2.03773) are these to
    their printed digits (2.03773 absorbs the next order, 0.09 %)."""
    d_exact = 1.5 * 2.0 ** (1.0 / 3.0)
    ne = 1e30  # keeps e^(-1.89 * 220) above the double range's floor
    a, d = _large_x_fit(lambda X: _dexter_reduced(X, ne=ne)[0], 0)
    assert d == pytest.approx(d_exact, abs=3e-5)
    assert a == pytest.approx(math.pi * math.sqrt(2.0 / 3.0), rel=1e-4, abs=0.0)
    a, d = _large_x_fit(lambda X: _dexter_reduced(X, ne=ne)[1], 0)
    assert d == pytest.approx(d_exact, abs=3e-5)
    assert a == pytest.approx(math.pi * math.sqrt(2.0 / 3.0), rel=1e-4, abs=0.0)
    a, d = _large_x_fit(lambda X: _dexter_reduced(X, ne=ne)[2], 1)
    assert d == pytest.approx(d_exact, abs=3e-5)
    assert a == pytest.approx(math.pi * 2.0 ** (1.0 / 6.0) / math.sqrt(3.0), rel=2e-3, abs=0.0)


# ------------------------------------------------ thermal rotativities

def _omega(ne, b):
    wp2 = 4.0 * math.pi * ne * EE * EE / ME
    wb = EE * b / (ME * CL)
    return wp2, wb


@pytest.mark.parametrize("thetae", [0.05, 0.3, 1.0, 3.0, 10.0, 100.0])
@pytest.mark.parametrize("theta", [0.4, 1.2, 2.5])
def test_thermal_rotativities_high_frequency_limit(thetae, theta):
    """Where x = Theta_e sqrt(sqrt(2) sin(theta) 1e3 omega_B / omega) -> 0 the
    fits reduce to the exact relativistic thermal forms:
    rho_V = omega_p^2 omega_B cos(theta) / (c omega^2) K0(1/Theta)/K2(1/Theta),
    rho_Q = omega_p^2 omega_B^2 sin^2(theta) / (2 c omega^3) (K1/K2 + 6 Theta)
    (Shcherbakov 2008; the conversion fit's jffunc(0) = 2.011 - 1 - 0.011 = 1).
    Both depend on the Bessel ratios, the powers of the frequency and the
    angle factors, so a swapped ratio, a wrong power or a dropped 6 Theta
    fails."""
    ne, b, nu = 1e4, 1e-10, 1e13
    c = _thermal(abi.EMISSION_THERMAL_PANDYA, nu, ne, thetae, b, theta)
    wp2, wb = _omega(ne, b)
    w = 2.0 * math.pi * nu
    z = 1.0 / thetae
    k0, k1, k2 = sp.kve(0, z), sp.kve(1, z), sp.kve(2, z)  # the common e^-z cancels in the ratios
    rho_v = wp2 * wb * math.cos(theta) / (CL * w * w) * k0 / k2
    rho_q = wp2 * wb * wb * math.sin(theta) ** 2 / (2.0 * CL * w ** 3) * (k1 / k2 + 6.0 * thetae)
    # the invariants are rho * nu; the oracle's Bessel polynomials are good to ~2e-6
    assert c.rV / nu == pytest.approx(rho_v, rel=1e-5, abs=0.0)
    assert c.rQ / nu == pytest.approx(rho_q, rel=1e-5, abs=0.0)


def test_thermal_rotativities_cold_plasma_limit():
    """Theta_e -> 0 at high frequency (Richardson extrapolation over
    Theta_e = 0.02, 0.01, 0.005): the textbook cold-plasma Faraday rotation
    e^3 n B cos(theta) / (pi m^2 c^2 nu^2) and conversion
    e^4 n B^2 sin^2(theta) / (4 pi^2 m^3 c^3 nu^3) (Rybicki & Lightman 8.2;
    rho = 2 d chi / ds for the polarisation angle chi).  Below Theta_e ~ 1e-3
    the fits' Bessel functions underflow and the reference's floor on K2 makes
    rho_V zero; the limit is taken above that."""
    ne, b, nu, theta = 1e3, 1e-10, 1e12, 0.7
    rv = EE ** 3 * ne * b * math.cos(theta) / (math.pi * ME * ME * CL * CL * nu * nu)
    rq = EE ** 4 * ne * b * b * math.sin(theta) ** 2 / (4.0 * math.pi ** 2 * ME ** 3 * CL ** 3 * nu ** 3)
    c1, c2, c4 = (_thermal(abi.EMISSION_THERMAL_PANDYA, nu, ne, t, b, theta) for t in (0.02, 0.01, 0.005))
    ext = lambda a, b_, c: (8.0 * c - 6.0 * b_ + a) / 3.0
    assert ext(c1.rV, c2.rV, c4.rV) / nu == pytest.approx(rv, rel=1e-4, abs=0.0)
    assert ext(c1.rQ, c2.rQ, c4.rQ) / nu == pytest.approx(rq, rel=1e-4, abs=0.0)
    assert _thermal(abi.EMISSION_THERMAL_PANDYA, nu, ne, 1e-4, b, theta).rV == 0.0


def test_thermal_conversion_large_x_closed_form():
    """For x >> 120 the conversion fit's tanh switch at ln 120 is on and its
    shape tends to the published large-x form -2^(-1/3) 3^(-23/6) 1e4 pi
    x^(-8/3) (Dexter 2016, eq. A13): rho_Q x^(8/3) / (its x -> 0 scale)
    approaches that constant, with the other terms below 1e-9 at x > 2e4."""
    ne, nu, thetae, theta = 1e4, 1e9, 300.0, 1.0
    closed = -(2.0 ** (-1.0 / 3.0)) * 3.0 ** (-23.0 / 6.0) * 1e4 * math.pi
    for b in (1e5, 1e6):
        c = _thermal(abi.EMISSION_THERMAL_PANDYA, nu, ne, thetae, b, theta)
        wp2, wb = _omega(ne, b)
        w = 2.0 * math.pi * nu
        x = thetae * math.sqrt(math.sqrt(2.0) * math.sin(theta) * (1e3 * wb / w))
        assert x > 2e4
        z = 1.0 / thetae
        scale = wp2 * wb * wb * math.sin(theta) ** 2 / (2.0 * CL * w ** 3) * (sp.kve(1, z) / sp.kve(2, z) + 6 * thetae)
        shape = (c.rQ / nu) / scale
        assert shape * x ** (8.0 / 3.0) == pytest.approx(closed, rel=1e-5, abs=0.0)


# ------------------------------------------------ kappa rotativity anchors

@pytest.mark.parametrize("kappa", [3.5, 4.0, 4.5, 5.0])
def test_kappa_rotativity_anchor_cold_high_frequency_limit(kappa):
    """Each anchor's rho_V fit (Marszewski et al. 2021 as the reference
    carries it) reduces, for a cold (w -> 0) distribution at high frequency
    (X -> infinity), to the textbook cold Faraday rotation
    e^3 n B cos(theta) / (pi m^2 c^2 nu^2) times K0(1/w)/K2(1/w); the limit is
    taken by Richardson extrapolation over w = 0.01, 0.005, 0.0025.  The
    conversion fit's X shape fx tends to 1 at large X and to 0 at small X."""
    ee, me = 4.80320680e-10, 9.1093826e-28  # the nonthermal file's constants
    ne, b, nu, theta = 1e3, 1e-3, 1e14, 0.6
    tb = ee ** 3 * ne * b * math.cos(theta) / (math.pi * me * me * CL * CL * nu * nu)

    def ratio(w):
        p = oracle.Nonthermal(2, kappa=kappa, kappa_width=w)
        c = oracle.nonthermal_coefficients(p, nu, ne, 10.0, b, math.sin(theta), math.cos(theta))
        z = 1.0 / w
        return (c.rV / nu) / (tb * sp.kve(0, z) / sp.kve(2, z))

    r1, r2, r4 = ratio(0.01), ratio(0.005), ratio(0.0025)
    assert (8.0 * r4 - 6.0 * r2 + r1) / 3.0 == pytest.approx(1.0, abs=1e-4)
    assert abs(r1 - 1.0) > abs(r4 - 1.0)  # the relativistic correction shrinks with w
    # conversion: rho_Q / (cold conversion x wterm) -> 1 at large X, -> 0 at small X
    w = 1.0
    ww = math.sqrt(w)
    wterm = {3.5: 17 * w - 3 * ww + 7 * ww * math.exp(-5 * w),
             4.0: 46 / 3 * w - 5 / 3 * ww + 17 / 3 * ww * math.exp(-5 * w),
             4.5: 14 * w - 13 / 8 * ww + 9 / 2 * ww * math.exp(-5 * w),
             5.0: 25 / 2 * w - ww + 5 * ww * math.exp(-5 * w)}[kappa]
    p = oracle.Nonthermal(2, kappa=kappa, kappa_width=w)
    hi = oracle.nonthermal_coefficients(p, 1e16, ne, 10.0, b, math.sin(theta), math.cos(theta))
    rq_hi = ee ** 4 * ne * b * b * math.sin(theta) ** 2 / (4 * math.pi ** 2 * me ** 3 * CL ** 3 * 1e48)
    assert -(hi.rQ / 1e16) / (rq_hi * wterm) == pytest.approx(1.0, abs=1e-6)
    lo = oracle.nonthermal_coefficients(p, 10.0, ne, 10.0, b, math.sin(theta), math.cos(theta))
    rq_lo = ee ** 4 * ne * b * b * math.sin(theta) ** 2 / (4 * math.pi ** 2 * me ** 3 * CL ** 3 * 1e3)
    assert abs((lo.rQ / 10.0) / (rq_lo * wterm)) < 1e-3


# ------------------------------------------------ GRMHD four-vectors

def _fmks_metric(mm, x1, x2):
    """g_{mu nu} and g^{mu nu} of the model's FMKS coordinates at (x1, x2)."""
    L = oracle.lib()

    class Fm(C.Structure):
        _fields_ = [(n, C.c_double) for n in ("startx1", "hslope", "mks_smooth", "poly_alpha", "poly_xt",
                                               "poly_norm")]
    fm = Fm(mm.startx1, mm.hslope, mm.mks_smooth, mm.poly_alpha, mm.poly_xt, mm.poly_norm)
    dp = C.POINTER(C.c_double)
    L.kpo_fmks_gcov.argtypes = [C.c_double, C.c_double, C.POINTER(Fm), dp, dp]
    L.kpo_fmks_gcon.argtypes = [C.c_double, C.c_double, C.POINTER(Fm), dp, dp]
    L.kpo_fmks_theta.argtypes = [C.POINTER(Fm), C.c_double, C.c_double]
    L.kpo_fmks_theta.restype = C.c_double
    x = np.array([0.0, x1, x2, 0.0])
    gcov = np.zeros(16)
    gcon = np.zeros(16)
    L.kpo_fmks_gcov(1.0, mm.spin, C.byref(fm), x.ctypes.data_as(dp), gcov.ctypes.data_as(dp))
    L.kpo_fmks_gcon(1.0, mm.spin, C.byref(fm), x.ctypes.data_as(dp), gcon.ctypes.data_as(dp))
    th = L.kpo_fmks_theta(C.byref(fm), x1, x2)
    return gcov.reshape(4, 4), gcon.reshape(4, 4), th


def test_grmhd_four_vectors_from_known_velocity():
    """A fluid in circular rotation u = u^t (1, 0, 0, Omega) in Boyer-Lindquist
    (the same components in Kerr-Schild and FMKS, since u^r = u^theta = 0) is
    written into primitives with the iharm normal-observer convention
    U^i = u^i - u^t g^{ti} / g^{tt} (ipole ucon_calc inverted), together with
    a field B^i; the oracle must return that u^mu and the magnetic
    four-vector b^t = B^i u_i, b^i = (B^i + b^t u^i) / u^t.  A sign slip in
    the primitive velocity or in the shift changes u^mu at O(1)."""
    n1, n2, n3 = 6, 5, 2
    cells, m = gs.make_snapshot(n1, n2, n3, seed=1)
    mm = abi.GRMHDModel.from_buffer_copy(m)
    L = oracle.lib()
    dp = C.POINTER(C.c_double)
    L.kpo_grmhd_cell_four_vectors.argtypes = [C.POINTER(abi.GRMHDModel), C.c_void_p, C.c_int32, C.c_int32,
                                              C.c_int32, dp, dp]
    prim = cells["prim"].reshape(n1, n2, n3, 8)
    want = {}
    for i in range(n1):
        for j in range(n2):
            x1 = m.startx1 + (i + 0.5) * m.dx1
            x2 = m.startx2 + (j + 0.5) * m.dx2
            g, gi, th = _fmks_metric(m, x1, x2)
            r = math.exp(x1)
            omega = 1.0 / (r ** 1.5 + m.spin) * (0.7 + 0.1 * j)
            norm = g[0, 0] + 2.0 * omega * g[0, 3] + omega * omega * g[3, 3]
            assert norm < 0.0
            ut = 1.0 / math.sqrt(-norm)
            u = np.array([ut, 0.0, 0.0, ut * omega])
            U = u[1:] - u[0] * gi[0, 1:] / gi[0, 0]
            B = np.array([0.3 / r ** 2, 0.02 * (j - 2), 0.05 / r])
            for k in range(n3):
                prim[i, j, k, 2:5] = U
                prim[i, j, k, 5:8] = B
            want[(i, j)] = (g, u)
    for (i, j), (g, u_exact) in want.items():
        # the primitives are stored in FP32: the exact u of the stored values
        # differs from u_exact by their rounding
        uc = np.zeros(4)
        bc = np.zeros(4)
        L.kpo_grmhd_cell_four_vectors(C.byref(mm), C.c_void_p(cells.ctypes.data), i, j, 1, uc.ctypes.data_as(dp),
                                      bc.ctypes.data_as(dp))
        np.testing.assert_allclose(uc, u_exact, rtol=2e-6, atol=2e-7 * abs(u_exact[0]))
        assert uc @ g @ uc == pytest.approx(-1.0, abs=1e-12)
        Bs = prim[i, j, 1, 5:8].astype(np.float64)
        ucov = g @ uc
        bt = Bs @ ucov[1:]
        bexp = np.concatenate([[bt], (Bs + bt * uc[1:]) / uc[0]])
        np.testing.assert_allclose(bc, bexp, rtol=1e-12, atol=1e-14 * np.abs(bexp).max())
        # b^2 in the closed form (g_ij B^i B^j + (B^i u_i)^2) / (u^t)^2
        bsq = bc @ g @ bc
        assert bsq == pytest.approx((Bs @ g[1:, 1:] @ Bs + bt * bt) / uc[0] ** 2, rel=1e-10, abs=0.0)


def test_grmhd_derived_scalars_special_values():
    """Derived scalars on cells with a known b^2: plasma beta = p_gas / p_mag,
    sigma = b^2 / rho, and the electron temperature of the R_high/R_low
    energy partition u = n k (T_p / (gam_p - 1) + T_e / (gam_e - 1)) with
    T_p / T_e = R: R -> R_high as beta -> infinity (no field), R -> R_low as
    beta -> 0, R = (R_high + R_low) / 2 at beta = beta_crit."""
    n1, n2, n3 = 4, 3, 2
    cells, m = gs.make_snapshot(n1, n2, n3, seed=2)
    m.trat_large, m.trat_small, m.beta_crit = 40.0, 1.0, 1.0
    prim = cells["prim"].reshape(n1, n2, n3, 8)
    prim[..., 2:5] = 0.0  # the normal observer: u^mu = n^mu
    rho, uu = 0.5, 0.02
    prim[..., 0] = rho
    prim[..., 1] = uu
    L = oracle.lib()
    dp = C.POINTER(C.c_double)
    L.kpo_grmhd_cell_four_vectors.argtypes = [C.POINTER(abi.GRMHDModel), C.c_void_p, C.c_int32, C.c_int32,
                                              C.c_int32, dp, dp]
    mp, me = 1.67262171e-24, 9.1093826e-28

    def thetae_of(R):
        # k T_e / (m_e c^2) from the energy partition, with n = rho / m_p
        return (mp / me) * (float(np.float32(uu)) / float(np.float32(rho))) / (R / (5.0 / 3.0 - 1.0) + 1.0 / (4.0 / 3.0 - 1.0))

    mm = abi.GRMHDModel.from_buffer_copy(m)
    for case, bfac in (("no field", 0.0), ("beta_crit", None), ("strong field", 1e4)):
        prim[..., 5:8] = 0.0
        if bfac is not None:
            prim[..., 5] = bfac
        if case == "beta_crit":
            # choose B^1 so that beta = 1 in cell (1, 1): b^2 = 2 (gam - 1) u
            uc = np.zeros(4)
            bc = np.zeros(4)
            prim[..., 5] = 1.0
            L.kpo_grmhd_cell_four_vectors(C.byref(mm), C.c_void_p(cells.ctypes.data), 1, 1, 0,
                                          uc.ctypes.data_as(dp), bc.ctypes.data_as(dp))
            x1 = m.startx1 + 1.5 * m.dx1
            x2 = m.startx2 + 1.5 * m.dx2
            g, _, _ = _fmks_metric(m, x1, x2)
            bsq1 = bc @ g @ bc
            prim[..., 5] = math.sqrt(2.0 * (m.gam - 1.0) * float(np.float32(uu)) / bsq1)
        oracle.grmhd_derive_scalars(m, cells)
        d = cells["derived"].reshape(n1, n2, n3, 5)[1, 1, 0]
        uc = np.zeros(4)
        bc = np.zeros(4)
        L.kpo_grmhd_cell_four_vectors(C.byref(mm), C.c_void_p(cells.ctypes.data), 1, 1, 0, uc.ctypes.data_as(dp),
                                      bc.ctypes.data_as(dp))
        x1 = m.startx1 + 1.5 * m.dx1
        x2 = m.startx2 + 1.5 * m.dx2
        g, gi, _ = _fmks_metric(m, x1, x2)
        # u = n: u^t = 1 / alpha = sqrt(-g^tt)
        assert uc[0] == pytest.approx(math.sqrt(-gi[0, 0]), rel=1e-13, abs=0.0)
        bsq = max(abs(bc @ g @ bc), 1e-40)
        if case == "no field":
            assert d[1] == pytest.approx(thetae_of(m.trat_large), rel=1e-12, abs=0.0)
        elif case == "strong field":
            assert d[1] == pytest.approx(thetae_of(m.trat_small), rel=1e-6, abs=0.0)
        else:
            assert d[4] == pytest.approx(1.0, rel=1e-6, abs=0.0)
            assert d[1] == pytest.approx(thetae_of(0.5 * (m.trat_large + m.trat_small)), rel=1e-5)
        assert d[3] == pytest.approx(bsq / float(np.float32(rho)), rel=1e-12)
        assert d[4] == pytest.approx((m.gam - 1.0) * float(np.float32(uu)) / (0.5 * bsq), rel=1e-12)
