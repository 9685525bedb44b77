"""Pins of the CPU oracle (oracle/kp_oracle.c) against the reference's own
golden vectors and known-answer tests, and against what the mathematics fixes.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand) in the oracle fails it.
"""
from __future__ import annotations

import copy
import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi
from oracle import host_rules as hr

GOLDEN = Path(__file__).resolve().parent / "golden"


def _resolve(params_list):
    opt = hr.new_options()
    for kv in params_list:
        k, _, v = kv.partition("=")
        hr.apply_option(opt, k, v)
    return opt


# ---------------------------------------------------------------- end to end

def test_riaf_bl_8x8_golden_fluxes():
    """End-to-end image golden: tests/test_hdf5_contract.py:671-681."""
    g = json.loads((GOLDEN / "riaf_bl_8x8_golden.json").read_text())
    opt = _resolve(g["parameters"])
    p, m, scale, qs = hr.resolve_riaf(opt)
    img, stats = oracle.run_riaf_image(p, m)
    I, Q, U, V, flux = hr.pack(img, opt["freq"], qs, scale)
    I2 = I.reshape(8, 8)
    got = {
        "flux_I_Jy": flux[0], "flux_Q_Jy": flux[1], "flux_U_Jy": flux[2], "flux_V_Jy": flux[3],
        "I_sum": float(I2.sum()), "I_max": float(I2.max()), "I_4_4": float(I2[4, 4]),
    }
    rel = g["rel_tolerance"]
    for key, (expected, abs_tol) in g["expected"].items():
        assert math.isclose(got[key], expected, rel_tol=rel, abs_tol=abs_tol), (key, got[key], expected)
    assert np.all(img["reason"] == abi.REASON_REACHED_CAMERA)
    # flux bookkeeping identity of the image contract (test_hdf5_contract.py:664)
    assert abs(float(I2.sum()) * scale - flux[0]) <= max(1e-10, abs(flux[0]) * 1e-10)


# ---------------------------------------------------------------- transfer

def _stokes(i=0.0, q=0.0, u=0.0, v=0.0):
    s = oracle.Stokes()
    s.I, s.Q, s.U, s.V = i, q, u, v
    return s


def _coeffs(**kw):
    c = oracle.Coeffs()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_transfer_pure_emission():
    """tests/test_core.cpp:247-253."""
    L = oracle.lib()
    s = _stokes()
    L.kpo_semi_analytic_step(C.byref(s), C.byref(_coeffs(jI=2.0)), 0.25)
    assert s.I == pytest.approx(0.5, abs=1e-12)


def test_transfer_absorption_emission():
    """tests/test_core.cpp:255-266."""
    L = oracle.lib()
    s = _stokes(3.0)
    L.kpo_semi_analytic_step(C.byref(s), C.byref(_coeffs(jI=4.0, aI=2.0)), 0.5)
    expected = 3.0 * math.exp(-1.0) + 2.0 * (1.0 - math.exp(-1.0))
    assert s.I == pytest.approx(expected, abs=1e-12)


def test_transfer_faraday_quarter_turn():
    """tests/test_core.cpp:268-277: rho_V = pi/2 over unit length maps Q -> U."""
    L = oracle.lib()
    s = _stokes(0.0, 1.0)
    L.kpo_semi_analytic_step(C.byref(s), C.byref(_coeffs(rV=1.57079632679489661923)), 1.0)
    assert (s.Q, s.U, s.V) == pytest.approx((0.0, 1.0, 0.0), abs=1e-9)


def test_transfer_polarized_absorption_matches_matrix_exponential():
    """With jQ..=0 and rho=0 the exact solution of dS/dl = -K S is expm(-K l) S:
    pins the eigen-split of absorb_emit (semi_analytic.hpp:39-91) including the
    perpendicular part attenuated by alpha_I."""
    from scipy.linalg import expm
    L = oracle.lib()
    aI, aQ, aU, aV = 1.3, 0.4, -0.3, 0.2
    K = np.array([[aI, aQ, aU, aV], [aQ, aI, 0, 0], [aU, 0, aI, 0], [aV, 0, 0, aI]])
    s0 = np.array([1.0, 0.2, -0.1, 0.3])
    dl = 0.7
    s = _stokes(*s0)
    L.kpo_semi_analytic_step(C.byref(s), C.byref(_coeffs(aI=aI, aQ=aQ, aU=aU, aV=aV)), dl)
    np.testing.assert_allclose([s.I, s.Q, s.U, s.V], expm(-K * dl) @ s0, rtol=1e-12, atol=1e-14)


def test_transfer_faraday_rotation_matches_matrix_exponential():
    """Pure Faraday: dP/dl = rho x P (Rodrigues about rho); exact for any angle."""
    from scipy.linalg import expm
    L = oracle.lib()
    rho = np.array([0.3, -1.1, 0.7])
    R = np.array([[0, -rho[2], rho[1]], [rho[2], 0, -rho[0]], [-rho[1], rho[0], 0]])
    p0 = np.array([0.2, 0.5, -0.4])
    dl = 1.9
    s = _stokes(1.0, *p0)
    L.kpo_semi_analytic_step(C.byref(s), C.byref(_coeffs(rQ=rho[0], rU=rho[1], rV=rho[2])), dl)
    np.testing.assert_allclose([s.Q, s.U, s.V], expm(R * dl) @ p0, rtol=1e-12, atol=1e-14)


def test_observer_basis_jones_contract():
    """tests/test_core.cpp:309-335: rotate/reflect complex field amplitudes."""
    L = oracle.lib()

    def stokes(x, y):
        return np.array([abs(x) ** 2 + abs(y) ** 2, abs(x) ** 2 - abs(y) ** 2,
                         2 * (x * np.conj(y)).real, -2 * (x * np.conj(y)).imag])

    x, y = complex(0.7, 0.2), complex(-0.3, 0.4)
    initial = stokes(x, y)
    for angle in (0.0, 0.37, 1.5707963267948966):
        for parity in (1.0, -1.0):
            r11, r12 = math.cos(angle), math.sin(angle)
            r21, r22 = -parity * math.sin(angle), parity * math.cos(angle)
            expected = stokes(r11 * x + r12 * y, r21 * x + r22 * y)
            out = oracle.Stokes()
            L.kpo_transform_to_observer_basis(C.byref(_stokes(*initial)), r11, r12, r21, r22,
                                              C.byref(out))
            np.testing.assert_allclose([out.I, out.Q, out.U, out.V], expected, atol=1e-12)


def test_radiation_substep_count_doubling():
    """pass_b.hpp:147-179: doubling until every limit holds, capped at 4096."""
    L = oracle.lib()
    c = _coeffs(aI=3.0, rV=0.0)
    assert L.kpo_radiation_substep_count(C.byref(c), 0.5, 1.0, 1, 1.0, 1.0, 1.0, 4.0) == 2
    c = _coeffs(aI=0.0, rV=20.0)
    assert L.kpo_radiation_substep_count(C.byref(c), 1.0, 1.0, 1, 1.0, 1.0, 1.0, 4.0) == 8
    c = _coeffs(aI=1e9)
    assert L.kpo_radiation_substep_count(C.byref(c), 1.0, 1.0, 1, 1.0, 1.0, 1.0, 4.0) == 4096
    c = _coeffs()
    assert L.kpo_radiation_substep_count(C.byref(c), 3.0, 1.0, 1, 1.0, 1.0, 1.0, 4.0) == 4


# ---------------------------------------------------------------- synchrotron

def _thermal(fit, nu, ne, thetae, b, theta, max_pol=0.99):
    L = oracle.lib()
    c = oracle.Coeffs()
    L.kpo_thermal_coefficients(fit, nu, ne, thetae, b, theta, math.sin(theta), math.cos(theta),
                               max_pol, 1e-4, 1.0, 1.0, 1.0, C.byref(c))
    return c


def test_symphony_pandya_checkpoints():
    """tests/test_core.cpp:2188-2237 Symphony fast-fit values (rel 1e-5)."""
    g = json.loads((GOLDEN / "symphony_checkpoints.json").read_text())
    st = g["state"]
    nu = st["nu"]
    c = _thermal(abi.EMISSION_THERMAL_PANDYA, nu, st["ne"], st["thetae"], st["b_cgs"], st["theta"])
    exp = g["thermal_pandya"]
    tol = g["rel_tolerance"]
    nu2 = nu * nu
    assert c.jI * nu2 == pytest.approx(exp["jI_phys"], rel=tol)
    assert -c.jQ * nu2 == pytest.approx(exp["minus_jQ_phys"], rel=tol)
    assert c.aI / nu == pytest.approx(exp["alphaI_phys"], rel=tol)
    assert -c.aQ / nu == pytest.approx(exp["minus_alphaQ_phys"], rel=tol)
    assert c.rQ / nu == pytest.approx(exp["rhoQ_phys"], rel=tol)
    assert abs(c.jV * nu2) == pytest.approx(exp["abs_jV_phys"], rel=tol)
    assert abs(c.aV / nu) == pytest.approx(exp["abs_alphaV_phys"], rel=tol)


@pytest.mark.parametrize("fit", [abi.EMISSION_THERMAL_PANDYA, abi.EMISSION_THERMAL_DEXTER])
def test_thermal_kirchhoff_and_symmetry(fit):
    """Kirchhoff's law alpha = j / B_nu for every Stokes component, the
    polarisation-fraction cap, and the theta -> pi - theta symmetry (jI, jQ,
    rho_Q even; jV, rho_V odd)."""
    h, me, c = 6.62607015e-27, 9.1093837015e-28, 2.99792458e10
    nu, ne, thetae, b = 2.3e11, 3e5, 8.0, 12.0
    for theta in (0.3, 1.0, 1.4):
        a = _thermal(fit, nu, ne, thetae, b, theta)
        x = h * nu / (me * c * c * thetae)
        bnu_inv = 2 * h / math.expm1(x) / (c * c)
        assert a.aI == pytest.approx(a.jI / bnu_inv, rel=1e-12)
        assert a.aQ == pytest.approx(a.jQ / bnu_inv, rel=1e-12)
        assert a.aV == pytest.approx(a.jV / bnu_inv, rel=1e-12)
        assert math.hypot(a.jQ, a.jV) <= 0.99 * a.jI * (1 + 1e-12)
        m = _thermal(fit, nu, ne, thetae, b, math.pi - theta)
        assert m.jI == pytest.approx(a.jI, rel=1e-9)
        assert m.rQ == pytest.approx(a.rQ, rel=1e-9)
        assert m.jV == pytest.approx(-a.jV, rel=1e-9)
        assert m.rV == pytest.approx(-a.rV, rel=1e-9)
        assert a.jU == 0.0 and a.aU == 0.0 and a.rU == 0.0


def test_dexter_fit_limits():
    """Dexter (2016) thermal fit: jI > 0, jQ of opposite sign to the
    ipole-convention Q axis (jQ_phys_ipole > 0 -> jQ_inv > 0 here), and
    jV -> 0 as theta -> pi/2 (cos theta -> 0)."""
    c = _thermal(abi.EMISSION_THERMAL_DEXTER, 2.3e11, 1e5, 20.0, 10.0, math.pi / 2 - 1e-9)
    assert c.jI > 0 and c.jQ > 0
    assert abs(c.jV) < 1e-6 * c.jI


def test_bessel_k_polynomials_against_scipy():
    """NR polynomial K0/K1 (thermal_synchrotron.hpp:113-165) to their ~1e-7
    published accuracy and K2 by recurrence."""
    import scipy.special as sp
    L = oracle.lib()
    for x in (0.05, 0.3, 1.0, 1.99, 2.01, 5.0, 30.0):
        k0, k1, k2 = C.c_double(), C.c_double(), C.c_double()
        L.kpo_bessel_k012(x, C.byref(k0), C.byref(k1), C.byref(k2))
        assert k0.value == pytest.approx(sp.kv(0, x), rel=2e-6)
        assert k1.value == pytest.approx(sp.kv(1, x), rel=2e-6)
        assert k2.value == pytest.approx(sp.kv(2, x), rel=5e-6)


# ---------------------------------------------------------------- geometry

def _mat(fn, mass, spin, x, n):
    L = oracle.lib()
    out = np.zeros(n)
    xa = np.asarray(x, dtype=np.float64)
    getattr(L, fn)(mass, spin, oracle.dptr(xa), oracle.dptr(out))
    return out


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.94])
def test_bl_metric_inverse_and_derivatives(spin):
    x = [0.0, 7.3, 1.1, 0.4]
    g = _mat("kpo_bl_gcov", 1.0, spin, x, 16).reshape(4, 4)
    gi = _mat("kpo_bl_gcon", 1.0, spin, x, 16).reshape(4, 4)
    np.testing.assert_allclose(g @ gi, np.eye(4), atol=1e-13)
    dg = _mat("kpo_bl_dgcov", 1.0, spin, x, 64).reshape(4, 4, 4)
    for al in (1, 2):
        hstep = 1e-5
        xp = list(x); xp[al] += hstep
        xm = list(x); xm[al] -= hstep
        fd = (_mat("kpo_bl_gcov", 1.0, spin, xp, 16) - _mat("kpo_bl_gcov", 1.0, spin, xm, 16)) / (2 * hstep)
        np.testing.assert_allclose(dg[:, :, al], fd.reshape(4, 4), rtol=1e-7, atol=1e-9)
    np.testing.assert_array_equal(dg[:, :, 0], 0.0)
    np.testing.assert_array_equal(dg[:, :, 3], 0.0)


def _ray_state(spin, r=6.0):
    """A null ray with a screen basis at radius r (from the camera construction)."""
    p = abi.pass_a_defaults()
    p.coordinate_system = abi.COORD_BOYER_LINDQUIST
    p.spin = spin
    p.camera.radius = r
    p.camera.inclination = 1.1
    p.camera.nx = p.camera.ny = 4
    p.camera.fov = 0.6
    s = oracle.State()
    oracle.lib().kpo_bl_camera_ray(C.byref(p), 1, C.byref(s))
    return p, s


@pytest.mark.parametrize("spin", [0.0, 0.7, 0.94])
def test_specialised_rhs_equals_generic_connection(spin):
    """kerr_boyer_lindquist.hpp:306-322 vs rk4.hpp:7-34 with connection.hpp:7-24."""
    L = oracle.lib()
    _, s = _ray_state(spin)
    a, b = oracle.State(), oracle.State()
    L.kpo_bl_transport_rhs(1.0, spin, C.byref(s), C.byref(a))
    L.kpo_bl_transport_rhs_generic(1.0, spin, C.byref(s), C.byref(b))
    np.testing.assert_allclose(oracle.state_array(a), oracle.state_array(b), rtol=1e-12, atol=1e-14)


def test_connection_symmetric_and_christoffel_identity():
    """Gamma^mu_ab symmetric in (a,b) and Gamma^a_{a r} = d_r ln sqrt(-g)."""
    spin = 0.8
    x = [0.0, 5.0, 0.9, 0.0]
    gam = _mat("kpo_bl_connection", 1.0, spin, x, 64).reshape(4, 4, 4)
    np.testing.assert_allclose(gam, gam.transpose(0, 2, 1), atol=1e-15)

    def lnsqrtg(xx):
        return 0.5 * math.log(-np.linalg.det(_mat("kpo_bl_gcov", 1.0, spin, xx, 16).reshape(4, 4)))

    for al in (1, 2):
        hstep = 1e-5
        xp = list(x); xp[al] += hstep
        xm = list(x); xm[al] -= hstep
        fd = (lnsqrtg(xp) - lnsqrtg(xm)) / (2 * hstep)
        assert sum(gam[a, a, al] for a in range(4)) == pytest.approx(fd, rel=1e-7)


@pytest.mark.parametrize("spin", [0.0, 0.94])
def test_camera_ray_normalisation(spin):
    """camera.hpp:484-513: k null, -e0.k = 1 for the camera observer, screen
    basis orthonormal and orthogonal to k (tests/test_core.cpp:222-233,742-761)."""
    L = oracle.lib()
    p = abi.pass_a_defaults()
    p.coordinate_system = abi.COORD_BOYER_LINDQUIST
    p.spin = spin
    p.camera.radius = 1000.0
    p.camera.inclination = 163.0 * math.pi / 180
    p.camera.nx = p.camera.ny = 16
    p.camera.fov = 0.02
    tet = np.zeros(16)
    L.kpo_bl_camera_tetrad(C.byref(p), oracle.dptr(tet))
    x = np.array([0.0, 1000.0, p.camera.inclination, 0.0])
    g = _mat("kpo_bl_gcov", 1.0, spin, x, 16).reshape(4, 4)
    e = tet.reshape(4, 4)
    np.testing.assert_allclose(e @ g @ e.T, np.diag([-1.0, 1.0, 1.0, 1.0]), atol=1e-11)
    for pix in (0, 37, 255):
        s = oracle.State()
        L.kpo_bl_camera_ray(C.byref(p), pix, C.byref(s))
        v = oracle.state_array(s)
        k, e1, e2 = v[4:8], v[8:12], v[12:16]
        assert abs(k @ g @ k) < 1e-12
        assert -(e[0] @ g @ k) == pytest.approx(1.0, abs=1e-12)
        assert e1 @ g @ e1 == pytest.approx(1.0, abs=1e-12)
        assert e2 @ g @ e2 == pytest.approx(1.0, abs=1e-12)
        assert abs(e1 @ g @ e2) < 1e-12 and abs(e1 @ g @ k) < 1e-12 and abs(e2 @ g @ k) < 1e-12


def _integrate(spin, s, h, n):
    L = oracle.lib()
    cur = s
    for _ in range(n):
        out = oracle.State()
        L.kpo_bl_rk4_step(1.0, spin, C.byref(cur), h, C.byref(out))
        cur = out
    return oracle.state_array(cur)


@pytest.mark.parametrize("spin", [0.0, 0.7])
def test_rk4_fourth_order_convergence(spin):
    """numerical_validation.cpp:254-292: observed order >= 3.7."""
    _, s = _ray_state(spin, r=20.0)
    total = -4.0
    a = _integrate(spin, s, total / 8, 8)
    b = _integrate(spin, s, total / 16, 16)
    c = _integrate(spin, s, total / 32, 32)
    order = math.log2(np.max(np.abs(a - b)) / np.max(np.abs(b - c)))
    assert order >= 3.7


def test_transport_preserves_null_and_frame():
    """Parallel transport keeps k.k = 0, e_i.k = 0, e_i.e_j = delta_ij (state.hpp:54-76)."""
    spin = 0.9
    _, s = _ray_state(spin, r=15.0)
    v = _integrate(spin, s, -0.01, 400)
    g = _mat("kpo_bl_gcov", 1.0, spin, v[:4], 16).reshape(4, 4)
    k, e1, e2 = v[4:8], v[8:12], v[12:16]
    assert abs(k @ g @ k) < 1e-9
    assert abs(e1 @ g @ k) < 1e-9 and abs(e2 @ g @ k) < 1e-9
    assert abs(e1 @ g @ e1 - 1) < 1e-9 and abs(e2 @ g @ e2 - 1) < 1e-9 and abs(e1 @ g @ e2) < 1e-9


def test_adaptive_step_doubling_contract():
    """rk4.hpp:109-161: accepted state is the two-half-step state, mid_state the
    first half, and the error estimate shrinks ~h^5 (tolerance monotonicity,
    numerical_validation.cpp:295-349)."""
    L = oracle.lib()
    spin = 0.5
    _, s = _ray_state(spin, r=10.0)
    st, mid = oracle.State(), oracle.State()
    used, nxt, err = C.c_double(), C.c_double(), C.c_double()
    ok = L.kpo_bl_adaptive_rk4_step(1.0, spin, C.byref(s), -0.2, 1e-3, 1e-5, 1.0, 10,
                                    C.byref(st), C.byref(mid), C.byref(used), C.byref(nxt), C.byref(err))
    assert ok == 1 and used.value == -0.2
    half = oracle.State()
    L.kpo_bl_rk4_step(1.0, spin, C.byref(s), -0.1, C.byref(half))
    np.testing.assert_array_equal(oracle.state_array(mid), oracle.state_array(half))
    two = oracle.State()
    L.kpo_bl_rk4_step(1.0, spin, C.byref(half), -0.1, C.byref(two))
    np.testing.assert_array_equal(oracle.state_array(st), oracle.state_array(two))
    errs = []
    for hh in (-0.4, -0.2, -0.1):
        L.kpo_bl_adaptive_rk4_step(1.0, spin, C.byref(s), hh, 1.0, 1e-5, 1.0, 10,
                                   C.byref(st), C.byref(mid), C.byref(used), C.byref(nxt), C.byref(err))
        errs.append(err.value)
    assert errs[0] > errs[1] > errs[2]
    assert math.log2(errs[1] / errs[2]) > 4.0
    # a tight tolerance forces a rejection: the used step is smaller than requested
    ok = L.kpo_bl_adaptive_rk4_step(1.0, spin, C.byref(s), -1.0, 1e-12, 1e-5, 1.0, 10,
                                    C.byref(st), C.byref(mid), C.byref(used), C.byref(nxt), C.byref(err))
    assert ok == 1 and abs(used.value) < 1.0 and err.value <= 1e-12


def test_isco_radius_closed_forms():
    """Bardeen-Press-Teukolsky ISCO: 6 for a=0, 1 for a->1 prograde, 9 retrograde."""
    L = oracle.lib()
    assert L.kpo_riaf_isco_radius(0.0) == pytest.approx(6.0, abs=1e-12)
    assert L.kpo_riaf_isco_radius(-1.0) == pytest.approx(9.0, abs=1e-12)
    assert L.kpo_riaf_isco_radius(0.999999) == pytest.approx(1.0, abs=0.03)


def test_riaf_fluid_velocity_normalised():
    """riaf.hpp:210-273: u.u = -1 in all three branches (plunging, Keplerian)."""
    L = oracle.lib()
    p = abi.pass_a_defaults()
    p.spin = 0.94
    m = abi.riaf_defaults()
    m.keplerian_factor, m.infall_factor = 0.8, 0.2
    for r in (1.5, 1.9, 4.0, 20.0):
        x = np.array([0.0, r, 1.2, 0.0])
        u = np.zeros(4)
        L.kpo_riaf_fluid_velocity(C.byref(p), C.byref(m), oracle.dptr(x), oracle.dptr(u))
        g = _mat("kpo_bl_gcov", 1.0, 0.94, x, 16).reshape(4, 4)
        assert u @ g @ u == pytest.approx(-1.0, abs=1e-10)


# ---------------------------------------------------------------------------
# Multi-frequency pass B on shared geodesics (pass_b.hpp:1130-1895)
# ---------------------------------------------------------------------------
def _mf_opts(**kw):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(Path(__file__).resolve().parent.parent / "kpolaris_amd" / "params"
                                         / "riaf_m87_c2.par"))
    for k, v in kw.items():
        o.set(k, str(v))
    return o.resolve_riaf()


def test_oracle_multifrequency_single_list_is_the_single_frequency_image():
    p, m, _, _ = _mf_opts(nx=6, ny=6)
    single, s1 = oracle.run_riaf_image(p, m)
    mf, s2 = oracle.run_riaf_multifrequency_image(p, m, [m.freq_cgs])
    for k in single:
        assert np.array_equal(single[k], mf[k]), k
    assert s1 == s2


def test_oracle_multifrequency_slice_equals_single_frequency_run_under_its_own_control():
    """With the step control on frequency f_k alone, slice k is the
    single-frequency image at f_k, bit for bit, even with depth limits that
    make the step sequence frequency dependent."""
    freqs = [86e9, 230e9, 345e9]
    p, m, _, _ = _mf_opts(nx=5, ny=5, max_absorption_depth=1e-3, max_faraday_depth=1e-2)
    npix = 25
    for k, fk in enumerate(freqs):
        mf, _ = oracle.run_riaf_multifrequency_image(p, m, freqs, [fk])
        mk = copy.deepcopy(m)
        mk.freq_cgs = fk
        single, _ = oracle.run_riaf_image(p, mk)
        for n in ("image_i", "image_q", "image_u", "image_v"):
            assert np.array_equal(mf[n][k * npix:(k + 1) * npix], single[n]), (k, n)
        assert np.array_equal(mf["steps"], single["steps"])


def test_oracle_multifrequency_control_is_the_max_over_the_set():
    """The control list enters only through the largest substep count: a
    reordered list (the recompute path) equals the output list itself (the
    reuse path), and the 86 GHz depth limits dominate this model."""
    freqs = [86e9, 230e9, 345e9]
    p, m, _, _ = _mf_opts(nx=5, ny=5, max_absorption_depth=1e-3, max_faraday_depth=1e-2)
    a, _ = oracle.run_riaf_multifrequency_image(p, m, freqs)
    b, _ = oracle.run_riaf_multifrequency_image(p, m, freqs, freqs[::-1])
    c, _ = oracle.run_riaf_multifrequency_image(p, m, freqs, [86e9])
    for k in a:
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k
    d, _ = oracle.run_riaf_multifrequency_image(p, m, freqs, [345e9])
    assert d["steps"].sum() < a["steps"].sum()
