"""direct_only (pass A ends at the ray's first vertical turn, arXiv:2512.09641)
on the CPU: the oracle's pieces pinned by the reference's own checks
(tests/test_direct_only.cpp: the bracket's truth table, dz/ds against its
closed form and a finite difference of the height in BL and FMKS, Kerr
turn endpoints with dz/ds ~ 0 that head back toward the midplane, rays
without a turn unchanged bit for bit), an image-level inequality the
truncation must satisfy, and the kernel's lanes (tests/emul) against the
oracle."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, grmhd_synth as gs, image
from tests import parity

P = C.POINTER
dp = P(C.c_double)


class Fm(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("startx1", "hslope", "mks_smooth", "poly_alpha", "poly_xt", "poly_norm")]


@pytest.fixture(scope="module")
def olib():
    L = oracle.lib()
    L.kpo_vertical_pv.argtypes = [P(Fm), P(oracle.State), dp, dp]
    L.kpo_vertical_turn_bracket.argtypes = [C.c_double] * 6
    L.kpo_vertical_turn_bracket.restype = C.c_int32
    L.kpo_rk4_step.argtypes = [C.c_double, C.c_double, P(Fm), P(oracle.State), C.c_double, P(oracle.State)]
    L.kpo_bl_pass_a_endpoint.argtypes = [P(abi.PassAParams), C.c_int32, P(oracle.State), P(C.c_int32),
                                         P(C.c_int32)]
    L.kpo_bl_pass_a_endpoint.restype = C.c_int32
    L.kpo_fmks_theta.argtypes = [P(Fm), C.c_double, C.c_double]
    L.kpo_fmks_theta.restype = C.c_double
    L.kpo_fmks_x2_from_theta.argtypes = [P(Fm), C.c_double, C.c_double]
    L.kpo_fmks_x2_from_theta.restype = C.c_double
    return L


def _state(v):
    s = oracle.State()
    for i, x in enumerate(v):
        s.v[i] = x
    return s


def test_vertical_turn_bracket_truth_table(olib):
    """test_direct_only.cpp:21-26: maxima and minima of z in the tracing
    direction are turns; an equatorial crossing, the wrong reversal and an
    equatorial ray are not."""
    b = olib.kpo_vertical_turn_bracket
    assert b(2., 1., 2., -1., 2., 5.)
    assert b(-2., -1., -2., 1., 2., 5.)
    assert b(2., -1., 2., 1., -2., 5.)
    assert not b(1., -1., -1., -1., 2., 5.)
    assert not b(2., -1., 2., 1., 2., 5.)
    assert not b(0., 0., 0., 0., 2., 5.)
    # heights within 64 ulp of the radius scale are on the midplane
    assert not b(1e-14, 1., 1e-14, -1., 2., 1e3)


def test_vertical_velocity_closed_form_and_finite_difference(olib):
    """test_direct_only.cpp:39-58: dz/ds = cos(theta) dr - r sin(theta) dtheta
    in BL; in FMKS the height r cos(theta(x1, x2)) differentiated along the
    native wavevector, against the closed form and a central difference of
    the height along the straight line x + k s."""
    r, th, dr, dth = 7.0, 1.1, -0.6, 0.08
    expected = math.cos(th) * dr - r * math.sin(th) * dth
    z, v = C.c_double(), C.c_double()
    olib.kpo_vertical_pv(None, _state([0, r, th, .4, 1, dr, dth, .03] + [0] * 8), C.byref(z), C.byref(v))
    assert z.value == pytest.approx(r * math.cos(th), rel=1e-15)
    assert v.value == pytest.approx(expected, rel=1e-14)
    fm = Fm(0.0, 0.3, 0.5, 14.0, 0.82, 0.75782)
    x1 = math.log(r)
    x2 = olib.kpo_fmks_x2_from_theta(C.byref(fm), x1, th)
    e = 1e-6
    t = lambda a, b: olib.kpo_fmks_theta(C.byref(fm), a, b)  # noqa: E731
    d1 = (t(x1 + e, x2) - t(x1 - e, x2)) / (2 * e)
    d2 = (t(x1, x2 + e) - t(x1, x2 - e)) / (2 * e)
    k1 = dr / r
    k2 = (dth - d1 * k1) / d2
    olib.kpo_vertical_pv(C.byref(fm), _state([0, x1, x2, .4, 1, k1, k2, .03] + [0] * 8), C.byref(z), C.byref(v))
    assert z.value == pytest.approx(r * math.cos(th), rel=1e-12)
    assert v.value == pytest.approx(expected, rel=1e-7)
    eps = 1e-5
    height = lambda s: math.exp(x1 + k1 * s) * math.cos(t(x1 + k1 * s, x2 + k2 * s))  # noqa: E731
    assert v.value == pytest.approx((height(eps) - height(-eps)) / (2 * eps), rel=1e-8)


def _kerr_params(inclination_deg):
    p = abi.pass_a_defaults()
    p.camera.nx = p.camera.ny = 6
    p.camera.radius = 50.0
    p.camera.fov = 0.4
    p.camera.inclination = math.radians(inclination_deg)
    p.inner_radius, p.outer_radius = 1.4, 30.0
    p.spin = 0.9375
    p.adaptive = 1
    p.adaptive_tolerance = 1e-12
    p.min_step, p.max_step, p.max_steps = 1e-10, 0.25, 200000
    p.coordinate_system = abi.COORD_BOYER_LINDQUIST
    return p


def test_kerr_turn_endpoints(olib):
    """test_direct_only.cpp:93-135 on Boyer-Lindquist geodesics: every
    endpoint is valid; at a turn endpoint dz/ds vanishes (1e-10) and a step
    toward the camera moves back toward the midplane; rays without a turn
    keep the full segment's reason, step count and endpoint bit for bit."""
    turns = no_turns = 0
    for inc in (17.0, 85.0, 163.0):
        p = _kerr_params(inc)
        full = _kerr_params(inc)
        p.direct_only = 1
        for pix in range(36):
            s, steps, reason = oracle.State(), C.c_int32(), C.c_int32()
            assert olib.kpo_bl_pass_a_endpoint(C.byref(p), pix, C.byref(s), C.byref(steps), C.byref(reason))
            if reason.value == abi.REASON_REACHED_DIRECT_TURN:
                turns += 1
                z, v = C.c_double(), C.c_double()
                olib.kpo_vertical_pv(None, C.byref(s), C.byref(z), C.byref(v))
                assert abs(v.value) < 1e-10, (inc, pix, v.value)
                after = oracle.State()
                olib.kpo_rk4_step(p.mass, p.spin, None, C.byref(s), 0.001, C.byref(after))
                z1, v1 = C.c_double(), C.c_double()
                olib.kpo_vertical_pv(None, C.byref(after), C.byref(z1), C.byref(v1))
                assert z1.value * v1.value < 0.0, (inc, pix)
            else:
                no_turns += 1
                s2, steps2, reason2 = oracle.State(), C.c_int32(), C.c_int32()
                olib.kpo_bl_pass_a_endpoint(C.byref(full), pix, C.byref(s2), C.byref(steps2), C.byref(reason2))
                assert (reason.value, steps.value) == (reason2.value, steps2.value)
                assert all(s.v[i] == s2.v[i] for i in range(16))
    assert turns > 0 and no_turns > 0, (turns, no_turns)


def _riaf(nx):
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", nx)
    p, m, _, _ = o.resolve_riaf()
    return p, m


def test_direct_only_image_bounded_by_the_full_image():
    """Dropping the emission behind the first vertical turn keeps the Stokes
    vector physical and can only remove intensity: per pixel the direct-only
    I is at most the full image's (up to the two runs' different pass-B step
    sequences), and the total is smaller."""
    p, m = _riaf(24)
    full, _ = oracle.run_riaf_image(p, m)
    p.direct_only = 1
    d, _ = oracle.run_riaf_image(p, m)
    peak = float(full["image_i"].max())
    assert np.all(d["image_i"] <= full["image_i"] * 1.002 + 1e-8 * peak)
    assert d["image_i"].sum() < full["image_i"].sum()
    pol = np.sqrt(d["image_q"] ** 2 + d["image_u"] ** 2 + d["image_v"] ** 2)
    assert np.all(pol <= d["image_i"] * (1 + 1e-9) + 1e-12 * peak)
    assert (d["pass_a_steps"] <= full["pass_a_steps"]).all() and (d["pass_a_steps"] < full["pass_a_steps"]).any()


@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_run_riaf_image.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), P(abi.ImageData), P(C.c_int64)]
    L.kp_emul_run_riaf_image.restype = C.c_int32
    L.kp_emul_run_grmhd_image.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), dp, C.c_int32, dp, C.c_int32,
                                          P(abi.ImageData), P(C.c_int64)]
    L.kp_emul_run_grmhd_image.restype = C.c_int32
    return L


def _close(got, ref, nfreq, npix):
    assert np.array_equal(got["reason"], ref["reason"])
    assert np.array_equal(got["steps"], ref["steps"])
    assert np.array_equal(got["pass_a_steps"], ref["pass_a_steps"])
    for f in range(nfreq):
        sl = slice(f * npix, (f + 1) * npix)
        peak = np.abs(ref["image_i"][sl]).max()
        for n in parity.STOKES:
            err = np.abs(got[n][sl] - ref[n][sl])
            assert np.all(err <= 1e-10 * np.abs(ref[n][sl]) + 1e-8 * peak), (f, n, err.max() / peak)


def test_emulated_direct_only_lanes_bl(emul):
    """The kernel's pass-A lanes with the turn bisection spread over ticks,
    against the oracle on the RIAF image (Boyer-Lindquist)."""
    p, m = _riaf(16)
    p.direct_only = 1
    ref, _ = oracle.run_riaf_image(p, m)
    arrays = oracle.alloc_image(256)
    rhs = C.c_int64()
    assert emul.kp_emul_run_riaf_image(C.byref(p), C.byref(m), C.byref(oracle.image_struct(arrays)),
                                       C.byref(rhs)) == 0
    _close(arrays, ref, 1, 256)


def test_emulated_direct_only_lanes_fmks(emul):
    """The same on FMKS geodesics (GRMHD snapshot, two frequencies)."""
    cells, m = gs.make_snapshot(40, 24, 24, seed=3)
    oracle.grmhd_derive_scalars(m, cells)
    p = gs.pass_a_params(m, nx=8, ny=8)
    p.direct_only = 1
    F = np.array([86e9, 230e9])
    ref, _ = oracle.run_grmhd_image(p, m, cells, F)
    mm = abi.GRMHDModel.from_buffer_copy(m)
    mm.cells = C.c_void_p(cells.ctypes.data)
    arrays = oracle.alloc_image(64)
    for n in parity.STOKES:
        arrays[n] = np.zeros(2 * 64)
    rhs = C.c_int64()
    assert emul.kp_emul_run_grmhd_image(C.byref(p), C.byref(mm), F.ctypes.data_as(dp), 2, F.ctypes.data_as(dp), 2,
                                        C.byref(oracle.image_struct(arrays)), C.byref(rhs)) == 0
    _close(arrays, ref, 2, 64)
