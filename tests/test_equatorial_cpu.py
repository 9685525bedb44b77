"""Equatorial emission wedge (emission outside |cos theta| <= h/r removed,
absorption and Faraday terms kept; emission_selection.hpp:11-122) on the
CPU: the oracle's pieces pinned by the reference's own checks
(tests/test_equatorial.cpp: the analytic emitting slab behind an absorbing
screen on a straight test path, the boundary split at a step's midpoint,
cos theta of the physical angle in BL and FMKS, what the selection removes
and keeps), the image-level properties of tests/test_equatorial.py (h/r 0
and 1 change nothing, a thin wedge keeps the geodesics and a fraction of the
flux, finer sampling converges, direct_only inside the wedge takes less),
and the kernel's lanes (tests/emul) against the oracle."""
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
FLAT = 2  # the oracle's flat test path: zero connection, radius 1, height ratio x[3]


class Fm(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("startx1", "hslope", "mks_smooth", "poly_alpha", "poly_xt", "poly_norm")]


class Coeffs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("jI", "jQ", "jU", "jV", "aI", "aQ", "aU", "aV", "rQ", "rU", "rV")]


class Stokes(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("I", "Q", "U", "V")]


@pytest.fixture(scope="module")
def olib():
    L = oracle.lib()
    Sel, S = P(abi.EmissionSelection), P(oracle.State)
    L.kpo_wedge_substeps.argtypes = [Sel, C.c_int32, P(Fm), S, S, S, C.c_double]
    L.kpo_wedge_substeps.restype = C.c_int32
    L.kpo_refine_wedge.argtypes = [Sel, C.c_int32, P(Fm), S, S, S, dp, C.c_int32, C.c_double]
    L.kpo_refine_wedge.restype = C.c_int32
    L.kpo_select_emission.argtypes = [Sel, C.c_int32, P(Fm), S, P(Coeffs)]
    L.kpo_height_ratio.argtypes = [C.c_int32, P(Fm), S]
    L.kpo_height_ratio.restype = C.c_double
    L.kpo_path_state.argtypes = [C.c_int32, C.c_double, C.c_double, P(Fm), S, C.c_double, C.c_int32, S]
    L.kpo_semi_analytic_step.argtypes = [P(Stokes), P(Coeffs), C.c_double]
    L.kpo_fmks_x2_from_theta.argtypes = [P(Fm), C.c_double, C.c_double]
    L.kpo_fmks_x2_from_theta.restype = C.c_double
    return L


def _state(v):
    s = oracle.State()
    for i, x in enumerate(v):
        s.v[i] = x
    return s


def _sel(hr, samples=8, faraday=1):
    s = abi.EmissionSelection()
    s.equatorial_h_over_r, s.equatorial_samples, s.faraday_rotation = hr, samples, faraday
    return s


def _path(L, first, h, composed):
    out = oracle.State()
    L.kpo_path_state(FLAT, 1.0, 0.0, None, C.byref(first), h, composed, C.byref(out))
    return out


@pytest.mark.parametrize("composed", [0, 1])
def test_analytic_slab(olib, composed):
    """test_equatorial.cpp:22-46: a unit emitter inside |z| <= eta behind an
    absorbing screen of depth 9, traced in both directions with every sample
    count: the step control and the boundary split give the analytic
    (1 - e^{-2}) e^{-9} / alpha to 2e-10."""
    for eta in (0.01, 1e-5):
        for samples in (2, 8, 16):
            for direction in (-1.0, 1.0):
                sel = _sel(eta, samples)
                state = _state([0, 1, 0, -10 * eta * direction, 1, 0, 0, direction] + [0] * 8)
                st = Stokes()
                length, absorption = 20 * eta, 1 / eta
                distance, steps = 0.0, 0
                while distance < length * (1 - 1e-13):
                    h = min(13.73 * eta, length - distance)
                    last = _path(olib, state, h, composed)
                    mid = _path(olib, state, h * 0.5, 0)
                    n = olib.kpo_wedge_substeps(C.byref(sel), FLAT, None, C.byref(state), C.byref(mid),
                                                C.byref(last), 2.0)
                    h /= n
                    last = _path(olib, state, h, composed)
                    mid = _path(olib, state, h * 0.5, 0)
                    hh = C.c_double(h)
                    olib.kpo_refine_wedge(C.byref(sel), FLAT, None, C.byref(state), C.byref(mid), C.byref(last),
                                          C.byref(hh), composed, 2.0)
                    h = hh.value
                    c = Coeffs(jI=1.0, aI=absorption)
                    olib.kpo_select_emission(C.byref(sel), FLAT, None, C.byref(mid), C.byref(c))
                    olib.kpo_semi_analytic_step(C.byref(st), C.byref(c), h)
                    state = last
                    distance += h
                    steps += 1
                    assert steps < 10000
                expected = (1 - math.exp(-2 * eta * absorption)) / absorption * math.exp(-9 * eta * absorption)
                assert st.I / expected == pytest.approx(1.0, abs=2e-10), (eta, samples, direction)


def test_boundary_at_the_step_midpoint(olib):
    """test_equatorial.cpp:47-52: a surface exactly at a step's midpoint splits
    the step there."""
    sel = _sel(0.01)
    first = _state([0, 1, 0, -0.03, 1, 0, 0, 1] + [0] * 8)
    mid, last = _path(olib, first, 0.02, 0), _path(olib, first, 0.04, 0)
    h = C.c_double(0.04)
    assert olib.kpo_refine_wedge(C.byref(sel), FLAT, None, C.byref(first), C.byref(mid), C.byref(last),
                                 C.byref(h), 0, 2.0)
    assert h.value == pytest.approx(0.02, abs=1e-14)


def test_height_ratio_and_selection(olib):
    """test_equatorial.cpp:54-85: cos of the physical angle in BL and FMKS;
    outside the wedge every emissivity is removed and absorption and Faraday
    terms are kept; h/r 0 and 1 change nothing and ask for no substeps."""
    r, th = 7.0, 1.1
    s = _state([0, r, th, 0.4] + [0] * 12)
    assert olib.kpo_height_ratio(0, None, C.byref(s)) == pytest.approx(math.cos(th), abs=1e-14)
    fm = Fm(0.0, 0.3, 0.5, 14.0, 0.82, 0.75782)
    x1 = math.log(r)
    sf = _state([0, x1, olib.kpo_fmks_x2_from_theta(C.byref(fm), x1, th), 0.4] + [0] * 12)
    assert olib.kpo_height_ratio(1, C.byref(fm), C.byref(sf)) == pytest.approx(math.cos(th), abs=1e-12)
    base = dict(jI=1, jQ=.2, jU=.3, jV=.1, aI=2, aQ=.02, aU=.03, aV=.01, rQ=.4, rU=.5, rV=.6)
    c = Coeffs(**base)
    olib.kpo_select_emission(C.byref(_sel(0.01)), 0, None, C.byref(s), C.byref(c))
    assert (c.jI, c.jQ, c.jU, c.jV) == (0, 0, 0, 0)
    assert all(getattr(c, k) == base[k] for k in ("aI", "aQ", "aU", "aV", "rQ", "rU", "rV"))
    for width in (0.0, 1.0):
        c = Coeffs(**base)
        olib.kpo_select_emission(C.byref(_sel(width)), 0, None, C.byref(s), C.byref(c))
        assert all(getattr(c, k) == base[k] for k in base)
        assert olib.kpo_wedge_substeps(C.byref(_sel(width)), 0, None, C.byref(s), C.byref(s), C.byref(s),
                                       100.0) == 1


def _riaf(nx=6, **kw):
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", nx)
    for k, v in kw.items():
        o.set(k, v)
    p, m, _, _ = o.resolve_riaf()
    p.adaptive_tolerance, p.max_step, p.max_radiation_step = 1e-12, 0.25, 0.05
    return p, m


def _with(p, hr=None, samples=None, direct=None, max_rs=None):
    q = abi.PassAParams.from_buffer_copy(p)
    if hr is not None:
        q.emission_selection.equatorial_h_over_r = hr
    if samples is not None:
        q.emission_selection.equatorial_samples = samples
    if direct is not None:
        q.direct_only = direct
    if max_rs is not None:
        q.max_radiation_step = max_rs
    return q


def _stokes(img):
    return np.stack([img[k] for k in parity.STOKES])


def test_oracle_wedge_images():
    """tests/test_equatorial.py:27-80 on the RIAF image at 17 degrees: h/r 0
    and 1 give the full image bit for bit; h/r 0.01 keeps every geodesic
    (pass-A steps) and 0 < flux fraction < 0.9; twice the samples with half
    the radiation step agree within 1.5 % (L1); direct_only inside the wedge
    takes at most the wedge image's flux."""
    p, m = _riaf(6, inclination_deg=17.0, radius=100.0, xspan=10.0, riaf_r_max=30.0, outer_radius=34.5)
    full, _ = oracle.run_riaf_image(p, m)
    assert (full["reason"] == abi.REASON_REACHED_CAMERA).all()
    for hr in (0.0, 1.0):
        other, _ = oracle.run_riaf_image(_with(p, hr=hr), m)
        assert np.array_equal(_stokes(other), _stokes(full)), hr
    thin, _ = oracle.run_riaf_image(_with(p, hr=0.01), m)
    assert (thin["reason"] == abi.REASON_REACHED_CAMERA).all()
    assert np.array_equal(thin["pass_a_steps"], full["pass_a_steps"])
    fraction = thin["image_i"].sum() / full["image_i"].sum()
    assert 0.0 < fraction < 0.9, fraction
    refined, _ = oracle.run_riaf_image(_with(p, hr=0.01, samples=16, max_rs=0.025), m)
    l1 = np.abs(_stokes(refined) - _stokes(thin)).sum() / np.abs(_stokes(thin)).sum()
    assert l1 < 0.015, l1
    direct, _ = oracle.run_riaf_image(_with(p, hr=0.01, direct=1), m)
    assert 0.0 < direct["image_i"].sum() <= thin["image_i"].sum() * 1.005


@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_run_riaf_image.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), P(abi.ImageData), P(C.c_int64)]
    L.kp_emul_run_riaf_image.restype = C.c_int32
    L.kp_emul_run_riaf_image_mixed.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), C.c_double,
                                               P(abi.ImageData), P(C.c_int64), P(C.c_int64)]
    L.kp_emul_run_riaf_image_mixed.restype = C.c_int32
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
        assert peak > 0.0
        for n in parity.STOKES:
            err = np.abs(got[n][sl] - ref[n][sl])
            assert np.all(err <= 1e-10 * np.abs(ref[n][sl]) + 1e-8 * peak), (f, n, err.max() / peak)


@pytest.mark.parametrize("adaptive", [1, 0])
def test_emulated_wedge_lanes_bl(emul, adaptive):
    """The kernel's pass-B lanes with the wedge (substep rule, boundary split,
    emission removed outside; fixed steps sample the step midpoint) against
    the oracle, adaptive and fixed-step, and with direct_only."""
    p, m = _riaf(12)
    p = _with(p, hr=0.05)
    p.adaptive = adaptive
    if not adaptive:
        # fixed steps from an inner boundary next to the horizon leave the
        # geodesic in Boyer-Lindquist coordinates (the state diverges in the
        # oracle and the kernel alike); stop rays at 2.5 M instead
        p.step = 0.05
        p.inner_radius = 2.5
    for direct in (0, 1):
        q = _with(p, direct=direct)
        ref = parity.oracle_with_sensitivity(q, m)
        arrays = oracle.alloc_image(144)
        rhs = C.c_int64()
        assert emul.kp_emul_run_riaf_image(C.byref(q), C.byref(m), C.byref(oracle.image_struct(arrays)),
                                           C.byref(rhs)) == 0
        print(parity.assert_parity(arrays, ref, f"emulated wedge adaptive {adaptive} direct {direct}",
                                   max_ill=parity.WEDGE_MAX_ILL, step_rel=parity.WEDGE_STEP_REL))


def test_emulated_wedge_lanes_mixed(emul):
    """The mixed-precision lanes with the wedge at the mixed bars (the
    production tolerance 1e-6: FP32-derived geodesic steps cannot meet the
    1e-12 of the FP64 cases)."""
    p, m = _riaf(16)
    p = _with(p, hr=0.05)
    p.adaptive_tolerance = 1e-6
    ref, _ = oracle.run_riaf_image(p, m)
    arrays = oracle.alloc_image(256)
    rhs, guard = C.c_int64(), C.c_int64()
    assert emul.kp_emul_run_riaf_image_mixed(C.byref(p), C.byref(m), abi.FP64_RADIUS_DEFAULT,
                                             C.byref(oracle.image_struct(arrays)), C.byref(rhs),
                                             C.byref(guard)) == 0
    parity.assert_mixed(arrays, ref, "emulated mixed wedge", pol="absolute")


def test_emulated_wedge_lanes_fmks(emul):
    """The wedge on FMKS geodesics (GRMHD snapshot, two frequencies on the
    single-pass loop)."""
    cells, m = gs.make_snapshot(40, 24, 24, seed=3)
    oracle.grmhd_derive_scalars(m, cells)
    p = _with(gs.pass_a_params(m, nx=8, ny=8), hr=0.1)
    F = np.array([86e9, 230e9])
    run = lambda q: oracle.run_grmhd_image(q, m, cells, F)[0]  # noqa: E731
    ref = run(p)
    cond = parity.sensitivity(run, p, ref)
    mm = abi.GRMHDModel.from_buffer_copy(m)
    mm.cells = C.c_void_p(cells.ctypes.data)
    arrays = oracle.alloc_image(64)
    for n in parity.STOKES:
        arrays[n] = np.zeros(2 * 64)
    rhs = C.c_int64()
    assert emul.kp_emul_run_grmhd_image(C.byref(p), C.byref(mm), F.ctypes.data_as(dp), 2, F.ctypes.data_as(dp), 2,
                                        C.byref(oracle.image_struct(arrays)), C.byref(rhs)) == 0
    print(parity.assert_parity_grid(arrays, ref, 2, "emulated wedge fmks", cond, max_ill=parity.WEDGE_MAX_ILL,
                                    step_rel=parity.WEDGE_STEP_REL))
