"""GRMHD path on the CPU: the oracle's FMKS geometry and GRMHD grid model
pinned by closed forms and finite differences, the synthetic snapshot
generator, and the kernel's lane code (tests/emul) against the oracle."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, grmhd_synth as gs
from tests import parity

dp = C.POINTER(C.c_double)


class Fm(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("startx1", "hslope", "mks_smooth", "poly_alpha", "poly_xt", "poly_norm")]


FM = Fm(math.log(1.2), 0.3, 0.5, 14.0, 0.82, gs.poly_norm(14.0, 0.82))


@pytest.fixture(scope="module")
def olib():
    L = oracle.lib()
    L.kpo_fmks_transport_rhs.argtypes = [C.c_double, C.c_double, C.POINTER(Fm), C.POINTER(oracle.State),
                                         C.POINTER(oracle.State)]
    L.kpo_fmks_transport_rhs_fd.argtypes = [C.c_double, C.c_double, C.POINTER(Fm), C.POINTER(oracle.State),
                                            C.c_double, C.POINTER(oracle.State)]
    L.kpo_fmks_gcov.argtypes = [C.c_double, C.c_double, C.POINTER(Fm), dp, dp]
    L.kpo_fmks_gcon.argtypes = [C.c_double, C.c_double, C.POINTER(Fm), dp, dp]
    L.kpo_fmks_theta.argtypes = [C.POINTER(Fm), C.c_double, C.c_double]
    L.kpo_fmks_theta.restype = C.c_double
    L.kpo_fmks_x2_from_theta.argtypes = [C.POINTER(Fm), C.c_double, C.c_double]
    L.kpo_fmks_x2_from_theta.restype = C.c_double
    L.kpo_grmhd_sample_prim.argtypes = [C.POINTER(abi.GRMHDModel), C.c_void_p, C.c_int, C.c_double, C.c_double,
                                        C.c_double]
    L.kpo_grmhd_sample_prim.restype = C.c_double
    return L


@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_fmks_rhs.argtypes = [C.c_double, C.c_double, dp, dp, dp]
    L.kp_emul_grmhd_derive_scalars.argtypes = [C.POINTER(abi.GRMHDModel), C.c_void_p]
    L.kp_emul_grmhd_derive_scalars.restype = C.c_int32
    L.kp_emul_run_grmhd_image.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.GRMHDModel), dp, C.c_int32, dp,
                                          C.c_int32, C.POINTER(abi.ImageData), C.POINTER(C.c_int64)]
    L.kp_emul_run_grmhd_image.restype = C.c_int32
    return L


def _state(rng):
    v = np.concatenate([[0.0, math.log(rng.uniform(1.6, 60.0)), rng.uniform(0.05, 0.95), rng.uniform(0, 6.2)],
                        rng.normal(size=12)])
    s = oracle.State()
    for i in range(16):
        s.v[i] = v[i]
    return s, v


def test_fmks_theta_map_and_inverse(olib):
    """theta(x1, x2) is the MKS map far from the hole, monotonic in x2, and
    native_x2_from_theta inverts it (64 bisections: 2^-64)."""
    for x1 in (math.log(1.5), math.log(10.0), math.log(900.0)):
        th = [olib.kpo_fmks_theta(C.byref(FM), x1, x2) for x2 in np.linspace(0.0, 1.0, 101)]
        assert np.all(np.diff(th) > 0.0)
        assert th[0] == pytest.approx(0.0, abs=1e-12) and th[-1] == pytest.approx(math.pi, abs=1e-12)
        for t in (0.3, 1.2, 2.8):
            x2 = olib.kpo_fmks_x2_from_theta(C.byref(FM), x1, t)
            assert olib.kpo_fmks_theta(C.byref(FM), x1, x2) == pytest.approx(t, abs=1e-13)
    # far from the hole the jet term is suppressed by exp(mks_smooth (startx1 - x1))
    x1 = math.log(1e6)
    x2 = 0.3
    mks = math.pi * x2 + 0.35 * math.sin(2 * math.pi * x2)
    assert olib.kpo_fmks_theta(C.byref(FM), x1, x2) == pytest.approx(mks, abs=1e-3)


def test_fmks_metric_inverse_and_ks_congruence(olib):
    rng = np.random.default_rng(5)
    for _ in range(5):
        _, v = _state(rng)
        x = (C.c_double * 4)(*v[:4])
        g, gi = (C.c_double * 16)(), (C.c_double * 16)()
        olib.kpo_fmks_gcov(1.0, 0.9375, C.byref(FM), x, g)
        olib.kpo_fmks_gcon(1.0, 0.9375, C.byref(FM), x, gi)
        G, Gi = np.array(g[:]).reshape(4, 4), np.array(gi[:]).reshape(4, 4)
        assert np.abs(G @ Gi - np.eye(4)).max() < 1e-10
        assert np.allclose(G, G.T, rtol=0, atol=1e-12 * np.abs(G).max())
        # det g = det g_KS (r dtheta/dx2)^2, det g_KS = -rho^4 sin^2 theta
        r = math.exp(v[1])
        th = olib.kpo_fmks_theta(C.byref(FM), v[1], v[2])
        h = 1e-7
        dth2 = (olib.kpo_fmks_theta(C.byref(FM), v[1], v[2] + h) - olib.kpo_fmks_theta(C.byref(FM), v[1], v[2] - h)) / (2 * h)
        rho2 = r * r + 0.9375 ** 2 * math.cos(th) ** 2
        assert np.linalg.det(G) == pytest.approx(-(rho2 * math.sin(th) * r * dth2) ** 2, rel=1e-6)


def test_fmks_specialised_rhs_equals_christoffel_from_metric(olib, emul):
    """The spherical-KS route of kerr_fmks.hpp:455-479 against Gamma from
    central differences of gcov; the device restatement against the oracle."""
    rng = np.random.default_rng(6)
    for _ in range(6):
        s, v = _state(rng)
        a, b = oracle.State(), oracle.State()
        olib.kpo_fmks_transport_rhs(1.0, 0.9375, C.byref(FM), C.byref(s), C.byref(a))
        olib.kpo_fmks_transport_rhs_fd(1.0, 0.9375, C.byref(FM), C.byref(s), 1e-5, C.byref(b))
        A, B = np.array(a.v[:]), np.array(b.v[:])
        assert np.abs(A - B).max() <= 1e-7 * np.abs(A).max()
        fm = np.array([FM.startx1, FM.hslope, FM.mks_smooth, FM.poly_alpha, FM.poly_xt, FM.poly_norm])
        d = np.zeros(16)
        emul.kp_emul_fmks_rhs(1.0, 0.9375, fm.ctypes.data_as(dp), v.ctypes.data_as(dp), d.ctypes.data_as(dp))
        assert np.abs(d - A).max() <= 1e-14 * np.abs(A).max()


def _snapshot(n1=32, n2=16, n3=16, seed=3):
    cells, m = gs.make_snapshot(n1, n2, n3, seed=seed)
    return cells, m


def test_trilinear_sampling_exact_on_linear_fields_and_ghosts(olib):
    """Cell-centred trilinear interpolation reproduces a field linear in the
    cell indices inside the grid (grmhd_grid.hpp:379-430); outside the cell
    centres the radial ghosts clamp and phi wraps periodically (:309-321)."""
    n1, n2, n3 = 12, 10, 8
    cells, m = _snapshot(n1, n2, n3)
    i, j, k = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    lin = (1.0 + 0.5 * i + 0.25 * j + 0.125 * k).astype(np.float32).reshape(-1)
    cells["prim"][:, 3] = lin
    mm = abi.GRMHDModel.from_buffer_copy(m)
    rng = np.random.default_rng(0)
    for _ in range(50):
        fi, fj = rng.uniform(0.0, n1 - 1), rng.uniform(0.0, n2 - 1)
        fk = rng.uniform(0.0, n3 - 1)
        x1 = m.startx1 + (fi + 0.5) * m.dx1
        x2 = m.startx2 + (fj + 0.5) * m.dx2
        ph = m.startx3 + (fk + 0.5) * m.dx3
        got = olib.kpo_grmhd_sample_prim(C.byref(mm), C.c_void_p(cells.ctypes.data), 3, x1, x2, ph)
        assert got == pytest.approx(1.0 + 0.5 * fi + 0.25 * fj + 0.125 * fk, rel=1e-6)
    # phi periodicity: sample at phi and phi + 2 pi agree
    x1 = m.startx1 + 3.3 * m.dx1
    x2 = m.startx2 + 4.1 * m.dx2
    a = olib.kpo_grmhd_sample_prim(C.byref(mm), C.c_void_p(cells.ctypes.data), 3, x1, x2, 0.1)
    b = olib.kpo_grmhd_sample_prim(C.byref(mm), C.c_void_p(cells.ctypes.data), 3, x1, x2, 0.1 + 2 * math.pi)
    assert a == pytest.approx(b, rel=1e-14)
    # inner radial ghost: below the first cell centre the value is the clamped edge value
    lo = olib.kpo_grmhd_sample_prim(C.byref(mm), C.c_void_p(cells.ctypes.data), 3, m.startx1 + 0.1 * m.dx1,
                                    m.startx2 + 4.5 * m.dx2, m.startx3 + 2.5 * m.dx3)
    assert lo == pytest.approx(1.0 + 0.25 * 4 + 0.125 * 2, rel=1e-6)


def test_derived_scalars_closed_forms_and_device_formulas(emul):
    """n_e = rho rho_unit / (m_p + m_e), B = sqrt(b^2) B_unit, sigma = b^2 / rho,
    beta = (gam - 1) u / (b^2 / 2) and the R_high/R_low electron temperature
    (iharm_loader.cpp:420-505); the device formulas equal the oracle's."""
    cells, m = _snapshot()
    dev = cells.copy()
    oracle.grmhd_derive_scalars(m, cells)
    assert emul.kp_emul_grmhd_derive_scalars(C.byref(m), C.c_void_p(dev.ctypes.data)) == 0
    d, e = cells["derived"], dev["derived"]
    np.testing.assert_allclose(e, d, rtol=1e-13, atol=0)
    rho = cells["prim"][:, 0].astype(np.float64)
    uu = cells["prim"][:, 1].astype(np.float64)
    L = 6.67430e-8 * m.mbh_solar * 1.98847e33 / 2.99792458e10 ** 2
    rho_unit = m.M_unit / L ** 3
    np.testing.assert_allclose(d[:, 0], rho * rho_unit / (1.67262171e-24 + 9.1093826e-28), rtol=1e-12)
    bsq = d[:, 3] * rho
    np.testing.assert_allclose(d[:, 4], (m.gam - 1.0) * uu / (0.5 * bsq), rtol=1e-10)
    b_unit = 2.99792458e10 * math.sqrt(4 * math.pi * rho_unit)
    np.testing.assert_allclose(d[:, 2], np.sqrt(bsq) * b_unit, rtol=1e-10)
    beta2 = d[:, 4] ** 2 / m.beta_crit ** 2
    trat = (m.trat_large * beta2 + m.trat_small) / (1.0 + beta2)
    mp_me = 1.67262171e-24 / 9.1093826e-28
    th_unit = mp_me * (1.0 / 3.0) * (2.0 / 3.0) / ((2.0 / 3.0) + (1.0 / 3.0) * trat)
    np.testing.assert_allclose(d[:, 1], th_unit * uu / rho, rtol=1e-12)


def test_synthetic_snapshot_is_seeded_and_physical():
    a, ma = _snapshot(24, 16, 16, seed=9)
    b, _ = _snapshot(24, 16, 16, seed=9)
    assert np.array_equal(a["prim"], b["prim"])
    p = a["prim"].reshape(24, 16, 16, 8)
    assert np.all(np.isfinite(p)) and np.all(p[..., 0] > 0) and np.all(p[..., 1] > 0)
    assert p[..., 0].max() > 0.5  # the torus is resolved
    oracle.grmhd_derive_scalars(ma, a)
    beta = a["derived"][:, 4].reshape(24, 16, 16)
    # the generator normalises min beta over cells above a tenth of the peak density
    torus = p[..., 0] > 0.1 * p[..., 0].max()
    assert beta[torus].min() == pytest.approx(100.0, rel=1e-3)


def test_emulated_grmhd_image_against_oracle(emul):
    """The kernel's FMKS/GRMHD lanes (pass A, pass B with three frequencies
    sharing geodesics) against the oracle at the parity bar."""
    cells, m = _snapshot(40, 24, 24)
    oracle.grmhd_derive_scalars(m, cells)
    p = gs.pass_a_params(m, nx=8, ny=8)
    F = np.array([86e9, 230e9, 345e9])
    ref, st = oracle.run_grmhd_image(p, m, cells, F)
    mm = abi.GRMHDModel.from_buffer_copy(m)
    mm.cells = C.c_void_p(cells.ctypes.data)
    arrays = oracle.alloc_image(64)
    for n in parity.STOKES:
        arrays[n] = np.zeros(3 * 64)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    assert emul.kp_emul_run_grmhd_image(C.byref(p), C.byref(mm), F.ctypes.data_as(dp), 3, F.ctypes.data_as(dp), 3,
                                        C.byref(img), C.byref(rhs)) == 0
    assert rhs.value == st["rhs_evals"]
    assert np.array_equal(arrays["reason"], ref["reason"]) and (ref["reason"] == 1).all()
    assert np.array_equal(arrays["steps"], ref["steps"])
    for f in range(3):
        sl = slice(f * 64, (f + 1) * 64)
        peak = np.abs(ref["image_i"][sl]).max()
        assert peak > 0.0
        for n in parity.STOKES:
            err = np.abs(arrays[n][sl] - ref[n][sl])
            assert np.all(err <= 1e-10 * np.abs(ref[n][sl]) + 1e-8 * peak), (f, n, err.max() / peak)


def test_emulated_grmhd_mixed_against_oracle(emul):
    """The mixed-precision GRMHD lanes (FP32 thermal coefficients and transfer
    on the FP64 geodesics, single-pass frequency loop) against the FP64
    oracle at the mixed-precision image bars, per frequency.  The image is
    coarse, so the net polarisation fractions are held in absolute terms."""
    emul.kp_emul_run_grmhd_image_mixed.argtypes = emul.kp_emul_run_grmhd_image.argtypes
    emul.kp_emul_run_grmhd_image_mixed.restype = C.c_int32
    cells, m = _snapshot(40, 24, 24)
    oracle.grmhd_derive_scalars(m, cells)
    p = gs.pass_a_params(m, nx=12, ny=12)
    F = np.array([86e9, 230e9, 345e9])
    ref, _ = oracle.run_grmhd_image(p, m, cells, F)
    mm = abi.GRMHDModel.from_buffer_copy(m)
    mm.cells = C.c_void_p(cells.ctypes.data)
    n = 144
    arrays = oracle.alloc_image(n)
    for k in parity.STOKES:
        arrays[k] = np.zeros(3 * n)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    assert emul.kp_emul_run_grmhd_image_mixed(C.byref(p), C.byref(mm), F.ctypes.data_as(dp), 3,
                                              F.ctypes.data_as(dp), 3, C.byref(img), C.byref(rhs)) == 0
    got = parity.split_frequencies(arrays, 3)
    exp = parity.split_frequencies(ref, 3)
    for f in range(3):
        e = parity.assert_mixed(got[f], exp[f], f"emulated mixed GRMHD f{f}", pol="absolute")
        assert e["reason_mismatch"] == 0, e
