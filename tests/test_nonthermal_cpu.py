"""Nonthermal synchrotron on the CPU: the oracle's kappa and power-law fits
(src/radiation/nonthermal_synchrotron.hpp) pinned by the reference's own
Symphony checkpoints (tests/test_core.cpp:2188-2288) and checks
(tests/test_core.cpp:2147-2185), by an independent 2F1 (scipy), and by the
textbook synchrotron power laws of the fits' limits; then the kernel's lane
code (tests/emul) against the oracle for RIAF and GRMHD images."""
from __future__ import annotations

import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, grmhd_synth as gs, image
from tests import parity

GOLDEN = Path(__file__).resolve().parent / "golden"
PARAMS = Path(__file__).resolve().parent.parent / "kpolaris_amd" / "params"
dp = C.POINTER(C.c_double)
ME, CL = 9.1093826e-28, 2.99792458e10   # the nonthermal header's constants


def _nt(dist, **kw):
    return oracle.Nonthermal(dist, **kw)


def _coeffs(params, nu=230e9, ne=1.0, thetae=10.0, b=30.0, theta=math.pi / 3):
    return oracle.nonthermal_coefficients(params, nu, ne, thetae, b, math.sin(theta), math.cos(theta))


def _eta_unit_density(p, b, gmin=1.0):
    """eta giving n_pl = 1 (the checkpoint's normalisation)."""
    return 2.0 * (p - 1.0) * ME * CL * CL * gmin / (b * b * (p - 2.0))


# ------------------------------------------------------------ checkpoints

def test_symphony_kappa_checkpoint():
    """tests/test_core.cpp:2239-2266: kappa = 3.5, w = 10, all eight
    coefficients at rel 1e-5 (the reference's tolerance)."""
    g = json.loads((GOLDEN / "symphony_checkpoints.json").read_text())
    st, exp, tol = g["state"], g["kappa"], g["rel_tolerance"]
    nu = st["nu"]
    c = _coeffs(_nt(abi.EMISSION_KAPPA, kappa=3.5, kappa_width=10.0, kappa_interp_begin=1e20, kappa_interp_end=1e20),
                nu, st["ne"], st["thetae"], st["b_cgs"], st["theta"])
    nu2 = nu * nu
    got = {"jI_phys": c.jI * nu2, "minus_jQ_phys": -c.jQ * nu2, "jV_phys": c.jV * nu2, "alphaI_phys": c.aI / nu,
           "minus_alphaQ_phys": -c.aQ / nu, "alphaV_phys": c.aV / nu, "rhoQ_phys": c.rQ / nu,
           "rhoV_phys": c.rV / nu}
    for k, v in got.items():
        assert v == pytest.approx(exp[k], rel=tol), k
    assert c.jU == c.aU == c.rU == 0.0


def test_symphony_power_law_checkpoint():
    """tests/test_core.cpp:2268-2286: p = 2.5, gamma in [1, 1000], n_pl = 1."""
    g = json.loads((GOLDEN / "symphony_checkpoints.json").read_text())
    st, exp, tol = g["state"], g["power_law"], g["rel_tolerance"]
    nu = st["nu"]
    prm = _nt(abi.EMISSION_POWER_LAW, power_law_p=2.5, power_law_gamma_min=1.0, power_law_gamma_max=1000.0,
              power_law_eta=_eta_unit_density(2.5, st["b_cgs"]))
    c = _coeffs(prm, nu, st["ne"], st["thetae"], st["b_cgs"], st["theta"])
    nu2 = nu * nu
    got = {"jI_phys": c.jI * nu2, "minus_jQ_phys": -c.jQ * nu2, "jV_phys": c.jV * nu2, "alphaI_phys": c.aI / nu,
           "minus_alphaQ_phys": -c.aQ / nu, "alphaV_phys": c.aV / nu}
    for k, v in got.items():
        assert v == pytest.approx(exp[k], rel=tol), k
    assert c.rQ == 0.0 and c.rV == 0.0


def test_reference_nonthermal_coefficient_checks():
    """tests/test_core.cpp:2147-2185: positive and finite kappa and power-law
    coefficients; no kappa rotativity outside the fit range [3.5, 5]; no
    power-law rotativity."""
    s = 0.7
    args = (230e9, 1e6, 20.0, 30.0, s, math.sqrt(1 - s * s))
    k = oracle.nonthermal_coefficients(_nt(abi.EMISSION_KAPPA, kappa=3.5), *args)
    assert k.jI > 0 and k.aI >= 0
    assert all(math.isfinite(v) for v in k.as_array())
    k6 = oracle.nonthermal_coefficients(_nt(abi.EMISSION_KAPPA, kappa=6.0), *args)
    assert k6.rQ == 0.0 and k6.rV == 0.0
    for kap in (3.4, 5.1):
        kk = oracle.nonthermal_coefficients(_nt(abi.EMISSION_KAPPA, kappa=kap), *args)
        assert kk.rQ == 0.0 and kk.rV == 0.0 and kk.jI > 0
    pl = oracle.nonthermal_coefficients(_nt(abi.EMISSION_POWER_LAW, kappa=6.0), *args)
    assert pl.jI > 0 and pl.aI >= 0 and all(math.isfinite(v) for v in pl.as_array())
    assert pl.rQ == 0.0 and pl.rV == 0.0


def test_reference_edge_cases_kept():
    """Two properties of the reference's fits that the restatement keeps:
    at exactly perpendicular propagation (sin theta = 1, which a field-free
    sample produces) the kappa jV fit divides 0 by 0 (both angle factors
    vanish, :282-294); above kappa ~ 171 the Gamma functions overflow and
    every kappa coefficient is NaN, and the blend toward thermal does not
    remove it ((1 - 1) * NaN, :585-592)."""
    c = oracle.nonthermal_coefficients(_nt(abi.EMISSION_KAPPA, kappa=3.5), 230e9, 1e5, 5.0, 10.0, 1.0, 0.0)
    assert math.isnan(c.jV) and math.isfinite(c.jI) and math.isfinite(c.aV)
    c = oracle.nonthermal_coefficients(_nt(abi.EMISSION_KAPPA, kappa=200.0, kappa_interp_begin=5.0,
                                           kappa_interp_end=7.0), 230e9, 1e5, 5.0, 10.0, 0.6, 0.8)
    assert math.isnan(c.jI)


# ----------------------------------------------------------- 2F1 and limits

@pytest.mark.parametrize("kappa", [3.1, 3.5, 4.25, 5.0, 7.0])
@pytest.mark.parametrize("z", [-0.5, -0.2, 0.0, 0.3, 0.5, -1.0, -1.7, -5.0, -35.0, -400.0])
def test_hyp2f1_against_scipy(kappa, z):
    """stable_hyp2f1 (:83-96) for the kappa absorption's 2F1(k-1/3, k+1; k+2/3; z)
    against scipy's: the Gauss series for |z| <= 1/2 and the 1/(1-z)
    transformation for z <= -1, where both 200-term series converge.  (For
    z in (-1, -1/2) the reference's 200-term series is truncated early and
    the oracle keeps that.)"""
    from scipy.special import hyp2f1
    a, b, c = kappa - 1 / 3, kappa + 1, kappa + 2 / 3
    got = oracle.lib().kpo_hyp2f1(a, b, c, z)
    assert got == pytest.approx(hyp2f1(a, b, c, z), rel=2e-12)


def test_variable_kappa_limits():
    """variable_kappa_from_sigma_beta (:493-502): kappa = 2.8 + 0.7/sqrt(sigma)
    at beta = 0, floored at 3.1; beta -> infinity adds 3.7 sigma^-0.19; an
    invalid sigma or beta returns the given default."""
    f = oracle.lib().kpo_variable_kappa
    assert f(1.0, 0.0, 9.0) == pytest.approx(3.5, abs=1e-15)
    assert f(0.25, 0.0, 9.0) == pytest.approx(4.2, abs=1e-15)
    assert f(100.0, 0.0, 9.0) == 3.1
    assert f(4.0, 1e6, 9.0) == pytest.approx(2.8 + 0.35 + 3.7 * 4.0 ** -0.19, rel=1e-15)
    assert f(0.0, 1.0, 9.0) == 9.0 and f(1.0, -1.0, 9.0) == 9.0


def _phys(c, nu):
    return c.jI * nu * nu, c.aI / nu


@pytest.mark.parametrize("p", [2.1, 2.5, 3.25, 3.5])
def test_power_law_spectral_indices(p):
    """Optically thin synchrotron of a power law: j_nu ~ nu^-(p-1)/2 and
    alpha_nu ~ nu^-(p+4)/2 (Rybicki & Lightman 6.36, 6.53); both linear in
    the nonthermal density (eta)."""
    prm = _nt(abi.EMISSION_POWER_LAW, power_law_p=p)
    j1, a1 = _phys(_coeffs(prm, 100e9), 100e9)
    j2, a2 = _phys(_coeffs(prm, 200e9), 200e9)
    assert j2 / j1 == pytest.approx(2.0 ** (-(p - 1) / 2), rel=1e-12)
    assert a2 / a1 == pytest.approx(2.0 ** (-(p + 4) / 2), rel=1e-12)
    prm2 = _nt(abi.EMISSION_POWER_LAW, power_law_p=p, power_law_eta=0.04)
    j3, a3 = _phys(_coeffs(prm2, 100e9), 100e9)
    assert j3 / j1 == pytest.approx(2.0, rel=1e-13) and a3 / a1 == pytest.approx(2.0, rel=1e-13)


@pytest.mark.parametrize("kappa", [3.5, 4.5, 6.0])
def test_kappa_fit_asymptotes(kappa):
    """The kappa fits interpolate between the low-frequency synchrotron
    spectrum (j ~ nu^1/3, alpha ~ nu^-5/3) and the power-law tail of index
    p = kappa - 1 (j ~ nu^-(kappa-2)/2, alpha ~ nu^-(kappa+3)/2).  Local
    slopes over a factor of two, far on either side of X = 1 (w = 1, B = 10 G:
    X ~ nu / (3e8 Hz); frequencies stay above the coefficients' 1 Hz floor)."""
    prm = _nt(abi.EMISSION_KAPPA, kappa=kappa, kappa_width=1.0)

    def slope(nu):
        ja, aa = _phys(_coeffs(prm, nu, b=10.0), nu)
        jb, ab = _phys(_coeffs(prm, 2 * nu, b=10.0), 2 * nu)
        return math.log2(jb / ja), math.log2(ab / aa)

    sj, sa = slope(4.0)    # X ~ 1e-8
    assert sj == pytest.approx(1 / 3, abs=2e-3) and sa == pytest.approx(-5 / 3, abs=2e-3)
    sj, sa = slope(2e17)   # X ~ 1e9
    assert sj == pytest.approx(-(kappa - 2) / 2, abs=2e-3)
    assert sa == pytest.approx(-(kappa + 3) / 2, abs=2e-3)


def test_kappa_rotativity_interpolates_the_anchor_fits():
    """interpolate_kappa_rho_{Q,V} (:462-489): linear in kappa between the
    anchor fits at kappa = 3.5, 4, 4.5, 5.  The fits depend on kappa through
    X = nu / (w^2 kappa^2 nu_c sin) as well, so the frequency is scaled with
    kappa^2 to hold X fixed, and the fit factor is the rotativity divided by
    its prefactor (-n e^2 nu_c^2 sin^2 / (m c nu^3) for rho_Q,
    2 n e^2 nu_c cos / (m c nu^2) for rho_V, :407-409, :437-438)."""
    ee, b, th, ne = 4.80320680e-10, 20.0, math.pi / 3, 1.0
    nuc = ee * b / (2 * math.pi * ME * CL)

    def fit(k):
        nu = 86e9 * (k / 4.0) ** 2
        c = _coeffs(_nt(abi.EMISSION_KAPPA, kappa=k, kappa_width=3.0), nu, ne=ne, b=b, theta=th)
        pq = -(ne * ee * ee * nuc * nuc * math.sin(th) ** 2) / (ME * CL * nu ** 3)
        pv = 2 * ne * ee * ee * nuc * math.cos(th) / (ME * CL * nu * nu)
        return np.array([c.rQ / nu / pq, c.rV / nu / pv])

    for lo in (3.5, 4.0, 4.5):
        a, bb = fit(lo), fit(lo + 0.5)
        np.testing.assert_allclose(fit(lo + 0.25), 0.5 * (a + bb), rtol=1e-10)
        np.testing.assert_allclose(fit(lo + 0.5 / 3), (2 * a + bb) / 3, rtol=1e-10)
        assert np.all(a != bb)
    assert np.all(fit(5.0) != 0.0)


def test_kappa_blend_toward_thermal_and_caps():
    """Between kappa_interp_begin and kappa_interp_end j and alpha blend
    linearly toward the thermal Pandya fit (unit scales, :522-540); rho stays
    the kappa fit's; the polarised-fraction caps hold."""
    args = dict(nu=230e9, ne=3e5, thetae=12.0, b=25.0, theta=1.1)
    th = oracle.Coeffs()
    oracle.lib().kpo_thermal_coefficients(1, args["nu"], args["ne"], args["thetae"], args["b"], 0.0,
                                          math.sin(args["theta"]), math.cos(args["theta"]), 0.99, 1e-4, 1.0, 1.0,
                                          1.0, C.byref(th))
    pure = _coeffs(_nt(abi.EMISSION_KAPPA, kappa=4.5, kappa_interp_begin=1e20), **args)
    full = _coeffs(_nt(abi.EMISSION_KAPPA, kappa=4.5, kappa_interp_begin=3.0, kappa_interp_end=4.0), **args)
    half = _coeffs(_nt(abi.EMISSION_KAPPA, kappa=4.5, kappa_interp_begin=4.0, kappa_interp_end=5.0), **args)
    for n in ("jI", "jQ", "jV", "aI", "aQ", "aV"):
        assert getattr(full, n) == pytest.approx(getattr(th, n), rel=1e-14), n
        assert getattr(half, n) == pytest.approx(0.5 * (getattr(pure, n) + getattr(th, n)), rel=1e-12), n
    assert full.rQ == pure.rQ and full.rV == pure.rV
    capped = _coeffs(_nt(abi.EMISSION_KAPPA, kappa=4.5, max_pol_frac_emission=0.3, max_pol_frac_absorption=0.2),
                     **args)
    assert math.hypot(capped.jQ, capped.jV) <= 0.3 * capped.jI * (1 + 1e-13)
    assert math.hypot(capped.aQ, capped.aV) <= 0.2 * capped.aI * (1 + 1e-13)


# ------------------------------------------------------ emulated kernel lanes

@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_run_riaf_image.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel),
                                         C.POINTER(abi.ImageData), C.POINTER(C.c_int64)]
    L.kp_emul_run_riaf_image.restype = C.c_int32
    L.kp_emul_run_grmhd_image.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.GRMHDModel), dp, C.c_int32, dp,
                                          C.c_int32, C.POINTER(abi.ImageData), C.POINTER(C.c_int64)]
    L.kp_emul_run_grmhd_image.restype = C.c_int32
    return L


def riaf_nonthermal(dist, nx=10, ny=10, **kw):
    o = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", ny)
    p, m, _, _ = o.resolve_riaf()
    m.emission_type = dist
    for k, v in kw.items():
        setattr(m, k, v)
    return p, m


@pytest.mark.parametrize("dist,kw", [(abi.EMISSION_KAPPA, {}),
                                     (abi.EMISSION_KAPPA, {"nonthermal_kappa": 4.2,
                                                           "variable_kappa_interp_start": 4.0,
                                                           "variable_kappa_max": 5.0}),
                                     (abi.EMISSION_POWER_LAW, {})])
def test_emulated_riaf_nonthermal_image(emul, dist, kw):
    p, m = riaf_nonthermal(dist, **kw)
    npix = p.camera.nx * p.camera.ny
    arrays = oracle.alloc_image(npix)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    assert emul.kp_emul_run_riaf_image(C.byref(p), C.byref(m), C.byref(img), C.byref(rhs)) == 0
    r = parity.oracle_with_sensitivity(p, m)
    parity.assert_parity(arrays, r, f"emulated RIAF emission {dist}")
    assert np.max(r["image_i"]) > 0.0
    _, st = oracle.run_riaf_image(p, m)
    assert rhs.value == st["rhs_evals"]


@pytest.mark.parametrize("variable", [0, 1])
@pytest.mark.parametrize("dist", [abi.EMISSION_KAPPA, abi.EMISSION_POWER_LAW])
def test_emulated_grmhd_nonthermal_image(emul, dist, variable):
    """GRMHD lanes with the kappa (fixed, or from sigma and beta) and power-law
    distributions against the oracle, two frequencies on shared geodesics.
    The snapshot is magnetised everywhere (beta_min 1, a weak uniform
    toroidal field) so that no sample meets the fits' edge cases above
    (sigma >= 8e-5, so kappa <= 103)."""
    if dist == abi.EMISSION_POWER_LAW and variable:
        pytest.skip("variable kappa applies to the kappa distribution only")
    cells, m = gs.make_snapshot(40, 24, 24, seed=3, beta_min=1.0, background_field=1e-2)
    m.emission_type = dist
    m.variable_kappa = variable
    m.variable_kappa_interp_start = 5.5
    m.variable_kappa_max = 7.0
    oracle.grmhd_derive_scalars(m, cells)
    p = gs.pass_a_params(m, nx=8, ny=8)
    F = np.array([86e9, 345e9])
    ref, st = oracle.run_grmhd_image(p, m, cells, F)
    mm = abi.GRMHDModel.from_buffer_copy(m)
    mm.cells = C.c_void_p(cells.ctypes.data)
    arrays = oracle.alloc_image(64)
    for n in parity.STOKES:
        arrays[n] = np.zeros(2 * 64)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    assert emul.kp_emul_run_grmhd_image(C.byref(p), C.byref(mm), F.ctypes.data_as(dp), 2, F.ctypes.data_as(dp), 2,
                                        C.byref(img), C.byref(rhs)) == 0
    assert rhs.value == st["rhs_evals"]
    assert np.array_equal(arrays["reason"], ref["reason"])
    assert np.array_equal(arrays["steps"], ref["steps"])
    for f in range(2):
        sl = slice(f * 64, (f + 1) * 64)
        peak = np.abs(ref["image_i"][sl]).max()
        assert peak > 0.0
        for n in parity.STOKES:
            err = np.abs(arrays[n][sl] - ref[n][sl])
            assert np.all(err <= 1e-10 * np.abs(ref[n][sl]) + 1e-8 * peak), (f, n, err.max() / peak)
