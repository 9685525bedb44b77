"""CPU build of the HIP kernel's lane state machines (tests/emul) against the
oracle: catches control-flow slips of the product kernels without a GPU.
On the CPU both sides use the C library's math, so agreement is far tighter
than the GPU bar."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, image
from tests import parity

ROOT = Path(__file__).resolve().parent.parent
PARAMS = ROOT / "kpolaris_amd" / "params"


@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_run_riaf_image.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel),
                                         C.POINTER(abi.ImageData), C.POINTER(C.c_int64)]
    L.kp_emul_run_riaf_image.restype = C.c_int32
    L.kp_emul_run_riaf_image_mixed.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel), C.c_double,
                                               C.POINTER(abi.ImageData), C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64)]
    L.kp_emul_run_riaf_image_mixed.restype = C.c_int32
    dp = C.POINTER(C.c_double)
    L.kp_emul_run_riaf_multifrequency_image.argtypes = [
        C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel), C.c_int32, C.c_double, dp, C.c_int32, dp,
        C.c_int32, C.POINTER(abi.ImageData), C.POINTER(C.c_int64)]
    L.kp_emul_run_riaf_multifrequency_image.restype = C.c_int32
    L.kp_emul_run_riaf_image_gradient.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel), dp,
                                                  C.POINTER(abi.ImageData), dp]
    L.kp_emul_run_riaf_image_gradient.restype = C.c_int32
    L.kp_emul_run_riaf_image_legs.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel), C.c_int32,
                                              C.c_int32, C.POINTER(abi.ImageData), C.POINTER(C.c_int64)]
    L.kp_emul_run_riaf_image_legs.restype = C.c_int32
    return L


def _run_emul(L, p, m):
    npix = p.camera.nx * p.camera.ny
    arrays = oracle.alloc_image(npix)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    assert L.kp_emul_run_riaf_image(C.byref(p), C.byref(m), C.byref(img), C.byref(rhs)) == 0
    return arrays, rhs.value


def _run_emul_mixed(L, p, m, fp64_radius=abi.FP64_RADIUS_DEFAULT):
    npix = p.camera.nx * p.camera.ny
    arrays = oracle.alloc_image(npix)
    img = oracle.image_struct(arrays)
    rhs, guard = C.c_int64(), C.c_int64()
    assert L.kp_emul_run_riaf_image_mixed(C.byref(p), C.byref(m), fp64_radius, C.byref(img), C.byref(rhs),
                                          C.byref(guard)) == 0
    return arrays, rhs.value, guard.value


def _compare(a, b):
    """b: oracle_with_sensitivity result; the parity bar of tests/parity.py."""
    parity.assert_parity(a, b, "emulated kernel")


def _opts(par, **kw):
    o = image.Options(parameter_file=str(PARAMS / par))
    for k, v in kw.items():
        o.set(k, v)
    return o.resolve_riaf()


def test_emulated_kernel_golden_8x8(emul):
    import json
    g = json.loads((Path(__file__).parent / "golden" / "riaf_bl_8x8_golden.json").read_text())
    o = image.Options()
    for kv in g["parameters"]:
        k, _, v = kv.partition("=")
        o.set(k, v)
    p, m, scale, sign = o.resolve_riaf()
    e, _ = _run_emul(emul, p, m)
    r = parity.oracle_with_sensitivity(p, m)
    _compare(e, r)
    flux = image.pack_image(e, m.freq_cgs, sign, scale)[4]
    assert flux[0] == pytest.approx(2.3672187308355426, rel=2e-6)


def test_emulated_kernel_adaptive_c2(emul):
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=18, ny=18)
    e, rhs = _run_emul(emul, p, m)
    r = parity.oracle_with_sensitivity(p, m)
    _compare(e, r)
    _, st = oracle.run_riaf_image(p, m)
    assert rhs == st["rhs_evals"]


def test_emulated_kernel_fixed_c1(emul):
    p, m, _, _ = _opts("riaf_m87_c1.par", nx=5, ny=5)
    e, _ = _run_emul(emul, p, m)
    r = parity.oracle_with_sensitivity(p, m)
    _compare(e, r)


def test_emulated_kernel_edge_7x3(emul):
    """A ray that winds once around the hole (ill-conditioned) and ragged shape."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=7, ny=3)
    e, _ = _run_emul(emul, p, m)
    r = parity.oracle_with_sensitivity(p, m)
    _compare(e, r)


def test_emulated_kernel_dexter_faraday_off_and_max_steps(emul):
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=9, ny=7, max_steps=250)
    m.emission_type = abi.EMISSION_THERMAL_DEXTER
    p.emission_selection.faraday_rotation = 0
    e, _ = _run_emul(emul, p, m)
    r = parity.oracle_with_sensitivity(p, m)
    _compare(e, r)
    assert (r["reason"] == abi.REASON_MAX_STEPS).any()


def test_emulated_mixed_precision_c2(emul):
    """The mixed-precision lanes against the FP64 oracle: image-level bars."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=20, ny=20)
    e, _, guard = _run_emul_mixed(emul, p, m)
    r, _ = oracle.run_riaf_image(p, m)
    err = parity.assert_mixed(e, r, "emulated mixed c2")
    print(err)
    assert guard > 0  # rays reach r < 2.5 and take FP64 attempts there


def test_emulated_mixed_precision_dexter_fixed_step(emul):
    """Dexter fit, Faraday rotation off, fixed-step RK4: the other branches."""
    p, m, _, _ = _opts("riaf_m87_c1.par", nx=6, ny=6)
    m.emission_type = abi.EMISSION_THERMAL_DEXTER
    p.emission_selection.faraday_rotation = 0
    e, _, _ = _run_emul_mixed(emul, p, m)
    r, _ = oracle.run_riaf_image(p, m)
    parity.assert_mixed(e, r, "emulated mixed dexter c1")


def test_emulated_mixed_precision_unguarded(emul):
    """With fp64_radius 0 (every FP64 guard off) every attempt is FP32-derived;
    still in bounds on this coarse image."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=12, ny=12)
    e, _, guard = _run_emul_mixed(emul, p, m, fp64_radius=0.0)
    assert guard == 0
    r, _ = oracle.run_riaf_image(p, m)
    parity.assert_mixed(e, r, "emulated mixed unguarded")


def test_emulated_mixed_turning_point_guard(emul):
    """With a guard radius no ray reaches (1e-9 M) the FP64 attempts are the
    turning-point retakes alone (an FP32 attempt across which dtheta/dlambda
    changes sign is taken again in FP64): present, a few per ray (under 1 % of
    the attempts), fewer than with the default radius, and the image stays
    within the mixed bars."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=20, ny=20)
    r, _ = oracle.run_riaf_image(p, m)
    _, _, g0 = _run_emul_mixed(emul, p, m, fp64_radius=0.0)
    turn, rhs, g1 = _run_emul_mixed(emul, p, m, fp64_radius=1e-9)
    _, _, g2 = _run_emul_mixed(emul, p, m)
    assert g0 == 0
    assert 0 < g1 < 0.01 * rhs / 11
    assert g1 < g2
    parity.assert_mixed(turn, r, "emulated mixed, turning-point guard only")


def _run_emul_mf(L, p, m, freqs, ctrl=None, mixed=False):
    npix = p.camera.nx * p.camera.ny
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    c = np.ascontiguousarray(freqs if ctrl is None else ctrl, dtype=np.float64)
    arrays = oracle.alloc_image(npix)
    for n in parity.STOKES:
        arrays[n] = np.zeros(len(f) * npix)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    dp = C.POINTER(C.c_double)
    assert L.kp_emul_run_riaf_multifrequency_image(
        C.byref(p), C.byref(m), int(mixed), abi.FP64_RADIUS_DEFAULT, f.ctypes.data_as(dp), len(f),
        c.ctypes.data_as(dp), len(c), C.byref(img), C.byref(rhs)) == 0
    return arrays


MF_FREQS = [86e9, 230e9, 345e9]


def test_emulated_multifrequency_depth_limited(emul):
    """Three frequencies sharing geodesics, with absorption/Faraday depth limits
    tight enough that the step control depends on the frequency."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=10, ny=10, max_absorption_depth=1e-3, max_faraday_depth=1e-2)
    e = _run_emul_mf(emul, p, m, MF_FREQS)
    r, _ = oracle.run_riaf_multifrequency_image(p, m, MF_FREQS)
    single = parity.oracle_with_sensitivity(p, m)
    parity.assert_parity_multifrequency(e, r, 3, single, "emulated mf")


def test_emulated_multifrequency_separate_control(emul):
    """Output frequencies with a different step-control list (the control variant)."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=8, ny=8, max_absorption_depth=1e-3, max_faraday_depth=1e-2)
    e = _run_emul_mf(emul, p, m, MF_FREQS, ctrl=[86e9])
    r, _ = oracle.run_riaf_multifrequency_image(p, m, MF_FREQS, [86e9])
    single = parity.oracle_with_sensitivity(p, m)
    parity.assert_parity_multifrequency(e, r, 3, single, "emulated mf ctrl")


def test_emulated_multifrequency_mixed(emul):
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=12, ny=12)
    e = _run_emul_mf(emul, p, m, MF_FREQS, mixed=True)
    r, _ = oracle.run_riaf_multifrequency_image(p, m, MF_FREQS)
    for f, (gf, rf) in enumerate(zip(parity.split_frequencies(e, 3), parity.split_frequencies(r, 3))):
        parity.assert_mixed(gf, rf, f"emulated mixed mf f{f}")


def oracle_central_difference(p, m, parameter: int, h: float = 1e-4):
    """(S(q=+h) - S(q=-h)) / 2h of the oracle for one PlasmaPerturbation family,
    and the mask of pixels whose step sequence is the same in both runs."""
    pp, pm = oracle.Perturbation(), oracle.Perturbation()
    pp.parameter = pm.parameter = parameter
    pp.log_scale, pm.log_scale = h, -h
    rp, _ = oracle.run_riaf_image(p, m, perturbation=pp)
    rm, _ = oracle.run_riaf_image(p, m, perturbation=pm)
    fd = {n: (rp[n] - rm[n]) / (2 * h) for n in parity.STOKES}
    same = (rp["steps"] == rm["steps"]) & (rp["pass_a_steps"] == rm["pass_a_steps"])
    return fd, same


def test_emulated_tangent_matches_oracle_central_differences(emul):
    """The tangent lanes' dS/dq against central differences of the oracle
    (h = 1e-4 in q: truncation ~h^2 = 1e-8 relative)."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=10, ny=10)
    n = 100
    arrays = oracle.alloc_image(n)
    img = oracle.image_struct(arrays)
    g = np.zeros(3 * 4 * n)
    q = np.zeros(3)
    dp = C.POINTER(C.c_double)
    assert emul.kp_emul_run_riaf_image_gradient(C.byref(p), C.byref(m), q.ctypes.data_as(dp), C.byref(img),
                                                g.ctypes.data_as(dp)) == 0
    ref = parity.oracle_with_sensitivity(p, m)
    parity.assert_parity(arrays, ref, "emulated tangent value")
    g = g.reshape(3, 4, n)
    for k in range(3):
        fd, same = oracle_central_difference(p, m, k + 1)
        assert same.all()
        for s, name in enumerate(parity.STOKES):
            scale = np.abs(fd[name]).max()
            np.testing.assert_allclose(g[k, s], fd[name], rtol=0, atol=1e-6 * scale, err_msg=f"q{k} {name}")


def test_emulated_mixed_coefficients_stay_finite_at_underflow(emul):
    """The FP32 rescaled source function S' = bconst / expm1(x) of the mixed
    lanes is subnormal for x between ln(bconst / FLT_MIN) and 88.7 (a very
    cold sample below the default temperature floor): the absorption and the
    polarisation cap must stay finite there (S' is treated as underflowed, as
    the FP64 path treats B_nu), not multiply zero emissivity by an infinite
    reciprocal."""
    import math
    L = emul
    L.kp_emul_thermal_coeffs_scaled.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel)] + \
        [C.c_float] * 7 + [C.POINTER(C.c_float)]
    L.kp_emul_thermal_coeffs_scaled.restype = C.c_int32
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=4, ny=4)
    nu = m.freq_cgs
    h, me, c = 6.62607015e-27, 9.1093837015e-28, 2.99792458e10
    bconst = 2 * h * nu ** 3 / c ** 2
    for x in (62.0, 70.0, 80.0, 88.0, 95.0):
        assert x > math.log(bconst / 1.17549435e-38) or x > 88.0
        thetae = h * nu / (me * c * c * x)
        out = (C.c_float * 11)()
        assert L.kp_emul_thermal_coeffs_scaled(C.byref(p), C.byref(m), nu, 1.0, 1e6, thetae, 10.0, 0.8, 0.6,
                                               out) == 0
        vals = list(out)
        assert all(math.isfinite(v) for v in vals), (x, vals)


@pytest.mark.parametrize("mixed", [0, 1])
def test_emulated_pass_b_legs_equal_the_one_leg_lanes(emul, mixed):
    """Pass B in two legs (the radiating lanes park where the ray leaves the
    emitting region outward, the vacuum lanes take it to the camera with no
    radiation code): outside the region every coefficient is zero, the substep
    counts are 1 and the transfer step is the identity, so the image equals
    the one-leg lanes' bit for bit, FP64 and mixed."""
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=24, ny=24)
    npix = p.camera.nx * p.camera.ny
    res = []
    for legs in (0, 1):
        arrays = oracle.alloc_image(npix)
        img = oracle.image_struct(arrays)
        parked = C.c_int64()
        assert emul.kp_emul_run_riaf_image_legs(C.byref(p), C.byref(m), mixed, legs, C.byref(img),
                                                C.byref(parked)) == 0
        res.append((arrays, parked.value))
    (one, _), (two, parked) = res
    assert parked > 0.5 * npix
    for k in one:
        assert np.array_equal(one[k], two[k]), k
