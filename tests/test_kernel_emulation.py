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
    return L


def _run_emul(L, p, m):
    npix = p.camera.nx * p.camera.ny
    arrays = oracle.alloc_image(npix)
    img = oracle.image_struct(arrays)
    rhs = C.c_int64()
    assert L.kp_emul_run_riaf_image(C.byref(p), C.byref(m), C.byref(img), C.byref(rhs)) == 0
    return arrays, rhs.value


def _compare(a, b, rel=1e-10):
    np.testing.assert_array_equal(a["reason"], b["reason"])
    np.testing.assert_array_equal(a["pass_a_steps"], b["pass_a_steps"])
    np.testing.assert_array_equal(a["steps"], b["steps"])
    peak = float(np.max(np.abs(b["image_i"])))
    for n in ("image_i", "image_q", "image_u", "image_v"):
        np.testing.assert_array_less(np.abs(a[n] - b[n]), rel * np.abs(b[n]) + 1e-10 * peak + 1e-300, err_msg=n)
    for n in ("closure_x", "closure_k", "final_null", "frame_error", "overlap_r11", "overlap_r12",
              "overlap_r21", "overlap_r22", "det_r", "basis_rotation_angle", "basis_identity_error"):
        # diagnostics of rays that wind near the photon ring are chaotic in the
        # end position; they are checked loosely (the parity bar is on I/Q/U/V)
        np.testing.assert_allclose(a[n], b[n], rtol=1e-4, atol=1e-8, err_msg=n)


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
    r, _ = oracle.run_riaf_image(p, m)
    _compare(e, r)
    flux = image.pack_image(e, m.freq_cgs, sign, scale)[4]
    assert flux[0] == pytest.approx(2.3672187308355426, rel=2e-6)


def test_emulated_kernel_adaptive_c2(emul):
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=18, ny=18)
    e, rhs = _run_emul(emul, p, m)
    r, st = oracle.run_riaf_image(p, m)
    _compare(e, r)
    assert rhs == st["rhs_evals"]


def test_emulated_kernel_fixed_c1(emul):
    p, m, _, _ = _opts("riaf_m87_c1.par", nx=5, ny=5)
    e, _ = _run_emul(emul, p, m)
    r, _ = oracle.run_riaf_image(p, m)
    _compare(e, r)


def test_emulated_kernel_dexter_faraday_off_and_max_steps(emul):
    p, m, _, _ = _opts("riaf_m87_c2.par", nx=9, ny=7, max_steps=250)
    m.emission_type = abi.EMISSION_THERMAL_DEXTER
    p.emission_selection.faraday_rotation = 0
    e, _ = _run_emul(emul, p, m)
    r, _ = oracle.run_riaf_image(p, m)
    _compare(e, r)
    assert (r["reason"] == abi.REASON_MAX_STEPS).any()
