"""Parameter responses on the CPU.  The oracle restates the reference's
response accumulation (src/diagnostics/response.hpp: the homogeneous step
matrix, the four mechanism forcings by central differences of coefficient
blocks, propagation by the baseline operator, the four full reruns), pinned
here by the reference's own closed forms (tests/test_physical_response.cpp:
the emitting slab, ordered rotation and conversion screens, a mixed
transfer against full-rerun differences).  The kernel's exact forward-mode
tangents (tests/emul lanes) are then held to the oracle's response maps of
the same geodesics for the density, temperature and magnetic families."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, image

P = C.POINTER


class Coeffs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("jI", "jQ", "jU", "jV", "aI", "aQ", "aU", "aV", "rQ", "rU", "rV")]


class Stokes(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("I", "Q", "U", "V")]


@pytest.fixture(scope="module")
def olib():
    L = oracle.lib()
    L.kpo_response_transfer_step.argtypes = [C.c_int32, C.c_int32, C.c_double, P(C.c_double), C.c_int32,
                                             P(Stokes), P(Coeffs), P(Coeffs), C.c_double]
    L.kpo_response_scale_coefficients.argtypes = [P(Coeffs), C.c_double, P(Coeffs)]
    L.kpo_semi_analytic_step.argtypes = [P(Stokes), P(Coeffs), C.c_double]
    L.kpo_run_riaf_response_image.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), C.c_int32, C.c_double,
                                              C.c_int32, P(abi.ImageData), P(C.c_double), P(oracle.Stats)]
    L.kpo_run_riaf_response_image.restype = C.c_int32
    return L


class Ray:
    """One ray's channels under the "coefficients" family (every
    coefficient scaled by e^q), as the reference's test fixture builds them."""

    def __init__(self, L, bins=2, step=1e-4):
        self.L, self.bins, self.step = L, bins, step
        self.data = np.zeros(4 * (5 * bins + 4))
        self.s = Stokes()

    def advance(self, c, dl, b):
        var = (Coeffs * 4)()
        for v in range(4):
            q = self.step * (-1 if v % 2 else 1) * (1 if v < 2 else 0.5)
            self.L.kpo_response_scale_coefficients(C.byref(c), q, C.byref(var[v]))
        self.L.kpo_response_transfer_step(self.bins, 1, self.step, self.data.ctypes.data_as(P(C.c_double)), b,
                                          C.byref(self.s), C.byref(c), var, dl)
        self.L.kpo_semi_analytic_step(C.byref(self.s), C.byref(c), dl)

    def channel(self, ch):
        return self.data[4 * ch:4 * ch + 4]

    def source(self, b):
        return self.channel(b)

    def response(self, m, b):
        return self.channel(self.bins + m * self.bins + b)

    def total(self):
        return sum(self.response(m, b) for m in range(4) for b in range(self.bins))

    def difference(self, first):
        a, b = self.channel(5 * self.bins + first), self.channel(5 * self.bins + first + 1)
        return (a - b) / (2 * self.step if first == 0 else self.step)


def _close(value, expected, tol=2e-6):
    assert math.isfinite(value) and abs(value - expected) <= tol * max(1.0, abs(expected)), (value, expected)


def test_slab(olib):
    """test_physical_response.cpp:44-56: a uniform emitting slab of depth 1.4
    split over two bins."""
    ray = Ray(olib)
    c = Coeffs(jI=2.0, aI=0.7)
    for k in range(100):
        ray.advance(c, 0.02, 0 if k < 40 else 1)
    tau, source = 1.4, 2 / 0.7
    _close(ray.s.I, source * (1 - math.exp(-tau)), 1e-12)
    _close(ray.source(0)[0] + ray.source(1)[0], ray.s.I, 1e-12)
    _close(ray.total()[0], source * tau * math.exp(-tau))
    _close(ray.response(0, 0)[0] + ray.response(0, 1)[0], ray.s.I)
    _close(ray.response(1, 0)[0] + ray.response(1, 1)[0], source * ((tau + 1) * math.exp(-tau) - 1))
    _close(ray.total()[0], ray.difference(0)[0])
    _close(ray.total()[0], ray.difference(2)[0])


def test_ordered_screens(olib):
    """test_physical_response.cpp:58-73: a rotation screen in front of a
    polarised source rotates it (its response at the screen's bin), a screen
    behind it has no influence, a conversion screen converts U into V."""
    src = Coeffs(jI=2.0, jQ=1.0)
    screen = Coeffs(rV=0.8)
    fg = Ray(olib)
    fg.advance(src, 1.0, 0)
    fg.advance(screen, 1.0, 1)
    _close(fg.response(2, 1)[1], -0.8 * math.sin(0.8))
    _close(fg.response(2, 1)[2], 0.8 * math.cos(0.8))
    _close(fg.source(1)[0], 0.0, 1e-15)
    bg = Ray(olib)
    bg.advance(screen, 1.0, 1)
    bg.advance(src, 1.0, 0)
    _close(bg.response(2, 1)[2], 0.0, 1e-15)
    src = Coeffs(jI=2.0, jU=1.0)
    conv = Coeffs(rQ=0.6)
    cv = Ray(olib)
    cv.advance(src, 1.0, 0)
    cv.advance(conv, 1.0, 1)
    _close(cv.response(3, 1)[3], 0.6 * math.cos(0.6))
    _close(cv.source(0)[3], math.sin(0.6), 1e-12)


def test_mixed_transfer(olib):
    """test_physical_response.cpp:75-90: every coefficient nonzero and
    varying over three bins; the source tags sum to the Stokes vector and the
    summed mechanism derivatives equal the half-step rerun difference."""
    ray = Ray(olib, bins=3)
    for k in range(75):
        c = Coeffs(jI=1 + 0.2 * math.sin(k * 0.2), jQ=0.3, jU=-0.1, jV=0.05, aI=0.4, aQ=0.1, aU=0.04, aV=-0.02,
                   rQ=0.6 * math.cos(k * 0.1), rU=-0.3, rV=0.8)
        ray.advance(c, 0.03, k // 25)
    total_src = ray.source(0) + ray.source(1) + ray.source(2)
    _close(total_src[0], ray.s.I, 1e-12)
    _close(total_src[3], ray.s.V, 1e-12)
    a, b = ray.total(), ray.difference(2)
    for k in range(4):
        _close(a[k], b[k])


def _riaf(nx):
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", nx)
    p, m, _, _ = o.resolve_riaf()
    return p, m


def _response_image(L, p, m, parameter, step=1e-3):
    npix = p.camera.nx * p.camera.ny
    arrays = oracle.alloc_image(npix)
    resp = np.zeros(36 * npix)
    st = oracle.Stats()
    assert L.kpo_run_riaf_response_image(C.byref(p), C.byref(m), parameter, step, 0,
                                         C.byref(oracle.image_struct(arrays)), resp.ctypes.data_as(P(C.c_double)),
                                         C.byref(st)) == 0
    ch = resp.reshape(9, 4, npix)
    return arrays, ch


@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_run_riaf_image_gradient.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), P(C.c_double),
                                                  P(abi.ImageData), P(C.c_double)]
    L.kp_emul_run_riaf_image_gradient.restype = C.c_int32
    return L


def test_response_image_identities(olib):
    """On the RIAF image: the source channel is the image (before the
    observer transform both are the same Stokes vector), and the summed
    mechanism derivatives agree with the full-rerun differences at h and at
    h/2 (the O(h^2) convergence of the reference's response check)."""
    p, m = _riaf(12)
    img, ch = _response_image(olib, p, m, 1)
    for k, n in enumerate(("image_i", "image_q", "image_u", "image_v")):
        np.testing.assert_allclose(ch[0, k], img[n], rtol=1e-12, atol=1e-12 * np.abs(img["image_i"]).max())
    total = ch[1:5].sum(axis=0)
    half = (ch[7] - ch[8]) / 1e-3
    full = (ch[5] - ch[6]) / 2e-3
    scale = np.abs(total[0]).max()
    assert np.abs(total - half).max() <= 1e-5 * scale
    assert np.abs(total - full).max() <= 1e-5 * scale


@pytest.mark.parametrize("parameter", [1, 2, 3])
def test_kernel_tangents_match_the_oracle_responses(olib, emul, parameter):
    """The kernel's exact tangent d(I,Q,U,V)/dq along the fixed geodesics (the
    emulated lanes of the gradient kernel) against the reference's response
    maps restated in the oracle (step 1e-3: central differences of each
    coefficient block, O(h^2) ~ 1e-6), family by family."""
    p, m = _riaf(12)
    npix = 144
    arrays = oracle.alloc_image(npix)
    grad = np.zeros(3 * 4 * npix)
    q = (C.c_double * 3)(0.0, 0.0, 0.0)
    assert emul.kp_emul_run_riaf_image_gradient(C.byref(p), C.byref(m), q, C.byref(oracle.image_struct(arrays)),
                                                grad.ctypes.data_as(P(C.c_double))) == 0
    g = grad.reshape(3, 4, npix)[parameter - 1]
    _, ch = _response_image(olib, p, m, parameter)
    total = ch[1:5].sum(axis=0)
    scale = np.abs(total).max(axis=1, keepdims=True)
    err = np.abs(g - total)
    assert np.all(err <= 1e-5 * np.abs(total) + 1e-6 * scale), (parameter, (err / scale).max())
    print(parameter, "max |tangent - response| / peak", float((err / scale).max()))
