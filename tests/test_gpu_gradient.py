"""Plasma-parameter derivatives of the image (kp_run_riaf_image_gradient) and
the torch.autograd.Function over them, against central finite differences of
the oracle (h = 1e-4 in the log-scale, the reference's PlasmaPerturbation
families) and of the device's own FP64 image."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

PARAMS = Path(__file__).resolve().parent.parent / "kpolaris_amd" / "params"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _opts(**kw):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    for k, v in kw.items():
        o.set(k, str(v))
    return o.resolve_riaf()


def _fd_oracle(p, m, parameter, h=1e-4):
    import oracle
    pp, pm = oracle.Perturbation(), oracle.Perturbation()
    pp.parameter = pm.parameter = parameter
    pp.log_scale, pm.log_scale = h, -h
    rp, _ = oracle.run_riaf_image(p, m, perturbation=pp)
    rm, _ = oracle.run_riaf_image(p, m, perturbation=pm)
    same = (rp["steps"] == rm["steps"]) & (rp["pass_a_steps"] == rm["pass_a_steps"])
    return {n: (rp[n] - rm[n]) / (2 * h) for n in parity.STOKES}, same


def test_gpu_gradient_against_oracle_central_differences():
    import torch
    from kpolaris_amd import image
    print("case: gradient 16x16 at q=0 vs oracle central differences", flush=True)
    p, m, _, _ = _opts(nx=16, ny=16)
    out, grad, st = image.run_riaf_image_gradient(p, m, (0.0, 0.0, 0.0))
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity(got, ref, "gradient value"), flush=True)
    g = grad.cpu().numpy()
    smooth = parity.smooth_for_derivatives(ref, ref["_cond"])
    for k in range(3):
        fd, same = _fd_oracle(p, m, k + 1)
        ok = same & smooth
        print(f"gradient check on {int(ok.sum())}/{ok.size} pixels", flush=True)
        assert ok.sum() >= parity.MIN_CHECKED * ok.size
        for s, name in enumerate(parity.STOKES):
            scale = np.abs(fd[name]).max()
            err = np.abs(g[k, s][ok] - fd[name][ok]).max() / scale
            print(f"q{k} {name} max rel err {err:.2e}", flush=True)
            assert err < 1e-6, (k, name, err)


def test_gpu_gradient_against_oracle_responses():
    """The device tangents against the reference's parameter-response maps
    restated in the oracle (src/diagnostics/response.hpp: mechanism
    forcings by central differences of coefficient blocks at step 1e-3,
    propagated along the same geodesic and step sequence), for the density,
    temperature and magnetic families, on every well-conditioned pixel (an
    ill-conditioned ray's device geodesic is not the oracle's)."""
    import ctypes as C
    import oracle
    from kpolaris_amd import abi, image
    print("case: gradient 16x16 vs oracle response maps", flush=True)
    p, m, _, _ = _opts(nx=16, ny=16)
    _, grad, _ = image.run_riaf_image_gradient(p, m, (0.0, 0.0, 0.0))
    g = grad.cpu().numpy()
    ref = parity.oracle_with_sensitivity(p, m)
    well = ~parity.ill_conditioned(ref, ref["_cond"], 1)
    L = oracle.lib()
    L.kpo_run_riaf_response_image.argtypes = [C.POINTER(abi.PassAParams), C.POINTER(abi.RIAFModel), C.c_int32,
                                              C.c_double, C.c_int32, C.POINTER(abi.ImageData),
                                              C.POINTER(C.c_double), C.POINTER(oracle.Stats)]
    L.kpo_run_riaf_response_image.restype = C.c_int32
    npix = 256
    for k in range(3):
        arrays = oracle.alloc_image(npix)
        resp = np.zeros(36 * npix)
        assert L.kpo_run_riaf_response_image(C.byref(p), C.byref(m), k + 1, 1e-3, 0,
                                             C.byref(oracle.image_struct(arrays)),
                                             resp.ctypes.data_as(C.POINTER(C.c_double)), C.byref(oracle.Stats())) == 0
        total = resp.reshape(9, 4, npix)[1:5].sum(axis=0)
        scale = np.abs(total).max(axis=1, keepdims=True)
        err = np.abs(g[k] - total)
        ok = err <= 1e-5 * np.abs(total) + 1e-6 * scale
        print(f"q{k} max |tangent - response| / peak on {int(well.sum())} well-conditioned pixels "
              f"{float((err[:, well] / scale).max()):.2e}", flush=True)
        assert ok[:, well].all(), (k, np.flatnonzero(~ok.all(axis=0) & well)[:8])


def test_gpu_gradient_at_a_perturbed_point():
    """Value at q = (0.3, 0, 0) equals the oracle's density-perturbed image;
    derivative there equals central differences of the device image."""
    import oracle
    import torch
    from kpolaris_amd import image
    print("case: gradient 12x12 at q_density = 0.3", flush=True)
    p, m, _, _ = _opts(nx=12, ny=12)
    q = 0.3
    out, grad, _ = image.run_riaf_image_gradient(p, m, (q, 0.0, 0.0))
    pert = oracle.Perturbation()
    pert.parameter, pert.log_scale = 1, q
    ref, _ = oracle.run_riaf_image(p, m, perturbation=pert)
    got = {k: v.cpu().numpy() for k, v in out.items()}
    peak = float(ref["image_i"].max())
    for n in parity.STOKES:
        np.testing.assert_allclose(got[n], ref[n], rtol=1e-9, atol=1e-8 * peak, err_msg=n)
    h = 1e-4
    op, _, _ = image.run_riaf_image_gradient(p, m, (q + h, 0.0, 0.0))
    om, _, _ = image.run_riaf_image_gradient(p, m, (q - h, 0.0, 0.0))
    torch.cuda.synchronize()
    g = grad.cpu().numpy()
    for s, n in enumerate(parity.STOKES):
        fd = (op[n].cpu().numpy() - om[n].cpu().numpy()) / (2 * h)
        scale = np.abs(fd).max()
        assert np.abs(g[0, s] - fd).max() <= 1e-6 * scale, n


def test_gpu_autograd_function_backward():
    """loss = sum(w * S): backward through the autograd Function equals the
    finite difference of the loss, for all three parameters at once."""
    import torch
    from kpolaris_amd.autograd import riaf_stokes
    print("case: autograd backward 12x12", flush=True)
    p, m, _, _ = _opts(nx=12, ny=12)
    q = torch.tensor([0.1, -0.2, 0.05], dtype=torch.float64, requires_grad=True)
    gen = torch.Generator(device="cuda").manual_seed(1)
    w = torch.randn((4, 144), dtype=torch.float64, device="cuda", generator=gen)
    S = riaf_stokes(q, p, m)
    scale = float(S[0].abs().max())
    loss = (w * S / scale).sum()
    loss.backward()
    h = 1e-4
    fd = []
    for k in range(3):
        dq = torch.zeros(3, dtype=torch.float64)
        dq[k] = h
        with torch.no_grad():
            lp = (w * riaf_stokes(q.detach() + dq, p, m) / scale).sum()
            lm = (w * riaf_stokes(q.detach() - dq, p, m) / scale).sum()
        fd.append(float((lp - lm) / (2 * h)))
    np.testing.assert_allclose(q.grad.numpy(), np.array(fd), rtol=1e-5, atol=1e-9)


def test_gpu_gradient_rejects_mixed_precision():
    from kpolaris_amd import image
    p, m, _, _ = _opts(nx=4, ny=4)
    with pytest.raises(ValueError):
        image.run_riaf_image_gradient(p, m, (0.0, 0.0))


def test_gpu_gradient_with_the_emission_wedge():
    """The tangent kernel with the equatorial emission wedge (h/r = 0.1)
    against the oracle's central differences of wedge images."""
    import torch
    from kpolaris_amd import image
    print("case: gradient 16x16 with the wedge h/r 0.1 vs oracle central differences", flush=True)
    p, m, _, _ = _opts(nx=16, ny=16)
    p.emission_selection.equatorial_h_over_r = 0.1
    out, grad, st = image.run_riaf_image_gradient(p, m, (0.0, 0.0, 0.0))
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity(got, ref, "wedge gradient value", max_ill=parity.WEDGE_MAX_ILL,
                               step_rel=parity.WEDGE_STEP_REL), flush=True)
    g = grad.cpu().numpy()
    smooth = parity.smooth_for_derivatives(ref, ref["_cond"])
    for k in range(3):
        fd, same = _fd_oracle(p, m, k + 1)
        ok = same & smooth
        print(f"wedge gradient check on {int(ok.sum())}/{ok.size} pixels", flush=True)
        assert ok.sum() >= parity.MIN_CHECKED * ok.size
        for s, name in enumerate(parity.STOKES):
            scale = np.abs(fd[name]).max()
            err = np.abs(g[k, s][ok] - fd[name][ok]).max() / scale
            print(f"wedge q{k} {name} max rel err {err:.2e}", flush=True)
            assert err < 1e-6, (k, name, err)
