"""Static GRMHD gradients on the GPU (kp_run_grmhd_image_gradient) and the
torch.autograd.Function over them: the value is the static GRMHD image, the
derivatives with respect to (q_density, R_high, q_magnetic) along the fixed
geodesics match Richardson-combined central differences of the oracle's
static image (h = 1e-4 and 5e-5), backward is the Jacobian contraction."""
from __future__ import annotations

import numpy as np
import pytest

from tests import parity

pytestmark = pytest.mark.gpu
STOKES = parity.STOKES


@pytest.fixture(scope="module")
def snap():
    from kpolaris_amd import grmhd_synth as gs, image
    cells, m = gs.make_snapshot(40, 24, 24, seed=3)
    grid = image.GRMHDGrid(m, cells)
    host = grid.host_cells()
    p = gs.pass_a_params(m, nx=12, ny=12)
    return grid, host, m, p


def _fd(p, model, cells, kind, h):
    import oracle
    from kpolaris_amd import abi
    if kind == 1:
        res = []
        for sgn in (1.0, -1.0):
            m2 = abi.GRMHDModel.from_buffer_copy(model)
            m2.trat_large = model.trat_large * (1.0 + sgn * h)
            c2 = cells.copy()
            oracle.grmhd_derive_scalars(m2, c2)
            res.append(oracle.run_grmhd_image(p, m2, c2)[0])
        step = 2.0 * h * model.trat_large
    else:
        param = 1 if kind == 0 else 3
        res = [oracle.run_grmhd_image(p, model, cells, perturbation=(param, s * h))[0] for s in (1.0, -1.0)]
        step = 2.0 * h
    same = (res[0]["steps"] == res[1]["steps"]) & (res[0]["pass_a_steps"] == res[1]["pass_a_steps"])
    return {n: (res[0][n] - res[1][n]) / step for n in STOKES}, same


def test_gpu_grmhd_gradient_value_and_oracle_differences(snap):
    import oracle
    import torch
    from kpolaris_amd import image
    grid, host, m, p = snap
    print("case: static GRMHD gradient 12x12", flush=True)
    out, grad, st = image.run_grmhd_image_gradient(p, grid)
    static, _ = image.run_grmhd_image(p, grid)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref_static = {k: v.cpu().numpy() for k, v in static.items()}
    assert np.array_equal(got["reason"], ref_static["reason"])
    assert np.array_equal(got["steps"], ref_static["steps"])
    peak = float(np.max(ref_static["image_i"]))
    for k in STOKES:
        assert np.max(np.abs(got[k] - ref_static[k])) <= 1e-12 * peak, k
    g = grad.cpu().numpy()
    ref, _ = oracle.run_grmhd_image(p, m, host)
    cond = parity.sensitivity(lambda q: oracle.run_grmhd_image(q, m, host)[0], p, ref)
    print(parity.assert_parity_grid(got, ref, 1, "gpu static GRMHD tangent value", cond), flush=True)
    # the derivatives are compared on the pixels that the oracle's own
    # sensitivity leaves well-conditioned (every component) and whose step
    # counts equal the oracle's
    same_steps = got["steps"] == ref["steps"]
    for kind in range(3):
        fd, same = _fd(p, m, host, kind, 1e-4)
        fd2, same2 = _fd(p, m, host, kind, 5e-5)
        for s, name in enumerate(STOKES):
            ok = same & same2 & same_steps & parity.smooth_for_derivatives(ref, cond, 0)
            print(f"gradient check of {name} on {int(ok.sum())}/{ok.size} pixels", flush=True)
            assert ok.sum() >= parity.MIN_CHECKED * ok.size
            scale = np.abs(fd[name][ok]).max()
            rich = (4.0 * fd2[name] - fd[name]) / 3.0
            err = np.abs(g[kind, s][ok] - fd[name][ok]).max() / scale
            err_r = np.abs(g[kind, s][ok] - rich[ok]).max() / scale
            print(f"d/d{('q_density', 'R_high', 'q_magnetic')[kind]} {name}: {err:.2e} (Richardson {err_r:.2e})",
                  flush=True)
            # plain central differences carry an O(h^2) truncation (up to ~1e-4 of the
            # largest derivative); the Richardson value removes it
            assert err < 1e-3 and err_r < 1e-7, (kind, name, err, err_r)
            # most pixels agree to rounding; a rounding-level difference can flip one
            # accept/reject decision along a ray without changing its step count, and
            # the derivative of the discretised transfer then moves by O(rtol)
            per_pixel = np.abs(g[kind, s][ok] - rich[ok]) / scale
            assert np.percentile(per_pixel, 95) < 1e-8, (kind, name, np.percentile(per_pixel, 95))



def test_gpu_grmhd_autograd_contraction(snap):
    import torch
    from kpolaris_amd import autograd, image
    grid, host, m, p = snap
    print("case: static GRMHD autograd", flush=True)
    r0 = float(grid.model.trat_large)
    npix = p.camera.nx * p.camera.ny
    theta0 = [0.1, r0, -0.05]
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    w = torch.randn(4, npix, dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(5))
    S = autograd.grmhd_stokes(theta, p, grid)
    (S * w).sum().backward()
    _, jac, _ = image.run_grmhd_image_gradient(p, grid, (theta0[0], theta0[2]), with_stats=False)
    torch.cuda.synchronize()
    contraction = (jac * w.unsqueeze(0)).sum(dim=(1, 2)).cpu().numpy()
    assert np.allclose(theta.grad.detach().numpy(), contraction, rtol=1e-12, atol=0.0)
