"""Slow-light gradients on the GPU (kp_run_grmhd_slow_light_image_gradient)
and the torch.autograd.Function over them: derivatives with respect to
(q_density, R_high, q_magnetic) along the fixed geodesics against central
differences of the oracle (h = 1e-4, Richardson-combined with h/2 to remove
the O(h^2) truncation), the value against the slow-light image."""
from __future__ import annotations

import numpy as np
import pytest

from tests import parity

pytestmark = pytest.mark.gpu

T_OBS = 275.0
STOKES = parity.STOKES


@pytest.fixture(scope="module")
def seq():
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32)
    grids = [image.GRMHDGrid(model, c) for c in cells]
    host = [g.host_cells() for g in grids]
    p = gs.pass_a_params(model, nx=12, ny=12, camera_radius=200.0, outer_radius=30.0)
    return grids, host, times, model, p


def _fd(p, model, cells, times, kind, h):
    import oracle
    from kpolaris_amd import abi
    if kind == 1:
        res = []
        for sgn in (1.0, -1.0):
            m2 = abi.GRMHDModel.from_buffer_copy(model)
            m2.trat_large = model.trat_large * (1.0 + sgn * h)
            cs = []
            for c in cells:
                c2 = c.copy()
                oracle.grmhd_derive_scalars(m2, c2)
                cs.append(c2)
            res.append(oracle.run_grmhd_slow_light_image(p, m2, cs, times, T_OBS)[0])
        step = 2.0 * h * model.trat_large
    else:
        param = 1 if kind == 0 else 3
        res = [oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS, perturbation=(param, s * h))[0]
               for s in (1.0, -1.0)]
        step = 2.0 * h
    same = (res[0]["steps"] == res[1]["steps"]) & (res[0]["pass_a_steps"] == res[1]["pass_a_steps"])
    return {n: (res[0][n] - res[1][n]) / step for n in STOKES}, same


@pytest.mark.parametrize("hr", [0.0, 0.1])
def test_gpu_slow_light_gradient_against_oracle(seq, hr):
    """hr > 0: with the equatorial emission wedge (its kernel; the wedge
    images' parity constants for the value).  The wedge keeps the midplane's
    emission, where the Faraday depth is largest, so Q and U curve more with
    the density scale and R_high: its central differences use steps ten times
    smaller (h = 1e-5 and 5e-6; at h = 1e-4 the density derivative's O(h^2)
    term is 7.5e-4 of Q's scale and the Richardson value's O(h^4) term 2e-7,
    at 2e-5 the R_high derivative's O(h^2) term is 1.1e-4 and the Richardson
    value agrees to 1.2e-8)."""
    import oracle
    import torch
    from kpolaris_amd import abi, image
    grids, host, times, model, p = seq
    p = abi.PassAParams.from_buffer_copy(p)
    p.emission_selection.equatorial_h_over_r = hr
    wedge = dict(max_ill=parity.WEDGE_MAX_ILL, step_rel=parity.WEDGE_STEP_REL) if hr > 0.0 else {}
    print(f"case: slow-light gradient 12x12 wedge {hr} vs oracle central differences", flush=True)
    out, grad, st = image.run_grmhd_slow_light_image_gradient(p, grids, times, T_OBS)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    g = grad.cpu().numpy()
    ref, _, _ = oracle.run_grmhd_slow_light_image(p, model, host, times, T_OBS)
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(q, model, host, times, t)[0]  # noqa: E731
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got, ref, 1, "gpu slow-light tangent value", cond, time_cut_steps=1, **wedge),
          flush=True)
    # the derivatives are compared on the pixels that the oracle's own
    # sensitivity leaves well-conditioned (every component) and whose step
    # counts equal the oracle's
    same_steps = got["steps"] == ref["steps"]
    for kind in range(3):
        h = 1e-4 if hr == 0.0 else 1e-5
        fd, same = _fd(p, model, host, times, kind, h)
        fd2, same2 = _fd(p, model, host, times, kind, 0.5 * h)
        for s, name in enumerate(STOKES):
            ok = same & same2 & same_steps & parity.smooth_for_derivatives(ref, cond, 1)
            print(f"gradient check of {name} on {int(ok.sum())}/{ok.size} pixels", flush=True)
            assert ok.sum() >= parity.MIN_CHECKED * ok.size
            scale = np.abs(fd[name][ok]).max()
            rich = (4.0 * fd2[name] - fd[name]) / 3.0
            err = np.abs(g[kind, s][ok] - fd[name][ok]).max() / scale
            err_r = np.abs(g[kind, s][ok] - rich[ok]).max() / scale
            print(f"d/d{('q_density', 'R_high', 'q_magnetic')[kind]} {name}: {err:.2e} (Richardson {err_r:.2e})",
                  flush=True)
            assert err < 1e-4 and err_r < 1e-7, (kind, name, err, err_r)
            # most pixels agree to rounding; a rounding-level difference can flip one
            # accept/reject decision along a ray without changing its step count, and
            # the derivative of the discretised transfer then moves by O(rtol)
            per_pixel = np.abs(g[kind, s][ok] - rich[ok]) / scale
            assert np.percentile(per_pixel, 95) < 1e-8, (kind, name, np.percentile(per_pixel, 95))



def test_gpu_slow_light_autograd(seq):
    """backward of a weighted sum of the Stokes image is the contraction of
    the incoming gradient with the Jacobian of the gradient entry point, and
    that Jacobian at a point away from the unit scales matches central
    differences of the device's own forward passes (R_high through re-derived
    snapshot scalars).  Step-size control makes the image a piecewise-smooth
    function of the parameters: a different accept or reject somewhere along a
    ray moves the pixel by a discretisation-sized jump, and a difference
    quotient over a window holding a jump grows as the window shrinks.  With
    D(h), D(h/2), D(h/4), a smooth pixel has |D(h/2) - D(h/4)| about a quarter
    of |D(h) - D(h/2)| (O(h^2) truncation), a pixel with a jump in the two
    smaller windows about twice; the comparison keeps pixels with ratio <= 0.3
    and checks them against the Richardson value (4 D(h/4) - D(h/2)) / 3."""
    import torch
    from kpolaris_amd import autograd, image
    grids, host, times, model, p = seq
    print("case: slow-light autograd at theta = (0.1, R_high, -0.05)", flush=True)
    r0 = float(model.trat_large)
    theta0 = [0.1, r0, -0.05]
    npix = p.camera.nx * p.camera.ny
    try:
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        w = torch.randn(4, npix, dtype=torch.float64, device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(3))
        S = autograd.slow_light_stokes(theta, p, grids, times, T_OBS)
        (S * w).sum().backward()
        g = theta.grad.detach().numpy()
        _, jac, _ = image.run_grmhd_slow_light_image_gradient(p, grids, times, T_OBS, (theta0[0], theta0[2]),
                                                              with_stats=False)
        torch.cuda.synchronize()
        contraction = (jac * w.unsqueeze(0)).sum(dim=(1, 2)).cpu().numpy()
        assert np.allclose(g, contraction, rtol=1e-12, atol=0.0), (g, contraction)
        J = jac.cpu().numpy()

        def forward(t):
            with torch.no_grad():
                return autograd.slow_light_stokes(torch.tensor(t, dtype=torch.float64), p, grids, times,
                                                  T_OBS).cpu().numpy()

        for k in range(3):
            h = 1e-4 * (r0 if k == 1 else 1.0)
            d = {}
            for hh in (h, h / 2, h / 4):
                vals = []
                for sgn in (1.0, -1.0):
                    t2 = list(theta0)
                    t2[k] += sgn * hh
                    vals.append(forward(t2))
                d[hh] = (vals[0] - vals[1]) / (2.0 * hh)
            rich = (4.0 * d[h / 4] - d[h / 2]) / 3.0
            for s, name in enumerate(STOKES):
                scale = np.abs(d[h][s]).max()
                a = np.abs(d[h][s] - d[h / 2][s])
                b = np.abs(d[h / 2][s] - d[h / 4][s])
                smooth = b <= 0.3 * a + 1e-9 * scale
                assert smooth.sum() >= npix // 2, (k, name, smooth.sum())
                err = np.abs(J[k, s][smooth] - rich[s][smooth]).max() / scale
                print(f"autograd d/d{('q_density', 'R_high', 'q_magnetic')[k]} {name}: {err:.2e} "
                      f"on {int(smooth.sum())} of {npix} pixels", flush=True)
                assert err < 1e-6, (k, name, err)
    finally:
        # restore the grids' R_high
        for gr in grids:
            if gr.model.trat_large != r0:
                gr.model.trat_large = r0
                gr.derive()
