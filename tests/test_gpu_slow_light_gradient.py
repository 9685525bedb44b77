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


def test_gpu_slow_light_gradient_against_oracle(seq):
    import oracle
    import torch
    from kpolaris_amd import image
    grids, host, times, model, p = seq
    print("case: slow-light gradient 12x12 vs oracle central differences", flush=True)
    out, grad, st = image.run_grmhd_slow_light_image_gradient(p, grids, times, T_OBS)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    g = grad.cpu().numpy()
    ref, _, _ = oracle.run_grmhd_slow_light_image(p, model, host, times, T_OBS)
    ill = parity.sensitivity(lambda q: oracle.run_grmhd_slow_light_image(q, model, host, times, T_OBS)[0], p, ref)
    print(parity.assert_parity_grid(got, ref, 1, "gpu slow-light tangent value", max_step_diff=1, ill=ill), flush=True)
    moved = ill | (got["steps"] != ref["steps"])
    for kind in range(3):
        fd, same = _fd(p, model, host, times, kind, 1e-4)
        fd2, same2 = _fd(p, model, host, times, kind, 5e-5)
        ok = same & same2 & ~moved
        assert ok.sum() >= 30
        for s, name in enumerate(STOKES):
            scale = np.abs(fd[name][ok]).max()
            rich = (4.0 * fd2[name] - fd[name]) / 3.0
            err = np.abs(g[kind, s][ok] - fd[name][ok]).max() / scale
            err_r = np.abs(g[kind, s][ok] - rich[ok]).max() / scale
            print(f"d/d{('q_density', 'R_high', 'q_magnetic')[kind]} {name}: {err:.2e} (Richardson {err_r:.2e})",
                  flush=True)
            assert err < 1e-4 and err_r < 1e-7, (kind, name, err, err_r)


def test_gpu_slow_light_autograd(seq):
    """backward of a weighted sum of the Stokes image equals the Jacobian
    contraction, and matches central differences of the device's own forward
    passes (R_high through re-derived snapshot scalars)."""
    import torch
    from kpolaris_amd import autograd
    grids, host, times, model, p = seq
    print("case: slow-light autograd", flush=True)
    r0 = float(model.trat_large)
    theta = torch.tensor([0.1, r0, -0.05], dtype=torch.float64, requires_grad=True)
    w = torch.randn(4, p.camera.nx * p.camera.ny, dtype=torch.float64, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(3))
    S = autograd.slow_light_stokes(theta, p, grids, times, T_OBS)
    loss = (S * w).sum() / S[0].abs().max().detach()
    loss.backward()
    g = theta.grad.detach().numpy()
    steps = (1e-4, 1e-4 * r0, 1e-4)
    for k in range(3):
        vals = []
        for sgn in (1.0, -1.0):
            t2 = theta.detach().clone()
            t2[k] += sgn * steps[k]
            S2 = autograd.slow_light_stokes(t2, p, grids, times, T_OBS)
            vals.append(float(((S2 * w).sum() / S[0].abs().max()).item()))
        fd = (vals[0] - vals[1]) / (2.0 * steps[k])
        print(k, g[k], fd, flush=True)
        assert abs(g[k] - fd) <= 1e-4 * max(abs(fd), 1e-12) + 1e-9
    # restore the grids' R_high
    for gr in grids:
        if gr.model.trat_large != r0:
            gr.model.trat_large = r0
            gr.derive()
