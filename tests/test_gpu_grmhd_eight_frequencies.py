"""The GRMHD multi-frequency configuration: eight frequencies log-spaced
86-345 GHz on shared FMKS geodesics (the single-pass kernel with two
LDS Stokes buffers of 8 x 4 doubles per lane, 64 KB per 128-lane block),
against the oracle on a reduced grid, and the full 288x128x128 snapshot /
512x512 image with property checks and an oracle spot sample; the RIAF
image at the same eight frequencies against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

FREQS = list(np.geomspace(86e9, 345e9, 8))


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def test_gpu_grmhd_eight_frequencies_parity():
    import oracle
    import torch
    from kpolaris_amd import grmhd_synth as gs, image
    print("case: GRMHD 48x32x32, 10x10 image, 8 frequencies 86-345 GHz", flush=True)
    cells, m = gs.make_snapshot(48, 32, 32, seed=20261202)
    grid = image.GRMHDGrid(m, cells)
    host = grid.host_cells()
    p = gs.pass_a_params(m, nx=10, ny=10)
    out, st = image.run_grmhd_image(p, grid, FREQS)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref, _ = oracle.run_grmhd_image(p, m, host, FREQS)
    cond = parity.sensitivity(lambda q: oracle.run_grmhd_image(q, m, host, FREQS)[0], p, ref)
    print(parity.assert_parity_grid(got, ref, 8, "gpu GRMHD 8 freqs", cond), flush=True)
    n = p.camera.nx * p.camera.ny
    assert all(np.max(ref["image_i"][f * n:(f + 1) * n]) > 0.0 for f in range(8))


def test_gpu_riaf_eight_frequencies_parity():
    import oracle
    import torch
    from kpolaris_amd import image
    print("case: RIAF 16x16, 8 frequencies 86-345 GHz", flush=True)
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    o.set("nx", 16)
    o.set("ny", 16)
    p, m, _, _ = o.resolve_riaf()
    out, _ = image.run_riaf_multifrequency_image(p, m, FREQS)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref_mf, _ = oracle.run_riaf_multifrequency_image(p, m, FREQS)
    single = parity.oracle_with_sensitivity(p, m)
    for r in parity.assert_parity_multifrequency(got, ref_mf, 8, single, "gpu riaf 8 freqs"):
        print(r, flush=True)


def test_gpu_grmhd_full_size_eight_frequencies():
    """288x128x128 snapshot, 512x512 image, 8 frequencies: finite outputs,
    the reason histogram, |P| <= I on rays that return to the camera, the
    spectrum, and 16 pixels against the oracle."""
    import oracle
    import torch
    from kpolaris_amd import grmhd_synth as gs, image
    print("case: GRMHD 288x128x128, 512x512, 8 frequencies", flush=True)
    cells, m = gs.make_snapshot(288, 128, 128)
    grid = image.GRMHDGrid(m, cells)
    p = gs.pass_a_params(m, nx=512, ny=512)
    out, st = image.run_grmhd_image(p, grid, FREQS)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    n = 512 * 512
    print(f"kernel {st['kernel_ms']:.1f} ms, {n / (st['kernel_ms'] * 1e-3):.3e} rays/s", flush=True)
    reasons = np.bincount(got["reason"], minlength=10)
    print("reasons", reasons.tolist(), flush=True)
    assert reasons[1] + reasons[2] == n, "rays end at the camera or at the inner boundary"
    assert reasons[1] > 0.5 * n
    for name in parity.STOKES:
        assert np.all(np.isfinite(got[name])), name
    ok = got["reason"] == 1
    flux = []
    for f in range(8):
        sl = slice(f * n, (f + 1) * n)
        I = got["image_i"][sl]
        P = np.sqrt(got["image_q"][sl] ** 2 + got["image_u"][sl] ** 2 + got["image_v"][sl] ** 2)
        peak = float(I.max())
        assert np.all(I[ok] >= -1e-12 * peak)
        assert np.all(P[ok] <= I[ok] * (1 + 1e-9) + 1e-12 * peak), f
        flux.append(float(I[ok].sum()) * FREQS[f] ** 3)
    print("flux spectrum (nu^3 sum I_inv)", flux, flush=True)
    assert all(v > 0.0 for v in flux)
    # oracle spot sample: 16 pixels spread over the image, with the oracle's
    # own conditioning of those pixels
    host = grid.host_cells()
    pix = (np.arange(16) * (n // 16) + 12345) % n

    def sample(img_full):
        d = {k: np.concatenate([img_full[k][f * n + pix] for f in range(8)]) for k in parity.STOKES}
        for k in ("reason", "steps", "pass_a_steps"):
            d[k] = img_full[k][pix]
        return d

    def oracle_pixels(q):
        full = None
        for k in pix:
            r1, _ = oracle.run_grmhd_image(q, m, host, FREQS, pixel_begin=int(k), pixel_end=int(k) + 1)
            if full is None:
                full = r1
            else:
                for name in parity.STOKES:
                    for f in range(8):
                        full[name][f * n + k] = r1[name][f * n + k]
                for name in ("reason", "steps", "pass_a_steps"):
                    full[name][k] = r1[name][k]
        return sample(full)

    ref_s = oracle_pixels(p)
    cond = parity.sensitivity(oracle_pixels, p, ref_s)
    peaks = [float(got["image_i"][f * n:(f + 1) * n].max()) for f in range(8)]
    print(parity.check(sample(got), ref_s, cond, "GRMHD 512x512 spot", nfreq=8, peak=peaks, diagnostics=False,
                       max_ill=0.25), flush=True)
