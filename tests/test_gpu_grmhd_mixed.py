"""Mixed-precision GRMHD images on the device: the thermal coefficients and
the Stokes transfer in FP32 (nu^3-scaled units) on FP64 FMKS geodesics and
FP64 grid sampling.  Held to the mixed-precision image bars (tests/parity.py)
against the FP64 oracle on reduced grids, and against the FP64 device image
at the full 288x128x128 snapshot / 512x512 image."""
from __future__ import annotations

import numpy as np
import pytest

from kpolaris_amd import abi
from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

FREQS = list(np.geomspace(86e9, 345e9, 8))


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def small():
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    cells, m = gs.make_snapshot(48, 32, 32, seed=20261202)
    grid = image.GRMHDGrid(m, cells)
    host = grid.host_cells()
    p = gs.pass_a_params(m, nx=16, ny=16)
    ref, _ = oracle.run_grmhd_image(p, m, host, FREQS)
    return grid, host, m, p, ref


def _np(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _per_frequency(got, ref, nfreq, label, pol):
    g = parity.split_frequencies(got, nfreq)
    r = parity.split_frequencies(ref, nfreq)
    return [parity.assert_mixed(g[f], r[f], f"{label} f{f}", pol=pol) for f in range(nfreq)]


@pytest.mark.parametrize("persistent,split", [(True, 1), (False, 1), (True, 0)])
def test_gpu_grmhd_mixed_against_oracle(small, persistent, split):
    """Eight frequencies on one list (the single-pass frequency loop), split
    persistent and per-pixel, and unsplit; the coarse image's net
    polarisation fractions are held in absolute terms."""
    from kpolaris_amd import image
    grid, host, m, p, ref = small
    print(f"case: mixed GRMHD 16x16 8 freqs persistent={persistent} split={split}", flush=True)
    out, st = image.run_grmhd_image(p, grid, FREQS, precision=abi.PRECISION_MIXED, persistent=persistent,
                                    split_transport=split)
    got = _np(out)
    for e in _per_frequency(got, ref, 8, "gpu mixed GRMHD", "absolute"):
        assert e["reason_mismatch"] == 0, e
    print({k: st[k] for k in ("kernel_ms", "pass_b_steps", "coefficient_evals", "fp64_guard_steps")}, flush=True)
    # every geodesic attempt on FMKS geodesics is FP64 and counted as such
    assert st["fp64_guard_steps"] >= st["pass_a_steps"] + st["pass_b_steps"]


def test_gpu_grmhd_mixed_separate_control_list(small):
    """Step control on two frequencies, output at three others: the two-pass
    loop (control list, then the transfer list) of the FP32 radiation."""
    import oracle
    from kpolaris_amd import image
    grid, host, m, p, _ = small
    F, ctrl = [100e9, 230e9, 300e9], [86e9, 345e9]
    print("case: mixed GRMHD 16x16 output 3 freqs, control 2 freqs", flush=True)
    out, _ = image.run_grmhd_image(p, grid, F, ctrl, precision=abi.PRECISION_MIXED)
    ref, _ = oracle.run_grmhd_image(p, m, host, F, ctrl)
    for e in _per_frequency(_np(out), ref, 3, "gpu mixed GRMHD separate control", "absolute"):
        assert e["reason_mismatch"] == 0, e


def test_gpu_grmhd_mixed_full_size_against_fp64():
    """288x128x128 snapshot, 512x512 image, 8 frequencies: the mixed image
    against the FP64 device image at the full-size bars."""
    from kpolaris_amd import grmhd_synth as gs, image
    print("case: mixed GRMHD 288x128x128, 512x512, 8 frequencies", flush=True)
    cells, m = gs.make_snapshot(288, 128, 128)
    grid = image.GRMHDGrid(m, cells)
    p = gs.pass_a_params(m, nx=512, ny=512)
    ref_out, st64 = image.run_grmhd_image(p, grid, FREQS)
    ref = _np(ref_out)
    del ref_out
    out, st = image.run_grmhd_image(p, grid, FREQS, precision=abi.PRECISION_MIXED)
    got = _np(out)
    print(f"fp64 pass B {st64['pass_b_ms']:.1f} ms, mixed pass B {st['pass_b_ms']:.1f} ms", flush=True)
    for e in _per_frequency(got, ref, 8, "gpu mixed GRMHD 512x512", "relative"):
        print({k: e[k] for k in ("total_i_rel", "lp", "cp", "pixel_max_rel_peak", "reason_mismatch")}, flush=True)
    assert np.array_equal(got["pass_a_steps"], ref["pass_a_steps"])


def test_gpu_grmhd_mixed_refuses_nonthermal(small):
    """The kappa distribution stays FP64: mixed precision is refused with
    KP_UNSUPPORTED, and the library stays usable."""
    import types
    from kpolaris_amd import image
    grid, host, m, p, _ = small
    print("case: mixed GRMHD refuses kappa", flush=True)
    mk = abi.GRMHDModel.from_buffer_copy(grid.model)
    mk.emission_type = abi.EMISSION_KAPPA
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_image(p, types.SimpleNamespace(model=mk), FREQS[:2], precision=abi.PRECISION_MIXED)
    assert e.value.status == abi.KP_UNSUPPORTED
    out, _ = image.run_grmhd_image(p, grid, FREQS[:2], precision=abi.PRECISION_MIXED)
    got = _np(out)
    assert np.sum(got["reason"] == abi.REASON_REACHED_CAMERA) > 0.5 * got["reason"].size
