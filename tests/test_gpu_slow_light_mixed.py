"""Mixed-precision slow light on the device: FP32 thermal coefficients (mixed
in FP32 between the two snapshots under coefficient interpolation) and FP32
transfer on the FP64 geodesics, window logic and fluid reconstruction.  Held
to the mixed-precision image bars (tests/parity.py) against the FP64 oracle;
the block runs and the streamed sequence equal the resident mixed run bit for
bit, as they do in FP64."""
from __future__ import annotations

import numpy as np
import pytest

from kpolaris_amd import abi
from tests import parity

pytestmark = pytest.mark.gpu

T_OBS = 275.0
KEYS = ("image_i", "image_q", "image_u", "image_v", "steps", "reason", "pass_a_steps", "closure_x", "final_null")
MIXED = abi.PRECISION_MIXED


@pytest.fixture(scope="module")
def seq():
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32)
    grids = [image.GRMHDGrid(model, c) for c in cells]
    host = [g.host_cells() for g in grids]
    p = gs.pass_a_params(model, nx=16, ny=16, camera_radius=200.0, outer_radius=30.0)
    return grids, host, times, model, p


def _np(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _resident(grids, times, p, freqs=None, **kw):
    from kpolaris_amd import image
    out, st = image.run_grmhd_slow_light_image(p, grids, times, T_OBS, freqs, **kw)
    return _np(out), st


@pytest.mark.parametrize("mode,interp,persistent,split", [
    (abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID, True, 1),
    (abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID, False, 1),
    (abi.SLOW_LIGHT_SNAPSHOT, abi.SLOW_LIGHT_COEFFICIENTS, True, 1),
    (abi.SLOW_LIGHT_BLOCK, abi.SLOW_LIGHT_FLUID, True, 0),
])
def test_gpu_mixed_slow_light_against_oracle(seq, mode, interp, persistent, split):
    """Two frequencies on shared geodesics (the single-pass loop) against the
    FP64 oracle; the coarse image's net polarisation fractions are held in
    absolute terms."""
    import oracle
    grids, host, times, model, p = seq
    freqs = [model.freq_cgs, 345e9]
    print(f"case: mixed slow light mode {mode} interp {interp} persistent {persistent} split {split}", flush=True)
    got, st = _resident(grids, times, p, freqs, step_mode=mode, interpolation=interp, persistent=persistent,
                        split_transport=split, precision=MIXED)
    ref, _, _ = oracle.run_grmhd_slow_light_image(p, model, host, times, T_OBS, freqs, step_mode=mode,
                                                  interpolation=interp)
    assert np.array_equal(got["reason"], ref["reason"])
    g = parity.split_frequencies(got, 2)
    r = parity.split_frequencies(ref, 2)
    for k in range(2):
        e = parity.assert_mixed(g[k], r[k], f"gpu mixed slow light f{k}", pol="absolute")
        print({n: e[n] for n in ("total_i_rel", "lp", "cp", "pixel_max_rel_peak")}, flush=True)


@pytest.mark.parametrize("mode,wpb", [(abi.SLOW_LIGHT_DECOUPLED, 2), (abi.SLOW_LIGHT_SNAPSHOT, 1)])
def test_gpu_mixed_slow_light_blocks_and_stream(seq, mode, wpb):
    """The windowed mixed kernels over blocks, resident and streamed through
    the loader, equal the single-block mixed run bit for bit."""
    from kpolaris_amd import image
    grids, host, times, model, p = seq
    print(f"case: mixed slow light blocks mode {mode} windows_per_block {wpb}", flush=True)
    one, _ = _resident(grids, times, p, step_mode=mode, precision=MIXED)
    blk, _ = _resident(grids, times, p, step_mode=mode, windows_per_block=wpb, precision=MIXED)
    out, _, rep = image.run_grmhd_slow_light_stream(p, grids[0].model, times, T_OBS, image.host_cell_loader(host),
                                                    step_mode=mode, windows_per_block=wpb, precision=MIXED)
    got = _np(out)
    for k in KEYS:
        assert np.array_equal(blk[k], one[k]), ("blocks", k)
        assert np.array_equal(got[k], one[k]), ("stream", k)
    assert rep["loads"] == len(times)
    assert np.sum(one["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * one["reason"].size


def test_gpu_mixed_slow_light_refusals(seq):
    """The kappa distribution is refused under mixed precision, and the
    library stays usable."""
    from kpolaris_amd import image
    grids, host, times, model, p = seq
    print("case: mixed slow light refusals", flush=True)
    mk = abi.GRMHDModel.from_buffer_copy(model)
    mk.emission_type = abi.EMISSION_KAPPA
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_slow_light_stream(p, mk, times, T_OBS, image.host_cell_loader(host), windows_per_block=2,
                                          precision=MIXED)
    assert e.value.status == abi.KP_UNSUPPORTED
    ok, _ = _resident(grids, times, p, precision=MIXED)
    assert np.sum(ok["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * ok["reason"].size
