"""Slow-light GRMHD images on the GPU through the C ABI against the oracle:
every stepping and interpolation mode, both launch modes and the unsplit
kernel, two frequencies on shared geodesics, the time-support error, and the
static-image limit of a repeated snapshot."""
from __future__ import annotations

import numpy as np
import pytest

from tests import parity

pytestmark = pytest.mark.gpu

T_OBS = 275.0


@pytest.fixture(scope="module")
def seq():
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32)
    grids = [image.GRMHDGrid(model, c) for c in cells]
    host = [g.host_cells() for g in grids]
    p = gs.pass_a_params(model, nx=16, ny=16, camera_radius=200.0, outer_radius=30.0)
    return grids, host, times, model, p


def _gpu(grids, times, p, t_obs, **kw):
    import torch
    from kpolaris_amd import image
    out, st = image.run_grmhd_slow_light_image(p, grids, times, t_obs, **kw)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}, st


def test_gpu_device_derived_scalars_feed_the_oracle(seq):
    """The oracle runs on the device-derived snapshot records it is given."""
    grids, host, times, model, p = seq
    assert all(np.all(np.isfinite(h["derived"])) for h in host)
    assert all(np.any(h["derived"][:, 0] > 0.0) for h in host)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("interp", [1, 0])
def test_gpu_slow_light_parity(seq, mode, interp):
    import oracle
    grids, host, times, model, p = seq
    print(f"case: slow light 16x16 mode {mode} interp {interp}", flush=True)
    got, st = _gpu(grids, times, p, T_OBS, step_mode=mode, interpolation=interp)
    ref, ost, w = oracle.run_grmhd_slow_light_image(p, model, host, times, T_OBS, step_mode=mode,
                                                    interpolation=interp)
    assert w[0] == 0
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(  # noqa: E731
        q, model, host, times, t, step_mode=mode, interpolation=interp)[0]
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got, ref, 1, f"gpu slow light {mode}/{interp}", cond, time_cut_steps=1),
          st["pass_b_ms"], flush=True)
    assert np.max(ref["image_i"]) > 0.0


@pytest.mark.parametrize("persistent,split", [(False, 1), (True, 0)])
def test_gpu_slow_light_launch_modes_and_frequencies(seq, persistent, split):
    """One-ray-per-thread and unsplit launches, two frequencies."""
    import oracle
    grids, host, times, model, p = seq
    F = [86e9, 230e9]
    print(f"case: slow light 2 freqs persistent={persistent} split={split}", flush=True)
    got, _ = _gpu(grids, times, p, T_OBS, frequencies=F, persistent=persistent, split_transport=split)
    ref, _, w = oracle.run_grmhd_slow_light_image(p, model, host, times, T_OBS, frequencies=F)
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(q, model, host, times, t, frequencies=F)[0]  # noqa: E731
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got, ref, 2, "gpu slow light 2 freqs", cond, time_cut_steps=1), flush=True)


def test_gpu_slow_light_time_support_error(seq):
    """Decoupled stepping that needs times before the first snapshot raises
    KP_TIME_SUPPORT naming the requested range; the image is written with the
    waiting rays as slow_light_time_exhausted, as the oracle has them."""
    import oracle
    from kpolaris_amd import abi, image
    grids, host, times, model, p = seq
    print("case: slow light time-support error", flush=True)
    out = image.alloc_device_image(p.camera.nx * p.camera.ny)
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_slow_light_image(p, grids, times, 230.0, out=out)
    assert e.value.status == abi.KP_TIME_SUPPORT
    assert "outside supplied physical time support" in str(e.value)
    import torch
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref, _, w = oracle.run_grmhd_slow_light_image(p, model, host, times, 230.0)
    assert w[0] > 0
    assert np.array_equal(got["reason"], ref["reason"])
    assert f"{int(w[0])} rays" in str(e.value)


def test_gpu_repeated_snapshot_is_the_static_image(seq):
    """A repeated snapshot with decoupled stepping images as the static grid
    (same geodesics, same coefficients; the kernels differ only in code
    layout, so rounding may differ in the last bits)."""
    from kpolaris_amd import image
    grids, host, times, model, p = seq
    print("case: slow light repeated snapshot vs static", flush=True)
    got, _ = _gpu([grids[0]] * 3, [-2000.0, 0.0, 2000.0], p, 0.0)
    import torch
    out, _ = image.run_grmhd_image(p, grids[0])
    torch.cuda.synchronize()
    ref = {k: v.cpu().numpy() for k, v in out.items()}
    assert np.array_equal(got["reason"], ref["reason"])
    assert np.array_equal(got["steps"], ref["steps"])
    peak = float(np.max(ref["image_i"]))
    for k in parity.STOKES:
        assert np.max(np.abs(got[k] - ref[k])) <= 1e-12 * peak, k


def test_gpu_slow_light_rejects_bad_sequences(seq):
    from kpolaris_amd import abi, image
    grids, host, times, model, p = seq
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_slow_light_image(p, grids[:1], times[:1], T_OBS)
    assert e.value.status == abi.KP_INVALID_ARGUMENT
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_slow_light_image(p, grids[:3], [0.0, 20.0, 20.0], T_OBS)
    assert e.value.status == abi.KP_INVALID_ARGUMENT
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_slow_light_image(p, grids[:2], times[:2], T_OBS, step_mode=7)
    assert e.value.status == abi.KP_INVALID_ARGUMENT


def test_gpu_time_probe_brackets_the_radiating_times(seq):
    """The probe's x^0 range (geometric stepping, no radiation cap) brackets
    the fluid times the image needs: a sequence spanning it with a margin of
    1 M leaves no ray waiting, one that starts a third of the way in does."""
    from kpolaris_amd import abi, image
    grids, host, times, model, p = seq
    print("case: slow light time probe", flush=True)
    pr = image.slow_light_time_probe(p, model)
    assert pr["radiating_pixels"] > 0 and pr["returned_pixels"] > 0
    lo, hi = T_OBS + pr["min_x0"], T_OBS + pr["max_x0"]
    assert hi > lo
    _gpu(grids, np.linspace(lo - 1.0, hi + 1.0, len(grids)), p, T_OBS)
    with pytest.raises(image.KPError) as e:
        _gpu(grids, np.linspace(lo + (hi - lo) / 3.0, hi + 1.0, len(grids)), p, T_OBS)
    assert e.value.status == abi.KP_TIME_SUPPORT
