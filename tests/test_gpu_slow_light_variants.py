"""Slow light with the kernel variants the block loop selects: the kappa
distribution (the nonthermal windowed kernel) and the equatorial emission
wedge (the wedge windowed kernels, FP64 and mixed).  The one-block image is
held to the oracle; blocks, a streamed sequence and observation-time jobs
equal it bit for bit (decoupled stepping does not depend on where blocks
fall)."""
from __future__ import annotations

import numpy as np
import pytest

from kpolaris_amd import abi
from tests import parity

pytestmark = pytest.mark.gpu

T_OBS = 275.0
KEYS = ("image_i", "image_q", "image_u", "image_v", "steps", "reason", "pass_a_steps", "closure_x", "final_null")


def _np(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _same(a, b, label):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (label, k)


def _sequence(**kw):
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32, **kw)
    return cells, times, model


def _grids(model, cells):
    from kpolaris_amd import image
    grids = [image.GRMHDGrid(model, c) for c in cells]
    return grids, [g.host_cells() for g in grids]


def _blocks_stream_jobs(p, grids, host, times, one, label, **kw):
    """Blocks of two pairs, the streamed sequence and two jobs (one at T_OBS)
    against the one-block image `one`."""
    from kpolaris_amd import image
    blk, _ = image.run_grmhd_slow_light_image(p, grids, times, T_OBS, windows_per_block=2, **kw)
    _same(_np(blk), one, f"{label} blocks")
    loader = image.host_cell_loader(host)
    out, _, _ = image.run_grmhd_slow_light_stream(p, grids[0].model, times, T_OBS, loader, windows_per_block=2, **kw)
    _same(_np(out), one, f"{label} streamed")
    probe = image.slow_light_time_probe(p, grids[0].model)
    t2 = times[-1] - probe["max_x0"] - 1.0
    assert times[0] - probe["min_x0"] < t2
    outs, _, _ = image.run_grmhd_slow_light_stream_jobs(p, grids[0].model, times, [(t2, 0, len(times) - 1),
                                                                                   (T_OBS, 0, len(times) - 1)],
                                                        loader, windows_per_block=2, **kw)
    _same(_np(outs[1]), one, f"{label} jobs")
    ref2, _ = image.run_grmhd_slow_light_image(p, grids, times, t2, **kw)
    _same(_np(outs[0]), _np(ref2), f"{label} jobs, second time")


def test_gpu_slow_light_kappa():
    """Kappa electrons (magnetised everywhere: no field-free sample) against
    the oracle, then in blocks, streamed and as jobs."""
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = _sequence(beta_min=1.0, background_field=1e-2)
    model.emission_type = abi.EMISSION_KAPPA
    grids, host = _grids(model, cells)
    p = gs.pass_a_params(model, nx=12, ny=12, camera_radius=200.0, outer_radius=30.0)
    print("case: slow light kappa 12x12", flush=True)
    out, _ = image.run_grmhd_slow_light_image(p, grids, times, T_OBS)
    one = _np(out)
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(q, model, host, times, t)[0]  # noqa: E731
    ref = run(p)
    assert np.max(ref["image_i"]) > 0.0
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(one, ref, 1, "gpu slow light kappa", cond, time_cut_steps=1), flush=True)
    print("case: slow light kappa blocks, stream, jobs", flush=True)
    _blocks_stream_jobs(p, grids, host, times, one, "kappa")


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_gpu_slow_light_wedge_blocks(precision):
    """The wedge (h/r = 0.1) windowed kernels: blocks, streamed and jobs equal
    the one-block wedge image; the wedge removes flux."""
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = _sequence()
    grids, host = _grids(model, cells)
    p = gs.pass_a_params(model, nx=12, ny=12, camera_radius=200.0, outer_radius=30.0)
    full, _ = image.run_grmhd_slow_light_image(p, grids, times, T_OBS, precision=precision)
    plain_probe = image.slow_light_time_probe(p, grids[0].model)
    p.emission_selection.equatorial_h_over_r = 0.1
    # the time probe is geometric: the wedge does not change it
    assert image.slow_light_time_probe(p, grids[0].model) == plain_probe
    print(f"case: slow light wedge 0.1 precision {precision}", flush=True)
    out, _ = image.run_grmhd_slow_light_image(p, grids, times, T_OBS, precision=precision)
    one = _np(out)
    assert 0.0 < one["image_i"].sum() < _np(full)["image_i"].sum()
    _blocks_stream_jobs(p, grids, host, times, one, f"wedge {precision}", precision=precision)


def test_gpu_slow_light_three_frequencies():
    """Three frequencies sharing the slow-light geodesics (the single-pass
    frequency loop of the windowed kernel), FP64 and mixed: blocks, streamed
    and jobs equal the one-block image; the FP64 image against the oracle."""
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = _sequence()
    grids, host = _grids(model, cells)
    p = gs.pass_a_params(model, nx=10, ny=10, camera_radius=200.0, outer_radius=30.0)
    F = [86e9, 230e9, 345e9]
    for precision in (abi.PRECISION_FP64, abi.PRECISION_MIXED):
        print(f"case: slow light 3 frequencies precision {precision}", flush=True)
        out, _ = image.run_grmhd_slow_light_image(p, grids, times, T_OBS, F, precision=precision)
        one = _np(out)
        if precision == abi.PRECISION_FP64:
            run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(q, model, host, times, t, F)[0]  # noqa: E731
            ref = run(p)
            cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
            print(parity.assert_parity_grid(one, ref, 3, "gpu slow light 3 frequencies", cond, time_cut_steps=1),
                  flush=True)
        _blocks_stream_jobs(p, grids, host, times, one, f"3 frequencies {precision}", frequencies=F,
                            precision=precision)
