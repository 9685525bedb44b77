"""Several observation times from one pass over a streamed sequence (the
reference's observer jobs, src/image/slow_light_batch_driver.hpp:105-340):
pass A once, every job's rays from those endpoints, each block running the
jobs whose snapshot support overlaps it.

Decoupled and snapshot stepping do not depend on where blocks fall, so each
job equals the single-observation-time run over its own snapshots
first..last bit for bit; block stepping does, and a job whose support starts
at the earliest job's first snapshot equals that run over first..last (the
blocks fall at the same snapshots).  The option rules are the reference's
(:116, 144-146)."""
from __future__ import annotations

import numpy as np
import pytest

from kpolaris_amd import abi

pytestmark = pytest.mark.gpu

KEYS = ("image_i", "image_q", "image_u", "image_v", "steps", "reason", "pass_a_steps", "closure_x", "final_null")


@pytest.fixture(scope="module")
def seq():
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=16, dt=10.0, t0=0.0, n1=48, n2=32, n3=32)
    grid = image.GRMHDGrid(model, cells[0])
    host = [image.GRMHDGrid(model, c).host_cells() for c in cells]
    p = gs.pass_a_params(model, nx=16, ny=16, camera_radius=200.0, outer_radius=30.0)
    probe = image.slow_light_time_probe(p, grid.model)
    return grid.model, host, np.asarray(times), p, probe


def _np(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _t_obs(times, probe, first, last, where=0.5):
    """An observation time whose radiating samples (t_obs + [min_x0, max_x0])
    lie inside snapshots first..last."""
    lo = times[first] - probe["min_x0"]
    hi = times[last] - probe["max_x0"]
    assert lo < hi, (first, last, probe)
    return lo + where * (hi - lo)


def _single(model, host, times, p, job, **kw):
    from kpolaris_amd import image
    t, first, last = job
    out, _, _ = image.run_grmhd_slow_light_stream(p, model, times[first:last + 1], t,
                                                  image.host_cell_loader(host[first:last + 1]), **kw)
    return _np(out)


@pytest.mark.parametrize("mode,wpb,precision", [(0, 2, abi.PRECISION_FP64), (1, 3, abi.PRECISION_FP64),
                                                (0, 1, abi.PRECISION_MIXED)])
def test_gpu_jobs_equal_single_observation_runs(seq, mode, wpb, precision):
    from kpolaris_amd import image
    model, host, times, p, probe = seq
    supports = ((0, 15, 0.5), (3, 15, 0.2), (0, 11, 0.8), (2, 13, 0.5))
    jobs = [(_t_obs(times, probe, a, b, w), a, b) for a, b, w in supports]
    print(f"case: observation jobs mode {mode} wpb {wpb} precision {precision} jobs {jobs}", flush=True)
    outs, st, rep = image.run_grmhd_slow_light_stream_jobs(p, model, times, jobs, image.host_cell_loader(host),
                                                           step_mode=mode, windows_per_block=wpb,
                                                           precision=precision)
    print(rep, flush=True)
    npix = p.camera.nx * p.camera.ny
    assert st["rays"] == npix * len(jobs)
    for j, job in enumerate(jobs):
        got = _np(outs[j])
        ref = _single(model, host, times, p, job, step_mode=mode, windows_per_block=wpb, precision=precision)
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), (j, k)
        assert np.sum(got["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * npix
    # the observation times differ, so do the images
    assert not np.array_equal(_np(outs[0])["image_i"], _np(outs[1])["image_i"])
    if rep["rewinds"] == 0:
        assert rep["loads"] == len(times)


def test_gpu_jobs_block_stepping(seq):
    """Block stepping: the jobs starting at snapshot 0 equal their single
    runs over 0..last in the same blocks."""
    from kpolaris_amd import image
    model, host, times, p, probe = seq
    jobs = [(_t_obs(times, probe, 0, 15), 0, 15), (_t_obs(times, probe, 0, 12, 0.3), 0, 12),
            (_t_obs(times, probe, 4, 15), 4, 15)]
    print(f"case: observation jobs block stepping {jobs}", flush=True)
    outs, _, rep = image.run_grmhd_slow_light_stream_jobs(p, model, times, jobs, image.host_cell_loader(host),
                                                          step_mode=abi.SLOW_LIGHT_BLOCK, windows_per_block=2)
    print(rep, flush=True)
    for j in (0, 1):
        got = _np(outs[j])
        ref = _single(model, host, times, p, jobs[j], step_mode=abi.SLOW_LIGHT_BLOCK, windows_per_block=2)
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), (j, k)
    assert np.sum(_np(outs[2])["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * p.camera.nx * p.camera.ny


def test_gpu_jobs_option_rules(seq):
    """No jobs, a non-finite time, first >= last, last >= nsnap or a negative
    first: KP_INVALID_ARGUMENT; a job whose decoupled samples fall outside
    its own support: KP_TIME_SUPPORT, naming that job's observation time."""
    from kpolaris_amd import image
    model, host, times, p, probe = seq
    loader = image.host_cell_loader(host)
    t = _t_obs(times, probe, 0, 15)
    print("case: observation jobs option rules", flush=True)
    for jobs in ([], [(float("nan"), 0, 15)], [(t, 5, 5)], [(t, 6, 2)], [(t, 0, 16)], [(t, -1, 15)]):
        with pytest.raises(image.KPError) as e:
            image.run_grmhd_slow_light_stream_jobs(p, model, times, jobs, loader)
        assert e.value.status == abi.KP_INVALID_ARGUMENT, jobs
    late = times[8] - probe["max_x0"] + 5.0
    with pytest.raises(image.KPError, match="observation time") as e:
        image.run_grmhd_slow_light_stream_jobs(p, model, times, [(t, 0, 15), (late, 0, 8)], loader,
                                               windows_per_block=2)
    assert e.value.status == abi.KP_TIME_SUPPORT


@pytest.mark.parametrize("wpb", [0, 2])
def test_gpu_resident_jobs_equal_streamed_jobs(seq, wpb):
    """kp_run_grmhd_slow_light_image_jobs over resident device snapshots (one
    block, or blocks of two pairs): each job equals the streamed jobs' image
    and the resident single-observation-time run over its own snapshots."""
    from kpolaris_amd import image
    model, host, times, p, probe = seq
    grids = [image.GRMHDGrid(model, c) for c in host]
    jobs = [(_t_obs(times, probe, 0, 15), 0, 15), (_t_obs(times, probe, 3, 14, 0.3), 3, 14)]
    print(f"case: resident observation jobs windows_per_block {wpb} {jobs}", flush=True)
    outs, st = image.run_grmhd_slow_light_image_jobs(p, grids, times, jobs, windows_per_block=wpb)
    assert st["rays"] == 2 * p.camera.nx * p.camera.ny
    streamed, _, _ = image.run_grmhd_slow_light_stream_jobs(p, model, times, jobs, image.host_cell_loader(host),
                                                            windows_per_block=max(wpb, 1))
    for j, (t, first, last) in enumerate(jobs):
        got = _np(outs[j])
        single, _ = image.run_grmhd_slow_light_image(p, grids[first:last + 1], times[first:last + 1], t)
        ref = _np(single)
        for k in KEYS:
            assert np.array_equal(got[k], ref[k]), (j, k)
            assert np.array_equal(got[k], _np(streamed[j])[k]), (j, k)
    with pytest.raises(image.KPError) as e:
        image.run_grmhd_slow_light_image_jobs(p, grids, times, [(jobs[0][0], 0, 16)])
    assert e.value.status == abi.KP_INVALID_ARGUMENT
