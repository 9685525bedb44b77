"""Slow light in blocks on the GPU: the batch scheduler's block loop
(src/image/slow_light_batch_driver.hpp:211-340) over resident snapshots and
over a sequence streamed from host memory or iHARM files through the
read-ahead thread and the device ring of snapshot slots
(src/image/slow_light_prefetch.hpp:15-88).

Decoupled and snapshot images do not depend on the block length, so every
blocked run equals the single-block image bit for bit; block stepping's image
does, and is held to the oracle's image of the same blocks.  Streaming equals
the resident run of the same blocks bit for bit, also when the ring is
shorter than the sequence and slots are reused."""
from __future__ import annotations

import numpy as np
import pytest

from kpolaris_amd import abi
from tests import parity

pytestmark = pytest.mark.gpu

T_OBS = 275.0
KEYS = ("image_i", "image_q", "image_u", "image_v", "steps", "reason", "pass_a_steps", "closure_x", "final_null")


@pytest.fixture(scope="module")
def seq():
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32)
    grids = [image.GRMHDGrid(model, c) for c in cells]
    host = [g.host_cells() for g in grids]
    p = gs.pass_a_params(model, nx=16, ny=16, camera_radius=200.0, outer_radius=30.0)
    return grids, host, cells, times, model, p


def _np(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _resident(grids, times, p, t_obs=T_OBS, **kw):
    from kpolaris_amd import image
    out, st = image.run_grmhd_slow_light_image(p, grids, times, t_obs, **kw)
    return _np(out), st


def _stream(model, times, p, loader, t_obs=T_OBS, **kw):
    from kpolaris_amd import image
    out, st, rep = image.run_grmhd_slow_light_stream(p, model, times, t_obs, loader, **kw)
    return _np(out), st, rep


def _same(a, b, label):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), (label, k)


@pytest.mark.parametrize("mode,wpb", [(0, 1), (0, 2), (1, 2)])
def test_gpu_resident_blocks_equal_the_single_block(seq, mode, wpb):
    grids, host, cells, times, model, p = seq
    print(f"case: resident blocks mode {mode} windows_per_block {wpb}", flush=True)
    one, _ = _resident(grids, times, p, step_mode=mode)
    got, st = _resident(grids, times, p, step_mode=mode, windows_per_block=wpb)
    _same(got, one, f"mode {mode} wpb {wpb}")
    assert np.sum(got["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * got["reason"].size


def test_gpu_block_stepping_in_blocks_matches_oracle(seq):
    """Block stepping in two-pair blocks against the oracle's two-pair
    blocks; in one-pair blocks it is snapshot stepping, bit for bit."""
    import oracle
    grids, host, cells, times, model, p = seq
    print("case: block stepping two-pair blocks", flush=True)
    got, _ = _resident(grids, times, p, step_mode=2, windows_per_block=2)
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(  # noqa: E731
        q, model, host, times, t, step_mode=2, windows_per_block=2)[0]
    ref = run(p)
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got, ref, 1, "gpu block stepping two-pair blocks", cond, time_cut_steps=1),
          flush=True)
    one, _ = _resident(grids, times, p, step_mode=2, windows_per_block=1)
    snap, _ = _resident(grids, times, p, step_mode=1)
    _same(one, snap, "block stepping one-pair blocks vs snapshot stepping")


@pytest.mark.parametrize("mode,wpb,prefetch,depth", [(0, 1, True, 0), (0, 2, False, 0), (2, 2, True, 1),
                                                     (1, 1, True, 3)])
def test_gpu_streamed_equals_resident_blocks(seq, mode, wpb, prefetch, depth):
    """Host snapshots streamed through the loader equal the resident run of
    the same blocks; the schedule reports one read per snapshot."""
    from kpolaris_amd import image
    grids, host, cells, times, model, p = seq
    print(f"case: streamed mode {mode} wpb {wpb} prefetch {prefetch} depth {depth}", flush=True)
    ref, _ = _resident(grids, times, p, step_mode=mode, windows_per_block=wpb)
    loader = image.host_cell_loader(host)
    got, st, rep = _stream(grids[0].model, times, p, loader, step_mode=mode, windows_per_block=wpb,
                           prefetch=prefetch, prefetch_snapshots=depth)
    _same(got, ref, f"streamed mode {mode} wpb {wpb}")
    halo = 4 if mode == 0 else 0
    assert rep["slots"] == min(2 * wpb + halo + 1, len(times))
    assert rep["slot_bytes"] == rep["slots"] * model.n1 * model.n2 * model.n3 * 72
    assert rep["loads"] == len(times) and rep["rewinds"] == 0
    assert rep["blocks"] == -(-(len(times) - 1) // wpb)
    assert rep["pass_b_ms"] > 0.0 and st["pass_b_ms"] == pytest.approx(rep["pass_b_ms"])
    print(rep, flush=True)


def test_gpu_streamed_primitives_derived_on_device(seq):
    """A loader that fills primitives only (derived scalars zero) with
    derive_scalars = 1 gives the image of the device-derived resident grids."""
    from kpolaris_amd import image
    grids, host, cells, times, model, p = seq
    print("case: streamed primitives, derived on the device", flush=True)
    prims = []
    for c in cells:
        c2 = c.copy()
        c2["derived"] = 0.0
        prims.append(c2)
    ref, _ = _resident(grids, times, p, windows_per_block=2)
    got, _, rep = _stream(grids[0].model, times, p, image.host_cell_loader(prims), windows_per_block=2,
                          derive_scalars=True)
    _same(got, ref, "derived on the device")


def test_gpu_long_sequence_reuses_ring_slots():
    """24 snapshots through a ring of 9 (decoupled, two-pair blocks) and 7
    (block stepping, three-pair blocks) slots: the same image as the resident
    sequence run in the same blocks."""
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=24, dt=5.0, t0=-10.0, n1=48, n2=32, n3=32)
    grids = [image.GRMHDGrid(model, c) for c in cells]
    host = [g.host_cells() for g in grids]
    p = gs.pass_a_params(model, nx=24, ny=24, camera_radius=200.0, outer_radius=30.0)
    for mode, wpb, slots in ((0, 2, 9), (2, 3, 7)):
        print(f"case: 24 snapshots mode {mode} wpb {wpb}", flush=True)
        ref, _ = _resident(grids, times, p, step_mode=mode, windows_per_block=wpb)
        if mode == 0:
            single, _ = _resident(grids, times, p, step_mode=mode)
            _same(ref, single, "decoupled blocks vs single block")
        got, _, rep = _stream(grids[0].model, times, p, image.host_cell_loader(host), step_mode=mode,
                              windows_per_block=wpb)
        _same(got, ref, f"24 snapshots mode {mode}")
        assert rep["slots"] == slots and rep["loads"] == len(times)
        assert np.sum(got["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * got["reason"].size
        print(rep, flush=True)


def test_gpu_streamed_iharm_dumps(seq, tmp_path):
    """iHARM dumps on disk streamed by the C loader kp_iharm_load_snapshot
    (primitives read as float, derived on the device) equal the resident
    grids built from the same files."""
    from kpolaris_amd import image, io
    grids, host, cells, times, model, p = seq
    print("case: streamed iHARM dumps", flush=True)
    paths = []
    for k, c in enumerate(cells):
        path = tmp_path / f"dump_{k:03d}.h5"
        io.write_iharm(path, model, c)
        paths.append(path)
    files = []
    for path in paths:
        _, c, m = io.read_iharm(path, model)
        files.append(image.GRMHDGrid(model, c))
    ref, _ = _resident(files, times, p, windows_per_block=2)
    stream = io.IharmStream(paths)
    got, _, rep = _stream(files[0].model, times, p, stream.loader, windows_per_block=2, derive_scalars=True,
                          user=stream.user)
    _same(got, ref, "iHARM stream")
    assert rep["loads"] == len(paths)


def test_gpu_streamed_errors(seq):
    """A failing loader ends the run with its status and the snapshot named;
    decoupled samples outside the sequence report KP_TIME_SUPPORT with the
    image written, as the resident run does; invalid residency options are
    rejected (slow_light_batch_driver.hpp:138-139)."""
    from kpolaris_amd import abi, image
    grids, host, cells, times, model, p = seq
    print("case: streamed errors", flush=True)
    good = image.host_cell_loader(host)

    def failing(user, index, out, n):
        if index == 3:
            return abi.KP_INVALID_ARGUMENT
        return good(user, index, out, n)

    bad = abi.SnapshotLoader(failing)
    with pytest.raises(image.KPError, match="snapshot 3"):
        _stream(grids[0].model, times, p, bad, windows_per_block=1)
    with pytest.raises(image.KPError, match="windows_per_block"):
        _stream(grids[0].model, times, p, good, windows_per_block=0)
    out_r = image.alloc_device_image(p.camera.nx * p.camera.ny)
    with pytest.raises(image.KPError) as e1:
        image.run_grmhd_slow_light_image(p, grids, times, 230.0, out=out_r)
    out_s = image.alloc_device_image(p.camera.nx * p.camera.ny)
    with pytest.raises(image.KPError) as e2:
        image.run_grmhd_slow_light_stream(p, grids[0].model, times, 230.0, good, windows_per_block=2, out=out_s)
    assert e1.value.status == e2.value.status == abi.KP_TIME_SUPPORT
    _same(_np(out_s), _np(out_r), "time support")
    # the library stays usable after the failures
    ok, _, _ = _stream(grids[0].model, times, p, good, windows_per_block=2)
    assert np.sum(ok["reason"] == abi.REASON_REACHED_CAMERA) > 0.9 * ok["reason"].size
