"""Batched images (kp_batch_*, kp_run_riaf_batch_host): many uneven RIAF jobs
in one persistent two-pass HIP graph.  The reference's contract is that a
batch equals serial dispatch bit for bit (tests/image_batch_checks.hpp:6-57);
here a batched job equals the single image of the same job and options, in
FP64 and in every mixed-precision mode, and FP64 jobs match the oracle.
The host-output entry runs the reference's window and budget contract
(tests/test_image_batch.cpp:54-76)."""
from __future__ import annotations

import numpy as np
import pytest

from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _host(out):
    return {k: v.cpu().numpy() for k, v in out.items()}


def _same_image(a, b, label):
    """require_same_image (tests/image_batch_checks.hpp:6-57): every field
    equal, every real field finite."""
    for k in a:
        assert np.array_equal(a[k], b[k]), (label, k, np.flatnonzero(a[k] != b[k])[:5])
        if a[k].dtype.kind == "f":
            assert np.all(np.isfinite(b[k])), (label, k)


def _single(p, m, **kw):
    import torch
    from kpolaris_amd import image
    out, _ = image.run_riaf_image(p, m, **kw)
    torch.cuda.synchronize()
    return _host(out)


def test_gpu_batch_fp64_equals_serial_and_oracle():
    import torch
    from kpolaris_amd import batch_jobs, image
    print("case: batch of 6 uneven jobs (8..24 px) fp64", flush=True)
    jobs = batch_jobs.jobs(seed=7, n=6, sizes=(8, 12, 16, 24))
    b = image.Batch(jobs)
    b.launch()
    torch.cuda.synchronize()
    st = b.stats()
    assert st["rays"] == sum(p.camera.nx * p.camera.ny for p, _ in jobs)
    for j, (p, m) in enumerate(jobs):
        got = _host(b.outputs[j])
        _same_image(_single(p, m), got, f"batch job {j}")
        ref = parity.oracle_with_sensitivity(p, m)
        print(j, parity.assert_parity(got, ref, f"batch job {j}"), flush=True)
    b.close()


@pytest.mark.parametrize("guard", [0, 1])
def test_gpu_batch_mixed_equals_serial_and_replays(guard):
    """Mixed batches step the geodesics as single images with the same
    guard_mode do; three graph replays with eager work between them all
    reproduce the eager single images."""
    import torch
    from kpolaris_amd import abi, batch_jobs, image
    print(f"case: batch mixed guard_mode {guard}, 3 graph replays with eager work between", flush=True)
    jobs = batch_jobs.jobs(seed=11, n=5, sizes=(16, 32))
    b = image.Batch(jobs, precision=abi.PRECISION_MIXED, guard_mode=guard)
    singles = [_single(p, m, precision=abi.PRECISION_MIXED, guard_mode=guard) for p, m in jobs]
    for rep in range(3):
        b.launch()
        torch.cuda.synchronize()
        for j in range(len(jobs)):
            _same_image(singles[j], _host(b.outputs[j]), f"replay {rep} job {j}")
        image.run_riaf_image(*jobs[0], precision=abi.PRECISION_MIXED)
        torch.cuda.synchronize()
        for o in b.outputs:
            for t in o.values():
                t.zero_()
        torch.cuda.synchronize()
    for j, (p, m) in enumerate(jobs):
        parity.assert_mixed(singles[j], _single(p, m), f"batch mixed job {j}", pol="absolute")
    b.close()


def test_gpu_batch_rejections():
    from kpolaris_amd import abi, batch_jobs, image
    jobs = batch_jobs.jobs(seed=3, n=2, sizes=(8,))
    with pytest.raises(image.KPError) as e:
        image.Batch(jobs, precision=abi.PRECISION_MIXED, guard_mode=2)
    assert e.value.status == abi.KP_UNSUPPORTED
    with pytest.raises(image.KPError) as e:
        image.Batch(jobs, precision=abi.PRECISION_MIXED, guard_mode=3)
    assert e.value.status == abi.KP_INVALID_ARGUMENT
    jobs[1][0].coordinate_system = abi.COORD_MKS
    with pytest.raises(image.KPError):
        image.Batch(jobs)


def test_gpu_batch_host_windows_and_budget():
    """tests/test_image_batch.cpp:28-76: five jobs of 2x2 and 3x2 pixels with
    at most 3 in flight and a budget of 10 pixels' outputs run in 3 windows
    within the budget and equal serial dispatch; an empty batch is valid;
    zero limits, a budget below one image, an overflowing or empty image and
    radiation_substeps = 0 are rejected before any work."""
    from kpolaris_amd import abi, image
    print("case: host batch windows 5 jobs, max_in_flight 3, budget 10 px", flush=True)
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    jobs = []
    for j in range(5):
        o.set("nx", 2 + j % 2)
        o.set("ny", 2)
        o.set("inclination_deg", 40.0 + 6.0 * j)
        jobs.append(o.resolve_riaf()[:2])
    budget = 10 * abi.BATCH_BYTES_PER_PIXEL
    outs, t = image.run_riaf_batch_host(jobs, max_in_flight=3, max_device_output_bytes=budget)
    print(t, flush=True)
    assert t["windows"] == 3 and 0 < t["peak_device_output_bytes"] <= budget
    for j, (p, m) in enumerate(jobs):
        _same_image(_single(p, m), outs[j], f"host batch job {j}")
    outs0, t0 = image.run_riaf_batch_host([])
    assert outs0 == [] and t0["windows"] == 0

    def rejects(js, **kw):
        with pytest.raises(image.KPError) as e:
            image.run_riaf_batch_host(js, **kw)
        assert e.value.status == abi.KP_INVALID_ARGUMENT, e.value

    rejects(jobs, max_in_flight=0)
    rejects(jobs, max_in_flight=2, max_device_output_bytes=1)
    import copy
    bad = copy.deepcopy(jobs)
    bad[-1][0].camera.nx = 2 ** 31 - 1
    rejects(bad)
    bad[-1][0].camera.nx = 0
    rejects(bad)
    bad[-1][0].camera.nx = 2
    rejects(bad, radiation_substeps=[1, 1, 1, 1, 0])


def test_gpu_batch_64_uneven_jobs_one_graph():
    """The 64-job draw (seed 20261202: 128^2..1024^2, spins {0, 0.5, 0.94},
    inclinations 10-90 deg, 86/230/345 GHz, density scales 0.3-3) in one
    graph launch per precision: mixed against the FP64 batch at the image
    bars for every job, and the FP64 batch against the oracle on a strided
    sample of 12 pixels per job."""
    import torch
    from kpolaris_amd import abi, batch_jobs, image
    print("case: 64-job batch, one graph launch, fp64 and mixed", flush=True)
    jobs = batch_jobs.jobs(seed=20261202, n=64)
    rays = sum(p.camera.nx * p.camera.ny for p, _ in jobs)
    fields = ("image_i", "image_q", "image_u", "image_v", "pass_a_steps", "steps", "reason")
    res = {}
    for prec in (abi.PRECISION_FP64, abi.PRECISION_MIXED):
        b = image.Batch(jobs, precision=prec, fields=fields)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        b.launch()
        end.record()
        torch.cuda.synchronize()
        st = b.stats()
        assert st["rays"] == rays
        ms = start.elapsed_time(end)
        print(f"precision {prec}: {rays} rays in {ms:.1f} ms = {rays / ms * 1e3:.3e} rays/s", flush=True)
        res[prec] = [_host(o) for o in b.outputs]
        b.close()
        del b
        torch.cuda.empty_cache()
    worst = {"total_i_rel": 0.0}
    for j in range(len(jobs)):
        e = parity.assert_mixed(res[abi.PRECISION_MIXED][j], res[abi.PRECISION_FP64][j], f"job {j}")
        worst["total_i_rel"] = max(worst["total_i_rel"], e["total_i_rel"])
    print("mixed vs fp64, worst total-I error over jobs", worst, flush=True)
    ill = checked = 0
    for j, (p, m) in enumerate(jobs):
        npix = p.camera.nx * p.camera.ny
        stride = npix // 12 + 1
        offset = (37 * j) % stride
        count = len(range(offset, npix, stride))
        ref = parity.oracle_with_sensitivity(p, m, offset, stride, count)
        got = {k: res[abi.PRECISION_FP64][j][k][offset::stride][:count] for k in fields}
        peak = float(np.max(np.abs(res[abi.PRECISION_FP64][j]["image_i"])))
        # twelve pixels per job: the ill-conditioned share is bounded over all jobs
        r = parity.check(got, ref, ref["_cond"], f"64-job batch job {j} spot", diagnostics=False, max_ill=1.0,
                         peak=peak)
        ill += r["ill_conditioned"]
        checked += r["pixels"]
    print(f"spot parity: {ill} of {checked} pixels ill-conditioned", flush=True)
    assert ill <= parity.MAX_ILL * checked
