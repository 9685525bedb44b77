"""Batched images (kp_batch_*): many uneven RIAF jobs in one persistent
two-pass HIP graph, against the oracle per job (FP64 at the parity bar) and
against single-image mixed runs (mixed bars); graph replays interleaved with
eager work must all reproduce the eager result."""
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


def test_gpu_batch_fp64_parity_uneven_jobs():
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
        ref = parity.oracle_with_sensitivity(p, m)
        print(j, parity.assert_parity(_host(b.outputs[j]), ref, f"batch job {j}"), flush=True)
    b.close()


def test_gpu_batch_graph_replays_and_mixed():
    import torch
    from kpolaris_amd import abi, batch_jobs, image
    print("case: batch mixed, 3 graph replays with eager work between", flush=True)
    jobs = batch_jobs.jobs(seed=11, n=5, sizes=(16, 32))
    b = image.Batch(jobs, precision=abi.PRECISION_MIXED)
    results = []
    for rep in range(3):
        b.launch()
        torch.cuda.synchronize()
        results.append([_host(o) for o in b.outputs])
        # eager work between replays: a single-image run on the same device
        image.run_riaf_image(*jobs[0], precision=abi.PRECISION_MIXED)
        torch.cuda.synchronize()
        for o in b.outputs:
            for t in o.values():
                t.zero_()
        torch.cuda.synchronize()
    for rep in range(1, 3):
        for j in range(len(jobs)):
            for k in results[0][j]:
                assert np.array_equal(results[rep][j][k], results[0][j][k]), (rep, j, k)
    for j, (p, m) in enumerate(jobs):
        single, _ = image.run_riaf_image(p, m, precision=abi.PRECISION_MIXED)
        parity.assert_mixed(results[0][j], _host(single), f"batch mixed job {j}")
    b.close()


def test_gpu_batch_invalid_job_fails_before_launch():
    from kpolaris_amd import abi, batch_jobs, image
    jobs = batch_jobs.jobs(seed=3, n=2, sizes=(8,))
    jobs[1][0].coordinate_system = abi.COORD_MKS
    with pytest.raises(image.KPError):
        image.Batch(jobs)
