"""Re-entrancy of the C ABI across streams and caller-owned workspaces.

The reference runs batch jobs on partitioned execution instances
(src/image/batch_driver.hpp:147) and its drivers allocate nothing the caller
can see; here a run is stream-ordered, each (device, stream) pair has its own
library workspace, and a caller may pass its own workspace
(kp_run_options.workspace, sized by kp_image_workspace_size).  Two images
launched concurrently on two streams must equal the same images run one after
the other, bit for bit; so must a run on a caller workspace.  The plain-C
caller of include/kp_image.h (tests/c_caller) must reproduce the Python
binding's image."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
PAR = ROOT / "kpolaris_amd" / "params" / "riaf_m87_c2.par"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _job(nx, incl):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(PAR))
    o.set("nx", nx)
    o.set("ny", nx)
    o.set("inclination_deg", incl)
    return o.resolve_riaf()[:2]


def _host(out):
    return {k: v.cpu().numpy() for k, v in out.items()}


def test_gpu_two_streams_equal_serial():
    import torch
    from kpolaris_amd import abi, image
    print("case: two images on two streams vs serial (fp64 48x48, mixed 40x40)", flush=True)
    ja, jb = _job(48, 163.0), _job(40, 60.0)
    serial_a = _host(image.run_riaf_image(*ja, with_stats=False)[0])
    serial_b = _host(image.run_riaf_image(*jb, precision=abi.PRECISION_MIXED, with_stats=False)[0])
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        with torch.cuda.stream(sa):
            oa, _ = image.run_riaf_image(*ja, with_stats=False)
        with torch.cuda.stream(sb):
            ob, _ = image.run_riaf_image(*jb, precision=abi.PRECISION_MIXED, with_stats=False)
        torch.cuda.synchronize()
        a, b = _host(oa), _host(ob)
        for k in serial_a:
            assert np.array_equal(a[k], serial_a[k], equal_nan=True), (rep, "a", k)
            assert np.array_equal(b[k], serial_b[k], equal_nan=True), (rep, "b", k)


def test_gpu_caller_workspace():
    import torch
    from kpolaris_amd import abi, image
    print("case: caller-owned workspace 32x32, fp64 and staged mixed", flush=True)
    p, m = _job(32, 163.0)
    npix = 32 * 32
    for kw in ({}, {"precision": abi.PRECISION_MIXED, "guard_mode": 2}):
        ref = _host(image.run_riaf_image(p, m, with_stats=False, **kw)[0])
        nbytes = image.workspace_size(npix, **kw)
        assert nbytes >= npix * (16 * 8 + 12)
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        got = _host(image.run_riaf_image(p, m, with_stats=False, workspace=ws, **kw)[0])
        for k in ref:
            assert np.array_equal(got[k], ref[k], equal_nan=True), (kw, k)
        small = torch.empty(nbytes - 256, dtype=torch.uint8, device="cuda")
        with pytest.raises(image.KPError) as e:
            image.run_riaf_image(p, m, with_stats=False, workspace=small, **kw)
        assert e.value.status == abi.KP_INVALID_ARGUMENT


def test_gpu_plain_c_caller():
    import torch
    from kpolaris_amd import image
    from tests.c_caller import build as cb
    print("case: plain-C caller 24x24", flush=True)
    assert cb.EXE.exists(), "build() compiles tests/c_caller"
    res = subprocess.run([str(cb.EXE), str(PAR), "nx=24", "ny=24"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    vals = res.stdout.split()
    got = {vals[i]: float(vals[i + 1]) for i in range(0, len(vals), 2)}
    o = image.Options(parameter_file=str(PAR))
    o.set("nx", 24)
    o.set("ny", 24)
    p, m, scale, sign = o.resolve_riaf()
    out, _ = image.run_riaf_image(p, m, with_stats=False)
    torch.cuda.synchronize()
    h = _host(out)
    assert got["npix"] == 576
    assert got["camera"] == int(np.sum(h["reason"] == 1))
    assert got["steps"] == int(np.sum(h["steps"]))
    total = 0.0
    for v in h["image_i"].tolist():  # the caller's summation order
        total += v
    assert got["sum_i"] == total
