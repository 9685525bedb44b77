"""The allocation-free entry point kp_run_riaf_image_noalloc (no reference
counterpart; for callers that cannot allocate device memory): through the C
ABI, two calls in a row, the second on a smaller image, each equal to
kp_run_riaf_image's image of the same inputs bit for bit (the same kernels run
on the library's static scratch array instead of a workspace), in FP64 and
mixed precision; images beyond the maximum, staged mixed precision and a
caller workspace are rejected."""
from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _job(nx, ny, incl=163.0):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", ny)
    o.set("inclination_deg", incl)
    return o.resolve_riaf()[:2]


def test_noalloc_equals_allocating_entry_point_twice():
    import torch
    from kpolaris_amd import abi, image
    for prec in (abi.PRECISION_FP64, abi.PRECISION_MIXED):
        for nx, ny, incl in ((96, 80, 163.0), (40, 36, 60.0)):
            print(f"case: noalloc precision {prec} {nx}x{ny}", flush=True)
            p, m = _job(nx, ny, incl)
            out = image.alloc_device_image(nx * ny)
            for t in out.values():
                t.fill_(7)  # the entry point writes every output
            image.run_riaf_image_noalloc(p, m, out, precision=prec)
            ref, _ = image.run_riaf_image(p, m, precision=prec, with_stats=False)
            torch.cuda.synchronize()
            for k in ref:
                a, b = out[k].cpu().numpy(), ref[k].cpu().numpy()
                assert np.array_equal(a, b, equal_nan=True), (prec, nx, k)
            assert (out["reason"].cpu().numpy() == 1).mean() > 0.9


def test_noalloc_with_selection_equals_allocating_entry_point():
    """The equatorial emission wedge and direct_only (the wedge kernels on the
    scratch array): equal to the allocating entry point's image bit for bit."""
    import torch
    from kpolaris_amd import abi, image
    for prec in (abi.PRECISION_FP64, abi.PRECISION_MIXED):
        print(f"case: noalloc wedge 0.1 + direct_only precision {prec} 48x40", flush=True)
        p, m = _job(48, 40)
        p.emission_selection.equatorial_h_over_r = 0.1
        p.direct_only = 1
        out = image.alloc_device_image(48 * 40)
        for t in out.values():
            t.fill_(7)
        image.run_riaf_image_noalloc(p, m, out, precision=prec)
        ref, _ = image.run_riaf_image(p, m, precision=prec, with_stats=False)
        torch.cuda.synchronize()
        for k in ref:
            assert np.array_equal(out[k].cpu().numpy(), ref[k].cpu().numpy(), equal_nan=True), (prec, k)


def test_noalloc_rejections():
    from kpolaris_amd import abi, image
    n = int(abi.NOALLOC_MAX_PIXELS ** 0.5) + 1
    p, m = _job(n, n)
    out = image.alloc_device_image(1)  # never written: rejected before any launch
    with pytest.raises(image.KPError) as e:
        image.run_riaf_image_noalloc(p, m, out)
    assert e.value.status == abi.KP_INVALID_ARGUMENT and "allocation-free maximum" in str(e.value)
    p, m = _job(8, 8)
    with pytest.raises(image.KPError) as e:
        image.run_riaf_image_noalloc(p, m, out, precision=abi.PRECISION_MIXED, guard_mode=2)
    assert e.value.status == abi.KP_UNSUPPORTED
