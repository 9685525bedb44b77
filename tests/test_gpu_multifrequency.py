"""Multi-frequency HIP pass B (one geodesic per pixel, nfreq Stokes vectors in
LDS) against the oracle's multi-frequency image, through the C ABI.

FP64: every frequency slice at the parity bar of tests/parity.py, with the
geodesic's conditioning taken from the single-frequency oracle.  Mixed: every
slice within the mixed-precision image bars.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

PARAMS = Path(__file__).resolve().parent.parent / "kpolaris_amd" / "params"
FREQS = [86e9, 230e9, 345e9]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _opts(**kw):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    for k, v in kw.items():
        o.set(k, str(v))
    return o.resolve_riaf()


def _run(p, m, freqs, ctrl=None, **kw):
    import torch
    from kpolaris_amd import image
    out, st = image.run_riaf_multifrequency_image(p, m, freqs, ctrl, **kw)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}, st


@pytest.mark.parametrize("persistent", [True, False])
def test_gpu_multifrequency_fp64_parity(persistent):
    import oracle
    print(f"case: multifrequency fp64 24x24 3 freqs persistent={persistent}", flush=True)
    p, m, _, _ = _opts(nx=24, ny=24, max_absorption_depth=1e-3, max_faraday_depth=1e-2)
    got, _ = _run(p, m, FREQS, persistent=persistent)
    ref, _ = oracle.run_riaf_multifrequency_image(p, m, FREQS)
    single = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity_multifrequency(got, ref, 3, single, "gpu mf"), flush=True)


def test_gpu_multifrequency_separate_control_and_single_list():
    import oracle
    from kpolaris_amd import image
    print("case: multifrequency control list [86 GHz], and a one-frequency list", flush=True)
    p, m, _, _ = _opts(nx=16, ny=16, max_absorption_depth=1e-3, max_faraday_depth=1e-2)
    got, _ = _run(p, m, FREQS, [86e9])
    ref, _ = oracle.run_riaf_multifrequency_image(p, m, FREQS, [86e9])
    single = parity.oracle_with_sensitivity(p, m)
    parity.assert_parity_multifrequency(got, ref, 3, single, "gpu mf ctrl")
    one, _ = _run(p, m, [m.freq_cgs])
    plain, _ = image.run_riaf_image(p, m)
    plain = {k: v.cpu().numpy() for k, v in plain.items()}
    # one output frequency runs the multi-frequency kernel, whose coefficient
    # code is split differently from the single-frequency kernel's, so the
    # compiler may fuse multiply-adds differently: identical step sequences
    # and termination, Stokes equal to rounding
    for k in plain:
        if k.startswith("image_"):
            peak = max(float(np.max(np.abs(plain["image_i"]))), 1e-300)
            assert np.max(np.abs(one[k] - plain[k])) <= 1e-12 * peak, k
        elif k in ("reason", "steps", "pass_a_steps"):
            assert np.array_equal(one[k], plain[k]), k
        else:
            assert np.allclose(one[k], plain[k], rtol=1e-12, atol=1e-12), k


def test_gpu_multifrequency_mixed():
    import oracle
    from kpolaris_amd import abi
    print("case: multifrequency mixed 32x32 3 freqs", flush=True)
    p, m, _, _ = _opts(nx=32, ny=32)
    got, st = _run(p, m, FREQS, precision=abi.PRECISION_MIXED)
    ref, _ = oracle.run_riaf_multifrequency_image(p, m, FREQS)
    for f, (g, r) in enumerate(zip(parity.split_frequencies(got, 3), parity.split_frequencies(ref, 3))):
        print(parity.assert_mixed(g, r, f"gpu mixed mf f{f}"), flush=True)


def test_gpu_multifrequency_sixteen_frequencies_and_errors():
    """The largest list (LDS-resident Stokes for 16 frequencies); invalid lists fail loudly."""
    import oracle
    from kpolaris_amd import image
    print("case: multifrequency 16 freqs 8x8", flush=True)
    p, m, _, _ = _opts(nx=8, ny=8)
    freqs = list(np.geomspace(86e9, 345e9, 16))
    got, _ = _run(p, m, freqs)
    ref, _ = oracle.run_riaf_multifrequency_image(p, m, freqs)
    single = parity.oracle_with_sensitivity(p, m)
    parity.assert_parity_multifrequency(got, ref, 16, single, "gpu mf16")
    with pytest.raises(image.KPError):
        image.run_riaf_multifrequency_image(p, m, list(np.geomspace(86e9, 345e9, 17)))
    with pytest.raises(image.KPError):
        image.run_riaf_multifrequency_image(p, m, [230e9, -1.0])
    with pytest.raises(image.KPError):
        image.run_riaf_multifrequency_image(p, m, [230e9], [])
