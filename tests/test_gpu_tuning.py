"""Launch-configuration tuning (kp_tune_riaf_launch) on the device: it times
every candidate, picks one of them, and an image run with the chosen shape is
bitwise the default-shape image (the launch shape never changes results)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu
PARAMS = Path(__file__).resolve().parent.parent / "kpolaris_amd" / "params"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_tuned_launch_is_a_candidate_and_changes_no_result(precision):
    import torch
    from kpolaris_amd import abi, image
    print(f"case: launch tuning 96x96 RIAF precision {precision}", flush=True)
    o = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    o.set("nx", 96)
    o.set("ny", 96)
    p, m, _, _ = o.resolve_riaf()
    prec = abi.PRECISION_FP64 if precision == 0 else abi.PRECISION_MIXED
    t = image.tune_riaf_launch(p, m, precision=prec, sample_divisor=2)
    print(t, flush=True)
    assert t["ncandidates"] == 9 and t["sample_pixels"] == 48 * 48
    assert t["block_size"] in (64, 128, 256) and t["waves_per_cu"] in (0, 2, 8)
    assert 0.0 < t["best_ms"] <= t["default_ms"]
    a, _ = image.run_riaf_image(p, m, precision=prec)
    b, _ = image.run_riaf_image(p, m, precision=prec, block_size=t["block_size"], waves_per_cu=t["waves_per_cu"])
    torch.cuda.synchronize()
    for k in ("image_i", "image_q", "image_u", "image_v", "steps", "reason", "pass_a_steps"):
        assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy()), k


def test_gpu_tuning_rejects_bad_arguments():
    from kpolaris_amd import abi, image
    o = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    p, m, _, _ = o.resolve_riaf()
    with pytest.raises(image.KPError) as e:
        image.tune_riaf_launch(p, m, sample_divisor=0)
    assert e.value.status == abi.KP_INVALID_ARGUMENT
