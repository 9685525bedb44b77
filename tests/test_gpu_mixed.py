"""Mixed-precision HIP kernel (FP64 ray state; FP32 derivatives, coefficients
and transfer; FP64 attempts inside fp64_radius) through the C ABI.

Mixed precision is held to image-level bars against the FP64 image of the
same inputs (tests/parity.py: total I within 1e-4, LP/CP fractions within 1e-3
relative, every pixel within 1e-2 of peak I, termination reasons matching on
all but 0.1% of pixels).  The FP64 kernel itself is pinned to the oracle by
test_gpu_riaf.py; the small cases here also compare with the oracle directly.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
PARAMS = ROOT / "kpolaris_amd" / "params"
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _options(par: str, **overrides):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(PARAMS / par))
    for k, v in overrides.items():
        o.set(k, v)
    return o


def _gpu_image(p, m, **kw):
    import torch
    from kpolaris_amd import image
    out, stats = image.run_riaf_image(p, m, **kw)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}, stats


def test_gpu_mixed_full_image_c2():
    """The bench workload (1024x1024 adaptive RIAF): mixed vs FP64 on the device."""
    from kpolaris_amd import abi
    print("case: mixed vs fp64, RIAF 1024x1024 adaptive", flush=True)
    p, m, _, _ = _options("riaf_m87_c2.par").resolve_riaf()
    ref, st64 = _gpu_image(p, m, precision=abi.PRECISION_FP64)
    got, stm = _gpu_image(p, m, precision=abi.PRECISION_MIXED)
    e = parity.assert_mixed(got, ref, "mixed 1024")
    print(e, "fp64 ms", st64["kernel_ms"], "mixed ms", stm["kernel_ms"], "guard", stm["fp64_guard_steps"],
          flush=True)
    assert stm["fp64_guard_steps"] > 0
    assert st64["fp64_guard_steps"] == 0


def test_gpu_mixed_golden_8x8():
    """The reference's 8x8 golden (tests/test_hdf5_contract.py:671-681) within the mixed bar."""
    from kpolaris_amd import abi, image
    print("case: mixed golden 8x8", flush=True)
    g = json.loads((GOLDEN / "riaf_bl_8x8_golden.json").read_text())
    o = image.Options()
    for kv in g["parameters"]:
        k, _, v = kv.partition("=")
        o.set(k, v)
    p, m, scale, sign = o.resolve_riaf()
    got, _ = _gpu_image(p, m, precision=abi.PRECISION_MIXED)
    flux = image.pack_image(got, m.freq_cgs, sign, scale)[4]
    assert math.isclose(flux[0], g["expected"]["flux_I_Jy"][0], rel_tol=parity.MIXED_TOTAL_I)
    import oracle
    ref, _ = oracle.run_riaf_image(p, m)
    parity.assert_mixed(got, ref, "mixed golden8x8", pol="absolute")


@pytest.mark.parametrize("persistent", [True, False])
def test_gpu_mixed_vs_oracle_adaptive(persistent):
    """40x40 adaptive RIAF against the FP64 oracle, both launch modes."""
    import oracle
    from kpolaris_amd import abi
    print(f"case: mixed vs oracle 40x40 persistent={persistent}", flush=True)
    p, m, _, _ = _options("riaf_m87_c2.par", nx=40, ny=40).resolve_riaf()
    got, _ = _gpu_image(p, m, precision=abi.PRECISION_MIXED, persistent=persistent)
    ref, _ = oracle.run_riaf_image(p, m)
    print(parity.assert_mixed(got, ref, "mixed40"), flush=True)


def test_gpu_mixed_dexter_guarded_and_unguarded():
    """Adaptive RK4 (rtol 1e-6), Dexter fit, Faraday rotation off.  The default
    mixed kernel meets the mixed bars; FP32 stepping with the guard off
    (guard_mode 1, fp64_radius 0: every attempt FP32-derived) runs no attempt
    in FP64 and keeps total I within the bar, the polarisation fractions
    within ten times it."""
    import oracle
    from kpolaris_amd import abi
    print("case: mixed dexter adaptive 16x12 guarded/unguarded", flush=True)
    p, m, _, _ = _options("riaf_m87_c2.par", nx=16, ny=12).resolve_riaf()
    m.emission_type = abi.EMISSION_THERMAL_DEXTER
    p.emission_selection.faraday_rotation = 0
    ref, _ = oracle.run_riaf_image(p, m)
    got, st = _gpu_image(p, m, precision=abi.PRECISION_MIXED)
    assert st["fp64_guard_steps"] > 0
    print(parity.assert_mixed(got, ref, "mixed dexter c2", pol="absolute"), flush=True)
    bare, st0 = _gpu_image(p, m, precision=abi.PRECISION_MIXED, fp64_radius=0.0, guard_mode=1)
    assert st0["fp64_guard_steps"] == 0
    e = parity.mixed_errors(bare, ref)
    print("unguarded", e, flush=True)
    assert e["total_i_rel"] <= parity.MIXED_TOTAL_I, e
    for key in ("lp", "cp"):
        assert abs(e[key][1] - e[key][0]) <= 10 * parity.MIXED_POL_FRAC * abs(e[key][0]) + 1e-6, (key, e)


def test_gpu_mixed_fixed_step_rk4():
    """Fixed-step RK4 (no error control) lies outside the mixed method, which
    leans on the step-doubling error estimate: the FP32 rounding of a large
    fixed step moves total I by about 2e-4 at 16x12 (the CPU emulation of
    the kernel gives the same).  Held to 1e-3 here, with termination
    reasons unchanged."""
    import oracle
    from kpolaris_amd import abi
    print("case: mixed fixed-step c1 16x12", flush=True)
    p, m, _, _ = _options("riaf_m87_c1.par", nx=16, ny=12).resolve_riaf()
    ref, _ = oracle.run_riaf_image(p, m)
    got, st = _gpu_image(p, m, precision=abi.PRECISION_MIXED)
    e = parity.mixed_errors(got, ref)
    print(e, flush=True)
    assert st["fp64_guard_steps"] > 0
    assert e["total_i_rel"] <= 10 * parity.MIXED_TOTAL_I, e
    assert e["reason_mismatch"] == 0, e
    assert e["pixel_max_rel_peak"] <= parity.MIXED_PIXEL, e


@pytest.mark.parametrize("guard_mode", [1, 2])
def test_gpu_mixed_guard_modes(guard_mode):
    """FP32 stepping with the per-lane FP64 guard in one kernel (1) and staged
    FP32/FP64 geodesic kernels (2) against the FP64 oracle at the mixed bars."""
    import oracle
    from kpolaris_amd import abi
    print(f"case: mixed guard_mode {guard_mode} 40x40", flush=True)
    p, m, _, _ = _options("riaf_m87_c2.par", nx=40, ny=40).resolve_riaf()
    got, st = _gpu_image(p, m, precision=abi.PRECISION_MIXED, guard_mode=guard_mode)
    ref, _ = oracle.run_riaf_image(p, m)
    print(parity.assert_mixed(got, ref, f"mixed guard_mode {guard_mode}"), st["fp64_guard_steps"], flush=True)
    assert st["fp64_guard_steps"] > 0


def test_gpu_mixed_turning_point_guard():
    """guard_mode 1 with a guard radius no ray reaches: the FP64 attempts are
    the turning-point retakes (dtheta/dlambda changes sign across an FP32
    attempt), a few per ray, fewer than with the default radius.  Total I and
    the pixels meet the mixed bars against the FP64 oracle; the LP and CP
    fractions of this 40x40 image are near-cancelling sums that any change of
    rounding along a few photon-ring rays moves by ~1e-3 relative (CPU lane
    emulation: 1.4e-5 with no FP64 guard, 2.6e-3 with the retakes alone,
    1.5e-4 with both guards), so they are held in absolute terms here; the
    default guards are held to the relative bars in test_gpu_mixed_guard_modes."""
    import oracle
    from kpolaris_amd import abi
    p, m, _, _ = _options("riaf_m87_c2.par", nx=40, ny=40).resolve_riaf()
    print("case: mixed guard_mode 1, turning-point guard only, 40x40", flush=True)
    ref, _ = oracle.run_riaf_image(p, m)
    got, st = _gpu_image(p, m, precision=abi.PRECISION_MIXED, guard_mode=1, fp64_radius=1e-9)
    _, st2 = _gpu_image(p, m, precision=abi.PRECISION_MIXED, guard_mode=1)
    print(parity.assert_mixed(got, ref, "mixed turning-point guard", pol="absolute"), st["fp64_guard_steps"],
          st2["fp64_guard_steps"], flush=True)
    assert 0 < st["fp64_guard_steps"] < 0.01 * st["rhs_evals"] / 11
    assert st["fp64_guard_steps"] < st2["fp64_guard_steps"]


def test_gpu_mixed_edge_shapes_and_repeat():
    """1x1 and ragged 7x3 images; a repeated run gives identical bits."""
    from kpolaris_amd import abi
    for nx, ny in ((1, 1), (7, 3)):
        print(f"case: mixed edge {nx}x{ny}", flush=True)
        p, m, _, _ = _options("riaf_m87_c2.par", nx=nx, ny=ny).resolve_riaf()
        a, _ = _gpu_image(p, m, precision=abi.PRECISION_MIXED)
        b, _ = _gpu_image(p, m, precision=abi.PRECISION_MIXED)
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        ref, _ = _gpu_image(p, m, precision=abi.PRECISION_FP64)
        assert np.array_equal(a["reason"], ref["reason"])
        assert np.all(np.isfinite(a["image_i"]))


def test_gpu_invalid_precision_fails_loudly():
    from kpolaris_amd import image
    p, m, _, _ = _options("riaf_m87_c2.par", nx=4, ny=4).resolve_riaf()
    with pytest.raises(image.KPError):
        image.run_riaf_image(p, m, precision=7)
    with pytest.raises(image.KPError):
        image.run_riaf_image(p, m, precision=1, fp64_radius=-1.0)
