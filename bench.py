#!/usr/bin/env python3
"""Benchmark: rays/s of the per-pixel imaging kernel on one MI355X.

Workload: the M87 analytic RIAF at 1024x1024 with step-doubling adaptive RK4
at tolerance 1e-6 (kpolaris_amd/params/riaf_m87_c2.par).  One "step" is one
whole image: pass A + pass B over all 1,048,576 rays, inputs resident on the
device.  Prints one JSON line.

The CPU baseline leg times the CPU oracle (a scalar C restatement of the
reference algorithm, OpenMP over pixels) on a strided pixel sample of the
same workload, on the host cores, for about ten seconds.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "rays/s on 1 MI355X: mixed FP32/FP64 vs plain FP64 at ≤1e-4 rel Stokes-I error"
PARAMS = ROOT / "kpolaris_amd" / "params" / "riaf_m87_c2.par"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed)


def opcounts() -> dict:
    from kpolaris_amd import build_lib
    return json.loads(build_lib.OPCOUNT.read_text())


def algorithmic_flops(stats: dict, adaptive: bool, fit: int) -> tuple[float, float]:
    """Algorithmic FLOPs of pass A and pass B from the kernel's unit counters and
    the per-unit table of kp_opcount (add/sub/mul/div/sqrt/transcendental = 1)."""
    oc = opcounts()
    per_rhs = oc["doubling_attempt"]["flops"] / 11.0 if adaptive else oc["rk4_step"]["flops"] / 4.0
    coeff = oc["coefficients_dexter" if fit == 4 else "coefficients_pandya"]["flops"]
    rhs_a = stats["pass_a_rhs_evals"]
    rhs_b = stats["rhs_evals"] - rhs_a
    flops_a = rhs_a * per_rhs
    flops_b = (rhs_b * per_rhs + stats["coefficient_evals"] * coeff
               + stats["transfer_steps"] * oc["transfer_step"]["flops"]
               + 2.0 * stats["pass_b_steps"] * oc["substep_count"]["flops"])
    return flops_a, flops_b


def cpu_baseline(p, m, target_s: float) -> dict:
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    npix = p.camera.nx * p.camera.ny
    calib = 64 * threads
    stride = max(1, npix // calib)
    t0 = time.perf_counter()
    oracle.run_riaf_image_strided(p, m, stride // 2, stride, calib, nthreads=threads)
    t_cal = time.perf_counter() - t0
    count = int(min(npix, max(calib, calib * target_s / max(t_cal, 1e-3))))
    stride = max(1, npix // count)
    count = min(count, (npix - stride // 2 + stride - 1) // stride)
    t0 = time.perf_counter()
    oracle.run_riaf_image_strided(p, m, stride // 2, stride, count, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": count / dt, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{count} rays (every {stride}th pixel of the 1024x1024 image), "
                      f"oracle C restatement, OpenMP static schedule, {dt:.1f} s"}


def run_precision(p, m, precision: int, steps: int, warmup: int):
    import torch
    from kpolaris_amd import image
    npix = p.camera.nx * p.camera.ny
    out = image.alloc_device_image(npix)
    for _ in range(warmup):
        image.run_riaf_image(p, m, precision=precision, out=out)
    torch.cuda.synchronize()
    per_step = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = image.run_riaf_image(p, m, precision=precision, out=out)
        per_step.append(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return out, per_step, wall


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="fp64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the imaging kernel runs on a single MI355X "
              "(pixels shard trivially across GPUs, but no multi-GPU path is built)", file=sys.stderr)
        return 2
    import torch
    from kpolaris_amd import abi, image
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    p, m, scale, sign = image.Options(parameter_file=str(PARAMS)).resolve_riaf()
    npix = p.camera.nx * p.camera.ny

    prec = abi.PRECISION_FP64
    out, per_step, wall = run_precision(p, m, prec, args.steps, args.warmup)
    kernel_ms = [s["kernel_ms"] for s in per_step]
    a_ms = statistics.median(s["pass_a_ms"] for s in per_step)
    b_ms = statistics.median(s["pass_b_ms"] for s in per_step)
    st = per_step[-1]
    flops_a, flops_b = algorithmic_flops(st, bool(p.adaptive), m.emission_type)
    dominant = "pass_b" if b_ms >= a_ms else "pass_a"
    dom_flops, dom_ms = (flops_b, b_ms) if dominant == "pass_b" else (flops_a, a_ms)
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    value = args.steps * npix / wall
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic RIAF model; no input data)",
        "config": {
            "workload": "analytic RIAF, M87 (6.5e9 Msun, 16.8 Mpc), spin 0.94, i=163 deg, "
                        "1024x1024 pinhole, 230 GHz, adaptive step-doubling RK4 rtol 1e-6",
            "rays_per_step": npix,
            "precision": "fp64",
            "kernel_ms_median": statistics.median(kernel_ms),
            "pass_a_ms_median": a_ms,
            "pass_b_ms_median": b_ms,
            "steps_per_s": (st["pass_a_steps"] + st["pass_b_steps"]) / (statistics.median(kernel_ms) * 1e-3),
            "rhs_per_s": st["rhs_evals"] / (statistics.median(kernel_ms) * 1e-3),
            "rays_per_s_kernel": npix / (statistics.median(kernel_ms) * 1e-3),
        },
        "roofline": {
            "bound": "valu",
            "kernel": dominant,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "algorithmic_flops_per_launch": dom_flops,
            "launch_ms": dom_ms,
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p, m, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
