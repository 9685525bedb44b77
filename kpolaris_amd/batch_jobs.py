"""Synthetic uneven batches of analytic-RIAF image solves.

Each job is the M87 RIAF parameter file (params/riaf_m87_c2.par) with its
resolution, spin, inclination, observing frequency and density scale drawn
from a seeded generator: resolution from {128, 256, 512, 1024}^2, spin from
{0, 0.5, 0.94}, inclination uniform in 10-90 degrees, frequency from
{86, 230, 345} GHz and riaf_ne_unit multiplied by a log-uniform factor in
[0.3, 3].  Inner and outer radii are re-resolved per job by the option layer
(1.05 r_h for the job's spin).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

PARAMS = Path(__file__).resolve().parent / "params" / "riaf_m87_c2.par"
SIZES = (128, 256, 512, 1024)
SPINS = (0.0, 0.5, 0.94)
FREQS = (86e9, 230e9, 345e9)


def draw(seed: int = 20261202, n: int = 64, sizes=SIZES) -> list[dict]:
    """Per-job option overrides (parameter-file keys)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        size = int(rng.choice(sizes))
        jobs.append({
            "nx": size, "ny": size,
            "spin": float(rng.choice(SPINS)),
            "inclination_deg": float(rng.uniform(10.0, 90.0)),
            "freq": float(rng.choice(FREQS)),
            "ne_scale": float(np.exp(rng.uniform(np.log(0.3), np.log(3.0)))),
        })
    return jobs


def resolve(overrides: dict, parameter_file: Path = PARAMS):
    """(PassAParams, RIAFModel) of one job."""
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(parameter_file))
    base_ne = None
    for k, v in overrides.items():
        if k == "ne_scale":
            continue
        o.set(k, str(v))
    p, m, _, _ = o.resolve_riaf()
    base_ne = m.ne_unit
    m.ne_unit = base_ne * overrides.get("ne_scale", 1.0)
    return p, m


def jobs(seed: int = 20261202, n: int = 64, sizes=SIZES, parameter_file: Path = PARAMS):
    return [resolve(j, parameter_file) for j in draw(seed, n, sizes)]
