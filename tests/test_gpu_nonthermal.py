"""Kappa and power-law synchrotron on the device through the C ABI against
the oracle: RIAF images (fixed kappa, kappa blended toward thermal, power
law), multi-frequency RIAF, GRMHD images with a fixed and a variable
(sigma, beta) kappa and with a power law; precisions and modes that do not
carry the nonthermal fits fail loudly."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu
PARAMS = Path(__file__).resolve().parent.parent / "kpolaris_amd" / "params"


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _riaf(dist, nx=24, ny=24, **kw):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(PARAMS / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", ny)
    p, m, _, _ = o.resolve_riaf()
    m.emission_type = dist
    for k, v in kw.items():
        setattr(m, k, v)
    return p, m


RIAF_CASES = [("kappa", 2, {}),
              ("kappa blended", 2, {"nonthermal_kappa": 4.2, "variable_kappa_interp_start": 4.0,
                                    "variable_kappa_max": 5.0}),
              ("power law", 3, {})]


@pytest.mark.parametrize("label,dist,kw", RIAF_CASES)
def test_gpu_riaf_nonthermal_parity(label, dist, kw):
    import torch
    from kpolaris_amd import image
    print(f"case: RIAF {label} 24x24", flush=True)
    p, m = _riaf(dist, **kw)
    out, st = image.run_riaf_image(p, m)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity(got, ref, f"riaf {label}"), flush=True)
    assert np.max(ref["image_i"]) > 0.0


def test_gpu_riaf_nonthermal_multifrequency_and_unsplit():
    import oracle
    import torch
    from kpolaris_amd import image
    print("case: RIAF kappa 3 frequencies, unsplit", flush=True)
    p, m = _riaf(2, nx=16, ny=16)
    F = [86e9, 230e9, 345e9]
    out, _ = image.run_riaf_multifrequency_image(p, m, F)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref, _ = oracle.run_riaf_multifrequency_image(p, m, F)
    cond = parity.sensitivity(lambda q: oracle.run_riaf_multifrequency_image(q, m, F)[0], p, ref)
    print(parity.assert_parity_grid(got, ref, 3, "riaf kappa mf", cond), flush=True)
    out2, _ = image.run_riaf_image(p, m, split_transport=0)
    torch.cuda.synchronize()
    got2 = {k: v.cpu().numpy() for k, v in out2.items()}
    ref2 = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity(got2, ref2, "riaf kappa unsplit"), flush=True)


@pytest.mark.parametrize("dist,variable", [(2, 0), (2, 1), (3, 0)])
def test_gpu_grmhd_nonthermal_parity(dist, variable):
    import oracle
    import torch
    from kpolaris_amd import grmhd_synth as gs, image
    # magnetised everywhere: no field-free sample (perpendicular-propagation
    # 0/0 of the kappa jV fit) and sigma >= 8e-5 (kappa <= 103, no Gamma overflow)
    print(f"case: GRMHD emission {dist} variable kappa {variable} 12x12, 2 freqs", flush=True)
    cells, m = gs.make_snapshot(40, 24, 24, seed=3, beta_min=1.0, background_field=1e-2)
    m.emission_type = dist
    m.variable_kappa = variable
    m.variable_kappa_interp_start = 5.5
    m.variable_kappa_max = 7.0
    grid = image.GRMHDGrid(m, cells)
    host = grid.host_cells()
    p = gs.pass_a_params(m, nx=12, ny=12)
    F = [86e9, 345e9]
    out, _ = image.run_grmhd_image(p, grid, F)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref, _ = oracle.run_grmhd_image(p, m, host, F)
    cond = parity.sensitivity(lambda q: oracle.run_grmhd_image(q, m, host, F)[0], p, ref)
    print(parity.assert_parity_grid(got, ref, 2, f"grmhd emission {dist}/{variable}", cond), flush=True)
    assert np.max(ref["image_i"]) > 0.0


def test_gpu_nonthermal_rejections():
    """Mixed precision, the tangent (gradient) entry points and batches carry
    only the thermal fits: they return KP_UNSUPPORTED for kappa or power law."""
    from kpolaris_amd import abi, image
    p, m = _riaf(2, nx=4, ny=4)
    with pytest.raises(image.KPError) as e:
        image.run_riaf_image(p, m, precision=abi.PRECISION_MIXED)
    assert e.value.status == abi.KP_UNSUPPORTED
    with pytest.raises(image.KPError) as e:
        image.run_riaf_image_gradient(p, m, (0.0, 0.0, 0.0))
    assert e.value.status == abi.KP_UNSUPPORTED
    with pytest.raises(image.KPError) as e:
        image.Batch([(p, m)])
    assert e.value.status == abi.KP_UNSUPPORTED
    m.emission_type = 9
    with pytest.raises(image.KPError) as e:
        image.run_riaf_image(p, m)
    assert e.value.status == abi.KP_INVALID_ARGUMENT
