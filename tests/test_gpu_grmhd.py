"""GRMHD path on the device (FMKS geodesics, trilinear sampling of 72-byte cell
records, derived scalars, Dexter coefficients) through the C ABI, against the
oracle on seeded synthetic Fishbone-Moncrief snapshots."""
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


@pytest.fixture(scope="module")
def snap():
    from kpolaris_amd import grmhd_synth as gs
    return gs.make_snapshot(40, 24, 24, seed=3)


def test_gpu_derived_scalars_match_oracle(snap):
    import oracle
    from kpolaris_amd import image
    print("case: derived scalars 40x24x24", flush=True)
    cells, m = snap
    grid = image.GRMHDGrid(m, cells)
    got = grid.host_cells()
    ref = cells.copy()
    oracle.grmhd_derive_scalars(m, ref)
    assert np.array_equal(got["prim"], ref["prim"])
    np.testing.assert_allclose(got["derived"], ref["derived"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("persistent", [True, False])
def test_gpu_grmhd_image_parity(snap, persistent):
    import oracle
    import torch
    from kpolaris_amd import grmhd_synth as gs, image
    print(f"case: GRMHD image 12x12, 3 freqs, persistent={persistent}", flush=True)
    cells, m = snap
    grid = image.GRMHDGrid(m, cells)
    host = grid.host_cells()
    p = gs.pass_a_params(m, nx=12, ny=12)
    F = [86e9, 230e9, 345e9]
    out, st = image.run_grmhd_image(p, grid, F, persistent=persistent)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref, ost = oracle.run_grmhd_image(p, m, host, F)
    cond = parity.sensitivity(lambda q: oracle.run_grmhd_image(q, m, host, F)[0], p, ref)
    print(parity.assert_parity_grid(got, ref, 3, f"gpu GRMHD 3 freqs persistent={persistent}", cond), flush=True)
    if np.array_equal(got["steps"], ref["steps"]) and np.array_equal(got["pass_a_steps"], ref["pass_a_steps"]):
        assert st["rhs_evals"] == ost["rhs_evals"]


def test_gpu_grmhd_unsplit_and_single_frequency(snap):
    """The unsplit kernel runs the split kernels' lane code (one
    contraction setting for every instantiation): identical images."""
    import torch
    from kpolaris_amd import grmhd_synth as gs, image
    print("case: GRMHD unsplit vs split 10x10", flush=True)
    cells, m = snap
    grid = image.GRMHDGrid(m, cells)
    p = gs.pass_a_params(m, nx=10, ny=10)
    a, _ = image.run_grmhd_image(p, grid)
    b, _ = image.run_grmhd_image(p, grid, split_transport=0)
    torch.cuda.synchronize()
    for k in a:
        assert np.array_equal(b[k].cpu().numpy(), a[k].cpu().numpy(), equal_nan=True), k


def test_gpu_grmhd_rejects_wrong_coordinates(snap):
    from kpolaris_amd import abi, grmhd_synth as gs, image
    cells, m = snap
    grid = image.GRMHDGrid(m, cells, derive=False)
    p = gs.pass_a_params(m, nx=4, ny=4)
    p.coordinate_system = abi.COORD_BOYER_LINDQUIST
    with pytest.raises(image.KPError):
        image.run_grmhd_image(p, grid)
