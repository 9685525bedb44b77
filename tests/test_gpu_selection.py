"""Ray and emission selection on the device: direct_only (pass A ends at the
first vertical turn, pass_a_core.hpp:240-275) and the equatorial emission
wedge (emission_selection.hpp:11-122) through the C ABI, against the oracle
on the same inputs: RIAF (split persistent with the two pass-B legs, one ray
per thread, unsplit, fixed step, mixed precision, batched), GRMHD and slow
light, and the option rules of kpolaris_model_image_options.cpp:196-209."""
from __future__ import annotations

import numpy as np
import pytest

from kpolaris_amd import abi
from tests import parity
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def _host(out):
    import torch
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _riaf(nx):
    from kpolaris_amd import image
    o = image.Options(parameter_file=str(image.PKG / "params" / "riaf_m87_c2.par"))
    o.set("nx", nx)
    o.set("ny", nx)
    p, m, _, _ = o.resolve_riaf()
    return p, m


def _with(p, hr=None, direct=None):
    q = abi.PassAParams.from_buffer_copy(p)
    if hr is not None:
        q.emission_selection.equatorial_h_over_r = hr
    if direct is not None:
        q.direct_only = direct
    return q


@pytest.mark.parametrize("persistent,split", [(True, 1), (False, 1), (True, 0)])
def test_gpu_direct_only_riaf(persistent, split):
    from kpolaris_amd import image
    p, m = _riaf(24)
    p = _with(p, direct=1)
    print(f"case: direct_only RIAF 24x24 persistent={persistent} split={split}", flush=True)
    out, st = image.run_riaf_image(p, m, persistent=persistent, split_transport=split)
    got = _host(out)
    ref = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity(got, ref, f"gpu direct_only persistent={persistent} split={split}"), flush=True)
    full, _ = image.run_riaf_image(_with(p, direct=0), m)
    assert (got["pass_a_steps"] <= _host(full)["pass_a_steps"]).all()
    assert (got["pass_a_steps"] < _host(full)["pass_a_steps"]).any()


@pytest.mark.parametrize("persistent,split,adaptive", [(True, 1, 1), (False, 1, 1), (True, 0, 1), (True, 1, 0)])
def test_gpu_wedge_riaf(persistent, split, adaptive):
    """h/r = 0.05 (with direct_only on the persistent split case); the wedge
    images keep the parity bar on the Stokes values with the wedge's
    step-count rule (tests/parity.py)."""
    from kpolaris_amd import image
    p, m = _riaf(16)
    p = _with(p, hr=0.05, direct=1 if (persistent and split and adaptive) else 0)
    p.adaptive = adaptive
    if not adaptive:
        p.step, p.inner_radius = 0.05, 2.5
    print(f"case: wedge RIAF 16x16 persistent={persistent} split={split} adaptive={adaptive}", flush=True)
    out, _ = image.run_riaf_image(p, m, persistent=persistent, split_transport=split)
    got = _host(out)
    ref = parity.oracle_with_sensitivity(p, m)
    print(parity.assert_parity(got, ref, f"gpu wedge persistent={persistent} split={split} adaptive={adaptive}",
                               max_ill=parity.WEDGE_MAX_ILL, step_rel=parity.WEDGE_STEP_REL), flush=True)
    full, _ = image.run_riaf_image(_with(p, hr=0.0), m)
    full = _host(full)
    assert np.array_equal(got["pass_a_steps"], full["pass_a_steps"])
    assert 0.0 < got["image_i"].sum() < full["image_i"].sum()


@pytest.mark.parametrize("guard_mode", [0, 1])
def test_gpu_selection_mixed(guard_mode):
    """direct_only and the wedge under mixed precision at the mixed bars."""
    import oracle
    from kpolaris_amd import image
    p, m = _riaf(24)
    p = _with(p, hr=0.1, direct=1)
    print(f"case: mixed direct_only + wedge RIAF 24x24 guard_mode {guard_mode}", flush=True)
    out, _ = image.run_riaf_image(p, m, precision=abi.PRECISION_MIXED, guard_mode=guard_mode)
    ref, _ = oracle.run_riaf_image(p, m)
    print(parity.assert_mixed(_host(out), ref, f"gpu mixed selection guard_mode {guard_mode}", pol="absolute"),
          flush=True)


def test_gpu_selection_batch_equals_single():
    """A batch of jobs with and without direct_only and the wedge: each job
    equals its single image bit for bit."""
    import torch
    from kpolaris_amd import batch_jobs, image
    print("case: batch with direct_only and wedge jobs", flush=True)
    jobs = batch_jobs.jobs(seed=11, n=4, sizes=(8, 12))
    jobs = [(_with(p, hr=[0.0, 0.05, 0.0, 0.2][j], direct=[0, 0, 1, 1][j]), m) for j, (p, m) in enumerate(jobs)]
    b = image.Batch(jobs)
    b.launch()
    torch.cuda.synchronize()
    for j, (p, m) in enumerate(jobs):
        got = _host(b.outputs[j])
        single, _ = image.run_riaf_image(p, m)
        single = _host(single)
        for k in got:
            assert np.array_equal(got[k], single[k]), (j, k)
    b.close()


def test_gpu_selection_grmhd():
    """GRMHD image on FMKS geodesics, three frequencies, with direct_only and
    with the wedge, against the oracle."""
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    cells, m = gs.make_snapshot(40, 24, 24, seed=3)
    grid = image.GRMHDGrid(m, cells)
    host = grid.host_cells()
    F = [86e9, 230e9, 345e9]
    for hr, direct in ((0.0, 1), (0.1, 0)):
        print(f"case: GRMHD 12x12 3 freqs wedge {hr} direct_only {direct}", flush=True)
        p = _with(gs.pass_a_params(m, nx=12, ny=12), hr=hr, direct=direct)
        out, _ = image.run_grmhd_image(p, grid, F)
        got = _host(out)
        run = lambda q: oracle.run_grmhd_image(q, m, host, F)[0]  # noqa: E731
        ref = run(p)
        cond = parity.sensitivity(run, p, ref)
        kw = dict(max_ill=parity.WEDGE_MAX_ILL, step_rel=parity.WEDGE_STEP_REL) if hr > 0.0 else {}
        print(parity.assert_parity_grid(got, ref, 3, f"gpu GRMHD wedge {hr} direct {direct}", cond, **kw),
              flush=True)


def test_gpu_selection_slow_light():
    """Slow light (decoupled, fluid interpolation) with the wedge and
    direct_only against the oracle."""
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32)
    grids = [image.GRMHDGrid(model, c) for c in cells]
    host = [g.host_cells() for g in grids]
    p = _with(gs.pass_a_params(model, nx=12, ny=12, camera_radius=200.0, outer_radius=30.0), hr=0.1, direct=1)
    print("case: slow light 12x12 wedge 0.1 direct_only", flush=True)
    out, _ = image.run_grmhd_slow_light_image(p, grids, times, 275.0)
    got = _host(out)
    run = lambda q, t=275.0: oracle.run_grmhd_slow_light_image(q, model, host, times, t)[0]  # noqa: E731
    ref = run(p)
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), 275.0))
    print(parity.assert_parity_grid(got, ref, 1, "gpu slow light wedge direct", cond, time_cut_steps=1,
                                    max_ill=parity.WEDGE_MAX_ILL, step_rel=parity.WEDGE_STEP_REL), flush=True)


def test_gpu_selection_option_rules():
    """equatorial_h_over_r outside {0} and [1e-6, 1], equatorial_samples
    outside [2, 1024] with a wedge, direct_only other than 0 or 1:
    KP_INVALID_ARGUMENT; h/r 1 (the whole sphere) equals no wedge."""
    from kpolaris_amd import image
    p, m = _riaf(8)
    print("case: selection option rules", flush=True)
    for hr, samples, direct in ((-0.1, 8, 0), (float("nan"), 8, 0), (1.1, 8, 0), (1e-9, 8, 0), (0.1, 1, 0),
                                (0.1, 2000, 0), (0.0, 8, 2)):
        q = _with(p, hr=hr, direct=direct)
        q.emission_selection.equatorial_samples = samples
        with pytest.raises(image.KPError) as e:
            image.run_riaf_image(q, m)
        assert e.value.status == abi.KP_INVALID_ARGUMENT, (hr, samples, direct)
    a, _ = image.run_riaf_image(_with(p, hr=1.0), m)
    b, _ = image.run_riaf_image(p, m)
    a, b = _host(a), _host(b)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_gpu_selection_multifrequency_and_mixed_grmhd():
    """The multi-frequency wedge kernels: RIAF at three frequencies in FP64
    against the oracle, and in mixed precision against the FP64 wedge image;
    the mixed GRMHD wedge image against the FP64 one, per frequency, at the
    mixed bars (net polarisation held in absolute terms on these coarse
    images)."""
    import oracle
    from kpolaris_amd import grmhd_synth as gs, image
    F = [86e9, 230e9, 345e9]
    p, m = _riaf(16)
    p = _with(p, hr=0.1)
    print("case: RIAF 16x16 3 freqs wedge 0.1 FP64 and mixed", flush=True)
    out, _ = image.run_riaf_multifrequency_image(p, m, F)
    got = _host(out)
    run = lambda q: oracle.run_riaf_multifrequency_image(q, m, F)[0]  # noqa: E731
    ref = run(p)
    cond = parity.sensitivity(run, p, ref)
    print(parity.assert_parity_grid(got, ref, 3, "gpu RIAF 3 freqs wedge", cond, max_ill=parity.WEDGE_MAX_ILL,
                                    step_rel=parity.WEDGE_STEP_REL), flush=True)
    mx, _ = image.run_riaf_multifrequency_image(p, m, F, precision=abi.PRECISION_MIXED)
    g, r = parity.split_frequencies(_host(mx), 3), parity.split_frequencies(got, 3)
    for f in range(3):
        print(parity.assert_mixed(g[f], r[f], f"gpu mixed RIAF wedge f{f}", pol="absolute"), flush=True)
    cells, gm = gs.make_snapshot(40, 24, 24, seed=3)
    grid = image.GRMHDGrid(gm, cells)
    q = _with(gs.pass_a_params(gm, nx=16, ny=16), hr=0.1)
    print("case: GRMHD 16x16 3 freqs wedge 0.1 mixed vs FP64", flush=True)
    a, _ = image.run_grmhd_image(q, grid, F)
    b, _ = image.run_grmhd_image(q, grid, F, precision=abi.PRECISION_MIXED)
    g, r = parity.split_frequencies(_host(b), 3), parity.split_frequencies(_host(a), 3)
    for f in range(3):
        print(parity.assert_mixed(g[f], r[f], f"gpu mixed GRMHD wedge f{f}", pol="absolute"), flush=True)
