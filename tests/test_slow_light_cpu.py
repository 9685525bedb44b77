"""Slow light on the CPU: the oracle's retarded-time sampling pinned by the
reference's own checks (fluid time interpolation: exact endpoints, the
mid-time state equals a snapshot built from mid values,
tests/test_fluid_time_interpolation.cpp:75-101) and by limits that fix the
algorithm (identical snapshots reproduce the static image; bisection over
the snapshot times against a brute-force search), then the kernel's lane code
(tests/emul) against the oracle in every stepping and interpolation mode."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, grmhd_synth as gs
from tests import parity

P = C.POINTER
STOKES = ("image_i", "image_q", "image_u", "image_v")
# grid, camera and sequence of the lane tests: radiating samples of the
# 12x12 image at camera radius 200 with outer radius 30 lie at fluid times
# 275 + x^0 in [about 5, 95]
T_OBS = 275.0


@pytest.fixture(scope="module")
def seq():
    cells, times, model = gs.make_sequence(nsnap=6, dt=20.0, t0=0.0, n1=48, n2=32, n3=32)
    for c in cells:
        oracle.grmhd_derive_scalars(model, c)
    p = gs.pass_a_params(model, nx=12, ny=12, camera_radius=200.0, outer_radius=30.0)
    return cells, times, model, p


@pytest.fixture(scope="module")
def emul():
    from tests.emul import build as emul_build
    if not emul_build.LIB.exists():
        emul_build.build()
    L = C.CDLL(str(emul_build.LIB))
    L.kp_emul_run_grmhd_slow_light_image.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLight),
                                                     P(C.c_double), C.c_int32, P(abi.ImageData), P(C.c_double),
                                                     P(C.c_ulonglong)]
    L.kp_emul_run_grmhd_slow_light_image.restype = C.c_int32
    L.kp_emul_run_grmhd_slow_light_blocks.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLight),
                                                      P(C.c_double), C.c_int32, P(abi.ImageData), P(C.c_double),
                                                      P(C.c_ulonglong), P(C.c_int32), P(C.c_int32)]
    L.kp_emul_run_grmhd_slow_light_blocks.restype = C.c_int32
    return L


def _emul_run(L, p, model, cells, times, t_obs, mode, interp, freqs=None):
    sq, keep = oracle.slow_light_sequence(cells, times, t_obs, mode, interp)
    npix = p.camera.nx * p.camera.ny
    f = np.array([model.freq_cgs] if freqs is None else freqs, dtype=np.float64)
    arrays = oracle.alloc_image(npix)
    for n in STOKES:
        arrays[n] = np.zeros(len(f) * npix)
    img = oracle.image_struct(arrays)
    req = np.zeros(npix)
    w = C.c_ulonglong()
    assert L.kp_emul_run_grmhd_slow_light_image(C.byref(p), C.byref(model), C.byref(sq), f.ctypes.data_as(P(C.c_double)),
                                                len(f), C.byref(img), req.ctypes.data_as(P(C.c_double)),
                                                C.byref(w)) == 0
    return arrays, req, w.value


def test_interpolate_time_value_and_fraction():
    """interpolate_time_value (time_interpolation.hpp:27-32): the endpoint
    values come back bit for bit, equal values stay exact, the interior is
    the linear interpolant; the fraction (slow_light_driver.hpp:62-69) is
    clamped to [0, 1] and 0 for a zero-length interval."""
    L = oracle.lib()
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, b = rng.normal(size=2) * 10.0 ** rng.integers(-30, 30, size=2)
        assert L.kpo_interpolate_time_value(a, b, 0.0) == a
        assert L.kpo_interpolate_time_value(a, b, -0.3) == a
        assert L.kpo_interpolate_time_value(a, b, 1.0) == b
        assert L.kpo_interpolate_time_value(a, b, 1.7) == b
        assert L.kpo_interpolate_time_value(a, a, 0.37) == a
        f = float(rng.uniform(0.01, 0.99))
        v = L.kpo_interpolate_time_value(a, b, f)
        assert abs(v - (a + f * (b - a))) <= 4e-16 * max(abs(a), abs(b))
    assert L.kpo_slow_light_fraction(10.0, 20.0, 100.0, -85.0) == 0.5
    assert L.kpo_slow_light_fraction(10.0, 20.0, 100.0, -95.0) == 0.0
    assert L.kpo_slow_light_fraction(10.0, 20.0, 100.0, -75.0) == 1.0
    assert L.kpo_slow_light_fraction(10.0, 10.0, 100.0, -85.0) == 0.0


def test_snapshot_pair_bisection_matches_brute_force():
    """sample_window (slow_light_driver.hpp:1000-1012) on nonuniform times:
    the largest k < nsnap - 1 with t_k <= t (times before the first give 0)."""
    L = oracle.lib()
    rng = np.random.default_rng(9)
    for n in (2, 3, 7, 32):
        times = np.cumsum(rng.uniform(0.1, 5.0, size=n))
        tp = times.ctypes.data_as(P(C.c_double))
        for t in np.concatenate([times, rng.uniform(times[0] - 3, times[-1] + 3, size=300)]):
            brute = int(np.clip(np.searchsorted(times, t, side="right") - 1, 0, n - 2))
            assert L.kpo_slow_light_pair(tp, n, float(t)) == brute, (n, t)


def _state_at(model, rng):
    """A transport state inside the torus with a random photon direction and basis."""
    s = np.zeros(16)
    s[1] = np.log(10.0 + rng.uniform(-2.0, 2.0))
    s[2] = 0.5 + rng.uniform(-0.05, 0.05)
    s[3] = rng.uniform(0.0, 2.0 * np.pi)
    s[4:8] = [1.0, *(rng.normal(size=3) * 0.05)]
    s[8:16] = rng.normal(size=8) * 0.1
    return s


def test_fluid_time_interpolation_mid_state_and_endpoints(seq):
    """The reference's fluid time-interpolation checks
    (tests/test_fluid_time_interpolation.cpp:75-101) on the oracle: with
    snapshot B = 3 A in the primitives (so the float32 mid-time values 2 A are
    exact) and C carrying the mid values of every primitive and derived
    scalar, the coefficients at time fraction 1/2 equal C's static
    coefficients; at the endpoints they are A's and B's bit for bit."""
    cells, times, model, p = seq
    L = oracle.lib()
    A = cells[0].copy()
    # 20-bit significands: 3 A and the mean 2 A are then exact in float32
    A["prim"] = (A["prim"].view(np.uint32) & np.uint32(0xFFFFFFF0)).view(np.float32)
    oracle.grmhd_derive_scalars(model, A)
    B = A.copy()
    B["prim"] = A["prim"] * np.float32(3.0)
    oracle.grmhd_derive_scalars(model, B)
    Cm = A.copy()
    Cm["prim"] = A["prim"] * np.float32(2.0)
    Cm["derived"] = 0.5 * (A["derived"] + B["derived"])
    assert np.array_equal(Cm["prim"].astype(np.float64), 0.5 * (A["prim"].astype(np.float64) + B["prim"].astype(np.float64)))
    rng = np.random.default_rng(3)
    faraday = 1
    nchecked = 0
    for interp in (abi.SLOW_LIGHT_FLUID, abi.SLOW_LIGHT_COEFFICIENTS):
        sq, keep = oracle.slow_light_sequence([A, B], [100.0, 140.0], 300.0, abi.SLOW_LIGHT_DECOUPLED, interp)
        for _ in range(40):
            s = _state_at(model, rng)
            st = oracle.state(s)
            out = {}
            for name, x0 in (("lo", -200.0), ("mid", -180.0), ("hi", -160.0)):
                st.v[0] = x0
                c = oracle.Coeffs()
                L.kpo_slow_light_coefficients(C.byref(model), C.byref(sq), 0, C.byref(st), None, faraday, C.byref(c))
                out[name] = c.as_array()
            ref = {}
            for name, cc in (("A", A), ("B", B), ("C", Cm)):
                m = abi.GRMHDModel.from_buffer_copy(model)
                m.cells = C.c_void_p(cc.ctypes.data)
                c = oracle.Coeffs()
                L.kpo_grmhd_coefficients(C.byref(m), C.byref(st), None, faraday, C.byref(c))
                ref[name] = c.as_array()
            assert np.array_equal(out["lo"], ref["A"]), interp
            assert np.array_equal(out["hi"], ref["B"]), interp
            if interp == abi.SLOW_LIGHT_FLUID and np.any(ref["C"] != 0.0):
                scale = np.max(np.abs(ref["C"]))
                assert np.max(np.abs(out["mid"] - ref["C"])) <= 1e-10 * scale
                nchecked += 1
            if interp == abi.SLOW_LIGHT_COEFFICIENTS:
                # the coefficient mode is the linear mix of the endpoint coefficients
                scale = max(np.max(np.abs(ref["A"])), np.max(np.abs(ref["B"])), 1e-300)
                assert np.max(np.abs(out["mid"] - 0.5 * (ref["A"] + ref["B"]))) <= 1e-14 * scale
    assert nchecked >= 10


def test_fluid_time_interpolation_four_velocity_normalisation(seq):
    """u.u = -1 of the four-velocity rebuilt from time-interpolated velocity
    primitives (test_fluid_time_interpolation.cpp:92-101): the oracle's fluid
    reconstruction of a mid-time state is the static reconstruction of the mid
    primitives, whose normalisation the GRMHD pins check; here the mid state
    must not lose the coefficients' structure (nonzero emission where both
    endpoint snapshots emit)."""
    cells, times, model, p = seq
    L = oracle.lib()
    sq, keep = oracle.slow_light_sequence(cells[:2], times[:2], 300.0, abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID)
    rng = np.random.default_rng(11)
    emitting = 0
    for _ in range(40):
        st = oracle.state(_state_at(model, rng))
        vals = []
        for x0 in (-300.0, -290.0, -280.0):
            st.v[0] = x0
            c = oracle.Coeffs()
            L.kpo_slow_light_coefficients(C.byref(model), C.byref(sq), 0, C.byref(st), None, 1, C.byref(c))
            vals.append(c.jI)
        if vals[0] > 0.0 and vals[2] > 0.0:
            assert vals[1] > 0.0
            assert min(vals[0], vals[2]) * 0.2 < vals[1] < max(vals[0], vals[2]) * 5.0
            emitting += 1
    assert emitting >= 10


def _emul_blocks(L, p, model, cells, times, t_obs, mode, interp, windows_per_block, freqs=None):
    sq, keep = oracle.slow_light_sequence(cells, times, t_obs, mode, interp, windows_per_block)
    npix = p.camera.nx * p.camera.ny
    f = np.array([model.freq_cgs] if freqs is None else freqs, dtype=np.float64)
    arrays = oracle.alloc_image(npix)
    for n in STOKES:
        arrays[n] = np.zeros(len(f) * npix)
    img = oracle.image_struct(arrays)
    req = np.zeros(npix)
    w = C.c_ulonglong()
    blocks, rewinds = C.c_int32(), C.c_int32()
    assert L.kp_emul_run_grmhd_slow_light_blocks(C.byref(p), C.byref(model), C.byref(sq),
                                                 f.ctypes.data_as(P(C.c_double)), len(f), C.byref(img),
                                                 req.ctypes.data_as(P(C.c_double)), C.byref(w), C.byref(blocks),
                                                 C.byref(rewinds)) == 0
    return arrays, req, w.value, blocks.value, rewinds.value


def test_oracle_block_stepping_in_blocks_of_one_pair_is_snapshot_stepping(seq):
    """Blocks of windows_per_block snapshot pairs (the batch scheduler,
    slow_light_batch_driver.hpp:211-340): block stepping in blocks of one pair
    cuts every step at each snapshot time and samples that pair, which is
    snapshot stepping, so the two images agree bit for bit; blocks as long as
    the sequence are the single block.  Two-pair blocks cut at alternate
    snapshot times only: a different image, whose Stokes I stays within
    1e-3 of the peak (the linear polarisation of this model moves far more
    with the step layout; the next test holds two-pair blocks to the oracle's
    own two-pair image instead)."""
    cells, times, model, p = seq
    run = lambda mode, w: oracle.run_grmhd_slow_light_image(  # noqa: E731
        p, model, cells, times, T_OBS, step_mode=mode, windows_per_block=w)[0]
    snap = run(abi.SLOW_LIGHT_SNAPSHOT, 0)
    one = run(abi.SLOW_LIGHT_BLOCK, 1)
    whole, longest = run(abi.SLOW_LIGHT_BLOCK, 0), run(abi.SLOW_LIGHT_BLOCK, len(times) - 1)
    two = run(abi.SLOW_LIGHT_BLOCK, 2)
    for k in (*STOKES, "steps", "reason", "pass_a_steps"):
        assert np.array_equal(one[k], snap[k]), k
        assert np.array_equal(longest[k], whole[k]), k
    assert not np.array_equal(two["steps"], whole["steps"])
    assert not np.array_equal(two["steps"], one["steps"])
    peak = float(np.max(whole["image_i"]))
    assert np.max(np.abs(two["image_i"] - whole["image_i"])) <= 1e-3 * peak


@pytest.mark.parametrize("mode,wpb", [(abi.SLOW_LIGHT_DECOUPLED, 1), (abi.SLOW_LIGHT_DECOUPLED, 2),
                                      (abi.SLOW_LIGHT_SNAPSHOT, 2), (abi.SLOW_LIGHT_SNAPSHOT, 1)])
def test_emulated_blocks_equal_the_single_block_lanes(seq, emul, mode, wpb):
    """The kernel's park and resume code under the library's block loop
    (CPU build): decoupled and snapshot images do not depend on the block
    length, so rays parked at block ends or waiting for a later block and
    resumed in the next launch give the single-block lanes' image bit for
    bit -- Stokes, step counts and reasons."""
    cells, times, model, p = seq
    ref, req0, w0 = _emul_run(emul, p, model, cells, times, T_OBS, mode, abi.SLOW_LIGHT_FLUID)
    got, req, w, blocks, rewinds = _emul_blocks(emul, p, model, cells, times, T_OBS, mode, abi.SLOW_LIGHT_FLUID, wpb)
    assert blocks == -(-(len(times) - 1) // wpb) and rewinds == 0
    assert w == w0 == 0
    for k in (*STOKES, "steps", "reason", "pass_a_steps", "closure_x", "final_null"):
        assert np.array_equal(got[k], ref[k]), k


def test_emulated_block_stepping_in_blocks_matches_oracle(seq, emul):
    """Block stepping in two-pair blocks: the emulated lanes against the
    oracle's blocks at the parity bar, and in one-pair blocks bit for bit
    equal to the emulated snapshot stepping."""
    cells, times, model, p = seq
    got, _, w, blocks, _ = _emul_blocks(emul, p, model, cells, times, T_OBS, abi.SLOW_LIGHT_BLOCK,
                                        abi.SLOW_LIGHT_FLUID, 2)
    assert w == 0 and blocks == 3
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(  # noqa: E731
        q, model, cells, times, t, step_mode=abi.SLOW_LIGHT_BLOCK, windows_per_block=2)[0]
    ref = run(p)
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got, ref, 1, "emulated block stepping, two-pair blocks", cond,
                                    time_cut_steps=1), flush=True)
    one, _, _, _, _ = _emul_blocks(emul, p, model, cells, times, T_OBS, abi.SLOW_LIGHT_BLOCK, abi.SLOW_LIGHT_FLUID, 1)
    snap, _, _ = _emul_run(emul, p, model, cells, times, T_OBS, abi.SLOW_LIGHT_SNAPSHOT, abi.SLOW_LIGHT_FLUID)
    for k in (*STOKES, "steps", "reason"):
        assert np.array_equal(one[k], snap[k]), k


def test_emulated_blocks_waiting_rays_outside_the_sequence(seq, emul):
    """Decoupled rays needing a time before the first snapshot end as
    slow_light_time_exhausted with the time they needed, in blocks as in the
    single block; two frequencies on shared geodesics park and resume with
    their Stokes vectors."""
    cells, times, model, p = seq
    ref, req0, w0 = _emul_run(emul, p, model, cells, times, 230.0, abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID)
    got, req, w, _, _ = _emul_blocks(emul, p, model, cells, times, 230.0, abi.SLOW_LIGHT_DECOUPLED,
                                     abi.SLOW_LIGHT_FLUID, 1)
    assert w == w0 > 0
    assert np.array_equal(np.isnan(req), np.isnan(req0))
    assert np.array_equal(req[~np.isnan(req)], req0[~np.isnan(req0)])
    for k in (*STOKES, "steps", "reason"):
        assert np.array_equal(got[k], ref[k]), k
    freqs = [86e9, 230e9]
    ref2, _, _ = _emul_run(emul, p, model, cells, times, T_OBS, abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID, freqs)
    got2, _, _, _, _ = _emul_blocks(emul, p, model, cells, times, T_OBS, abi.SLOW_LIGHT_DECOUPLED,
                                    abi.SLOW_LIGHT_FLUID, 2, freqs)
    for k in (*STOKES, "steps", "reason"):
        assert np.array_equal(got2[k], ref2[k]), k


def test_identical_snapshots_reproduce_the_static_image(seq):
    """A sequence of one repeated snapshot is the static model: decoupled and
    block stepping with fluid interpolation give the static GRMHD image bit
    for bit (interpolate_time_value returns equal values exactly and no step is
    cut); snapshot stepping cuts steps at the interior snapshot times, which
    moves the image by the integration tolerance only."""
    cells, times, model, p = seq
    same = [cells[0]] * 3
    ref, _ = oracle.run_grmhd_image(p, abi.GRMHDModel.from_buffer_copy(model), cells[0])
    peak = float(np.max(ref["image_i"]))
    for mode in (abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_BLOCK, abi.SLOW_LIGHT_SNAPSHOT):
        got, _, w = oracle.run_grmhd_slow_light_image(p, model, same, [-2000.0, 0.0, 2000.0], 0.0, step_mode=mode)
        assert w[0] == 0
        assert np.array_equal(got["reason"], ref["reason"]), mode
        if mode == abi.SLOW_LIGHT_SNAPSHOT:
            for k in STOKES:
                assert np.max(np.abs(got[k] - ref[k])) <= 1e-8 * peak, (mode, k)
        else:
            for k in (*STOKES, "steps"):
                assert np.array_equal(got[k], ref[k]), (mode, k)


def test_decoupled_time_support_error_and_snapshot_vacuum_finish(seq):
    """Decoupled stepping with radiating samples before the first snapshot
    leaves those rays waiting (reason 6) and reports the requested range
    (slow_light_batch_driver.hpp:282-290); snapshot stepping with the sequence
    ending before the rays leave the emitting region ends them as
    slow_light_time_exhausted after the last window (advance_slow_light_vacuum_to_camera,
    slow_light_driver.hpp:1580-1584)."""
    cells, times, model, p = seq
    got, _, w = oracle.run_grmhd_slow_light_image(p, model, cells, times, 230.0)
    assert w[0] > 0 and w[1] < times[0]
    assert np.all(got["reason"][got["reason"] != abi.REASON_REACHED_CAMERA] == abi.REASON_SLOW_LIGHT_TIME_EXHAUSTED)
    assert np.sum(got["reason"] == abi.REASON_SLOW_LIGHT_TIME_EXHAUSTED) == w[0]
    late, _, w2 = oracle.run_grmhd_slow_light_image(p, model, cells[:3], times[:3], T_OBS,
                                                    step_mode=abi.SLOW_LIGHT_SNAPSHOT)
    assert w2[0] == 0
    assert np.sum(late["reason"] == abi.REASON_SLOW_LIGHT_TIME_EXHAUSTED) > 0
    ok, _, w3 = oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS)
    assert w3[0] == 0 and np.all(ok["reason"] == abi.REASON_REACHED_CAMERA)


@pytest.mark.parametrize("mode", [abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_SNAPSHOT, abi.SLOW_LIGHT_BLOCK])
@pytest.mark.parametrize("interp", [abi.SLOW_LIGHT_FLUID, abi.SLOW_LIGHT_COEFFICIENTS])
def test_emulated_slow_light_lanes_match_oracle(seq, emul, mode, interp):
    """The kernel's slow-light lanes (window cuts, pair bisection, vacuum
    finish) against the oracle: identical termination reasons, step counts
    within one (a ray that approaches a snapshot time by ever smaller cut steps
    stops when the remaining step is below 1e-14; where that happens depends on
    the last bits of x^0), Stokes at the parity bar."""
    cells, times, model, p = seq
    got, req, w = _emul_run(emul, p, model, cells, times, T_OBS, mode, interp)
    ref, _, wo = oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS, step_mode=mode,
                                                  interpolation=interp)
    assert w == wo[0] == 0
    assert np.all(np.isnan(req))
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(  # noqa: E731
        q, model, cells, times, t, step_mode=mode, interpolation=interp)[0]
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got, ref, 1, f"emulated slow light mode {mode} interp {interp}", cond,
                                    time_cut_steps=1), flush=True)


def test_emulated_slow_light_waiting_rays_and_frequencies(seq, emul):
    """Waiting rays carry the time they needed; two frequencies on shared
    geodesics match the oracle."""
    cells, times, model, p = seq
    got, req, w = _emul_run(emul, p, model, cells, times, 230.0, abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID)
    ref, _, wo = oracle.run_grmhd_slow_light_image(p, model, cells, times, 230.0)
    assert w == wo[0] > 0
    waiting = ~np.isnan(req)
    assert waiting.sum() == w
    assert np.all(got["reason"][waiting] == abi.REASON_SLOW_LIGHT_TIME_EXHAUSTED)
    assert np.isclose(req[waiting].min(), wo[1], rtol=0, atol=1e-9)
    assert np.isclose(req[waiting].max(), wo[2], rtol=0, atol=1e-9)
    freqs = [86e9, 230e9]
    got2, _, _ = _emul_run(emul, p, model, cells, times, T_OBS, abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_FLUID, freqs)
    ref2, _, _ = oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS, frequencies=freqs)
    run = lambda q, t=T_OBS: oracle.run_grmhd_slow_light_image(q, model, cells, times, t, frequencies=freqs)[0]  # noqa: E731
    cond = parity.sensitivity(run, p, ref2, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(got2, ref2, 2, "emulated slow light 2 freqs", cond, time_cut_steps=1),
          flush=True)


def _slow_fd(p, model, cells, times, kind, h=1e-4):
    """Central differences of the oracle's slow-light image: kind 0 the density
    family and 2 the field scale (log-scale h, PlasmaPerturbation), kind 1
    R_high (relative step h, the derived scalars recomputed per snapshot)."""
    if kind == 1:
        res = []
        for sgn in (1.0, -1.0):
            m2 = abi.GRMHDModel.from_buffer_copy(model)
            m2.trat_large = model.trat_large * (1.0 + sgn * h)
            cs = []
            for c in cells:
                c2 = c.copy()
                oracle.grmhd_derive_scalars(m2, c2)
                cs.append(c2)
            res.append(oracle.run_grmhd_slow_light_image(p, m2, cs, times, T_OBS)[0])
        step = 2.0 * h * model.trat_large
    else:
        param = 1 if kind == 0 else 3
        res = [oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS, perturbation=(param, s * h))[0]
               for s in (1.0, -1.0)]
        step = 2.0 * h
    same = (res[0]["steps"] == res[1]["steps"]) & (res[0]["pass_a_steps"] == res[1]["pass_a_steps"])
    return {n: (res[0][n] - res[1][n]) / step for n in STOKES}, same


def test_emulated_slow_light_gradient_against_central_differences(seq, emul):
    """The kernel's slow-light tangent lanes: the value equals the slow-light
    image, and d(I,Q,U,V)/d(q_density, R_high, q_magnetic) along the fixed
    geodesics matches central differences of the oracle (h = 1e-4) on the rays
    whose step sequence the perturbation leaves unchanged."""
    cells, times, model, p = seq
    L = emul
    L.kp_emul_run_grmhd_slow_light_gradient.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLight),
                                                        P(C.c_double), P(abi.ImageData), P(C.c_double),
                                                        P(C.c_double)]
    L.kp_emul_run_grmhd_slow_light_gradient.restype = C.c_int32
    sq, keep = oracle.slow_light_sequence(cells, times, T_OBS)
    npix = p.camera.nx * p.camera.ny
    arrays = oracle.alloc_image(npix)
    img = oracle.image_struct(arrays)
    grad = np.zeros((3, 4, npix))
    req = np.zeros(npix)
    q = np.zeros(2)
    assert L.kp_emul_run_grmhd_slow_light_gradient(C.byref(p), C.byref(model), C.byref(sq), q.ctypes.data_as(P(C.c_double)),
                                                   C.byref(img), grad.ctypes.data_as(P(C.c_double)),
                                                   req.ctypes.data_as(P(C.c_double))) == 0
    ref, _, _ = oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS)
    run = lambda qq, t=T_OBS: oracle.run_grmhd_slow_light_image(qq, model, cells, times, t)[0]  # noqa: E731
    cond = parity.sensitivity(run, p, ref, parity.time_shifts(lambda t: run(p, t), T_OBS))
    print(parity.assert_parity_grid(arrays, ref, 1, "emulated slow-light tangent value", cond, time_cut_steps=1),
          flush=True)
    # central differences at h = 1e-4 carry an O(h^2) truncation error (up to
    # 3e-5 of the largest derivative on strongly Faraday-rotated rays); the
    # Richardson combination (4 D(h/2) - D(h)) / 3 removes it
    for kind in range(3):
        fd, same = _slow_fd(p, model, cells, times, kind, 1e-4)
        fd2, same2 = _slow_fd(p, model, cells, times, kind, 5e-5)
        for s, name in enumerate(STOKES):
            ok = same & same2 & (arrays["steps"] == ref["steps"]) & parity.smooth_for_derivatives(ref, cond, 1)
            print(f"gradient check of {name} on {int(ok.sum())}/{ok.size} pixels", flush=True)
            assert ok.sum() >= parity.MIN_CHECKED * ok.size
            scale = np.abs(fd[name][ok]).max()
            assert scale > 0.0
            err = np.abs(grad[kind, s][ok] - fd[name][ok]).max() / scale
            rich = (4.0 * fd2[name] - fd[name]) / 3.0
            err_r = np.abs(grad[kind, s][ok] - rich[ok]).max() / scale
            print(f"slow-light d/d{('q_density', 'R_high', 'q_magnetic')[kind]} {name}: max rel err {err:.2e}, "
                  f"against Richardson {err_r:.2e}", flush=True)
            assert err < 1e-4, (kind, name, err)
            assert err_r < 1e-7, (kind, name, err_r)
            # most pixels agree to rounding; a rounding-level difference can flip one
            # accept/reject decision along a ray without changing its step count, and
            # the derivative of the discretised transfer then moves by O(rtol)
            per_pixel = np.abs(grad[kind, s][ok] - rich[ok]) / scale
            assert np.percentile(per_pixel, 95) < 1e-8, (kind, name, np.percentile(per_pixel, 95))



@pytest.mark.parametrize("mode", [abi.SLOW_LIGHT_DECOUPLED, abi.SLOW_LIGHT_SNAPSHOT])
@pytest.mark.parametrize("interp", [abi.SLOW_LIGHT_FLUID, abi.SLOW_LIGHT_COEFFICIENTS])
def test_emulated_mixed_slow_light_against_oracle(seq, emul, mode, interp):
    """The mixed-precision slow-light lanes (FP32 thermal coefficients mixed
    in FP32 between snapshots under coefficient interpolation, FP32 transfer,
    FP64 geodesics and time logic) against the FP64 oracle at the
    mixed-precision image bars, two frequencies on shared geodesics; the
    coarse image's net polarisation fractions are held in absolute terms."""
    emul.kp_emul_run_grmhd_slow_light_image_mixed.argtypes = emul.kp_emul_run_grmhd_slow_light_image.argtypes
    emul.kp_emul_run_grmhd_slow_light_image_mixed.restype = C.c_int32
    cells, times, model, p = seq
    freqs = [model.freq_cgs, 345e9]
    sq, keep = oracle.slow_light_sequence(cells, times, T_OBS, mode, interp)
    npix = p.camera.nx * p.camera.ny
    f = np.array(freqs, dtype=np.float64)
    arrays = oracle.alloc_image(npix)
    for n in STOKES:
        arrays[n] = np.zeros(2 * npix)
    img = oracle.image_struct(arrays)
    req = np.zeros(npix)
    w = C.c_ulonglong()
    assert emul.kp_emul_run_grmhd_slow_light_image_mixed(C.byref(p), C.byref(model), C.byref(sq),
                                                         f.ctypes.data_as(P(C.c_double)), 2, C.byref(img),
                                                         req.ctypes.data_as(P(C.c_double)), C.byref(w)) == 0
    ref, _, _ = oracle.run_grmhd_slow_light_image(p, model, cells, times, T_OBS, freqs, step_mode=mode,
                                                  interpolation=interp)
    assert np.array_equal(arrays["reason"], ref["reason"])
    g = parity.split_frequencies(arrays, 2)
    r = parity.split_frequencies(ref, 2)
    for k in range(2):
        parity.assert_mixed(g[k], r[k], f"emulated mixed slow light mode {mode} interp {interp} f{k}",
                            pol="absolute")
