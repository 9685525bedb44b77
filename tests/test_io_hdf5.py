"""HDF5 input and output (libkpolaris_io.so) on the CPU.

* iHARM dumps: a synthetic snapshot written in the reference loader's
  layout reads back bit for bit, its derived scalars and its oracle image equal
  those of the in-memory snapshot, and the loader's header and primitive
  checks reject the corrupted fixtures of the reference's contract test
  (tests/test_hdf5_contract.py:1296-1424) with the same messages.
* kpolaris_image v1: the oracle's image of the reference's 8x8 golden case,
  written by the library, meets the reference's image contract
  (tests/test_hdf5_contract.py:610-681), including its golden fluxes; the
  comparison norms of scripts/compare_stokes_hdf5.py are zero for a file
  against itself and equal numpy's restatement for two different images.
"""
from __future__ import annotations

import json
import math
import shutil
from pathlib import Path

import numpy as np
import pytest

import oracle
from kpolaris_amd import abi, grmhd_synth as gs, image
from tests import h5lite
from oracle import host_rules as hr

GOLDEN = Path(__file__).resolve().parent / "golden"

pytest.importorskip("ctypes")


@pytest.fixture(scope="module")
def kio():
    from kpolaris_amd import io
    if not io.LIB_PATH.exists():
        pytest.skip("libkpolaris_io.so is not built")
    return io


@pytest.fixture(scope="module")
def snap():
    return gs.make_snapshot(24, 16, 16, seed=5)


def test_iharm_round_trip(tmp_path, kio, snap):
    cells, m = snap
    path = tmp_path / "dump.h5"
    kio.write_iharm(path, m, cells, extra_prims=2)
    h, back, m2 = kio.read_iharm(path)
    assert (h.n1, h.n2, h.n3, h.nprim) == (m.n1, m.n2, m.n3, 10)
    assert np.array_equal(back["prim"], cells["prim"])
    for n in ("spin", "gam", "startx1", "startx2", "startx3", "dx1", "dx2", "dx3", "hslope", "mks_smooth",
              "poly_alpha", "poly_xt", "r_in", "r_out"):
        assert getattr(m2, n) == getattr(m, n), n
    # poly_norm from the header (iharm_loader.cpp:322-326)
    assert m2.poly_norm == pytest.approx(0.5 * math.pi / (1.0 + 1.0 / (m.poly_alpha + 1.0) / m.poly_xt ** m.poly_alpha),
                                         rel=1e-15, abs=0.0)
    with h5lite.File(path) as f:
        prims, cls, size = f.dataset("/prims")
        assert prims.shape == (m.n1, m.n2, m.n3, 10) and cls == h5lite.H5T_FLOAT and size == 4


def test_iharm_sequence_loader(tmp_path, kio, snap):
    """kp_iharm_load_snapshot, the streamed slow-light scheduler's loader for
    dump files: each index reads its own file as kp_iharm_read_cells does; a
    dump on another grid, a wrong cell count and an index outside the
    sequence are rejected."""
    import ctypes as C
    cells, m = snap
    paths = []
    for k in range(3):
        c = cells.copy()
        c["prim"][:, 0] *= 1.0 + 0.25 * k
        paths.append(tmp_path / f"dump_{k}.h5")
        kio.write_iharm(paths[-1], m, c)
    stream = kio.IharmStream(paths)
    L = kio.library()
    n = m.n1 * m.n2 * m.n3
    for k in (2, 0, 1):
        got = np.zeros(n, dtype=cells.dtype)
        assert L.kp_iharm_load_snapshot(stream.user, k, C.c_void_p(got.ctypes.data), n) == abi.KP_OK
        _, ref, _ = kio.read_iharm(paths[k])
        assert np.array_equal(got["prim"], ref["prim"]) and np.all(got["derived"] == 0.0)
        assert np.array_equal(got["prim"][:, 0], (cells["prim"][:, 0] * (1.0 + 0.25 * k)).astype(np.float32))
    buf = np.zeros(n, dtype=cells.dtype)
    assert L.kp_iharm_load_snapshot(stream.user, 3, C.c_void_p(buf.ctypes.data), n) == abi.KP_INVALID_ARGUMENT
    assert L.kp_iharm_load_snapshot(stream.user, 0, C.c_void_p(buf.ctypes.data), n - 1) == abi.KP_INVALID_ARGUMENT
    m2 = abi.GRMHDModel.from_buffer_copy(m)
    m2.dx1 = m.dx1 * 1.5
    kio.write_iharm(paths[1], m2, cells)
    assert L.kp_iharm_load_snapshot(stream.user, 1, C.c_void_p(buf.ctypes.data), n) == abi.KP_INVALID_ARGUMENT
    assert b"common coordinate mapping" in image.library().kp_last_error()


def test_iharm_dump_feeds_the_same_image(tmp_path, kio, snap):
    """Derived scalars and the oracle GRMHD image from the dump equal those
    of the in-memory snapshot (the dump keeps the FP32 primitives exactly)."""
    cells, m = snap
    path = tmp_path / "dump.h5"
    kio.write_iharm(path, m, cells)
    _, back, m2 = kio.read_iharm(path, model=m)
    a = cells.copy()
    oracle.grmhd_derive_scalars(m, a)
    oracle.grmhd_derive_scalars(m2, back)
    assert np.array_equal(a["derived"], back["derived"])
    p = gs.pass_a_params(m, nx=6, ny=6)
    r1, _ = oracle.run_grmhd_image(p, m, a)
    r2, _ = oracle.run_grmhd_image(p, m2, back)
    for k in ("image_i", "image_q", "image_u", "image_v", "steps", "reason"):
        assert np.array_equal(r1[k], r2[k]), k


HEADER_CORRUPTIONS = (
    ("/header/n1", 0, "iharm header dimensions must be positive"),
    ("/header/a", math.inf, "iharm header spin must be finite and lie in [-1, 1]"),
    ("/header/gam", 1.0, "iharm header gamma must be finite and greater than 1"),
    ("/header/geom/startx1", math.nan, "iharm grid starts must be finite"),
    ("/header/geom/dx2", 0.0, "iharm grid spacings must be finite and positive"),
    ("/header/geom/dx3", -1.0, "iharm grid spacings must be finite and positive"),
    ("/header/geom/fmks/r_out", 2.0, "iharm radial bounds must be finite, positive, and increasing"),
    ("/header/geom/fmks/hslope", math.nan, "iharm FMKS/MMKS mapping parameters must be finite"),
    ("/header/geom/fmks/poly_xt", 0.0, "iharm FMKS/MMKS mapping parameters must be finite"),
)


@pytest.mark.parametrize("dataset,value,message", HEADER_CORRUPTIONS)
def test_iharm_header_corruptions(tmp_path, kio, dataset, value, message):
    """tests/test_hdf5_contract.py:1333-1405."""
    cells, m = gs.make_snapshot(8, 8, 8, seed=1)
    m.r_in = 3.0  # above the corrupted r_out of 2
    path = tmp_path / "dump.h5"
    kio.write_iharm(path, m, cells)
    with h5lite.File(path, write=True) as f:
        f.overwrite_scalar(dataset, value)
    with pytest.raises(image.KPError) as e:
        kio.read_iharm(path)
    assert e.value.status == abi.KP_INVALID_ARGUMENT
    assert message in str(e.value)


@pytest.mark.parametrize("shape,message", [
    ((8, 8, 8), "iharm /prims dataset must be rank 4"),
    ((8, 8, 8, 7), "iharm /prims dimensions do not match header or have fewer than 8 primitives"),
    ((7, 8, 8, 8), "iharm /prims dimensions do not match header or have fewer than 8 primitives"),
])
def test_iharm_prims_shape_corruptions(tmp_path, kio, shape, message):
    """tests/test_hdf5_contract.py:1296-1331."""
    cells, m = gs.make_snapshot(8, 8, 8, seed=1)
    path = tmp_path / "dump.h5"
    kio.write_iharm(path, m, cells)
    with h5lite.File(path, write=True) as f:
        f.replace_float32("/prims", np.zeros(shape, dtype=np.float32))
    with pytest.raises(image.KPError) as e:
        kio.read_iharm(path)
    assert message in str(e.value)


@pytest.mark.parametrize("index,value", [((0, 0, 0, 2), math.nan), ((0, 0, 0, 7), math.inf)])
def test_iharm_nonfinite_primitives(tmp_path, kio, index, value):
    """tests/test_hdf5_contract.py:1406-1424."""
    cells, m = gs.make_snapshot(8, 8, 8, seed=1)
    prims = cells["prim"].reshape(8, 8, 8, 8).copy()
    prims[index] = value
    path = tmp_path / "dump.h5"
    kio.write_iharm(path, m, cells)
    with h5lite.File(path, write=True) as f:
        f.replace_float32("/prims", prims)
    with pytest.raises(image.KPError) as e:
        kio.read_iharm(path)
    assert "iharm /prims first eight primitive fields must be finite" in str(e.value)


def _golden_image():
    g = json.loads((GOLDEN / "riaf_bl_8x8_golden.json").read_text())
    opt = hr.new_options()
    for kv in g["parameters"]:
        k, _, v = kv.partition("=")
        hr.apply_option(opt, k, v)
    p, m, scale, qs = hr.resolve_riaf(opt)
    img, _ = oracle.run_riaf_image(p, m)
    return g, opt, p, m, scale, qs, img


def test_image_contract_and_golden_fluxes(tmp_path, kio):
    """check_image_contract (tests/test_hdf5_contract.py:610-681) on the
    oracle's 8x8 golden image written by the library with evpa_0 = W (the
    golden case's option)."""
    g, opt, p, m, scale, qs, img = _golden_image()
    path = tmp_path / "golden.h5"
    evpa = "W" if qs == 1 else "N"
    kio.write_image(path, p, img, [m.freq_cgs], evpa_0=evpa, flux_scale=scale, mbh_solar=m.mbh_solar)
    with h5lite.File(path) as f:
        assert f.attr("/", "schema") == "kpolaris_image" and f.attr("/", "schema_version") == 1
        assert f.attr("/", "code") == "KPolaris"
        for n in ("code", "code_version", "code_revision", "code_source_dirty", "compiler_id", "compiler_version",
                  "build_type"):
            assert f.attr("/", n) == f.attr("/header", n), n
        assert f.attr("/", "code_source_dirty") in (-1, 0, 1)
        assert f.attr("/", "code_source_fingerprint") == "unknown"
        assert f.attr("/", "nfreq") == 1 and f.attr("/", "evpa_0") == evpa
        nx, ny = f.attr("/", "nx"), f.attr("/", "ny")
        for name, n in (("/grid/x", nx), ("/grid/y", ny), ("/grid/frequency_hz", 1)):
            a, cls, size = f.dataset(name)
            assert a.shape == (n,) and cls == h5lite.H5T_FLOAT and size == 8
        s = f.attr("/", "intensity_to_flux_jy_per_pixel")
        flux = f.attr("/results", "flux_I_Jy")
        I, _, _ = f.dataset("/frame_0/freq_0/I")
        for name in ("I", "Q", "U", "V", "I_inv", "Q_inv", "U_inv", "V_inv"):
            a, cls, size = f.dataset(f"/frame_0/freq_0/{name}")
            assert a.shape == (ny, nx) and cls == h5lite.H5T_FLOAT and size == 8
        ints = {"reason", "pass_a_steps", "steps", "total_steps"}
        for name in ("reason", "pass_a_steps", "steps", "total_steps", "closure_x", "closure_k", "final_null",
                     "frame_error", "det_r", "overlap_r11", "overlap_r12", "overlap_r21", "overlap_r22",
                     "basis_identity_error", "basis_rotation_angle"):
            a, cls, size = f.dataset(f"/frame_0/diagnostics/{name}")
            assert a.shape == I.shape
            assert (cls, size) == ((h5lite.H5T_INTEGER, 4) if name in ints else (h5lite.H5T_FLOAT, 8)), name
        assert math.isfinite(s) and s > 0.0 and math.isfinite(flux) and np.isfinite(I).all()
        assert abs(float(I.sum()) * s - flux) <= max(1e-10, abs(flux) * 1e-10)
        assert f.exists("/results/flux_I_Jy_by_freq") and not f.exists("/results/sum_I_nu_cgs_by_freq")
        gold = g["expected"]
        got = {"flux_I_Jy": flux, "flux_Q_Jy": f.attr("/results", "flux_Q_Jy"),
               "flux_U_Jy": f.attr("/results", "flux_U_Jy"), "flux_V_Jy": f.attr("/results", "flux_V_Jy"),
               "I_sum": float(I.sum()), "I_max": float(I.max()), "I_4_4": float(I[4, 4])}
        for k, (v, atol) in gold.items():
            assert got[k] == pytest.approx(v, rel=g["rel_tolerance"], abs=atol), k


def test_invalid_pixels_are_zero(tmp_path, kio):
    """check_invalid_pixel_contract (tests/test_hdf5_contract.py:684-692)."""
    g, opt, p, m, scale, qs, img = _golden_image()
    img = {k: v.copy() for k, v in img.items()}
    img["reason"][[3, 17, 40]] = 4
    path = tmp_path / "invalid.h5"
    kio.write_image(path, p, img, [m.freq_cgs], flux_scale=scale)
    with h5lite.File(path) as f:
        reason, _, _ = f.dataset("/frame_0/diagnostics/reason")
        bad = reason != 1
        assert bad.any()
        for name in ("I", "Q", "U", "V", "I_inv", "Q_inv", "U_inv", "V_inv"):
            a, _, _ = f.dataset(f"/frame_0/freq_0/{name}")
            assert np.count_nonzero(a[bad]) == 0


def test_compare_norms(tmp_path, kio):
    """compare_stokes_hdf5.py norms: zero against itself; against a changed
    image they equal numpy's evaluation of the script's formulas on the N-basis
    arrays (a W file's Q and U flip)."""
    g, opt, p, m, scale, qs, img = _golden_image()
    a = tmp_path / "a.h5"
    b = tmp_path / "b.h5"
    kio.write_image(a, p, img, [m.freq_cgs], evpa_0="W", flux_scale=scale)
    same = kio.compare_images(a, a)
    assert same["worst_relative_l1"] == 0.0 and same["worst_linf_over_peak"] == 0.0
    img2 = {k: v.copy() for k, v in img.items()}
    rng = np.random.default_rng(3)
    for n in ("image_i", "image_q", "image_u", "image_v"):
        img2[n] = img2[n] * (1.0 + 1e-3 * rng.standard_normal(img2[n].shape))
    img2["steps"] = img2["steps"] + (rng.random(img2["steps"].shape) < 0.2)
    kio.write_image(b, p, img2, [m.freq_cgs], evpa_0="N", flux_scale=scale)
    r = kio.compare_images(a, b)
    nu3 = m.freq_cgs ** 3
    ok = img["reason"] == 1
    for n, key, sgn_a, sgn_b in (("I", "image_i", 1, 1), ("Q", "image_q", -1, -1), ("U", "image_u", -1, -1),
                                  ("V", "image_v", 1, 1)):
        # both arrays in the N basis: the W file's Q,U flip on reading, the N
        # file was written flipped
        ref = np.where(ok, sgn_a * img[key] * nu3, 0.0)
        got = np.where(ok, sgn_b * img2[key] * nu3, 0.0)
        d = got - ref
        row = r["metrics"][n]
        assert row["l1_relative"] == pytest.approx(np.abs(d).sum() / np.abs(ref).sum(), rel=1e-12)
        assert row["rms_relative"] == pytest.approx(math.sqrt((d * d).sum() / (ref * ref).sum()), rel=1e-12)
        assert row["linf_over_peak"] == pytest.approx(np.abs(d).max() / np.abs(ref).max(), rel=1e-12)
        assert row["sum_relative"] == pytest.approx((got.sum() - ref.sum()) / ref.sum(), rel=1e-9)
    assert r["diagnostics"]["steps"]["mismatched_values"] == int(np.count_nonzero(img2["steps"] != img["steps"]))
    assert r["diagnostics"]["reason"]["mismatched_values"] == 0


def test_compare_rejects_mismatches(tmp_path, kio):
    g, opt, p, m, scale, qs, img = _golden_image()
    a = tmp_path / "a.h5"
    b = tmp_path / "b.h5"
    kio.write_image(a, p, img, [m.freq_cgs], flux_scale=scale)
    kio.write_image(b, p, img, [m.freq_cgs * 1.001], flux_scale=scale)
    with pytest.raises(image.KPError, match="frequency mismatch"):
        kio.compare_images(a, b)
    with pytest.raises(image.KPError, match="missing"):
        kio.compare_images(a, a, frequency_index=1)
