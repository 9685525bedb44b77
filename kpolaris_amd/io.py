"""HDF5 input and output (libkpolaris_io.so, include/kp_io.h): iHARM GRMHD
dumps (src/grmhd/iharm_loader.cpp:262-396) and kpolaris_image v1 products
(docs/hdf5_schema.md:81-200) with the reference's comparison norms
(scripts/compare_stokes_hdf5.py:66-143)."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import abi, image

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libkpolaris_io.so"
_lib = None


class IharmHeader(C.Structure):
    """kp_iharm_header."""
    _fields_ = [("n1", C.c_int32), ("n2", C.c_int32), ("n3", C.c_int32), ("nprim", C.c_int32)] + [
        (n, C.c_double) for n in ("a", "gam", "startx1", "startx2", "startx3", "dx1", "dx2", "dx3", "hslope",
                                  "mks_smooth", "poly_alpha", "poly_xt", "r_in", "r_out", "poly_norm")]


class IharmSequence(C.Structure):
    """kp_iharm_sequence."""
    _fields_ = [("nsnap", C.c_int32), ("reserved0", C.c_int32), ("paths", C.POINTER(C.c_char_p)),
                ("header", IharmHeader)]


class ImageFileMeta(C.Structure):
    """kp_image_file_meta."""
    _fields_ = [("model", C.c_char_p), ("camera", C.c_char_p), ("coordinate", C.c_char_p), ("metric", C.c_char_p),
                ("evpa_0", C.c_char_p), ("parameter_file", C.c_char_p), ("mbh_solar", C.c_double),
                ("dsource_pc", C.c_double), ("spin", C.c_double), ("intensity_to_flux_jy_per_pixel", C.c_double),
                ("nfreq", C.c_int32), ("reserved0", C.c_int32), ("frequency_hz", C.POINTER(C.c_double))]


class StokesNorms(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("reference_sum", "candidate_sum", "sum_relative", "l1_relative",
                                           "rms_relative", "linf_over_peak")]


class ImageComparison(C.Structure):
    """kp_image_comparison."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("frequency_hz", C.c_double), ("stokes", StokesNorms * 4),
                ("worst_relative_l1", C.c_double), ("worst_linf_over_peak", C.c_double),
                ("grid_max_abs_difference", C.c_double * 2), ("discrete_mismatched", C.c_int64 * 4),
                ("discrete_max_abs_difference", C.c_int64 * 4), ("continuous_l1_relative", C.c_double * 5),
                ("continuous_linf", C.c_double * 5)]

    def as_dict(self) -> dict:
        names = ("I", "Q", "U", "V")
        disc = ("reason", "pass_a_steps", "steps", "total_steps")
        cont = ("closure_x", "closure_k", "final_null", "frame_error", "basis_identity_error")
        return {
            "shape": (self.ny, self.nx), "frequency_hz": self.frequency_hz,
            "metrics": {n: {k: getattr(self.stokes[i], k) for k, _ in StokesNorms._fields_}
                        for i, n in enumerate(names)},
            "worst_relative_l1": self.worst_relative_l1, "worst_linf_over_peak": self.worst_linf_over_peak,
            "grid_max_abs_difference": tuple(self.grid_max_abs_difference),
            "diagnostics": {**{n: {"mismatched_values": self.discrete_mismatched[i],
                                   "max_absolute_difference": self.discrete_max_abs_difference[i]}
                               for i, n in enumerate(disc)},
                            **{n: {"l1_relative": self.continuous_l1_relative[i], "linf": self.continuous_linf[i]}
                               for i, n in enumerate(cont)}},
        }


def library():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m kpolaris_amd.build_lib`")
    image.library()  # the imaging library it links against, loaded first
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    L.kp_iharm_read_header.argtypes = [C.c_char_p, P(IharmHeader)]
    L.kp_iharm_read_cells.argtypes = [C.c_char_p, P(IharmHeader), C.c_void_p]
    L.kp_iharm_write.argtypes = [C.c_char_p, P(IharmHeader), C.c_void_p]
    L.kp_iharm_load_snapshot.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64]
    L.kp_iharm_load_snapshot.restype = C.c_int32
    L.kp_iharm_apply_header.argtypes = [P(IharmHeader), P(abi.GRMHDModel)]
    L.kp_iharm_apply_header.restype = None
    L.kp_image_write_hdf5.argtypes = [C.c_char_p, P(ImageFileMeta), P(abi.PassAParams), P(abi.ImageData)]
    L.kp_image_compare_hdf5.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, P(ImageComparison)]
    for fn in (L.kp_iharm_read_header, L.kp_iharm_read_cells, L.kp_iharm_write, L.kp_image_write_hdf5,
               L.kp_image_compare_hdf5):
        fn.restype = C.c_int32
    _lib = L
    return L


def _check(st: int):
    if st != abi.KP_OK:
        raise image.KPError(st, (image.library().kp_last_error() or b"").decode())


def header_of_model(model: abi.GRMHDModel, nprim: int = 8) -> IharmHeader:
    h = IharmHeader()
    h.n1, h.n2, h.n3, h.nprim = model.n1, model.n2, model.n3, nprim
    h.a, h.gam = model.spin, model.gam
    for n in ("startx1", "startx2", "startx3", "dx1", "dx2", "dx3", "hslope", "mks_smooth", "poly_alpha",
              "poly_xt", "r_in", "r_out", "poly_norm"):
        setattr(h, n, getattr(model, n))
    return h


def write_iharm(path, model: abi.GRMHDModel, cells: np.ndarray, extra_prims: int = 0) -> None:
    """Writes the cells' primitives (float32) as an iHARM dump; extra_prims
    appends zero-filled primitives beyond the eight the reader uses."""
    n = model.n1 * model.n2 * model.n3
    prims = np.zeros((n, 8 + extra_prims), dtype=np.float32)
    prims[:, :8] = cells["prim"]
    h = header_of_model(model, 8 + extra_prims)
    _check(library().kp_iharm_write(str(path).encode(), C.byref(h), C.c_void_p(prims.ctypes.data)))


def read_iharm(path, model: abi.GRMHDModel | None = None):
    """(header, host cells with primitives, model): the model is `model` (or
    kp_grmhd_model_default) with the dump's grid and geometry."""
    from .grmhd_synth import CELL_DTYPE
    L = library()
    h = IharmHeader()
    _check(L.kp_iharm_read_header(str(path).encode(), C.byref(h)))
    cells = np.zeros(h.n1 * h.n2 * h.n3, dtype=CELL_DTYPE)
    _check(L.kp_iharm_read_cells(str(path).encode(), C.byref(h), C.c_void_p(cells.ctypes.data)))
    m = abi.GRMHDModel()
    if model is not None:
        m = abi.GRMHDModel.from_buffer_copy(model)
    else:
        image.library().kp_grmhd_model_default(C.byref(m))
    L.kp_iharm_apply_header(C.byref(h), C.byref(m))
    m.cells = None
    return h, cells, m


class IharmStream:
    """The C loader kp_iharm_load_snapshot over dump files in time order, as
    (loader, user) for image.run_grmhd_slow_light_stream (with
    derive_scalars=True).  Keep the object alive for the run."""

    def __init__(self, paths):
        self.encoded = [str(p).encode() for p in paths]
        self.array = (C.c_char_p * len(self.encoded))(*self.encoded)
        self.seq = IharmSequence()
        self.seq.nsnap = len(self.encoded)
        self.seq.paths = C.cast(self.array, C.POINTER(C.c_char_p))
        _check(library().kp_iharm_read_header(self.encoded[0], C.byref(self.seq.header)))
        self.loader = abi.SnapshotLoader(C.cast(library().kp_iharm_load_snapshot, C.c_void_p).value)
        self.user = C.cast(C.byref(self.seq), C.c_void_p)


def write_image(path, params: abi.PassAParams, arrays: dict, frequencies, *, model_name="riaf",
                coordinate="bl", evpa_0="N", flux_scale=math.nan, mbh_solar=0.0, dsource_pc=0.0,
                parameter_file=None) -> None:
    """Writes host arrays (Stokes [nfreq * npix], diagnostics [npix]) as a
    kpolaris_image v1 file."""
    f = np.ascontiguousarray(frequencies, dtype=np.float64)
    keep = {}
    img = abi.ImageData()
    for k in abi.IMAGE_FIELDS:
        a = np.ascontiguousarray(arrays[k], dtype=np.int32 if k in abi.IMAGE_INT_FIELDS else np.float64)
        keep[k] = a
        setattr(img, k, a.ctypes.data_as(C.POINTER(C.c_int32 if k in abi.IMAGE_INT_FIELDS else C.c_double)))
    meta = ImageFileMeta(model_name.encode(), b"pinhole", coordinate.encode(), b"kerr", evpa_0.encode(),
                         parameter_file.encode() if parameter_file else None, mbh_solar, dsource_pc, params.spin,
                         flux_scale, len(f), 0, f.ctypes.data_as(C.POINTER(C.c_double)))
    _check(library().kp_image_write_hdf5(str(path).encode(), C.byref(meta), C.byref(params), C.byref(img)))
    del keep


def compare_images(reference, candidate, frame: int = 0, frequency_index: int = 0) -> dict:
    out = ImageComparison()
    _check(library().kp_image_compare_hdf5(str(reference).encode(), str(candidate).encode(), frame,
                                           frequency_index, C.byref(out)))
    return out.as_dict()
