"""Python host layer over the C ABI (include/kp_image.h).

Mirrors the reference's image entry point
``run_riaf_image(pass_a, model, radiation_substeps, timing, split_transport)``
(src/image/riaf.hpp:12-17) returning the ImageHostData arrays
(src/image/result.hpp:9-31), plus the image tool's option layer
(tools/kpolaris_model_image_options.cpp) through ``Options``.

Device memory and streams come from torch (plumbing only); every computation
runs in the HIP library.  If the library is missing this module raises: there
is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

import numpy as np

from . import abi

PKG = Path(__file__).resolve().parent
# KPOLARIS_AMD_LIB selects another build of the library (a tuning build);
# the default is the in-tree build.
LIB_PATH = Path(os.environ.get("KPOLARIS_AMD_LIB") or (PKG / "_lib" / "libkpolaris_amd.so"))
HEADER = PKG.parent / "include" / "kp_image.h"

_lib = None


class KPError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def library():
    """Load libkpolaris_amd.so (built by kpolaris_amd.build_lib)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m kpolaris_amd.build_lib` "
            "(or __graft_entry__.build()); the imaging kernel has no CPU fallback")
    L = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    L.kp_abi_version.restype = C.c_int32
    L.kp_last_error.restype = C.c_char_p
    L.kp_device_count.restype = C.c_int32
    for fn in (L.kp_run_riaf_image, L.kp_run_riaf_image_host):
        fn.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), C.c_int32, P(abi.RunOptions),
                       P(abi.ImageData), P(abi.ImageStats)]
        fn.restype = C.c_int32
    L.kp_run_riaf_image_noalloc.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), C.c_int32, P(abi.RunOptions),
                                            P(abi.ImageData)]
    L.kp_run_riaf_image_noalloc.restype = C.c_int32
    for fn in (L.kp_run_riaf_multifrequency_image, L.kp_run_riaf_multifrequency_image_host):
        fn.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), C.c_int32, P(C.c_double), C.c_int32,
                       P(C.c_double), C.c_int32, P(abi.RunOptions), P(abi.ImageData), P(abi.ImageStats)]
        fn.restype = C.c_int32
    L.kp_pass_a_params_default.argtypes = [P(abi.PassAParams)]
    L.kp_riaf_model_default.argtypes = [P(abi.RIAFModel)]
    L.kp_run_options_default.argtypes = [P(abi.RunOptions)]
    L.kp_options_create.restype = C.c_void_p
    L.kp_options_destroy.argtypes = [C.c_void_p]
    L.kp_options_set.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
    L.kp_options_set.restype = C.c_int32
    L.kp_options_load_file.argtypes = [C.c_void_p, C.c_char_p]
    L.kp_options_load_file.restype = C.c_int32
    L.kp_options_resolve_riaf.argtypes = [C.c_void_p, P(abi.PassAParams), P(abi.RIAFModel),
                                          P(C.c_double), P(C.c_int32)]
    L.kp_options_resolve_riaf.restype = C.c_int32
    dp, ip = P(C.c_double), P(C.c_int32)
    L.kp_pack_image.argtypes = [dp, dp, dp, dp, ip, C.c_int64, C.c_double, C.c_int32, C.c_double,
                                dp, dp, dp, dp, dp]
    L.kp_pack_image.restype = C.c_int32
    L.kp_run_riaf_image_gradient.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), C.c_int32, P(C.c_double),
                                             P(abi.RunOptions), P(abi.ImageData), P(C.c_double), P(abi.ImageStats)]
    L.kp_run_riaf_image_gradient.restype = C.c_int32
    L.kp_run_grmhd_image.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), C.c_int32, P(abi.RunOptions),
                                     P(abi.ImageData), P(abi.ImageStats)]
    L.kp_run_grmhd_image.restype = C.c_int32
    L.kp_run_grmhd_multifrequency_image.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), C.c_int32,
                                                    P(C.c_double), C.c_int32, P(C.c_double), C.c_int32,
                                                    P(abi.RunOptions), P(abi.ImageData), P(abi.ImageStats)]
    L.kp_run_grmhd_multifrequency_image.restype = C.c_int32
    L.kp_grmhd_derive_scalars.argtypes = [P(abi.GRMHDModel), C.c_void_p]
    L.kp_grmhd_derive_scalars.restype = C.c_int32
    L.kp_grmhd_model_default.argtypes = [P(abi.GRMHDModel)]
    L.kp_batch_create.argtypes = [P(abi.BatchJob), C.c_int32, P(abi.RunOptions), P(C.c_void_p)]
    L.kp_batch_create.restype = C.c_int32
    L.kp_batch_launch.argtypes = [C.c_void_p, C.c_void_p]
    L.kp_batch_launch.restype = C.c_int32
    L.kp_batch_read_stats.argtypes = [C.c_void_p, P(abi.ImageStats)]
    L.kp_batch_read_stats.restype = C.c_int32
    L.kp_batch_destroy.argtypes = [C.c_void_p]
    L.kp_batch_destroy.restype = None
    L.kp_run_riaf_batch.argtypes = [P(abi.BatchJob), C.c_int32, P(abi.RunOptions), P(abi.ImageStats)]
    L.kp_run_riaf_batch.restype = C.c_int32
    L.kp_image_workspace_size.argtypes = [C.c_int64, C.c_int32, C.c_int32, P(abi.RunOptions), P(C.c_uint64)]
    L.kp_image_workspace_size.restype = C.c_int32
    L.kp_batch_limits_default.argtypes = [P(abi.BatchLimits)]
    L.kp_batch_limits_default.restype = None
    L.kp_run_riaf_batch_host.argtypes = [P(abi.BatchJob), C.c_int32, P(abi.BatchLimits), P(abi.RunOptions),
                                         P(abi.BatchTiming)]
    L.kp_run_riaf_batch_host.restype = C.c_int32
    L.kp_run_grmhd_slow_light_image.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLight),
                                                P(C.c_double), C.c_int32, P(abi.RunOptions), P(abi.ImageData),
                                                P(abi.ImageStats)]
    L.kp_run_grmhd_slow_light_image.restype = C.c_int32
    L.kp_run_grmhd_slow_light_image_jobs.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLight),
                                                     P(abi.SlowLightJob), C.c_int32, P(C.c_double), C.c_int32,
                                                     P(abi.RunOptions), P(abi.ImageData), P(abi.ImageStats)]
    L.kp_run_grmhd_slow_light_image_jobs.restype = C.c_int32
    L.kp_slow_light_default.argtypes = [P(abi.SlowLight)]
    L.kp_slow_light_stream_default.argtypes = [P(abi.SlowLightStream)]
    L.kp_slow_light_stream_default.restype = None
    L.kp_run_grmhd_slow_light_stream.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLightStream),
                                                 P(C.c_double), C.c_int32, P(abi.RunOptions), P(abi.ImageData),
                                                 P(abi.ImageStats), P(abi.SlowLightSchedule)]
    L.kp_run_grmhd_slow_light_stream.restype = C.c_int32
    L.kp_run_grmhd_slow_light_stream_jobs.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLightStream),
                                                      P(abi.SlowLightJob), C.c_int32, P(C.c_double), C.c_int32,
                                                      P(abi.RunOptions), P(abi.ImageData), P(abi.ImageStats),
                                                      P(abi.SlowLightSchedule)]
    L.kp_run_grmhd_slow_light_stream_jobs.restype = C.c_int32
    L.kp_slow_light_time_probe.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.RunOptions),
                                           P(abi.SlowLightProbe)]
    L.kp_run_grmhd_slow_light_image_gradient.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(abi.SlowLight),
                                                         P(C.c_double), P(abi.RunOptions), P(abi.ImageData),
                                                         P(C.c_double), P(abi.ImageStats)]
    L.kp_run_grmhd_slow_light_image_gradient.restype = C.c_int32
    L.kp_run_grmhd_image_gradient.argtypes = [P(abi.PassAParams), P(abi.GRMHDModel), P(C.c_double),
                                              P(abi.RunOptions), P(abi.ImageData), P(C.c_double),
                                              P(abi.ImageStats)]
    L.kp_run_grmhd_image_gradient.restype = C.c_int32
    L.kp_slow_light_time_probe.restype = C.c_int32
    L.kp_tune_riaf_launch.argtypes = [P(abi.PassAParams), P(abi.RIAFModel), P(C.c_double), C.c_int32, C.c_int32,
                                      P(abi.RunOptions), P(abi.LaunchTuning)]
    L.kp_tune_riaf_launch.restype = C.c_int32
    if L.kp_abi_version() != abi.ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI version {L.kp_abi_version()}, this binding expects "
                           f"{abi.ABI_VERSION}: rebuild the library")
    _lib = L
    return L


def header_functions() -> list[str]:
    """Names of every function the public header declares."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # typedef kp_status (*kp_loader)(...) declares a pointer type, not a function
    text = re.sub(r"typedef[^;]*\(\s*\*[^;]*;", "", text)
    return sorted(set(re.findall(r"\b(kp_[a-z0-9_]+)\s*\(", text)))


def _check(status: int):
    if status != abi.KP_OK:
        msg = library().kp_last_error().decode(errors="replace")
        raise KPError(status, msg)


class Options:
    """The image tool's option set (key=value, parameter files)."""

    def __init__(self, params: dict | None = None, parameter_file: str | None = None):
        self._h = library().kp_options_create()
        if not self._h:
            raise MemoryError("kp_options_create")
        if parameter_file:
            _check(library().kp_options_load_file(self._h, str(parameter_file).encode()))
        for k, v in (params or {}).items():
            self.set(k, v)

    def set(self, key: str, value) -> None:
        _check(library().kp_options_set(self._h, str(key).encode(), str(value).encode()))

    def resolve_riaf(self):
        """Returns (PassAParams, RIAFModel, intensity_to_flux_jy_per_pixel, qu_sign)."""
        p, m = abi.PassAParams(), abi.RIAFModel()
        scale, sign = C.c_double(), C.c_int32()
        _check(library().kp_options_resolve_riaf(self._h, C.byref(p), C.byref(m), C.byref(scale),
                                                 C.byref(sign)))
        return p, m, scale.value, sign.value

    def __del__(self):
        try:
            if self._h:
                library().kp_options_destroy(self._h)
        except Exception:
            pass


def pack_image(image: dict, frequency: float, qu_sign: int, scale: float):
    """Host packing: (I_nu, Q_nu, U_nu, V_nu, [flux_I, Q, U, V]) as numpy arrays."""
    L = library()
    arrs = [np.ascontiguousarray(np.asarray(image[n], dtype=np.float64))
            for n in ("image_i", "image_q", "image_u", "image_v")]
    reason = np.ascontiguousarray(np.asarray(image["reason"], dtype=np.int32))
    n = reason.size
    outs = [np.zeros(n) for _ in range(4)]
    flux = np.zeros(4)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    _check(L.kp_pack_image(*(dp(a) for a in arrs), reason.ctypes.data_as(C.POINTER(C.c_int32)), n,
                           frequency, qu_sign, scale, *(dp(a) for a in outs), dp(flux)))
    return outs[0], outs[1], outs[2], outs[3], flux


def run_options(precision: int = abi.PRECISION_FP64, persistent: bool = True, block_size: int = 0,
                waves_per_cu: int = 0, stream=None, fp64_radius: float | None = None,
                split_transport: int = 1, guard_mode: int = 0, workspace=None) -> abi.RunOptions:
    """workspace: a caller-owned device tensor (any dtype) used as the run's
    workspace (kp_run_options.workspace); None lets the library keep one per
    stream."""
    o = abi.RunOptions()
    library().kp_run_options_default(C.byref(o))
    o.precision = precision
    o.persistent = 1 if persistent else 0
    o.block_size = block_size
    o.waves_per_cu = waves_per_cu
    if fp64_radius is not None:
        o.fp64_radius = fp64_radius
    o.split_transport = split_transport
    o.guard_mode = guard_mode
    o.stream = stream
    if workspace is not None:
        o.workspace = C.c_void_p(workspace.data_ptr())
        o.workspace_bytes = workspace.numel() * workspace.element_size()
    return o


def workspace_size(npix: int, nfreq: int = 1, nsnap: int = 0, **run_kw) -> int:
    """kp_image_workspace_size: device bytes of a run's workspace."""
    b = C.c_uint64()
    _check(library().kp_image_workspace_size(npix, nfreq, nsnap, C.byref(run_options(**run_kw)), C.byref(b)))
    return int(b.value)


def alloc_device_image(npix: int, device="cuda", fields=abi.IMAGE_FIELDS, nfreq: int = 1):
    """Device output tensors; Stokes fields hold nfreq * npix values ([nfreq][npix])."""
    import torch
    out = {}
    for n in fields:
        dt = torch.int32 if n in abi.IMAGE_INT_FIELDS else torch.float64
        size = npix * nfreq if n in abi.STOKES_FIELDS else npix
        out[n] = torch.empty(size, dtype=dt, device=device)
    return out


def image_struct_from_tensors(arrays: dict) -> abi.ImageData:
    img = abi.ImageData()
    for n, t in arrays.items():
        ptr = t.data_ptr()
        if n in abi.IMAGE_INT_FIELDS:
            setattr(img, n, C.cast(C.c_void_p(ptr), C.POINTER(C.c_int32)))
        else:
            setattr(img, n, C.cast(C.c_void_p(ptr), C.POINTER(C.c_double)))
    return img


def run_riaf_image(params: abi.PassAParams, model: abi.RIAFModel, *, radiation_substeps: int = 1,
                   precision: int = abi.PRECISION_FP64, persistent: bool = True,
                   block_size: int = 0, waves_per_cu: int = 0, out: dict | None = None,
                   fields=abi.IMAGE_FIELDS, with_stats: bool = True, fp64_radius: float | None = None,
                   split_transport: int = 1, guard_mode: int = 0, workspace=None):
    """Run pass A + pass B on the current HIP device; returns (dict of device
    tensors, stats dict or None).  Work is queued on torch's current stream;
    workspace: optional caller-owned device tensor (see run_options)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    if out is None:
        out = alloc_device_image(npix, fields=fields)
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, persistent, block_size, waves_per_cu, C.c_void_p(stream), fp64_radius,
                       split_transport, guard_mode, workspace)
    st = abi.ImageStats()
    _check(library().kp_run_riaf_image(C.byref(params), C.byref(model), radiation_substeps,
                                       C.byref(opts), C.byref(img),
                                       C.byref(st) if with_stats else None))
    return out, (st.as_dict() if with_stats else None)


def run_riaf_image_noalloc(params: abi.PassAParams, model: abi.RIAFModel, out: dict, *,
                           precision: int = abi.PRECISION_FP64, persistent: bool = True, block_size: int = 0,
                           fp64_radius: float | None = None, split_transport: int = 1, guard_mode: int = 0):
    """kp_run_riaf_image_noalloc into the caller's device tensors `out` (see
    alloc_device_image) on torch's current stream; the library allocates
    nothing (its workspace is a static device array of at most
    abi.NOALLOC_MAX_PIXELS pixels)."""
    import torch
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, persistent, block_size, 0, C.c_void_p(stream), fp64_radius, split_transport,
                       guard_mode)
    _check(library().kp_run_riaf_image_noalloc(C.byref(params), C.byref(model), 1, C.byref(opts), C.byref(img)))
    return out


def run_riaf_image_host(params: abi.PassAParams, model: abi.RIAFModel, *, radiation_substeps: int = 1,
                        precision: int = abi.PRECISION_FP64, persistent: bool = True,
                        block_size: int = 0, fp64_radius: float | None = None):
    """Same computation with host (numpy) outputs; synchronous."""
    npix = params.camera.nx * params.camera.ny
    arrays = {n: np.zeros(npix, dtype=np.int32 if n in abi.IMAGE_INT_FIELDS else np.float64)
              for n in abi.IMAGE_FIELDS}
    img = abi.ImageData()
    for n, a in arrays.items():
        ct = C.c_int32 if n in abi.IMAGE_INT_FIELDS else C.c_double
        setattr(img, n, a.ctypes.data_as(C.POINTER(ct)))
    opts = run_options(precision, persistent, block_size, fp64_radius=fp64_radius)
    st = abi.ImageStats()
    _check(library().kp_run_riaf_image_host(C.byref(params), C.byref(model), radiation_substeps,
                                            C.byref(opts), C.byref(img), C.byref(st)))
    return arrays, st.as_dict()


def run_riaf_multifrequency_image(params: abi.PassAParams, model: abi.RIAFModel, frequencies,
                                  step_control_frequencies=None, *, radiation_substeps: int = 1,
                                  precision: int = abi.PRECISION_FP64, persistent: bool = True,
                                  block_size: int = 0, waves_per_cu: int = 0, out: dict | None = None,
                                  fields=abi.IMAGE_FIELDS, with_stats: bool = True,
                                  fp64_radius: float | None = None, split_transport: int = 1,
                                  guard_mode: int = 0):
    """Multi-frequency image on shared geodesics on the current HIP device.
    Stokes tensors are [nfreq * npix] (frequency-major); step control uses
    step_control_frequencies (default: the output frequencies)."""
    import torch
    f = np.ascontiguousarray(frequencies, dtype=np.float64)
    c = np.ascontiguousarray(frequencies if step_control_frequencies is None else step_control_frequencies,
                             dtype=np.float64)
    npix = params.camera.nx * params.camera.ny
    if out is None:
        out = alloc_device_image(npix, fields=fields, nfreq=len(f))
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, persistent, block_size, waves_per_cu, C.c_void_p(stream), fp64_radius,
                       split_transport, guard_mode)
    st = abi.ImageStats()
    dp = C.POINTER(C.c_double)
    _check(library().kp_run_riaf_multifrequency_image(
        C.byref(params), C.byref(model), radiation_substeps, f.ctypes.data_as(dp), len(f), c.ctypes.data_as(dp),
        len(c), C.byref(opts), C.byref(img), C.byref(st) if with_stats else None))
    return out, (st.as_dict() if with_stats else None)


def run_riaf_batch_host(jobs, *, max_in_flight: int | None = None, max_device_output_bytes: int | None = None,
                        radiation_substeps=None, precision: int = abi.PRECISION_FP64, block_size: int = 0,
                        waves_per_cu: int = 0, fp64_radius: float | None = None, guard_mode: int = 0):
    """run_image_batch_metric (src/image/batch_driver.hpp:126-187) with host
    outputs: jobs run in windows bounded by max_in_flight and the device
    output budget.  Returns (list of host image dicts, timing dict).
    radiation_substeps: one value per job (default 1)."""
    import torch
    n = len(jobs)
    arr = (abi.BatchJob * max(n, 1))()
    outputs = []
    for j, (p, m) in enumerate(jobs):
        npix = max(p.camera.nx, 0) * max(p.camera.ny, 0) if p.camera.nx < 2 ** 20 and p.camera.ny < 2 ** 20 else 0
        a = {k: np.zeros(npix, dtype=np.int32 if k in abi.IMAGE_INT_FIELDS else np.float64) for k in abi.IMAGE_FIELDS}
        outputs.append(a)
        arr[j].params = p
        arr[j].model = m
        arr[j].radiation_substeps = 1 if radiation_substeps is None else int(radiation_substeps[j])
        img = abi.ImageData()
        for k, v in a.items():
            ct = C.c_int32 if k in abi.IMAGE_INT_FIELDS else C.c_double
            setattr(img, k, v.ctypes.data_as(C.POINTER(ct)) if v.size else None)
        arr[j].out = img
    lim = abi.BatchLimits()
    library().kp_batch_limits_default(C.byref(lim))
    if max_in_flight is not None:
        lim.max_in_flight = max_in_flight
    if max_device_output_bytes is not None:
        lim.max_device_output_bytes = max_device_output_bytes
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, True, block_size, waves_per_cu, C.c_void_p(stream), fp64_radius, 1, guard_mode)
    t = abi.BatchTiming()
    _check(library().kp_run_riaf_batch_host(arr, n, C.byref(lim), C.byref(opts), C.byref(t)))
    return outputs, t.as_dict()


class Batch:
    """Many RIAF images in one persistent two-pass HIP graph (kp_batch_*).

    jobs: list of (PassAParams, RIAFModel); outputs are allocated here as
    device tensors (``self.outputs[j]``) and stay owned by this object."""

    def __init__(self, jobs, *, precision: int = abi.PRECISION_FP64, block_size: int = 0,
                 waves_per_cu: int = 0, fp64_radius: float | None = None, fields=abi.IMAGE_FIELDS,
                 guard_mode: int = 0):
        self._h = None
        n = len(jobs)
        arr = (abi.BatchJob * n)()
        self.outputs = []
        for j, (p, m) in enumerate(jobs):
            out = alloc_device_image(p.camera.nx * p.camera.ny, fields=fields)
            self.outputs.append(out)
            arr[j].params = p
            arr[j].model = m
            arr[j].radiation_substeps = 1
            arr[j].out = image_struct_from_tensors(out)
        opts = run_options(precision, True, block_size, waves_per_cu, None, fp64_radius, 1, guard_mode)
        h = C.c_void_p()
        _check(library().kp_batch_create(arr, n, C.byref(opts), C.byref(h)))
        self._h = h

    def launch(self, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        _check(library().kp_batch_launch(self._h, C.c_void_p(s)))

    def stats(self) -> dict:
        st = abi.ImageStats()
        _check(library().kp_batch_read_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h is not None:
            library().kp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_riaf_image_gradient(params: abi.PassAParams, model: abi.RIAFModel, log_scales=(0.0, 0.0, 0.0), *,
                            persistent: bool = True, block_size: int = 0, out: dict | None = None,
                            fields=abi.IMAGE_FIELDS, with_stats: bool = True):
    """Image at plasma log-scales (q_density, q_temperature, q_magnetic) and its
    derivatives d(I,Q,U,V)/dq as a [3, 4, npix] device tensor (float64)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    if out is None:
        out = alloc_device_image(npix, fields=fields)
    grad = torch.empty((3, 4, npix), dtype=torch.float64, device="cuda")
    q = np.ascontiguousarray(log_scales, dtype=np.float64)
    if q.shape != (3,):
        raise ValueError("log_scales must hold 3 values")
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(abi.PRECISION_FP64, persistent, block_size, 0, C.c_void_p(stream))
    st = abi.ImageStats()
    dp = C.POINTER(C.c_double)
    _check(library().kp_run_riaf_image_gradient(
        C.byref(params), C.byref(model), 1, q.ctypes.data_as(dp), C.byref(opts), C.byref(img),
        C.cast(C.c_void_p(grad.data_ptr()), dp), C.byref(st) if with_stats else None))
    return out, grad, (st.as_dict() if with_stats else None)


class GRMHDGrid:
    """A GRMHD snapshot resident on the device: kp_grmhd_cell records
    (72 bytes each) in one uint8 tensor, with the model parameters.

    cells: numpy structured array of kp_grmhd_cell records (primitives set);
    the derived scalars are computed on the device (kp_grmhd_derive_scalars)."""

    def __init__(self, model: abi.GRMHDModel, cells: np.ndarray, derive: bool = True):
        import torch
        assert cells.dtype.itemsize == 72 and cells.size == model.n1 * model.n2 * model.n3
        raw = np.ascontiguousarray(cells).view(np.uint8).reshape(-1)
        self.buffer = torch.from_numpy(raw.copy()).to("cuda")
        self.model = abi.GRMHDModel.from_buffer_copy(model)
        self.model.cells = C.c_void_p(self.buffer.data_ptr())
        if derive:
            self.derive()

    def derive(self, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        _check(library().kp_grmhd_derive_scalars(C.byref(self.model), C.c_void_p(s)))

    def host_cells(self) -> np.ndarray:
        from kpolaris_amd.grmhd_synth import CELL_DTYPE
        return self.buffer.cpu().numpy().view(CELL_DTYPE)


def run_grmhd_image(params: abi.PassAParams, grid: "GRMHDGrid", frequencies=None, step_control_frequencies=None,
                    *, precision: int = abi.PRECISION_FP64, persistent: bool = True, block_size: int = 0,
                    waves_per_cu: int = 0, split_transport: int = 1, out: dict | None = None,
                    fields=abi.IMAGE_FIELDS, with_stats: bool = True):
    """GRMHD image on FMKS geodesics; frequencies None = the model's freq_cgs.
    Stokes tensors are [nfreq * npix].  precision PRECISION_MIXED evaluates
    the thermal coefficients and the transfer in FP32 (geodesics and the
    grid sampling stay FP64)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    f = np.ascontiguousarray([grid.model.freq_cgs] if frequencies is None else frequencies, dtype=np.float64)
    c = np.ascontiguousarray(f if step_control_frequencies is None else step_control_frequencies, dtype=np.float64)
    if out is None:
        out = alloc_device_image(npix, fields=fields, nfreq=len(f))
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, persistent, block_size, waves_per_cu, C.c_void_p(stream), None,
                       split_transport)
    st = abi.ImageStats()
    dp = C.POINTER(C.c_double)
    _check(library().kp_run_grmhd_multifrequency_image(
        C.byref(params), C.byref(grid.model), 1, f.ctypes.data_as(dp), len(f), c.ctypes.data_as(dp), len(c),
        C.byref(opts), C.byref(img), C.byref(st) if with_stats else None))
    return out, (st.as_dict() if with_stats else None)


def _sequence(grids, times, observation_time, step_mode, interpolation, windows_per_block=0):
    t = np.ascontiguousarray(times, dtype=np.float64)
    ptrs = (C.c_void_p * len(grids))(*[g.model.cells for g in grids])
    seq = abi.SlowLight()
    library().kp_slow_light_default(C.byref(seq))
    seq.nsnap = len(grids)
    seq.step_mode = step_mode
    seq.interpolation = interpolation
    seq.windows_per_block = windows_per_block
    seq.observation_time = observation_time
    seq.times = t.ctypes.data_as(C.POINTER(C.c_double))
    seq.cells = C.cast(ptrs, C.POINTER(C.c_void_p))
    return seq, (t, ptrs)


def run_grmhd_slow_light_image_gradient(params: abi.PassAParams, grids, times, observation_time: float,
                                        log_scales=(0.0, 0.0), *, step_mode: int = abi.SLOW_LIGHT_DECOUPLED,
                                        interpolation: int = abi.SLOW_LIGHT_FLUID, persistent: bool = True,
                                        split_transport: int = 1, out: dict | None = None,
                                        fields=abi.IMAGE_FIELDS, with_stats: bool = True):
    """Slow-light image at one frequency (the model's) and its derivatives
    along the fixed geodesics with respect to (q_density, R_high,
    q_magnetic); returns (outputs, grad [3, 4, npix] device tensor, stats).
    log_scales = (q_density, q_magnetic) is the evaluation point."""
    import torch
    npix = params.camera.nx * params.camera.ny
    model = grids[0].model
    seq, keep = _sequence(grids, times, observation_time, step_mode, interpolation)
    if out is None:
        out = alloc_device_image(npix, fields=fields)
    img = image_struct_from_tensors(out)
    grad = torch.empty((3, 4, npix), dtype=torch.float64, device="cuda")
    q = (C.c_double * 2)(*log_scales)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(abi.PRECISION_FP64, persistent, 0, 0, C.c_void_p(stream), None, split_transport)
    st = abi.ImageStats()
    _check(library().kp_run_grmhd_slow_light_image_gradient(
        C.byref(params), C.byref(model), C.byref(seq), q, C.byref(opts), C.byref(img),
        C.cast(C.c_void_p(grad.data_ptr()), C.POINTER(C.c_double)), C.byref(st) if with_stats else None))
    return out, grad, (st.as_dict() if with_stats else None)


def tune_riaf_launch(params: abi.PassAParams, model: abi.RIAFModel, frequencies=None, *,
                     precision: int = abi.PRECISION_FP64, persistent: bool = True, sample_divisor: int = 4,
                     split_transport: int = 1, guard_mode: int = 0) -> dict:
    """Time the image kernels over launch shapes on a reduced copy of the
    image (kp_tune_riaf_launch); returns the chosen block_size and
    waves_per_cu with the timings.  Pass the result to run_riaf_image."""
    import torch
    f = None if frequencies is None else np.ascontiguousarray(frequencies, dtype=np.float64)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, persistent, 0, 0, C.c_void_p(stream), None, split_transport, guard_mode)
    rep = abi.LaunchTuning()
    _check(library().kp_tune_riaf_launch(
        C.byref(params), C.byref(model), None if f is None else f.ctypes.data_as(C.POINTER(C.c_double)),
        0 if f is None else len(f), sample_divisor, C.byref(opts), C.byref(rep)))
    return {n: getattr(rep, n) for n, _ in abi.LaunchTuning._fields_}


def run_grmhd_image_gradient(params: abi.PassAParams, grid: "GRMHDGrid", log_scales=(0.0, 0.0), *,
                             persistent: bool = True, split_transport: int = 1, out: dict | None = None,
                             fields=abi.IMAGE_FIELDS, with_stats: bool = True):
    """Static GRMHD image at one frequency (the model's) and its derivatives
    along the fixed geodesics with respect to (q_density, R_high,
    q_magnetic); returns (outputs, grad [3, 4, npix] device tensor, stats)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    if out is None:
        out = alloc_device_image(npix, fields=fields)
    img = image_struct_from_tensors(out)
    grad = torch.empty((3, 4, npix), dtype=torch.float64, device="cuda")
    q = (C.c_double * 2)(*log_scales)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(abi.PRECISION_FP64, persistent, 0, 0, C.c_void_p(stream), None, split_transport)
    st = abi.ImageStats()
    _check(library().kp_run_grmhd_image_gradient(
        C.byref(params), C.byref(grid.model), q, C.byref(opts), C.byref(img),
        C.cast(C.c_void_p(grad.data_ptr()), C.POINTER(C.c_double)), C.byref(st) if with_stats else None))
    return out, grad, (st.as_dict() if with_stats else None)


def run_grmhd_slow_light_image(params: abi.PassAParams, grids, times, observation_time: float, frequencies=None,
                               *, step_mode: int = abi.SLOW_LIGHT_DECOUPLED,
                               interpolation: int = abi.SLOW_LIGHT_FLUID, persistent: bool = True,
                               block_size: int = 0, waves_per_cu: int = 0, split_transport: int = 1,
                               windows_per_block: int = 0, precision: int = abi.PRECISION_FP64,
                               out: dict | None = None, fields=abi.IMAGE_FIELDS, with_stats: bool = True):
    """Slow-light GRMHD image over a snapshot sequence (grids: GRMHDGrid per
    time, all on one grid geometry; times strictly increasing) sampled at
    retarded time observation_time + x^0.  windows_per_block > 0 runs pass B
    in blocks of that many snapshot pairs (the reference's batch scheduler).
    precision PRECISION_MIXED: FP32 thermal coefficients and transfer.
    Raises KPError(KP_TIME_SUPPORT) when decoupled stepping needs a time
    outside the sequence (the image is still written into `out`)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    model = grids[0].model
    f = np.ascontiguousarray([model.freq_cgs] if frequencies is None else frequencies, dtype=np.float64)
    seq, keep = _sequence(grids, times, observation_time, step_mode, interpolation, windows_per_block)
    if out is None:
        out = alloc_device_image(npix, fields=fields, nfreq=len(f))
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, persistent, block_size, waves_per_cu, C.c_void_p(stream), None,
                       split_transport)
    st = abi.ImageStats()
    _check(library().kp_run_grmhd_slow_light_image(
        C.byref(params), C.byref(model), C.byref(seq), f.ctypes.data_as(C.POINTER(C.c_double)), len(f),
        C.byref(opts), C.byref(img), C.byref(st) if with_stats else None))
    return out, (st.as_dict() if with_stats else None)


def run_grmhd_slow_light_image_jobs(params: abi.PassAParams, grids, times, jobs, frequencies=None, *,
                                    step_mode: int = abi.SLOW_LIGHT_DECOUPLED,
                                    interpolation: int = abi.SLOW_LIGHT_FLUID, windows_per_block: int = 0,
                                    block_size: int = 0, waves_per_cu: int = 0,
                                    precision: int = abi.PRECISION_FP64, fields=abi.IMAGE_FIELDS,
                                    with_stats: bool = True):
    """Images at several observation times over resident snapshots
    (kp_run_grmhd_slow_light_image_jobs): `jobs` is a list of
    (observation_time, first, last).  Returns (list of outputs, one per job,
    stats summed over the jobs)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    model = grids[0].model
    f = np.ascontiguousarray([model.freq_cgs] if frequencies is None else frequencies, dtype=np.float64)
    seq, keep = _sequence(grids, times, 0.0, step_mode, interpolation, windows_per_block)
    table = (abi.SlowLightJob * max(1, len(jobs)))()
    for j, (tobs, first, last) in enumerate(jobs):
        table[j] = abi.SlowLightJob(float(tobs), int(first), int(last))
    outs = [alloc_device_image(npix, fields=fields, nfreq=len(f)) for _ in jobs]
    imgs = (abi.ImageData * max(1, len(jobs)))(*[image_struct_from_tensors(o) for o in outs])
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, True, block_size, waves_per_cu, C.c_void_p(stream))
    st = abi.ImageStats()
    _check(library().kp_run_grmhd_slow_light_image_jobs(
        C.byref(params), C.byref(model), C.byref(seq), table, len(jobs), f.ctypes.data_as(C.POINTER(C.c_double)),
        len(f), C.byref(opts), imgs, C.byref(st) if with_stats else None))
    return outs, (st.as_dict() if with_stats else None)


def host_cell_loader(snapshots):
    """A kp_snapshot_loader over host snapshots (numpy arrays of 72-byte
    kp_grmhd_cell records, one per time); returns the ctypes callback, which
    the caller keeps alive for the run."""
    raws = [np.ascontiguousarray(c).view(np.uint8).reshape(-1) for c in snapshots]

    def load(user, index, cells, ncells):
        try:
            raw = raws[index]
            if raw.size != int(ncells) * 72:
                return abi.KP_INVALID_ARGUMENT
            C.memmove(cells, raw.ctypes.data, raw.size)
            return abi.KP_OK
        except Exception:  # an exception cannot cross the C boundary
            return abi.KP_INVALID_ARGUMENT

    return abi.SnapshotLoader(load)


def run_grmhd_slow_light_stream(params: abi.PassAParams, model: abi.GRMHDModel, times, observation_time: float,
                                loader, frequencies=None, *, step_mode: int = abi.SLOW_LIGHT_DECOUPLED,
                                interpolation: int = abi.SLOW_LIGHT_FLUID, windows_per_block: int = 1,
                                prefetch: bool = True, prefetch_snapshots: int = 0, derive_scalars: bool = False,
                                user=None, block_size: int = 0, waves_per_cu: int = 0,
                                precision: int = abi.PRECISION_FP64, out: dict | None = None,
                                fields=abi.IMAGE_FIELDS, with_stats: bool = True):
    """Slow-light GRMHD image of a sequence streamed from host storage
    (kp_run_grmhd_slow_light_stream): `loader` is an abi.SnapshotLoader
    (host_cell_loader, or a C loader such as kp_iharm_load_snapshot with its
    `user` pointer).  Returns (outputs, stats, schedule)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    f = np.ascontiguousarray([model.freq_cgs] if frequencies is None else frequencies, dtype=np.float64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    q = abi.SlowLightStream()
    library().kp_slow_light_stream_default(C.byref(q))
    q.nsnap = len(t)
    q.step_mode = step_mode
    q.interpolation = interpolation
    q.windows_per_block = windows_per_block
    q.prefetch = 1 if prefetch else 0
    q.prefetch_snapshots = prefetch_snapshots
    q.derive_scalars = 1 if derive_scalars else 0
    q.observation_time = observation_time
    q.times = t.ctypes.data_as(C.POINTER(C.c_double))
    q.load = loader
    q.user = user
    if out is None:
        out = alloc_device_image(npix, fields=fields, nfreq=len(f))
    img = image_struct_from_tensors(out)
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, True, block_size, waves_per_cu, C.c_void_p(stream))
    st = abi.ImageStats()
    rep = abi.SlowLightSchedule()
    _check(library().kp_run_grmhd_slow_light_stream(
        C.byref(params), C.byref(model), C.byref(q), f.ctypes.data_as(C.POINTER(C.c_double)), len(f),
        C.byref(opts), C.byref(img), C.byref(st) if with_stats else None, C.byref(rep)))
    return out, (st.as_dict() if with_stats else None), rep.as_dict()


def run_grmhd_slow_light_stream_jobs(params: abi.PassAParams, model: abi.GRMHDModel, times, jobs, loader,
                                     frequencies=None, *, step_mode: int = abi.SLOW_LIGHT_DECOUPLED,
                                     interpolation: int = abi.SLOW_LIGHT_FLUID, windows_per_block: int = 1,
                                     prefetch: bool = True, prefetch_snapshots: int = 0, derive_scalars: bool = False,
                                     user=None, block_size: int = 0, waves_per_cu: int = 0,
                                     precision: int = abi.PRECISION_FP64, fields=abi.IMAGE_FIELDS,
                                     with_stats: bool = True):
    """Images at several observation times from one pass over a streamed
    sequence (kp_run_grmhd_slow_light_stream_jobs): `jobs` is a list of
    (observation_time, first, last) with the inclusive snapshot support of
    each.  Returns (list of outputs, one per job, stats summed over the jobs,
    schedule)."""
    import torch
    npix = params.camera.nx * params.camera.ny
    f = np.ascontiguousarray([model.freq_cgs] if frequencies is None else frequencies, dtype=np.float64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    q = abi.SlowLightStream()
    library().kp_slow_light_stream_default(C.byref(q))
    q.nsnap = len(t)
    q.step_mode = step_mode
    q.interpolation = interpolation
    q.windows_per_block = windows_per_block
    q.prefetch = 1 if prefetch else 0
    q.prefetch_snapshots = prefetch_snapshots
    q.derive_scalars = 1 if derive_scalars else 0
    q.times = t.ctypes.data_as(C.POINTER(C.c_double))
    q.load = loader
    q.user = user
    table = (abi.SlowLightJob * max(1, len(jobs)))()
    for j, (tobs, first, last) in enumerate(jobs):
        table[j] = abi.SlowLightJob(float(tobs), int(first), int(last))
    outs = [alloc_device_image(npix, fields=fields, nfreq=len(f)) for _ in jobs]
    imgs = (abi.ImageData * max(1, len(jobs)))(*[image_struct_from_tensors(o) for o in outs])
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(precision, True, block_size, waves_per_cu, C.c_void_p(stream))
    st = abi.ImageStats()
    rep = abi.SlowLightSchedule()
    _check(library().kp_run_grmhd_slow_light_stream_jobs(
        C.byref(params), C.byref(model), C.byref(q), table, len(jobs), f.ctypes.data_as(C.POINTER(C.c_double)),
        len(f), C.byref(opts), imgs, C.byref(st) if with_stats else None, C.byref(rep)))
    return outs, (st.as_dict() if with_stats else None), rep.as_dict()


def slow_light_time_probe(params: abi.PassAParams, model: abi.GRMHDModel, *, persistent: bool = True) -> dict:
    """x^0 range of the radiating samples of an image (the fluid times a
    slow-light image needs are observation_time + [min_x0, max_x0]); runs
    synchronously on the current device."""
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    opts = run_options(abi.PRECISION_FP64, persistent, 0, 0, C.c_void_p(stream))
    out = abi.SlowLightProbe()
    _check(library().kp_slow_light_time_probe(C.byref(params), C.byref(model), C.byref(opts), C.byref(out)))
    return {n: getattr(out, n) for n, _ in abi.SlowLightProbe._fields_}
