"""ctypes mirror of include/kp_image.h (the C ABI of the imaging kernel).

Field order and types follow the C header exactly; the header in turn mirrors
the reference's by-value structs (CameraParams, PassAParams,
RIAFAnalyticRadiationModel; src/camera/camera.hpp:12-25,
src/geodesic/pass_a_core.hpp:10-36, src/model/riaf.hpp:17-52).
"""
from __future__ import annotations

import ctypes as C

KP_OK = 0
KP_INVALID_ARGUMENT = 1
KP_UNSUPPORTED = 2
KP_HIP_ERROR = 3
KP_OUT_OF_MEMORY = 4
KP_TIME_SUPPORT = 5

STATUS_NAMES = {
    KP_OK: "KP_OK",
    KP_INVALID_ARGUMENT: "KP_INVALID_ARGUMENT",
    KP_UNSUPPORTED: "KP_UNSUPPORTED",
    KP_HIP_ERROR: "KP_HIP_ERROR",
    KP_OUT_OF_MEMORY: "KP_OUT_OF_MEMORY",
    KP_TIME_SUPPORT: "KP_TIME_SUPPORT",
}

# SlowLightStepMode / SlowLightInterpolation (src/model/time_interpolation.hpp:9-25)
SLOW_LIGHT_DECOUPLED = 0
SLOW_LIGHT_SNAPSHOT = 1
SLOW_LIGHT_BLOCK = 2
SLOW_LIGHT_COEFFICIENTS = 0
SLOW_LIGHT_FLUID = 1

# TerminationReason (src/common/types.hpp:37-48)
REASON_NONE = 0
REASON_REACHED_CAMERA = 1
REASON_REACHED_INNER_BOUNDARY = 2
REASON_ESCAPED_DOMAIN = 3
REASON_MAX_STEPS = 4
REASON_INVALID_STATE = 5
REASON_SLOW_LIGHT_TIME_EXHAUSTED = 6
REASON_ADAPTIVE_STEP_UNDERFLOW = 7
REASON_METRIC_TIME_EXHAUSTED = 8
REASON_REACHED_DIRECT_TURN = 9

COORD_CARTESIAN_KS = 0
COORD_BOYER_LINDQUIST = 1
COORD_SPHERICAL_KS = 2
COORD_FMKS = 3
COORD_MKS = 4

CAMERA_PINHOLE = 0
CAMERA_PARALLEL_PLANE = 1

EMISSION_THERMAL_PANDYA = 1
EMISSION_KAPPA = 2
EMISSION_POWER_LAW = 3
EMISSION_THERMAL_DEXTER = 4

PRECISION_FP64 = 0
PRECISION_MIXED = 1
# mixed precision: integration attempts that start inside this radius run in FP64
FP64_RADIUS_DEFAULT = 2.5

# include/kp_image.h KP_ABI_VERSION this mirror matches
ABI_VERSION = 7

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)


class CameraParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("radius", C.c_double), ("inclination", C.c_double),
        ("fov", C.c_double), ("fov_y", C.c_double),
        ("xspan", C.c_double), ("yspan", C.c_double),
        ("x_offset", C.c_double), ("y_offset", C.c_double),
        ("model", C.c_int32), ("reserved0", C.c_int32),
    ]


class EmissionSelection(C.Structure):
    _fields_ = [
        ("equatorial_h_over_r", C.c_double),
        ("equatorial_samples", C.c_int32),
        ("faraday_rotation", C.c_int32),
    ]


class PassAParams(C.Structure):
    _fields_ = [
        ("camera", CameraParams),
        ("emission_selection", EmissionSelection),
        ("mass", C.c_double), ("spin", C.c_double),
        ("inner_radius", C.c_double), ("outer_radius", C.c_double),
        ("step", C.c_double),
        ("max_steps", C.c_int32), ("adaptive", C.c_int32),
        ("direct_only", C.c_int32), ("coordinate_system", C.c_int32),
        ("adaptive_tolerance", C.c_double), ("min_step", C.c_double),
        ("max_step", C.c_double), ("max_radiation_step", C.c_double),
        ("max_radiation_depth", C.c_double), ("max_absorption_depth", C.c_double),
        ("max_faraday_depth", C.c_double),
        ("fmks_startx1", C.c_double), ("fmks_hslope", C.c_double),
        ("fmks_mks_smooth", C.c_double), ("fmks_poly_alpha", C.c_double),
        ("fmks_poly_xt", C.c_double), ("fmks_poly_norm", C.c_double),
    ]


class RIAFModel(C.Structure):
    _fields_ = [
        ("nth0", C.c_double), ("Te0", C.c_double), ("disk_h", C.c_double),
        ("pow_nth", C.c_double), ("pow_T", C.c_double),
        ("ne_unit", C.c_double), ("te_unit", C.c_double), ("b_unit", C.c_double),
        ("freq_cgs", C.c_double), ("mbh_solar", C.c_double),
        ("r_min", C.c_double), ("r_max", C.c_double),
        ("horizon_buffer", C.c_double), ("emission_scale", C.c_double),
        ("absorption_scale", C.c_double), ("faraday_scale", C.c_double),
        ("max_pol_frac", C.c_double), ("nonthermal_kappa", C.c_double),
        ("variable_kappa", C.c_int32), ("emission_type", C.c_int32),
        ("variable_kappa_min", C.c_double), ("variable_kappa_interp_start", C.c_double),
        ("variable_kappa_max", C.c_double), ("powerlaw_p", C.c_double),
        ("powerlaw_eta", C.c_double), ("powerlaw_gamma_min", C.c_double),
        ("powerlaw_gamma_max", C.c_double), ("powerlaw_gamma_cutoff", C.c_double),
        ("keplerian_factor", C.c_double), ("infall_factor", C.c_double),
        ("min_frequency_scale", C.c_double), ("min_sin_theta", C.c_double),
        ("min_thetae", C.c_double), ("min_radius", C.c_double),
    ]


IMAGE_REAL_FIELDS = (
    "image_i", "image_q", "image_u", "image_v",
    "closure_x", "closure_k", "final_null", "frame_error", "det_r",
    "overlap_r11", "overlap_r12", "overlap_r21", "overlap_r22",
    "basis_identity_error", "basis_rotation_angle",
)
IMAGE_INT_FIELDS = ("pass_a_steps", "steps", "reason")
IMAGE_FIELDS = IMAGE_REAL_FIELDS + IMAGE_INT_FIELDS
# per-frequency outputs of a multi-frequency image ([nfreq][npix])
STOKES_FIELDS = ("image_i", "image_q", "image_u", "image_v")


class ImageData(C.Structure):
    _fields_ = [(n, c_double_p) for n in IMAGE_REAL_FIELDS] + [
        (n, c_int32_p) for n in IMAGE_INT_FIELDS
    ]


class GRMHDModel(C.Structure):
    _fields_ = [
        ("n1", C.c_int32), ("n2", C.c_int32), ("n3", C.c_int32), ("emission_type", C.c_int32),
        ("spin", C.c_double), ("gam", C.c_double),
        ("startx1", C.c_double), ("startx2", C.c_double), ("startx3", C.c_double),
        ("dx1", C.c_double), ("dx2", C.c_double), ("dx3", C.c_double),
        ("r_in", C.c_double), ("r_out", C.c_double),
        ("hslope", C.c_double), ("mks_smooth", C.c_double), ("poly_alpha", C.c_double),
        ("poly_xt", C.c_double), ("poly_norm", C.c_double),
        ("freq_cgs", C.c_double), ("mbh_solar", C.c_double), ("M_unit", C.c_double),
        ("trat_small", C.c_double), ("trat_large", C.c_double), ("beta_crit", C.c_double),
        ("sigma_cut", C.c_double), ("sigma_cut_high", C.c_double),
        ("ne_factor", C.c_double),
        ("emission_scale", C.c_double), ("absorption_scale", C.c_double),
        ("faraday_scale", C.c_double), ("max_pol_frac", C.c_double),
        ("nonthermal_kappa", C.c_double), ("variable_kappa", C.c_int32), ("reserved_nt", C.c_int32),
        ("variable_kappa_min", C.c_double), ("variable_kappa_interp_start", C.c_double),
        ("variable_kappa_max", C.c_double), ("powerlaw_p", C.c_double), ("powerlaw_eta", C.c_double),
        ("powerlaw_gamma_min", C.c_double), ("powerlaw_gamma_max", C.c_double),
        ("powerlaw_gamma_cutoff", C.c_double),
        ("cells", C.c_void_p),
    ]


class SlowLight(C.Structure):
    _fields_ = [
        ("nsnap", C.c_int32), ("step_mode", C.c_int32), ("interpolation", C.c_int32),
        ("windows_per_block", C.c_int32),
        ("observation_time", C.c_double),
        ("times", C.POINTER(C.c_double)),
        ("cells", C.POINTER(C.c_void_p)),
    ]


# kp_snapshot_loader: (user, index, cells, ncells) -> kp_status
SnapshotLoader = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64)


class SlowLightStream(C.Structure):
    """kp_slow_light_stream (SlowLightBatchOptions residency,
    src/image/slow_light_batch_driver.hpp:13-24, with a host loader)."""
    _fields_ = [
        ("nsnap", C.c_int32), ("step_mode", C.c_int32), ("interpolation", C.c_int32),
        ("windows_per_block", C.c_int32), ("prefetch", C.c_int32), ("prefetch_snapshots", C.c_int32),
        ("derive_scalars", C.c_int32), ("reserved0", C.c_int32),
        ("observation_time", C.c_double),
        ("times", C.POINTER(C.c_double)),
        ("load", SnapshotLoader),
        ("user", C.c_void_p),
    ]


class SlowLightJob(C.Structure):
    """kp_slow_light_job (SlowLightBatchJob, src/image/slow_light_batch_driver.hpp:13-17)."""
    _fields_ = [("observation_time", C.c_double), ("first", C.c_int32), ("last", C.c_int32)]


class SlowLightSchedule(C.Structure):
    _fields_ = [("blocks", C.c_int32), ("rewinds", C.c_int32), ("loads", C.c_int32), ("slots", C.c_int32),
                ("slot_bytes", C.c_int64), ("load_ms", C.c_double), ("wait_ms", C.c_double),
                ("pass_b_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LaunchTuning(C.Structure):
    _fields_ = [("ncandidates", C.c_int32), ("block_size", C.c_int32), ("waves_per_cu", C.c_int32),
                ("sample_pixels", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double)]


class SlowLightProbe(C.Structure):
    _fields_ = [("min_x0", C.c_double), ("max_x0", C.c_double),
                ("radiating_pixels", C.c_int64), ("returned_pixels", C.c_int64)]


class BatchJob(C.Structure):
    _fields_ = [
        ("params", PassAParams), ("model", RIAFModel),
        ("radiation_substeps", C.c_int32), ("reserved0", C.c_int32),
        ("out", ImageData),
    ]


class BatchLimits(C.Structure):
    """kp_batch_limits (ImageBatchOptions, src/image/batch_driver.hpp:23-27)."""
    _fields_ = [("max_in_flight", C.c_int64), ("max_device_output_bytes", C.c_int64)]


class BatchTiming(C.Structure):
    """kp_batch_timing (ImageBatchTiming, src/image/batch_driver.hpp:29-35)."""
    _fields_ = [("allocate_seconds", C.c_double), ("kernel_seconds", C.c_double), ("copy_seconds", C.c_double),
                ("windows", C.c_int64), ("peak_device_output_bytes", C.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


BATCH_BYTES_PER_PIXEL = 15 * 8 + 3 * 4
NOALLOC_MAX_PIXELS = 512 * 512  # KP_NOALLOC_MAX_PIXELS


class RunOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("split_transport", C.c_int32),
        ("block_size", C.c_int32), ("persistent", C.c_int32),
        ("waves_per_cu", C.c_int32), ("guard_mode", C.c_int32),
        ("fp64_radius", C.c_double),
        ("stream", C.c_void_p),
        ("workspace", C.c_void_p), ("workspace_bytes", C.c_uint64),
    ]


class ImageStats(C.Structure):
    _fields_ = [
        ("pass_a_ms", C.c_double), ("pass_b_ms", C.c_double), ("kernel_ms", C.c_double),
        ("rays", C.c_int64), ("pass_a_steps", C.c_int64), ("pass_b_steps", C.c_int64),
        ("rhs_evals", C.c_int64), ("pass_a_rhs_evals", C.c_int64),
        ("coefficient_evals", C.c_int64),
        ("transfer_steps", C.c_int64), ("fp64_guard_steps", C.c_int64), ("vacuum_steps", C.c_int64),
        ("vacuum_rhs_evals", C.c_int64), ("pass_b_vacuum_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def struct_to_dict(s: C.Structure) -> dict:
    out = {}
    for name, typ in s._fields_:
        v = getattr(s, name)
        out[name] = struct_to_dict(v) if isinstance(v, C.Structure) else v
    return out


def pass_a_defaults() -> PassAParams:
    """PassAParams / CameraParams / EmissionSelection field initialisers
    (src/geodesic/pass_a_core.hpp:10-36, src/camera/camera.hpp:12-25,
    src/radiation/emission_selection.hpp:11-19)."""
    p = PassAParams()
    c = p.camera
    c.nx, c.ny = 64, 64
    c.radius = 100.0
    c.inclination = 1.04719755119659774615
    c.fov = 10.0
    c.fov_y = -1.0
    c.xspan = -1.0
    c.yspan = -1.0
    c.x_offset = 0.0
    c.y_offset = 0.0
    c.model = CAMERA_PINHOLE
    p.emission_selection.equatorial_h_over_r = 0.0
    p.emission_selection.equatorial_samples = 8
    p.emission_selection.faraday_rotation = 1
    p.mass = 1.0
    p.spin = 0.0
    p.inner_radius = 2.2
    p.outer_radius = 200.0
    p.step = 0.05
    p.max_steps = 4096
    p.adaptive = 0
    p.direct_only = 0
    p.coordinate_system = COORD_CARTESIAN_KS
    p.adaptive_tolerance = 1e-7
    p.min_step = 1e-5
    p.max_step = 1.0
    p.max_radiation_step = 1.0
    p.max_radiation_depth = 1.0
    p.max_absorption_depth = -1.0
    p.max_faraday_depth = -1.0
    p.fmks_startx1 = 0.0
    p.fmks_hslope = 0.3
    p.fmks_mks_smooth = 0.5
    p.fmks_poly_alpha = 14.0
    p.fmks_poly_xt = 0.82
    p.fmks_poly_norm = 1.0
    return p


def riaf_defaults() -> RIAFModel:
    """RIAFAnalyticRadiationModel field initialisers (src/model/riaf.hpp:17-52)."""
    m = RIAFModel()
    vals = dict(
        nth0=1.0, Te0=1.0, disk_h=0.35, pow_nth=-1.1, pow_T=-0.84, ne_unit=5.0e6,
        te_unit=1.0e11, b_unit=30.0, freq_cgs=230.0e9, mbh_solar=4.3e6, r_min=1.0,
        r_max=100.0, horizon_buffer=0.1, emission_scale=1.0, absorption_scale=1.0,
        faraday_scale=1.0, max_pol_frac=0.99, nonthermal_kappa=3.5, variable_kappa=0,
        variable_kappa_min=3.1, variable_kappa_interp_start=1e20, variable_kappa_max=7.0,
        powerlaw_p=3.25, powerlaw_eta=0.02, powerlaw_gamma_min=1e2, powerlaw_gamma_max=1e5,
        powerlaw_gamma_cutoff=1e10, keplerian_factor=1.0, infall_factor=0.0,
        emission_type=1, min_frequency_scale=1e-6, min_sin_theta=1e-4, min_thetae=1e-3,
        min_radius=1e-6,
    )
    for k, v in vals.items():
        setattr(m, k, v)
    return m


def run_options_defaults() -> RunOptions:
    o = RunOptions()
    o.precision = PRECISION_FP64
    o.split_transport = 1
    o.block_size = 0
    o.persistent = 1
    o.waves_per_cu = 0
    o.fp64_radius = FP64_RADIUS_DEFAULT
    o.stream = None
    return o


def grmhd_defaults() -> GRMHDModel:
    """GRMHDGridRadiationModel field defaults (src/model/grmhd_grid.hpp:53-89)."""
    m = GRMHDModel()
    m.emission_type = EMISSION_THERMAL_DEXTER
    m.spin, m.gam = 0.5, 4.0 / 3.0
    m.dx1 = m.dx2 = m.dx3 = 1.0
    m.r_in, m.r_out = 1.0, 1000.0
    m.hslope, m.mks_smooth, m.poly_alpha, m.poly_xt, m.poly_norm = 0.3, 0.5, 14.0, 0.82, 1.0
    m.freq_cgs, m.mbh_solar, m.M_unit = 230.0e9, 6.2e9, 3.0e25
    m.trat_small, m.trat_large, m.beta_crit = 1.0, 20.0, 1.0
    m.sigma_cut, m.sigma_cut_high = 1.0, -1.0
    m.ne_factor = 1.0
    m.emission_scale = m.absorption_scale = m.faraday_scale = 1.0
    m.max_pol_frac = 0.99
    m.nonthermal_kappa, m.variable_kappa, m.variable_kappa_min = 3.5, 0, 3.1
    m.variable_kappa_interp_start, m.variable_kappa_max = 1e20, 7.0
    m.powerlaw_p, m.powerlaw_eta = 3.25, 0.02
    m.powerlaw_gamma_min, m.powerlaw_gamma_max, m.powerlaw_gamma_cutoff = 1e2, 1e5, 1e10
    m.cells = None
    return m
