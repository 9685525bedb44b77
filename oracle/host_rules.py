"""TEST INFRASTRUCTURE ONLY: the reference image tool's host rules, restated.

These are the option semantics the tests use to turn a parameter set into the
kernel's inputs and the kernel's outputs into fluxes, independently of the
product's C++ option layer (which the tests compare against this module):

* option defaults       tools/kpolaris_model_image_options.hpp:18-99,177-178
* key semantics         tools/kpolaris_model_image_options.cpp:489-600
* camera extents        tools/kpolaris_model_image_options.cpp:119-180
* inner/outer radius    tools/kpolaris_model_image.cpp:4925-4927, 4959-4961
* PassAParams assembly  tools/kpolaris_model_image.cpp:3940-3985
* flux scale            tools/kpolaris_model_image.cpp:558-595
* packing               tools/kpolaris_model_image.cpp:2668-2692, tools/kpolaris_evpa.hpp:15-17
"""
from __future__ import annotations

import math

import numpy as np

from kpolaris_amd import abi

PI = 3.141592653589793238462643383279502884

DEFAULTS = {
    "model": "riaf", "evpa_0": "N", "camera": "pinhole", "coordinate": "cartesian_ks",
    "nx": 96, "ny": 96, "radius": 35.0, "inclination": 1.04719755119659774615,
    "fov": 0.08, "fovy": -1.0, "xspan": -1.0, "yspan": -1.0, "dsource_pc": -1.0,
    "fovx_dsource": -1.0, "fovy_dsource": -1.0, "spin": 0.9375, "inner_radius": -1.0,
    "inner_radius_explicit": False, "outer_radius": -1.0, "step": 0.025, "max_steps": 4096,
    "adaptive": 1, "direct_only": 0, "equatorial_h_over_r": 0.0, "equatorial_samples": 8,
    "faraday_rotation": 1, "adaptive_tolerance": 1e-10, "min_step": 1e-5, "max_step": 1.0,
    "max_radiation_step": 1.0, "max_radiation_depth": 1.0, "max_absorption_depth": 1.0,
    "max_faraday_depth": 4.0, "freq": 230.0e9, "riaf_r_min": 1.0, "riaf_r_max": 100.0,
    "riaf_nth0": 1.0, "riaf_Te0": 1.0, "riaf_disk_h": 0.35, "riaf_pow_nth": -1.1,
    "riaf_pow_T": -0.84, "riaf_ne_unit": 5.0e6, "riaf_te_unit": 1.0e11,
    "riaf_mbh_solar": 4.3e6, "riaf_keplerian_factor": 1.0, "riaf_infall_factor": 0.0,
    "emission_type": 0, "x_offset": 0.0, "y_offset": 0.0, "use_pinhole_pixel_bias": 0,
    "pinhole_pixel_bias": -0.01, "radiation_substeps": 2, "split_transport": 1,
}

FLOAT_KEYS = {
    "radius", "fov", "fovy", "xspan", "yspan", "dsource_pc", "fovx_dsource", "fovy_dsource",
    "spin", "inner_radius", "outer_radius", "step", "equatorial_h_over_r",
    "adaptive_tolerance", "min_step", "max_step", "max_radiation_step",
    "max_radiation_depth", "max_absorption_depth", "max_faraday_depth", "freq",
    "riaf_r_min", "riaf_r_max", "riaf_nth0", "riaf_Te0", "riaf_disk_h", "riaf_pow_nth",
    "riaf_pow_T", "riaf_ne_unit", "riaf_te_unit", "riaf_mbh_solar",
    "riaf_keplerian_factor", "riaf_infall_factor", "x_offset", "y_offset",
    "pinhole_pixel_bias",
}
INT_KEYS = {
    "nx", "ny", "max_steps", "adaptive", "direct_only", "equatorial_samples",
    "faraday_rotation", "emission_type", "use_pinhole_pixel_bias", "radiation_substeps",
    "split_transport",
}
STRING_KEYS = {"model", "evpa_0", "camera", "coordinate"}
IGNORED_KEYS = {"format", "output", "parameter_output", "timing", "analysis_mode"}


def new_options() -> dict:
    return dict(DEFAULTS)


def apply_option(opt: dict, key: str, value: str) -> None:
    """apply_option key semantics (tools/kpolaris_model_image_options.cpp:489-600)."""
    key = key.strip().lstrip("-").replace("-", "_")
    value = value.strip()
    if key in ("dsource", "dsource_pc"):
        opt["dsource_pc"] = float(value)
    elif key == "inclination_deg":
        opt["inclination"] = float(value) * PI / 180.0
    elif key == "inclination_rad":
        opt["inclination"] = float(value)
    elif key == "fov":
        opt["fov"] = float(value)
        opt["fovy"] = -1.0
        opt["xspan"] = -1.0
        opt["yspan"] = -1.0
        opt["fovx_dsource"] = -1.0
        opt["fovy_dsource"] = -1.0
    elif key in ("fovy", "fov_y"):
        opt["fovy"] = float(value)
        opt["fovy_dsource"] = -1.0
    elif key == "xspan":
        opt["xspan"] = float(value)
        opt["fovx_dsource"] = -1.0
    elif key == "yspan":
        opt["yspan"] = float(value)
        opt["fovy_dsource"] = -1.0
    elif key in ("fovx_dsource", "fovx_muas", "fovx_uas"):
        opt["fovx_dsource"] = float(value)
        opt["xspan"] = -1.0
    elif key in ("fovy_dsource", "fovy_muas", "fovy_uas"):
        opt["fovy_dsource"] = float(value)
        opt["yspan"] = -1.0
    elif key == "inner_radius":
        opt["inner_radius"] = float(value)
        opt["inner_radius_explicit"] = True
    elif key in ("max_radiation_depth", "radiation_sampling_depth"):
        v = float(value)
        opt["max_radiation_depth"] = v
        opt["max_absorption_depth"] = v
        opt["max_faraday_depth"] = v
    elif key == "evpa_0":
        v = value.upper()
        opt["evpa_0"] = {"NORTH": "N", "WEST": "W"}.get(v, v)
    elif key == "coordinate":
        aliases = {"bl": "boyer_lindquist", "boyer_lindquist": "boyer_lindquist"}
        opt["coordinate"] = aliases.get(value.lower(), value.lower())
    elif key in FLOAT_KEYS:
        opt[key] = float(value)
    elif key in INT_KEYS:
        opt[key] = int(value)
    elif key in STRING_KEYS:
        opt[key] = value
    elif key in IGNORED_KEYS:
        pass
    else:
        raise KeyError(f"unsupported option {key!r}")


def parse_parameter_text(text: str, opt: dict | None = None) -> dict:
    opt = new_options() if opt is None else opt
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        key, _, value = line.partition("=")
        apply_option(opt, key, value)
    return opt


def length_unit_cgs_from_mbh_solar(mbh: float) -> float:
    """tools/kpolaris_model_image_options.cpp:119-124 (CGS constants differ from the model's)."""
    return 6.6742e-8 * mbh * 1.989e33 / (2.99792458e10 * 2.99792458e10)


def resolve_camera_extents(opt: dict) -> None:
    """resolve_camera_extents (tools/kpolaris_model_image_options.cpp:136-180)."""
    pc_cgs = 3.085678e18
    muas_per_rad = 2.06265e11
    source_fov = opt["fovx_dsource"] > 0 or opt["fovy_dsource"] > 0
    mbh = opt["riaf_mbh_solar"]
    if source_fov:
        if opt["dsource_pc"] <= 0:
            raise ValueError("fovx_dsource/fovy_dsource require dsource in parsec")
        if opt["fovx_dsource"] <= 0 < opt["fovy_dsource"]:
            opt["fovx_dsource"] = opt["fovy_dsource"]
        if opt["fovy_dsource"] <= 0 < opt["fovx_dsource"]:
            opt["fovy_dsource"] = opt["fovx_dsource"]
        l_unit = length_unit_cgs_from_mbh_solar(mbh)
        fov_to_m = opt["dsource_pc"] * pc_cgs / l_unit / muas_per_rad
        opt["image_width_x"] = opt["fovx_dsource"] * fov_to_m
        opt["image_width_y"] = opt["fovy_dsource"] * fov_to_m
        opt["fov"] = opt["image_width_x"] / opt["radius"]
        opt["fovy"] = opt["image_width_y"] / opt["radius"]
        if opt["camera"] == "parallel_plane":
            opt["xspan"] = opt["image_width_x"] / 2.0
            opt["yspan"] = opt["image_width_y"] / 2.0
        else:
            opt["xspan"] = -1.0
            opt["yspan"] = -1.0
        return
    if opt["camera"] != "parallel_plane" and opt["xspan"] > 0:
        opt["fov"] = 2.0 * opt["xspan"] / opt["radius"]
        if opt["yspan"] > 0:
            opt["fovy"] = 2.0 * opt["yspan"] / opt["radius"]
    if opt["fovy"] <= 0:
        opt["fovy"] = opt["fov"]
    if opt["camera"] == "parallel_plane":
        if opt["xspan"] > 0 and opt["yspan"] <= 0:
            opt["yspan"] = opt["xspan"] * opt["ny"] / opt["nx"]
        opt["image_width_x"] = 2.0 * opt["xspan"] if opt["xspan"] > 0 else opt["fov"]
        opt["image_width_y"] = 2.0 * opt["yspan"] if opt["yspan"] > 0 else opt["fovy"]
    else:
        opt["image_width_x"] = opt["fov"] * opt["radius"]
        opt["image_width_y"] = opt["fovy"] * opt["radius"]


def resolve_riaf(opt: dict):
    """Camera extents, radii, PassAParams and RIAF model as the image tool
    builds them for model=riaf.  Returns (pass_a, riaf, flux_scale, qu_sign)."""
    opt = dict(opt)
    resolve_camera_extents(opt)
    spin = opt["spin"]
    horizon = 1.0 + math.sqrt(max(0.0, 1.0 - spin * spin))
    if opt["inner_radius"] <= 0:
        opt["inner_radius"] = horizon * 1.05
    if opt["outer_radius"] <= 0:
        opt["outer_radius"] = opt["riaf_r_max"] * 1.15
    p = abi.pass_a_defaults()
    c = p.camera
    c.nx, c.ny = opt["nx"], opt["ny"]
    c.radius = opt["radius"]
    c.inclination = opt["inclination"]
    c.fov = opt["fov"]
    c.fov_y = opt["fovy"] if opt["fovy"] > 0 else opt["fov"]
    c.x_offset = opt["x_offset"]
    c.y_offset = opt["y_offset"]
    if opt["camera"] == "pinhole" and opt["use_pinhole_pixel_bias"]:
        c.x_offset += opt["pinhole_pixel_bias"] / float(opt["nx"]) * opt["fov"]
    c.xspan = opt["xspan"] if opt["camera"] == "parallel_plane" else -1.0
    if opt["camera"] == "parallel_plane" and opt["xspan"] > 0:
        c.yspan = opt["yspan"] if opt["yspan"] > 0 else opt["xspan"] * opt["ny"] / opt["nx"]
    else:
        c.yspan = -1.0
    c.model = abi.CAMERA_PARALLEL_PLANE if opt["camera"] == "parallel_plane" else abi.CAMERA_PINHOLE
    coord = {"boyer_lindquist": abi.COORD_BOYER_LINDQUIST, "cartesian_ks": abi.COORD_CARTESIAN_KS,
             "spherical_ks": abi.COORD_SPHERICAL_KS, "fmks": abi.COORD_FMKS, "mks": abi.COORD_MKS}
    p.coordinate_system = coord[opt["coordinate"]]
    p.mass = 1.0
    p.spin = spin
    p.inner_radius = opt["inner_radius"]
    p.outer_radius = opt["outer_radius"]
    p.step = opt["step"]
    p.max_steps = opt["max_steps"]
    p.adaptive = opt["adaptive"]
    p.direct_only = opt["direct_only"]
    p.emission_selection.equatorial_h_over_r = opt["equatorial_h_over_r"]
    p.emission_selection.equatorial_samples = opt["equatorial_samples"]
    p.emission_selection.faraday_rotation = opt["faraday_rotation"]
    for k in ("adaptive_tolerance", "min_step", "max_step", "max_radiation_step",
              "max_radiation_depth", "max_absorption_depth", "max_faraday_depth"):
        setattr(p, k, opt[k])
    m = abi.riaf_defaults()
    m.freq_cgs = opt["freq"]
    m.r_min = opt["riaf_r_min"]
    m.r_max = opt["riaf_r_max"]
    m.nth0 = opt["riaf_nth0"]
    m.Te0 = opt["riaf_Te0"]
    m.disk_h = opt["riaf_disk_h"]
    m.pow_nth = opt["riaf_pow_nth"]
    m.pow_T = opt["riaf_pow_T"]
    m.ne_unit = opt["riaf_ne_unit"]
    m.te_unit = opt["riaf_te_unit"]
    m.mbh_solar = opt["riaf_mbh_solar"]
    m.keplerian_factor = opt["riaf_keplerian_factor"]
    m.infall_factor = opt["riaf_infall_factor"]
    # compiled RIAF fit is Symphony/Pandya (CMakeLists.txt:188); a conflicting
    # runtime request is an error there (tools/kpolaris_model_image.cpp:504-516).
    if opt["emission_type"] not in (0, 1):
        raise ValueError("emission_type conflicts with the compiled RIAF fit (Pandya)")
    m.emission_type = abi.EMISSION_THERMAL_PANDYA
    scale = flux_scale(opt)
    qu_sign = -1 if opt["evpa_0"] == "N" else 1
    return p, m, scale, qu_sign


def flux_scale(opt: dict) -> float:
    """compute_flux_scale (tools/kpolaris_model_image.cpp:568-588)."""
    pc_cgs = 3.085678e18
    jy = 1.0e-23
    l_unit = length_unit_cgs_from_mbh_solar(opt["riaf_mbh_solar"])
    d_cm = opt["dsource_pc"] * pc_cgs
    if opt["dsource_pc"] > 0 and opt["image_width_x"] > 0 and opt["image_width_y"] > 0:
        px = opt["image_width_x"] * l_unit / opt["nx"]
        py = opt["image_width_y"] * l_unit / opt["ny"]
        sr = px * py / (d_cm * d_cm)
        return sr / jy
    return float("nan")


def pack(image: dict, frequency: float, qu_sign: int, scale: float):
    """pack_transport_frequency_image (tools/kpolaris_model_image.cpp:2668-2692)
    plus flux_density_jy (:590-595).  Returns (I_nu, Q_nu, U_nu, V_nu, fluxes)."""
    reached = image["reason"] == abi.REASON_REACHED_CAMERA
    nu3 = frequency * frequency * frequency
    out = []
    fluxes = []
    for name, sign in (("image_i", 1), ("image_q", qu_sign), ("image_u", qu_sign), ("image_v", 1)):
        inv = np.where(reached, sign * image[name], 0.0)
        out.append(inv * nu3)
        fluxes.append(float(inv.sum()) * frequency * frequency * frequency * scale)
    return out[0], out[1], out[2], out[3], fluxes
