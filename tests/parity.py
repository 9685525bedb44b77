"""Per-pixel parity of an image against the oracle, with the oracle's own
conditioning.

Bar (per pixel, on I/Q/U/V): |x - oracle| <= 1e-10 |oracle| + 1e-8 peak(I).

Some rays wind around the hole or return to the camera sphere far from the
camera point; they amplify last-bit differences chaotically, so no two
implementations with different rounding agree on them to 1e-10 (the oracle
disagrees with itself).  Such a pixel is identified by the oracle alone: the
oracle is re-run with the camera offset moved by +-1e-15 in x and in y (four
runs; about 1e-13 of the ray slope), and the per-pixel spread of those runs
around the unperturbed oracle is the pixel's measured sensitivity.

* A pixel is ill-conditioned when its step counts or reason change in one of
  the perturbed runs, or AMPLIFICATION times the spread of some Stokes
  component exceeds that component's tight bar.  Each test bounds the
  fraction of ill-conditioned pixels it accepts and reports it.
* Every pixel must meet |x - oracle| <= 1e-10 |oracle| + 1e-8 peak
  + AMPLIFICATION x spread: the bar widens by a fixed multiple of the
  oracle's own measured sensitivity, not to a fixed loose tolerance.  A pixel
  whose value matches one of the perturbed oracle runs to the tight bar (the
  kernel took the branch the oracle takes under a 1e-15 change) passes too.
* Termination reasons agree on every pixel; step counts agree on every
  well-conditioned pixel.  The one documented exception is slow light, where
  a ray that approaches a snapshot time by ever smaller cut steps (the
  reference stops when the remaining step is below 1e-14) may take one step
  more or less depending on the last bits of x^0; its Stokes values must
  still meet the bar (time_cut_steps=1).  The other is the equatorial
  emission wedge, whose boundary split (a 48-step bisection on the step's
  RK path) decides on the last bits of cos(theta): the oracle's own 1e-15
  perturbations move the step counts of about a third of a wedge image's
  pixels, and a kernel's rounding moves a few more by up to 0.2 % of their
  steps (WEDGE_MAX_ILL, WEDGE_STEP_REL); the Stokes bar is unchanged.
"""
from __future__ import annotations

import copy

import numpy as np

import oracle

REL = 1e-10
PEAK_FLOOR = 1e-8
SENSITIVITY = 1e-11
PERTURBATION = 1e-15
SENS_FLOOR = 1e-6        # floor (of the peak) of the ill-conditioning test
AMPLIFICATION = 10.0     # the bar's multiple of the oracle's measured spread
# bound on the ill-conditioned share of an image (measured on the GPU test
# images: at most 5 %, on ragged images of a few dozen photon-ring pixels)
MAX_ILL = 0.1
# share of pixels on which a Stokes component's derivative must be checkable
MIN_CHECKED = 0.35
# equatorial-wedge images (module docstring): ill-conditioned share and the
# relative step-count difference allowed on well-conditioned pixels
WEDGE_MAX_ILL = 0.5
WEDGE_STEP_REL = 2e-3
STOKES = ("image_i", "image_q", "image_u", "image_v")
OFFSETS = ((PERTURBATION, 0.0), (-PERTURBATION, 0.0), (0.0, PERTURBATION), (0.0, -PERTURBATION))


def conditioning(run, p, ref, extra=()):
    """Perturbed oracle runs of an image: run(params) -> image dict, plus the
    zero-argument runners in `extra` (other rounding-level perturbations of
    the same image, e.g. time_shifts).  Returns
    {"spread": {name: max |r_k - ref| per element}, "dsteps": the largest
    change of the pass-B step count, "moved": pass-A steps or reason changed
    in some run (per pixel), "samples": the perturbed images}."""
    spread = {n: np.zeros_like(ref[n]) for n in STOKES}
    moved = np.zeros(ref["reason"].shape, dtype=bool)
    dsteps = np.zeros(ref["reason"].shape, dtype=np.int64)
    samples = []
    def runs():
        for dx, dy in OFFSETS:
            q = copy.deepcopy(p)
            q.camera.x_offset += dx
            q.camera.y_offset += dy
            yield run(q)
        for e in extra:
            yield e()

    for r2 in runs():
        samples.append({k: r2[k] for k in (*STOKES, "steps", "pass_a_steps", "reason")})
        for n in STOKES:
            spread[n] = np.maximum(spread[n], np.abs(r2[n] - ref[n]))
        dsteps = np.maximum(dsteps, np.abs(r2["steps"].astype(np.int64) - ref["steps"]))
        moved |= (r2["pass_a_steps"] != ref["pass_a_steps"]) | (r2["reason"] != ref["reason"])
    return {"spread": spread, "dsteps": dsteps, "moved": moved, "samples": samples}


def oracle_with_sensitivity(p, m, offset=None, stride=None, count=None):
    """Oracle RIAF image with its conditioning under "_cond"."""
    def run(pp):
        if offset is None:
            return oracle.run_riaf_image(pp, m)[0]
        return oracle.run_riaf_image_strided(pp, m, offset, stride, count)[0]

    ref = run(p)
    ref["_cond"] = conditioning(run, p, ref)
    return ref


def sensitivity(run, p, ref, extra=()):
    """Conditioning of an oracle image (any model) from a runner."""
    return conditioning(run, p, ref, extra)


def time_shifts(run_at, t_obs, rel=1e-13):
    """Slow light: the image at observation times t_obs (1 +- rel).  Where a
    ray meets a snapshot time, the reference cuts its step to land on it and
    stops once the remaining step is below 1e-14, so the number and size of
    those cut steps depend on the last bits of x^0; shifting t_obs by 1e-13
    exposes the rays whose result depends on that."""
    return [lambda: run_at(t_obs * (1.0 + rel)), lambda: run_at(t_obs * (1.0 - rel))]


def _peaks(ref, nfreq):
    n = ref["reason"].size
    return [float(np.max(np.abs(ref["image_i"][f * n:(f + 1) * n]))) if n else 0.0 for f in range(nfreq)]


def ill_conditioned(ref, cond, nfreq=1, time_cut_steps=0, peaks=None):
    """Per pixel: the oracle's own perturbation changes its step counts
    (beyond the time-cut jitter slow light allows) or its reason, or moves a
    Stokes component by more than the tight bar / AMPLIFICATION, i.e. the
    conditioned bar of that component is wider than the tight one."""
    n = ref["reason"].size
    ill = cond["moved"] | (cond["dsteps"] > time_cut_steps)
    for f, peak in enumerate(peaks if peaks is not None else _peaks(ref, nfreq)):
        sl = slice(f * n, (f + 1) * n)
        for name in STOKES:
            tight = REL * np.abs(ref[name][sl]) + PEAK_FLOOR * peak
            ill |= AMPLIFICATION * cond["spread"][name][sl] > tight
    return ill


def smooth_for_derivatives(ref, cond, time_cut_steps=0):
    """Single-frequency pixels on which a derivative can be compared with
    difference quotients of the oracle: no step or reason change under the
    oracle's own perturbations, and no Stokes component moved by more than
    SENSITIVITY of its value (floor SENS_FLOOR of the peak), which excludes
    the rays whose transfer amplifies perturbations (large Faraday depth,
    photon-ring winding)."""
    peak = float(np.max(np.abs(ref["image_i"]))) if ref["image_i"].size else 0.0
    ok = ~cond["moved"] & (cond["dsteps"] <= time_cut_steps)
    for name in STOKES:
        ok &= cond["spread"][name] <= SENSITIVITY * (np.abs(ref[name]) + SENS_FLOOR * peak)
    return ok


def check(got, ref, cond, label, nfreq=1, max_ill=MAX_ILL, time_cut_steps=0, diagnostics=True, peak=None,
          step_rel=0.0):
    """The parity rule of this module for a [nfreq][npix]-Stokes image.
    peak: the image's peak I for a sample of its pixels, one value or one
    per frequency (default: the sample's own, per frequency).  Returns a summary (printed by the tests):
    pixels, ill-conditioned count, pixels that needed the widened bar, the
    worst error over the tight bar."""
    n = ref["reason"].size
    assert np.array_equal(got["reason"], ref["reason"]), (
        f"{label}: reasons differ at {np.flatnonzero(got['reason'] != ref['reason'])[:10]}")
    if peak is None:
        peaks = _peaks(ref, nfreq)
    else:
        peaks = list(peak) if np.ndim(peak) else [float(peak)] * nfreq
    ill = ill_conditioned(ref, cond, nfreq, time_cut_steps, peaks)
    frac = float(ill.mean()) if n else 0.0
    assert frac <= max_ill, f"{label}: {int(ill.sum())}/{n} ill-conditioned pixels exceed the bound {max_ill}"
    widened = np.zeros(n, dtype=bool)
    worst_tight = 0.0
    for f, peak in enumerate(peaks):
        sl = slice(f * n, (f + 1) * n)
        for name in STOKES:
            g, r = got[name][sl], ref[name][sl]
            err = np.abs(g - r)
            tight = REL * np.abs(r) + PEAK_FLOOR * peak
            # well-conditioned pixels: the tight bar alone
            bar = tight + np.where(ill, AMPLIFICATION * cond["spread"][name][sl], 0.0)
            ok = err <= bar
            if not ok.all():
                # an ill-conditioned pixel on the branch a perturbed oracle run takes
                for s in cond["samples"]:
                    sv = s[name][sl]
                    ok |= ill & (np.abs(g - sv) <= REL * np.abs(sv) + PEAK_FLOOR * peak)
            assert ok.all(), (
                f"{label} f{f} {name}: {int((~ok).sum())} pixels beyond the conditioned bar at "
                f"{np.flatnonzero(~ok)[:6]}, err/tight {(err / tight)[~ok][:6]}, ill {ill[~ok][:6]}")
            widened |= err > tight
            worst_tight = max(worst_tight, float(np.max(err / tight, initial=0.0)))
    dsteps = np.abs(got["steps"] - ref["steps"])
    da = got["pass_a_steps"] != ref["pass_a_steps"]
    steps_bad = ((dsteps > time_cut_steps + step_rel * ref["steps"]) | da) & ~ill
    assert not steps_bad.any(), (
        f"{label}: step counts differ on well-conditioned pixels {np.flatnonzero(steps_bad)[:8]}")
    if diagnostics:
        # end-state diagnostics on well-conditioned rays that reached the
        # camera (rays stopped by max_steps near the horizon, where the
        # Boyer-Lindquist k diverges, have no well-defined end state)
        wc = ~ill & (ref["reason"] == 1) & (dsteps == 0)
        for name in ("closure_x", "closure_k", "final_null", "frame_error", "det_r", "overlap_r11",
                     "overlap_r12", "overlap_r21", "overlap_r22", "basis_rotation_angle"):
            if name in got and name in ref:
                np.testing.assert_allclose(got[name][wc], ref[name][wc], rtol=1e-4, atol=1e-8,
                                           err_msg=f"{label} {name}")
    print(f"parity {label}: ill-conditioned {int(ill.sum())}/{n} (bound {max_ill}), needing the widened bar "
          f"{int(widened.sum())}, worst error / tight bar {worst_tight:.3g}", flush=True)
    return {"pixels": n, "ill_conditioned": int(ill.sum()), "ill_fraction": round(frac, 4),
            "widened": int(widened.sum()), "max_err_over_tight_bar": worst_tight,
            "step_count_differences": int(((dsteps > 0) | da).sum())}


def assert_parity(got, ref, label, peak=None, diagnostics=True, max_ill=MAX_ILL, step_rel=0.0):
    """Single-frequency RIAF image against oracle_with_sensitivity's result;
    peak: the full image's peak I when got/ref are a sample of its pixels."""
    return check(got, ref, ref["_cond"], label, 1, max_ill=max_ill, diagnostics=diagnostics, peak=peak,
                 step_rel=step_rel)


def assert_parity_grid(got, ref, nfreq, label, cond, max_ill=MAX_ILL, time_cut_steps=0, step_rel=0.0):
    """GRMHD, slow-light or multi-frequency image ([nfreq][npix] Stokes)
    against the oracle with the conditioning `cond` of the same runner."""
    return check(got, ref, cond, label, nfreq, max_ill=max_ill, time_cut_steps=time_cut_steps,
                 diagnostics=False, step_rel=step_rel)


def split_frequencies(img, nfreq):
    """Per-frequency views of a multi-frequency image ([nfreq][npix] Stokes,
    per-pixel diagnostics shared)."""
    npix = img["reason"].size
    out = []
    for f in range(nfreq):
        d = {k: v for k, v in img.items() if not k.startswith("_") and k not in STOKES}
        for n in STOKES:
            d[n] = img[n][f * npix:(f + 1) * npix]
        out.append(d)
    return out


# ---------------------------------------------------------------------------
# Mixed precision (FP32 coefficients and transfer; FP64 geodesics by default)
# is not a parity path: it is held to image-level error bars against an FP64
# image of the same inputs.
# ---------------------------------------------------------------------------
MIXED_TOTAL_I = 1e-4     # |sum I - sum I_ref| / sum I_ref
MIXED_POL_FRAC = 1e-3    # relative error of the image-integrated LP and CP fractions
MIXED_POL_FLOOR = 1e-6   # absolute floor on a fraction's error (CP can be ~0)
MIXED_PIXEL = 1e-2       # per pixel |I - I_ref| <= MIXED_PIXEL * peak(I_ref)
MIXED_REASONS = 1e-3     # at most this share of pixels may end for another reason


def polarisation_fractions(img):
    tot = float(np.sum(img["image_i"]))
    lp = float(np.hypot(np.sum(img["image_q"]), np.sum(img["image_u"]))) / tot
    cp = float(np.sum(img["image_v"])) / tot
    return tot, lp, cp


def mixed_errors(got, ref):
    t0, lp0, cp0 = polarisation_fractions(ref)
    t1, lp1, cp1 = polarisation_fractions(got)
    peak = float(np.max(ref["image_i"]))
    return {
        "total_i_rel": abs(t1 / t0 - 1.0),
        "lp_frac_rel": abs(lp1 - lp0) / max(abs(lp0), MIXED_POL_FLOOR),
        "cp_frac_rel": abs(cp1 - cp0) / max(abs(cp0), MIXED_POL_FLOOR),
        "lp": (lp0, lp1),
        "cp": (cp0, cp1),
        "pixel_max_rel_peak": float(np.max(np.abs(got["image_i"] - ref["image_i"]))) / peak,
        "reason_mismatch": int(np.sum(got["reason"] != ref["reason"])),
        "pixels": int(ref["reason"].size),
    }


def assert_mixed(got, ref, label, pol="relative"):
    """Mixed-precision image bars.  pol="relative" holds the image-integrated
    LP and CP fractions to MIXED_POL_FRAC of their FP64 values (full-size
    images); pol="absolute" holds them to MIXED_POL_FRAC in the fraction
    itself, for coarse images whose net polarisation is a near-cancelling sum
    over a few dozen pixels, where one photon-ring pixel moves the relative
    error of a 0.4 % LP fraction by a percent."""
    e = mixed_errors(got, ref)
    assert e["total_i_rel"] <= MIXED_TOTAL_I, (label, e)
    for key in ("lp", "cp"):
        f0, f1 = e[key]
        bar = MIXED_POL_FRAC * abs(f0) + MIXED_POL_FLOOR if pol == "relative" else MIXED_POL_FRAC
        assert abs(f1 - f0) <= bar, (label, key, e)
    assert e["pixel_max_rel_peak"] <= MIXED_PIXEL, (label, e)
    assert e["reason_mismatch"] <= MIXED_REASONS * e["pixels"], (label, e)
    for n in STOKES:
        assert np.all(np.isfinite(got[n])), (label, n)
    return e


def assert_parity_multifrequency(got, ref_mf, nfreq, ref_single, label, max_ill=MAX_ILL):
    """Each frequency slice of got against the oracle's multi-frequency image,
    with the conditioning of ref_single (an oracle_with_sensitivity run of the
    same geodesics: a ray's conditioning is a property of its geodesic, and
    each frequency's spread is bounded by the single-frequency one scaled to
    that frequency's values)."""
    g = split_frequencies(got, nfreq)
    r = split_frequencies(ref_mf, nfreq)
    single = ref_single["_cond"]
    res = []
    for f in range(nfreq):
        rf = dict(r[f])
        # the single-frequency spread, relative to its own values, carried to
        # this frequency's values
        scale = {}
        for n in STOKES:
            rel = single["spread"][n] / (np.abs(ref_single[n]) + SENS_FLOOR * float(
                np.max(np.abs(ref_single["image_i"]))) + 1e-300)
            scale[n] = rel * (np.abs(rf[n]) + SENS_FLOOR * float(np.max(np.abs(rf["image_i"]))))
        cond = {"spread": scale, "moved": single["moved"], "dsteps": single["dsteps"], "samples": []}
        res.append(check(g[f], rf, cond, f"{label} f{f}", 1, max_ill=max_ill, diagnostics=(f == 0)))
    return res
