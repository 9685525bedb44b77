"""Per-pixel parity of an image against the oracle, with the oracle's own
conditioning.

Bar (per pixel, on I/Q/U/V): |x - oracle| <= 1e-10 |oracle| + 1e-8 peak(I).

Some rays wind around the hole or return to the camera sphere far from the
camera point; they amplify last-bit differences chaotically, so no two
implementations with different rounding agree on them to 1e-10 (the oracle
disagrees with itself).  Such a pixel is identified by the oracle alone: it is
re-run with the ray slope moved by 1e-15 (about 1e-13 relative) in x and in
y, and if any Stokes component moves by more than 1e-11 of its value (floor
1e-6 peak) the pixel is ill-conditioned.  Pixels beyond the bar must be
ill-conditioned, and must still agree to 1e-3 (floor 1e-6 peak); termination
reasons must agree everywhere and step counts on well-conditioned pixels.
"""
from __future__ import annotations

import copy

import numpy as np

import oracle

REL = 1e-10
PEAK_FLOOR = 1e-8
SENSITIVITY = 1e-11
PERTURBATION = 1e-15
LOOSE_REL = 1e-3
LOOSE_FLOOR = 1e-6
STOKES = ("image_i", "image_q", "image_u", "image_v")


def oracle_with_sensitivity(p, m, offset=None, stride=None, count=None):
    """Oracle image plus per-field max |change| under the two perturbations."""
    def run(pp):
        if offset is None:
            return oracle.run_riaf_image(pp, m)[0]
        return oracle.run_riaf_image_strided(pp, m, offset, stride, count)[0]

    ref = run(p)
    sens = {n: np.zeros_like(ref[n]) for n in STOKES}
    moved_steps = np.zeros(ref["reason"].shape, dtype=bool)
    for dx, dy in ((PERTURBATION, 0.0), (0.0, PERTURBATION)):
        q = copy.deepcopy(p)
        q.camera.x_offset += dx
        q.camera.y_offset += dy
        r2 = run(q)
        for n in STOKES:
            sens[n] = np.maximum(sens[n], np.abs(r2[n] - ref[n]))
        moved_steps |= (r2["steps"] != ref["steps"]) | (r2["pass_a_steps"] != ref["pass_a_steps"])
    ref["_sens"] = sens
    ref["_moved_steps"] = moved_steps
    return ref


def ill_conditioned(ref, peak):
    s = np.zeros(ref["reason"].shape)
    for n in STOKES:
        s = np.maximum(s, ref["_sens"][n] / (np.abs(ref[n]) + LOOSE_FLOOR * peak + 1e-300))
    return (s > SENSITIVITY) | ref["_moved_steps"]


def assert_parity(got, ref, label, peak=None, diagnostics=True):
    """got/ref: dicts of per-pixel arrays; ref from oracle_with_sensitivity."""
    assert np.array_equal(got["reason"], ref["reason"]), (
        f"{label}: reasons differ at {np.flatnonzero(got['reason'] != ref['reason'])[:10]}")
    if peak is None:
        peak = float(np.max(np.abs(ref["image_i"]))) if ref["image_i"].size else 0.0
    ill = ill_conditioned(ref, peak)
    beyond = np.zeros(ill.shape, dtype=bool)
    for n in STOKES:
        err = np.abs(got[n] - ref[n])
        beyond |= err > REL * np.abs(ref[n]) + PEAK_FLOOR * peak
        loose = err > LOOSE_REL * np.abs(ref[n]) + LOOSE_FLOOR * peak
        assert not loose.any(), (
            f"{label} {n}: {int(loose.sum())} pixels beyond even the 1e-3 bound at "
            f"{np.flatnonzero(loose)[:5]}")
    bad = beyond & ~ill
    assert not bad.any(), (
        f"{label}: {int(bad.sum())} well-conditioned pixels beyond 1e-10 at {np.flatnonzero(bad)[:8]}")
    steps_bad = ((got["steps"] != ref["steps"]) | (got["pass_a_steps"] != ref["pass_a_steps"])) & ~ill
    assert not steps_bad.any(), (
        f"{label}: step counts differ on well-conditioned pixels {np.flatnonzero(steps_bad)[:8]}")
    if diagnostics:
        # end-state diagnostics are compared on well-conditioned rays that
        # reached the camera; rays stopped by max_steps close to the horizon
        # (where the Boyer-Lindquist k diverges) have no well-defined end state
        wc = ~ill & (ref["reason"] == 1)
        for n in ("closure_x", "closure_k", "final_null", "frame_error", "det_r", "overlap_r11",
                  "overlap_r12", "overlap_r21", "overlap_r22", "basis_rotation_angle"):
            np.testing.assert_allclose(got[n][wc], ref[n][wc], rtol=1e-4, atol=1e-8,
                                       err_msg=f"{label} {n}")
    return {"pixels": int(ill.size), "ill_conditioned": int(ill.sum()),
            "beyond_1e-10": int(beyond.sum())}


# ---------------------------------------------------------------------------
# Mixed precision (FP32 derivatives, coefficients and transfer; FP64 state) is
# not a parity path: it is held to image-level error bars against an FP64
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


def sensitivity(run, p, ref):
    """Per-pixel ill-conditioning of an oracle image from a runner
    run(params) -> image: the pixel moves by more than SENSITIVITY of its
    value (floor LOOSE_FLOOR of the frequency's peak) when the ray slope moves
    by PERTURBATION in x or y, or its step counts change."""
    n = ref["reason"].size
    nfreq = ref["image_i"].size // max(n, 1)
    ill = np.zeros(n, dtype=bool)
    for dx, dy in ((PERTURBATION, 0.0), (0.0, PERTURBATION)):
        q = copy.deepcopy(p)
        q.camera.x_offset += dx
        q.camera.y_offset += dy
        r2 = run(q)
        ill |= (r2["steps"] != ref["steps"]) | (r2["pass_a_steps"] != ref["pass_a_steps"])
        for f in range(nfreq):
            sl = slice(f * n, (f + 1) * n)
            peak = float(np.max(np.abs(ref["image_i"][sl]))) if n else 0.0
            for name in STOKES:
                d = np.abs(r2[name][sl] - ref[name][sl])
                ill |= d > SENSITIVITY * (np.abs(ref[name][sl]) + LOOSE_FLOOR * peak)
    return ill


def assert_parity_grid(got, ref, nfreq, label, max_step_diff=None, ill=None):
    """Parity of a GRMHD (grid or snapshot-sequence) image, [nfreq][npix]
    Stokes: identical termination reasons; on rays whose step counts agree,
    the per-pixel bar (1e-10 relative + 1e-8 of the frequency's peak I); rays
    whose step sequence moved (a chaotic near-photon-ring ray, or a slow-light
    ray approaching a snapshot time by ever smaller cut steps), and the
    ill-conditioned rays `ill` of the oracle's own sensitivity, within 1e-3
    (+1e-6 of peak).  max_step_diff bounds |steps - steps_ref| on the rays
    that are not ill-conditioned when given."""
    assert np.array_equal(got["reason"], ref["reason"]), (
        f"{label}: reasons differ at {np.flatnonzero(got['reason'] != ref['reason'])[:10]}")
    moved = (got["steps"] != ref["steps"]) | (got["pass_a_steps"] != ref["pass_a_steps"])
    if ill is not None:
        moved = moved | ill
    if max_step_diff is not None:
        # ill-conditioned rays change their step counts under rounding-level
        # perturbations of the oracle itself, so the bound holds on the others
        d = np.abs(got["steps"] - ref["steps"])
        if ill is not None:
            d = np.where(ill, 0, d)
        assert np.max(d, initial=0) <= max_step_diff, (label, np.flatnonzero(d > max_step_diff)[:8])
    n = ref["reason"].size
    worst = 0.0
    for f in range(nfreq):
        sl = slice(f * n, (f + 1) * n)
        peak = float(np.max(np.abs(ref["image_i"][sl]))) if n else 0.0
        for name in STOKES:
            err = np.abs(got[name][sl] - ref[name][sl])
            bad = err > REL * np.abs(ref[name][sl]) + PEAK_FLOOR * peak
            assert not (bad & ~moved).any(), (label, f, name, np.flatnonzero(bad & ~moved)[:8])
            loose = err > LOOSE_REL * np.abs(ref[name][sl]) + LOOSE_FLOOR * peak
            assert not (loose & moved).any(), (label, f, name, np.flatnonzero(loose & moved)[:8])
            if peak > 0.0:
                worst = max(worst, float(np.max(err[~moved], initial=0.0)) / peak)
    return {"pixels": n, "moved": int(moved.sum()), "max_err_over_peak_unmoved": worst}


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


def assert_parity_multifrequency(got, ref_mf, nfreq, ref_single, label):
    """Each frequency slice of got against the oracle's multi-frequency image at
    the parity bar, with ill-conditioned pixels taken from ref_single (an
    oracle_with_sensitivity run of the same geodesics: the conditioning of a
    ray is a property of its geodesic)."""
    g = split_frequencies(got, nfreq)
    r = split_frequencies(ref_mf, nfreq)
    res = []
    for f in range(nfreq):
        rf = dict(r[f])
        rf["_sens"] = ref_single["_sens"]
        rf["_moved_steps"] = ref_single["_moved_steps"]
        res.append(assert_parity(g[f], rf, f"{label} f{f}",
                                 peak=float(np.max(np.abs(rf["image_i"]))), diagnostics=(f == 0)))
    return res
