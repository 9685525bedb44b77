// kp_prepare.h -- host-side preparation of the kernels' parameter block
// (camera tetrad and extents, integrator scalars, model constants,
// frequency lists).  Included by the library's host translation unit
// and by the CPU build of the lane logic.
#pragma once

#include "kp_kernels.h"
#include "kp_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

// ---------------------------------------------------------------------------
// Parameter preparation (host, double): tetrad, extents, model constants.
// ---------------------------------------------------------------------------
// Camera extents and integrator scalars shared by every model.
static kp_status prepare_common(const kp_pass_a_params* p, Params& P) {
    if (p->camera.model != KP_CAMERA_PINHOLE)
        return kp_internal::fail(KP_UNSUPPORTED, "only the pinhole camera is built");
    // the option rules of kpolaris_model_image_options.cpp:196-209
    const double hr = p->emission_selection.equatorial_h_over_r;
    if (!std::isfinite(hr) || hr < 0.0 || hr > 1.0 || (hr > 0.0 && hr < 1e-6))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "equatorial_h_over_r must be 0 (disabled) or in [1e-6,1]");
    const int32_t samples = p->emission_selection.equatorial_samples;
    if (hr > 0.0 && hr < 1.0 && (samples < 2 || samples > 1024))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "equatorial_samples must be an integer in [2,1024]");
    if (p->direct_only != 0 && p->direct_only != 1)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "direct_only must be 0 or 1");
    if (p->camera.nx <= 0 || p->camera.ny <= 0)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "camera nx and ny must be positive");
    const long long npix = (long long)p->camera.nx * (long long)p->camera.ny;
    if (npix > (1ll << 30)) return kp_internal::fail(KP_INVALID_ARGUMENT, "image too large");
    if (p->max_steps < 0) return kp_internal::fail(KP_INVALID_ARGUMENT, "max_steps must be non-negative");
    std::memset(&P, 0, sizeof(P));
    P.stage = -1;
    P.park = PARK_NONE;
    P.nx = p->camera.nx;
    P.ny = p->camera.ny;
    P.npix = (int32_t)npix;
    P.adaptive = p->adaptive ? 1 : 0;
    P.max_steps = p->max_steps;
    P.radius = p->camera.radius;
    P.inclination = p->camera.inclination;
    const kp_camera_params& c = p->camera;
    P.x_extent = c.xspan > 0.0 ? 2.0 * c.xspan : c.fov;
    P.y_extent = c.yspan > 0.0 ? 2.0 * c.yspan
               : (c.xspan > 0.0 ? 2.0 * c.xspan * double(c.ny) / double(c.nx) : (c.fov_y > 0.0 ? c.fov_y : c.fov));
    P.x_offset = c.x_offset;
    P.y_offset = c.y_offset;
    P.mass = p->mass;
    P.spin = p->spin;
    P.inner = p->inner_radius;
    P.outer = p->outer_radius;
    P.step = p->step;
    P.tol = p->adaptive_tolerance;
    P.min_step = p->min_step;
    P.max_step = p->max_step;
    P.max_rs = p->max_radiation_step;
    P.max_rd = p->max_radiation_depth;
    P.max_ad = p->max_absorption_depth;
    P.max_fd = p->max_faraday_depth;
    P.direct_only = p->direct_only;
    // a wedge of 0 (disabled) or 1 (the whole sphere) selects nothing
    P.wedge_hr = hr > 0.0 && hr < 1.0 ? hr : 0.0;
    P.wedge_samples = samples;
    return KP_OK;
}

static bool emission_known(int t) {
    return t == KP_EMISSION_THERMAL_PANDYA || t == KP_EMISSION_THERMAL_DEXTER || t == KP_EMISSION_KAPPA ||
           t == KP_EMISSION_POWER_LAW;
}

static bool emission_nonthermal(int t) { return t == KP_EMISSION_KAPPA || t == KP_EMISSION_POWER_LAW; }

// NonthermalSynchrotronParams as the models assemble them (riaf.hpp:497-513,
// grmhd_grid.hpp:1176-1197), with the kappa- and slope-only constants of the
// fits evaluated once here (host libm) for a fixed kappa.
static void fill_nonthermal(NtDev& N, int dist, double kappa, double kappa_min, int variable, double interp_begin,
                            double interp_end, double max_pol, double min_sin, double es, double as, double fs,
                            double p, double eta, double gmin, double gmax) {
    std::memset(&N, 0, sizeof(N));
    N.dist = dist;
    N.variable_kappa = variable;
    N.kappa_min = kappa_min;
    N.kappa_width = -1.0;
    N.interp_begin = interp_begin;
    N.interp_end = interp_end;
    N.max_pol_em = N.max_pol_abs = max_pol;
    N.min_sin = min_sin;
    N.es = es;
    N.as = as;
    N.fs = fs;
    N.p = p;
    kappa_constants(kappa, N.kc);
    power_law_constants(p, eta, gmin, gmax, N.pl);
}

static kp_status prepare(const kp_pass_a_params* p, const kp_riaf_model* m, Params& P) {
    if (!p || !m) return kp_internal::fail(KP_INVALID_ARGUMENT, "null parameters");
    if (p->coordinate_system != KP_COORD_BOYER_LINDQUIST)
        return kp_internal::fail(KP_UNSUPPORTED, "the RIAF image is built only in Boyer-Lindquist coordinates");
    if (p->camera.model != KP_CAMERA_PINHOLE)
        return kp_internal::fail(KP_UNSUPPORTED, "only the pinhole camera is defined in Boyer-Lindquist coordinates");
    if (!emission_known(m->emission_type))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown emission_type");
    kp_status st = prepare_common(p, P);
    if (st != KP_OK) return st;
    double g5[5];
    kp_internal::bl_camera_tetrad(p->mass, p->spin, p->camera.radius, p->camera.inclination, P.tetrad, g5);
    P.xcam[0] = 0.0;
    P.xcam[1] = p->camera.radius;
    P.xcam[2] = p->camera.inclination;
    P.xcam[3] = 0.0;
    BLGcov<double> b;
    b.g00 = g5[0];
    b.g03 = g5[1];
    b.g11 = g5[2];
    b.g22 = g5[3];
    b.g33 = g5[4];
    P.gcam = metric4_of(b);
    kp_internal::riaf_constants(*p, *m, 0, 0.0, P.riaf);
    fill_nonthermal(P.nt, m->emission_type, m->nonthermal_kappa, 0.0, 0, m->variable_kappa_interp_start,
                    m->variable_kappa_max, m->max_pol_frac, m->min_sin_theta, m->emission_scale,
                    m->absorption_scale, m->faraday_scale, m->powerlaw_p, m->powerlaw_eta, m->powerlaw_gamma_min,
                    m->powerlaw_gamma_max);
    return KP_OK;
}

// GRMHD grid image on FMKS geodesics: camera in FMKS, model constants.
static kp_status prepare_grmhd(const kp_pass_a_params* p, const kp_grmhd_model* g, Params& P,
                               bool need_cells = true) {
    if (!p || !g) return kp_internal::fail(KP_INVALID_ARGUMENT, "null parameters");
    if (p->coordinate_system != KP_COORD_FMKS)
        return kp_internal::fail(KP_UNSUPPORTED, "GRMHD images are built in FMKS coordinates");
    if (!emission_known(g->emission_type))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown emission_type");
    if (g->n1 < 1 || g->n2 < 1 || g->n3 < 2 || (need_cells && !g->cells))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "GRMHD grid dimensions and cells are required");
    if (!(g->dx1 > 0.0) || !(g->dx2 > 0.0) || !(g->dx3 > 0.0))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "GRMHD grid spacings must be positive");
    kp_status st = prepare_common(p, P);
    if (st != KP_OK) return st;
    P.fmks = fmks_map(p->fmks_startx1, p->fmks_hslope, p->fmks_mks_smooth, p->fmks_poly_alpha, p->fmks_poly_xt,
                      p->fmks_poly_norm);
    double fm[6] = {p->fmks_startx1, p->fmks_hslope, p->fmks_mks_smooth, p->fmks_poly_alpha, p->fmks_poly_xt,
                    p->fmks_poly_norm};
    double g16[16];
    kp_internal::fmks_camera_tetrad(p->mass, p->spin, fm, p->camera.radius, p->camera.inclination, P.xcam, P.tetrad,
                                    g16);
    P.gcam = Metric4<double>{g16[0], g16[1], g16[2], g16[3], g16[5], g16[6], g16[7], g16[10], g16[11], g16[15]};
    GrmhdDev& G = P.grmhd;
    G.n1 = g->n1;
    G.n2 = g->n2;
    G.n3 = g->n3;
    G.fit = g->emission_type;
    G.startx1 = g->startx1;
    G.startx2 = g->startx2;
    G.startx3 = g->startx3;
    G.dx1 = g->dx1;
    G.dx2 = g->dx2;
    G.dx3 = g->dx3;
    G.r_in = g->r_in;
    G.r_out = g->r_out;
    G.spin = g->spin;
    G.gam = g->gam;
    G.map = fmks_map(g->startx1, g->hslope, g->mks_smooth, g->poly_alpha, g->poly_xt, g->poly_norm);
    G.sigma_cut = g->sigma_cut;
    G.sigma_cut_high = g->sigma_cut_high;
    G.emission_scale = g->emission_scale;
    G.absorption_scale = g->absorption_scale;
    G.faraday_scale = g->faraday_scale;
    G.max_pol_frac = g->max_pol_frac;
    G.freq_cgs = g->freq_cgs;
    // length_unit_cgs, rho_unit_cgs, b_unit_cgs (grmhd_grid.hpp:140-156)
    const double cl = KP_CL;
    G.length_unit = 6.67430e-8 * g->mbh_solar * 1.98847e33 / (cl * cl);
    G.rho_unit = g->M_unit / std::max(G.length_unit * G.length_unit * G.length_unit, 1e-300);
    G.b_unit = cl * std::sqrt(4.0 * PI_D * G.rho_unit);
    G.ne_factor = g->ne_factor;
    G.trat_small = g->trat_small;
    G.trat_large = g->trat_large;
    G.beta_crit = g->beta_crit;
    G.faraday_rotation = p->emission_selection.faraday_rotation ? 1 : 0;
    G.cells = static_cast<const kp_grmhd_cell*>(g->cells);
    // kappa floored at variable_kappa_min, fixed or from sigma and beta
    // (grmhd_grid.hpp:1185-1192); the blend starts at min(start, max)
    G.variable_kappa = g->variable_kappa ? 1 : 0;
    G.kappa_min = g->variable_kappa_min;
    G.kappa = g->nonthermal_kappa < g->variable_kappa_min ? g->variable_kappa_min : g->nonthermal_kappa;
    fill_nonthermal(P.nt, g->emission_type, G.kappa, g->variable_kappa_min, G.variable_kappa,
                    std::min(g->variable_kappa_interp_start, g->variable_kappa_max), g->variable_kappa_max,
                    g->max_pol_frac, 1e-4, g->emission_scale, g->absorption_scale, g->faraday_scale, g->powerlaw_p,
                    g->powerlaw_eta, g->powerlaw_gamma_min, g->powerlaw_gamma_max);
    return KP_OK;
}

// Single-precision model constants and the rescaled-unit factors of the
// mixed-precision radiation path (kp_mixed.h).
static void prepare_mixed(Params& P) {
    const RiafDev<double>& d = P.riaf;
    RiafDev<float>& f = P.rm.f;
    f.nth0 = float(d.nth0); f.Te0 = float(d.Te0); f.disk_h = float(d.disk_h);
    f.pow_nth = float(d.pow_nth); f.pow_T = float(d.pow_T); f.ne_unit = float(d.ne_unit);
    f.thetae_unit = float(d.thetae_unit); f.r_min = float(d.r_min); f.r_max = float(d.r_max);
    f.horizon_buffer = float(d.horizon_buffer); f.emission_scale = float(d.emission_scale);
    f.absorption_scale = float(d.absorption_scale); f.faraday_scale = float(d.faraday_scale);
    f.max_pol_frac = float(d.max_pol_frac); f.keplerian_factor = float(d.keplerian_factor);
    f.infall_factor = float(d.infall_factor); f.min_frequency_scale = float(d.min_frequency_scale);
    f.min_sin_theta = float(d.min_sin_theta); f.min_thetae = float(d.min_thetae);
    f.min_radius = float(d.min_radius); f.freq_cgs = float(d.freq_cgs); f.dlambda_scale = float(d.dlambda_scale);
    f.rh = float(d.rh); f.risco = float(d.risco); f.ne_factor = float(d.ne_factor);
    f.thetae_factor = float(d.thetae_factor); f.b_factor = float(d.b_factor);
    f.mass = float(d.mass); f.spin = float(d.spin); f.fit = d.fit; f.faraday_rotation = d.faraday_rotation;
    const double nu_obs = std::max(d.freq_cgs, 1.0);
    P.rm.L = float(d.dlambda_scale * nu_obs);
    P.rm.bconst = float(2.0 * KP_HPL * nu_obs * nu_obs * nu_obs / (KP_CL * KP_CL));
    P.rm.x0 = float(KP_HPL / (KP_ME * KP_CL * KP_CL));
    P.inv_nu3 = 1.0 / (nu_obs * nu_obs * nu_obs);
    P.mass_f = float(P.mass);
    P.spin_f = float(P.spin);
}

// The constants of the mixed-precision GRMHD coefficients (thermal fits in
// FP32, nu^3-scaled units): the GRMHD model's fit, scales and polarisation
// cap, the sin(theta) floor 1e-4 of its coefficient evaluation
// (grmhd_grid.hpp:1210-1240), and the length unit GM/c^2 of mbh_solar
// (the dl/dlambda of frequency nu is that length over nu).
static void prepare_mixed_grmhd(Params& P, double mbh_solar) {
    const GrmhdDev& G = P.grmhd;
    RiafDev<float>& f = P.rm.f;
    f = RiafDev<float>{};
    f.fit = G.fit;
    f.faraday_rotation = G.faraday_rotation;
    f.emission_scale = float(G.emission_scale);
    f.absorption_scale = float(G.absorption_scale);
    f.faraday_scale = float(G.faraday_scale);
    f.max_pol_frac = float(G.max_pol_frac);
    f.min_sin_theta = 1e-4f;
    const double cl = KP_CL;
    P.rm.L = float(6.67430e-8 * mbh_solar * 1.98847e33 / (cl * cl));
    P.rm.bconst = 0.0f;  // per frequency: P.bconst_f, P.bconst_c
    P.rm.x0 = float(KP_HPL / (KP_ME * KP_CL * KP_CL));
    P.mass_f = float(P.mass);
    P.spin_f = float(P.spin);
}

// Frequency lists of a multi-frequency image (host copies); nfreq = 0 for
// the single-frequency image at model->freq_cgs.
struct FreqLists {
    int nfreq = 0, nctrl = 0;
    const double* freqs = nullptr;
    const double* ctrl = nullptr;
};

// mbh_solar: the model's black-hole mass, which sets its dl/dlambda.
static kp_status prepare_frequencies(double mbh_solar, const FreqLists& fl, Params& P) {
    if (fl.nfreq < 1 || fl.nfreq > KP_MAX_FREQUENCIES || !fl.freqs)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "frequencies: 1 to KP_MAX_FREQUENCIES values required");
    if (fl.nctrl < 1 || fl.nctrl > KP_MAX_FREQUENCIES || !fl.ctrl)
        return kp_internal::fail(KP_INVALID_ARGUMENT,
                                 "step-control frequencies: 1 to KP_MAX_FREQUENCIES values required");
    P.nfreq = fl.nfreq;
    P.nctrl = fl.nctrl;
    // same control and output lists, at most 8 frequencies: two Stokes
    // buffers of 4 nfreq doubles per lane stay within the LDS of the
    // resident waves
    P.single_pass = 0;
    if (fl.nfreq == fl.nctrl && fl.nfreq <= 8) {
        bool same = true;
        for (int f = 0; f < fl.nfreq; ++f) same = same && fl.freqs[f] == fl.ctrl[f];
        P.single_pass = same ? 1 : 0;
    }
    const double cl = KP_CL;
    const double l_unit = 6.67430e-8 * mbh_solar * 1.98847e33 / (cl * cl);
    auto fill = [&](const double* src, int n, double* f, double* dls, float* bconst, double* inv3) {
        for (int i = 0; i < n; ++i) {
            if (!(src[i] > 0.0) || !std::isfinite(src[i])) return false;
            const double nu = std::max(src[i], 1.0);
            f[i] = src[i];
            dls[i] = l_unit / nu;  // the model copy's dlambda_scale (src/model/riaf.hpp:76-85)
            bconst[i] = float(2.0 * KP_HPL * nu * nu * nu / (cl * cl));
            if (inv3) inv3[i] = 1.0 / (nu * nu * nu);
        }
        return true;
    };
    if (!fill(fl.freqs, fl.nfreq, P.freqs, P.dls_f, P.bconst_f, P.inv_nu3_f) ||
        !fill(fl.ctrl, fl.nctrl, P.ctrl, P.dls_c, P.bconst_c, nullptr))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "frequencies must be positive and finite");
    return KP_OK;
}

// Tangent request: evaluation point (q_density, q_temperature, q_magnetic)
// and the [3][4][npix] device array for the derivatives.
struct Tangent {
    const double* log_scales = nullptr;
    double* grad = nullptr;
};

