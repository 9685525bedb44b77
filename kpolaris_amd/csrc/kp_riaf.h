// kp_riaf.h -- analytic RIAF plasma model on the device, Boyer-Lindquist metric.
// Restates RIAFAnalyticRadiationModel::coefficients and its helpers
// (src/model/riaf.hpp:54-96, 137-273, 320-386, 410-549) for one sample state.
#pragma once

#include "kp_device.h"
#include "kp_nonthermal.h"

namespace kpd {

// Model constants; the host precomputes every ray-independent quantity
// (units, horizon, ISCO, perturbation factors) in double.
template <class T> struct RiafDev {
    T nth0, Te0, disk_h, pow_nth, pow_T, ne_unit, thetae_unit;
    T r_min, r_max, horizon_buffer;
    T emission_scale, absorption_scale, faraday_scale, max_pol_frac;
    T keplerian_factor, infall_factor;
    T min_frequency_scale, min_sin_theta, min_thetae, min_radius;
    T freq_cgs, dlambda_scale;
    T rh, risco;                           // horizon and ISCO radii
    T ne_factor, thetae_factor, b_factor;  // plasma perturbation (1 = none)
    T mass, spin;
    int fit;
    int faraday_rotation;
};

template <class T> struct RiafMetric {
    T g00, g03, g11, g22, g33;   // gcov
    T c00, c03, c11, c22, c33;   // gcon
};

// bl_metric (riaf.hpp:137-166): the model's unit-mass BL metric with its own floors.
template <class T> KP_DEV RiafMetric<T> riaf_bl_metric(T r, T sn, T cs, T a) {
    RiafMetric<T> g;
    const T a2 = a * a;
    const T sin2 = maxv(sn * sn, T(1e-30));
    const T sigma = r * r + a2 * cs * cs;
    const T delta = r * r - T(2) * r + a2;
    const T isig = T(1) / sigma;
    g.g00 = -(T(1) - T(2) * r * isig);
    g.g03 = -T(2) * a * r * sin2 * isig;
    g.g11 = sigma / maxv(delta, tiny<T>());
    g.g22 = sigma;
    g.g33 = (r * r + a2 + T(2) * a2 * r * sin2 * isig) * sin2;
    const T isd = T(1) / maxv(sigma * delta, tiny<T>());
    g.c00 = -((r * r + a2) * (r * r + a2) - a2 * delta * sin2) * isd;
    g.c03 = -T(2) * a * r * isd;
    g.c11 = delta * isig;
    g.c22 = isig;
    g.c33 = (delta - a2 * sin2) / maxv(sigma * delta * sin2, tiny<T>());
    return g;
}

template <class T> KP_DEV T signed_nonzero(T x) {
    if (absv(x) > tiny<T>()) return x;
    return x < T(0) ? -tiny<T>() : tiny<T>();
}

// fluid_four_velocity_from_bl_coords, BL branch (riaf.hpp:210-273).
template <class T>
KP_DEV void riaf_fluid_u(const RiafDev<T>& M, T r, T sn, T cs, const BLGcov<T>& gx, T* u) {
    const RiafMetric<T> g = riaf_bl_metric(r, sn, cs, M.spin);
    T vt = T(1), vr = T(0), vphi = T(0);
    if (r < M.rh) {
        vt = T(1) / sqrt(maxv(-g.c00, tiny<T>()));
    } else if (r < M.risco) {
        const RiafMetric<T> gi = riaf_bl_metric(M.risco, sn, cs, M.spin);
        const T omega_isco = T(1) / (M.risco * sqrt(M.risco) + M.spin);
        const T qi = gi.g00 + T(2) * omega_isco * gi.g03 + omega_isco * omega_isco * gi.g33;
        const T ut_i = T(1) / sqrt(maxv(-qi, tiny<T>()));
        const T e = (gi.g00 + gi.g03 * omega_isco) * ut_i;
        const T l = (gi.g03 + gi.g33 * omega_isco) * ut_i;
        const T k_con = g.c00 * e * e + T(2) * g.c03 * e * l + g.c33 * l * l;
        const T ur_cov = -sqrt(maxv(-(T(1) + k_con) / maxv(g.c11, tiny<T>()), T(0)));
        const T vr_tmp = g.c11 * ur_cov;
        const T vt_tmp = g.c00 * e + g.c03 * l;
        const T vphi_tmp = g.c03 * e + g.c33 * l;
        const T omega_k = vphi_tmp / maxv(vt_tmp, tiny<T>());
        const T omega_ff = g.c03 / signed_nonzero(g.c00);
        const T omega = omega_k + (T(1) - M.keplerian_factor) * (omega_ff - omega_k);
        const T ur_ff = -sqrt(maxv(-(T(1) + g.c00) * g.c11, T(0)));
        vr = vr_tmp + M.infall_factor * (ur_ff - vr_tmp);
        const T q = g.g00 + T(2) * omega * g.g03 + omega * omega * g.g33;
        vt = sqrt(maxv(-(T(1) + vr * vr * g.g11) / q, T(0)));
        vphi = omega * vt;
    } else {
        const T r_eff = maxv(r, M.min_radius);
        const T omega_k = T(1) / (r_eff * sqrt(r_eff) + M.spin);
        const T omega_ff = g.c03 / signed_nonzero(g.c00);
        const T omega = omega_k + (T(1) - M.keplerian_factor) * (omega_ff - omega_k);
        vr = M.infall_factor * -sqrt(maxv(-(T(1) + g.c00) * g.c11, T(0)));
        const T q = g.g00 + T(2) * omega * g.g03 + omega * omega * g.g33;
        vt = sqrt(maxv(-(T(1) + vr * vr * g.g11) / q, T(0)));
        vphi = omega * vt;
    }
    u[0] = vt;
    u[1] = vr;
    u[2] = T(0);
    u[3] = vphi;
    const T n2 = gdot(gx, u, u);
    if (n2 < T(0)) {
        const T f = T(1) / sqrt(maxv(-n2, tiny<T>()));
#pragma unroll
        for (int i = 0; i < 4; ++i) u[i] = u[i] * f;
    }
}

// magnetic_unit_four_vector_with_gcov, BL branch (riaf.hpp:350-386):
// toroidal trial vector projected orthogonal to u, with fallbacks.
template <class T> KP_DEV void riaf_b_unit(const BLGcov<T>& g, const T* u, T* b) {
    T trial[4] = {T(0), T(0), T(0), T(1)};
    T udotb = gdot(g, u, trial);
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = trial[i] + u[i] * udotb;
    T bsq = gdot(g, b, b);
    if (!(bsq > T(1e-30))) {
        // the reference retries the same (0,0,0,1) trial once, then (0,1,0,0)
        const T tz[4] = {T(0), T(0), T(0), T(1)};
        const T uz = gdot(g, u, tz);
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = tz[i] + u[i] * uz;
        bsq = gdot(g, b, b);
    }
    if (!(bsq > T(1e-30))) {
        const T tx[4] = {T(0), T(1), T(0), T(0)};
        const T ux = gdot(g, u, tx);
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = tx[i] + u[i] * ux;
        bsq = gdot(g, b, b);
    }
    if (bsq > T(1e-30)) {
        const T f = T(1) / sqrt(bsq);
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = b[i] * f;
    }
}

// screen_inner_product_with_gcov (riaf.hpp:320-336).
template <class T>
KP_DEV T screen_inner(const BLGcov<T>& g, const T* u, const T* k, const T* a, const T* b, T uk) {
    if (absv(uk) <= tiny<T>()) return gdot(g, a, b);
    T eK[4];
    const T f = T(-1) / uk;
#pragma unroll
    for (int i = 0; i < 4; ++i) eK[i] = k[i] * f - u[i];
    return gdot(g, a, b) + gdot(g, u, a) * gdot(g, u, b) - gdot(g, eK, a) * gdot(g, eK, b);
}

// The frequency-independent part of RIAFAnalyticRadiationModel::coefficients
// at one sample state: plasma scalars, fluid-frame frequency ratio, field
// angle and the screen rotation of the field.  Only nu = freq * nu_scale and
// dlambda_scale depend on the observing frequency, so a multi-frequency step
// evaluates this once and riaf_coeffs_at once per frequency, with results
// identical to re-running the whole model per frequency (pass_b.hpp:1422-1430).
template <class T> struct RiafLocal {
    int valid;    // inside the model domain with positive density
    int aligned;  // k exactly along b: only rho_V survives (riaf.hpp:505-512)
    int has_chi;  // the field has a screen projection
    T ne, thetae, b, nu_scale, cos_theta, sin_theta, cos2chi, sin2chi;
    T kappa;      // GRMHD nonthermal kappa at the sample (from sigma and beta when variable)
};

// riaf.hpp:410-500 and 529-547 without the frequency.
template <class T> KP_DEV void riaf_local(const RiafDev<T>& M, const St<T>& s, RiafLocal<T>& L) {
    L.valid = 0;
    L.aligned = 0;
    L.has_chi = 0;
    const T* x = s.v;
    const T* k = s.v + 4;
    const T r = maxv(x[1], M.min_radius);
    if (r <= M.rh + M.horizon_buffer || r < M.r_min || r > M.r_max) return;
    T sn, cs;
    ksincos(x[2], &sn, &cs);
    // density_profile_bl (riaf.hpp:54-60)
    const T zc = r * cs;
    const T rc = maxv(r * sn, M.min_radius);
    const T hd = maxv(M.disk_h, T(1e-6));
    const T ne_norm = M.nth0 * exp(-(zc * zc) / (T(2) * rc * rc * hd * hd)) * pow(r, M.pow_nth);
    if (!(ne_norm > T(0))) return;
    T ne = ne_norm * M.ne_unit;
    // thetae_profile (riaf.hpp:67-74), magnetic_field_cgs (riaf.hpp:87-96)
    T thetae = maxv(M.Te0 * pow(r, M.pow_T) * M.thetae_unit, M.min_thetae);
    const T mp = T(1.67262192369e-24);
    T b = maxv(sqrt(maxv(T(8) * T(PI_D) * T(0.1) * ne * mp * T(KP_CL) * T(KP_CL) / (T(6) * r), T(0))),
               T(1e-12));
    // PlasmaPerturbation::apply (src/radiation/plasma_perturbation.hpp:17-32)
    ne *= M.ne_factor;
    thetae *= M.thetae_factor;
    b *= M.b_factor;

    const BLGcov<T> g = bl_gcov(M.mass, M.spin, x[1], sn, cs);
    T u[4];
    riaf_fluid_u(M, r, sn, cs, g, u);
    T ucov[4];
    glower(g, u, ucov);
    const T uk = ucov[0] * k[0] + ucov[1] * k[1] + ucov[2] * k[2] + ucov[3] * k[3];
    const T nu_scale = maxv(absv(-uk), M.min_frequency_scale);
    T bc[4], bcov[4];
    riaf_b_unit(g, u, bc);
    glower(g, bc, bcov);
    const T kdotb = k[0] * bcov[0] + k[1] * bcov[1] + k[2] * bcov[2] + k[3] * bcov[3];
    const T cos_theta = clampv(kdotb / maxv(nu_scale, T(1e-30)), T(-1), T(1));
    L.valid = 1;
    L.ne = ne;
    L.thetae = thetae;
    L.b = b;
    L.nu_scale = nu_scale;
    L.cos_theta = cos_theta;
    L.sin_theta = maxv(sqrt(maxv(T(0), T(1) - cos_theta * cos_theta)), M.min_sin_theta);
    // theta = acos(cos_theta) is 0 or pi exactly when cos_theta = +-1.
    L.aligned = cos_theta == T(1) || cos_theta == T(-1);
    if (L.aligned) return;
    // screen projections of the field: e2 along projected B (riaf.hpp:529-547)
    const T ukc = gdot(g, u, k);
    const T b1 = screen_inner(g, u, k, s.v + 8, bc, ukc);
    const T b2 = screen_inner(g, u, k, s.v + 12, bc, ukc);
    const T bp2 = b1 * b1 + b2 * b2;
    if (bp2 > T(1e-30)) {
        const T ib = T(1) / bp2;
        L.has_chi = 1;
        L.cos2chi = (b2 * b2 - b1 * b1) * ib;
        L.sin2chi = -T(2) * b1 * b2 * ib;
    }
}

// riaf.hpp:500-549 at observing frequency freq; N carries the nonthermal
// distribution (kappa, power law: riaf.hpp:497-522) in FP64.
template <class T>
KP_DEV void riaf_coeffs_at(const RiafDev<T>& M, const RiafLocal<T>& L, T freq, Coeffs<T>& c,
                           const NtDev* N = nullptr) {
    coeffs_zero(c);
    if (!L.valid) return;
    const T nu = maxv(freq * L.nu_scale, T(1));
    if (L.aligned) {
        const T theta = L.cos_theta == T(1) ? T(0) : T(PI_D);
        c.rV = M.faraday_scale * rho_v_aligned(nu, L.ne, L.thetae, L.b, theta, M.min_sin_theta) * nu;
        if (!M.faraday_rotation) c.rV = T(0);
        return;
    }
    Coeffs<T> mc;
    if constexpr (sizeof(T) == sizeof(double)) {
        if (N && (M.fit == KP_EMISSION_KAPPA || M.fit == KP_EMISSION_POWER_LAW)) {
            nonthermal_coeffs(*N, N->kc, nu, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, mc);
        } else {
            thermal_coeffs(M.fit, nu, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, M.max_pol_frac,
                           M.min_sin_theta, M.emission_scale, M.absorption_scale, M.faraday_scale, mc);
        }
    } else {
        thermal_coeffs(M.fit, nu, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, M.max_pol_frac, M.min_sin_theta,
                       M.emission_scale, M.absorption_scale, M.faraday_scale, mc);
    }
    c.jI = mc.jI;
    c.jV = mc.jV;
    c.aI = mc.aI;
    c.aV = mc.aV;
    c.rV = M.faraday_rotation ? mc.rV : T(0);
    if (L.has_chi) {
        c.jQ = mc.jQ * L.cos2chi;
        c.jU = mc.jQ * L.sin2chi;
        c.aQ = mc.aQ * L.cos2chi;
        c.aU = mc.aQ * L.sin2chi;
        c.rQ = mc.rQ * L.cos2chi;
        c.rU = mc.rQ * L.sin2chi;
    }
}

// RIAFAnalyticRadiationModel::coefficients for the BL metric (riaf.hpp:410-549).
template <class T>
KP_DEV void riaf_coefficients(const RiafDev<T>& M, const St<T>& s, Coeffs<T>& c, const NtDev* N = nullptr) {
    RiafLocal<T> L;
    riaf_local(M, s, L);
    riaf_coeffs_at(M, L, M.freq_cgs, c, N);
}

}  // namespace kpd
