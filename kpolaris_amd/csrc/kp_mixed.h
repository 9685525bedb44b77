// kp_mixed.h -- mixed FP32/FP64 arithmetic for the imaging kernels.
//
// * Geodesic attempts: the transport state stays FP64; every RHS evaluation
//   and the RK stage accumulators run in FP32, and each stage's increment is
//   added to the FP64 base.  FP32 derivative errors therefore enter the state
//   as h*f*eps32 per step, not as a loss of the state's own precision.  The
//   step-doubling error estimate compares the two FP64 end states.
// * Plasma coefficients and the Stokes transfer run in FP32 in rescaled units:
//   I' = nu_obs^3 I_inv (the observed specific intensity, ~1e-4 cgs) and
//   coefficients per unit affine parameter (j_h = nu_obs^3 j_inv dl/dlambda,
//   a_h = a_inv dl/dlambda, rho_h likewise), so the invariant quantities'
//   extreme exponents (j_inv ~ 1e-45 and below) never meet FP32's range.
//   Every expression is ordered so that no intermediate leaves that range.
// The algorithms are the reference's (src/geodesic/rk4.hpp:109-161,
// src/model/riaf.hpp:410-549, src/radiation/thermal_synchrotron.hpp:311-616,
// src/radiation/semi_analytic.hpp:9-100); only the arithmetic differs.
#pragma once

#include "kp_device.h"
#include "kp_riaf.h"

namespace kpd {

// FP32 stage helpers: y, acc in FP32, the final increment added to an FP64 base.
KP_DEV void stage_mid_f(float m, float a, St<float>& y, const St<float>& base, St<float>& acc, float c,
                        uint32_t* cnt) {
    ++*cnt;
    bl_rhs_apply(m, a, y, [&](int i, float f) {
        acc.v[i] = acc.v[i] + 2.0f * f;
        y.v[i] = base.v[i] + f * c;
    });
}

// One step-doubling attempt with FP32 derivatives and FP64 state.
// s: FP64 base (registers or LDS); half, two_half: FP64 outputs.
template <class S>
KP_DEV double doubling_attempt_mixed(float m, float a, const S& s, double h, St<double>& half,
                                     St<double>& two_half, uint32_t* cnt, bool full = true) {
    St<float> s32, k1, acc, y;
#pragma unroll
    for (int i = 0; i < 16; ++i) s32.v[i] = float(s[i]);
    const float hf = float(h);
    const float h2 = hf * 0.5f;
    const float q2 = 0.5f * h2;
    const double h2d = h * 0.5;
    ++*cnt;
    bl_rhs_apply(m, a, s32, [&](int i, float f) { k1.v[i] = f; });
    // first half step
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        acc.v[i] = k1.v[i];
        y.v[i] = s32.v[i] + k1.v[i] * q2;
    }
    stage_mid_f(m, a, y, s32, acc, q2, cnt);
    stage_mid_f(m, a, y, s32, acc, h2, cnt);
    ++*cnt;
    const float c6 = float(h2d / 6.0);
    bl_rhs_apply(m, a, y, [&](int i, float f) { half.v[i] = s[i] + double((acc.v[i] + f) * c6); });
    // second half step from the FP64 mid state
#pragma unroll
    for (int i = 0; i < 16; ++i) s32.v[i] = float(half.v[i]);
    ++*cnt;
    bl_rhs_apply(m, a, s32, [&](int i, float f) {
        acc.v[i] = f;
        y.v[i] = s32.v[i] + f * q2;
    });
    stage_mid_f(m, a, y, s32, acc, q2, cnt);
    stage_mid_f(m, a, y, s32, acc, h2, cnt);
    ++*cnt;
    bl_rhs_apply(m, a, y, [&](int i, float f) { two_half.v[i] = half.v[i] + double((acc.v[i] + f) * c6); });
    if (!full) return 0.0;
    // full step from s, compared with two_half in FP64
#pragma unroll
    for (int i = 0; i < 16; ++i) s32.v[i] = float(s[i]);
    const float hh = 0.5f * hf;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        acc.v[i] = k1.v[i];
        y.v[i] = s32.v[i] + k1.v[i] * hh;
    }
    stage_mid_f(m, a, y, s32, acc, hh, cnt);
    stage_mid_f(m, a, y, s32, acc, hf, cnt);
    double err = 0.0;
    const float c6f = float(h / 6.0);
    ++*cnt;
    bl_rhs_apply(m, a, y, [&](int i, float f) {
        const double full = s[i] + double((acc.v[i] + f) * c6f);
        const double two = two_half.v[i];
        const double sc = maxv(1.0, maxv(absv(full), absv(two)));
        err = maxv(err, absv(full - two) / sc);
    });
    return err;
}

// RK4 step with FP32 derivatives (crossing refinement in fixed-step mode).
template <class S> KP_DEV void rk4_mixed(float m, float a, const S& s, double h, St<double>& out, uint32_t* cnt) {
    St<float> s32, acc, y;
#pragma unroll
    for (int i = 0; i < 16; ++i) s32.v[i] = float(s[i]);
    const float hf = float(h), hh = 0.5f * hf;
    ++*cnt;
    bl_rhs_apply(m, a, s32, [&](int i, float f) {
        acc.v[i] = f;
        y.v[i] = s32.v[i] + f * hh;
    });
    stage_mid_f(m, a, y, s32, acc, hh, cnt);
    stage_mid_f(m, a, y, s32, acc, hf, cnt);
    const float c6 = float(h / 6.0);
    ++*cnt;
    bl_rhs_apply(m, a, y, [&](int i, float f) { out.v[i] = s[i] + double((acc.v[i] + f) * c6); });
}

// ---------------------------------------------------------------------------
// Rescaled FP32 plasma coefficients.
// ---------------------------------------------------------------------------
struct RiafMixed {
    RiafDev<float> f;        // model constants in FP32
    float L;        // length unit GM/c^2 (cm) = nu_obs * dl/dlambda
    float bconst;   // 2 h nu_obs^3 / c^2: the rescaled source function is bconst / expm1(x)
    float x0;       // h / (m_e c^2): x = h nu / (m_e c^2 theta_e) = x0 * nu / theta_e
};

// Thermal Pandya/Dexter coefficients in the magnetic basis, rescaled:
// j_h = j_phys * L / g^2, a_h = j_h / S' with S' = nu_obs^3 B_nu / nu^3 =
// bconst / expm1(x) (Kirchhoff's law in the rescaled units), and
// rho_h = rho_phys * g * L.  Follows
// thermal_synchrotron_magnetic_basis_coefficients_fit (thermal_synchrotron.hpp:512-616).
KP_DEV void thermal_coeffs_scaled(const RiafMixed& R, float freq, float bconst, float g, float ne,
                                  float thetae, float b,
                                  float sin_theta_in, float cos_theta, Coeffs<float>& c) {
    coeffs_zero(c);
    const RiafDev<float>& M = R.f;
    const float nu = fmaxf(freq * g, 1.0f);
    if (!(ne > 0.0f) || !(thetae > 0.0f) || !(b > 0.0f)) return;
    const float pi = float(PI_D);
    const float ee = float(KP_EE), me = float(KP_ME), cl = float(KP_CL);
    const float sin_theta = fmaxf(sin_theta_in, M.min_sin_theta);
    float jI, jQ, jV;  // physical emissivities (cgs), ~1e-15..1e-30 or 0 on underflow
    if (M.fit == 4) {
        const float nus = kdiv(3.0f * ee * b * sin_theta, 4.0f * pi * me * cl) * thetae * thetae + 1.0f;
        const float x = fmaxf(kdiv(nu, fmaxf(nus, 1e-30f)), 1e-30f);
        const float xm13 = krcp(cbrtf(x));
        const float exp_shape = expf(kdiv(-1.8899f, xm13));
        const float pref = kdiv(ne * ee * ee * nu, 2.0f * sqrtf(3.0f) * cl * thetae * thetae);
        jI = pref * (2.5651f * (1.0f + 1.92f * xm13 + 0.9977f * xm13 * xm13) * exp_shape);
        jQ = pref * (2.5651f * (1.0f + 0.93193f * xm13 + 0.499873f * xm13 * xm13) * exp_shape);
        const float iv = (kdiv(1.81384f, x) + 3.42319f * xm13 * xm13 + kdiv(0.0292545f, sqrtf(x)) + 2.03773f * xm13) * exp_shape;
        float cft = cos_theta;
        if (fabsf(cft) < 1e-10f) cft = cft >= 0.0f ? 1e-10f : -1e-10f;
        const float tan_theta = kdiv(sin_theta, cft);
        jV = kdiv(2.0f * ne * ee * ee * nu, tan_theta * 3.0f * sqrtf(3.0f) * cl * thetae * thetae * thetae) * iv;
    } else {
        const float nu_c = ee * b * (1.0f / (2.0f * pi * me * cl));
        const float nu_s = 2.0f / 9.0f * nu_c * sin_theta * thetae * thetae;
        const float x = fmaxf(kdiv(nu, fmaxf(nu_s, 1e-30f)), 1e-30f);
        const float pref = ne * ee * ee * nu_c * (1.0f / cl);
        const float term1 = sqrtf(2.0f) * pi / 27.0f * sin_theta;
        const float x13 = cbrtf(x);
        const float sx = sqrtf(x), sx13 = sqrtf(x13);
        const float ex13 = expf(-x13);
        const float shapeI = sx + 1.8877486253633869933f * sx13;
        jI = pref * term1 * (shapeI * shapeI) * ex13;
        const float t2425 = powf(thetae, 24.0f / 25.0f);
        const float term2 = kdiv(7.0f * t2425 + 35.0f, 10.0f * t2425 + 75.0f);
        const float shapeQ = sx + term2 * 1.8877486253633869933f * sx13;
        jQ = pref * term1 * (shapeQ * shapeQ) * ex13;
        const float sin_shift = sin_theta * 0.4356824462767121f - cos_theta * 0.9001004421765051f;
        const float v1 = kdiv(37.0f - 87.0f * sin_shift, 100.0f * (thetae + 1.0f));
        const float v2 = powf(1.0f + (powf(thetae, 0.6f) * 0.04f + 0.7f) * powf(x, 0.36f), 5.0f / 3.0f);
        jV = pref * v1 * v2 * ex13;
    }
    // rescaled emissivities and the polarisation cap
    const float jscale = kdiv(R.L, g * g);
    const float jIh = jI * jscale;
    float jQh = jQ * jscale;
    float jVh = jV * jscale;
    const float jps = fmaxf(fabsf(jQh), fabsf(jVh));
    // a subnormal jps has no finite reciprocal: the polarised emissivity has
    // underflowed and is left uncapped
    const bool jps_normal = jps >= 1.17549435e-38f;
    const float ijps = jps_normal ? krcp(jps) : 0.0f;
    const float jP = jps_normal ? jps * sqrtf(sq(jQh * ijps) + sq(jVh * ijps)) : 0.0f;
    if (jIh > 0.0f && jP > 0.0f && jIh < kdiv(jP, fmaxf(M.max_pol_frac, 1e-30f))) {
        const float scale = kdiv(jIh, jP) * M.max_pol_frac;
        jQh *= scale;
        jVh *= scale;
    }
    // Kirchhoff: a_h = j_h / S'.  S' below FLT_MIN (expm1 overflowed, or a
    // subnormal S' whose reciprocal is infinite) leaves a_h = 0, as the
    // double-precision path does once B_nu underflows.
    float aI = 0.0f, aQ = 0.0f, aV = 0.0f;
    {
        const float x = kdiv(R.x0 * nu, thetae);
        const float sp = kdiv(bconst, expm1f(x));
        if (sp >= 1.17549435e-38f) {
            const float isp = krcp(sp);
            aI = jIh * isp;
            aQ = jQh * isp;
            aV = jVh * isp;
        }
    }
    // Faraday rotation and conversion (symphony_rho_QV, thermal_synchrotron.hpp:461-506)
    float rhoQ, rhoV;
    {
        const float omega0 = ee * b * (1.0f / (me * cl));
        const float two_pi_nu = 2.0f * pi * nu;
        const float safe = fmaxf(two_pi_nu, 1e-30f);
        const float isafe = krcp(safe);
        const float wp2 = 4.0f * pi * ne * (ee * ee / me);
        const float xx = fmaxf(thetae * sqrtf(fmaxf(0.0f, sqrtf(2.0f) * sin_theta * (1e3f * omega0 * isafe))), 1e-12f);
        float k0, k1, k2;
        bessel_k012(krcp(thetae), k0, k1, k2);
        const float exp_x_47 = expf(-xx * (1.0f / 47.2f));
        const float extraterm = (0.011f * exp_x_47 - 369.69342869090815f * powf(xx + 1e-16f, -8.0f / 3.0f)) *
                                (0.5f + 0.5f * tanhf((logf(xx) - 4.787491742782046f) * 10.0f));
        const float jffunc = 2.011f * expf(-powf(xx, 1.035f) * (1.0f / 4.7f)) -
                             cosf(xx / 2.0f) * expf(-powf(xx, 1.2f) * (1.0f / 2.73f)) - 0.011f * exp_x_47 + extraterm;
        const float k1_ratio = k2 > 0.0f ? kdiv(k1, k2) : 1.0f;
        // eps11m22 = jffunc wp2 omega0^2 / (2 pi nu)^4, kept in range by scaling by 1/(2 pi nu) early
        const float w = omega0 * isafe;
        const float e11 = jffunc * (wp2 * isafe) * w * (w * isafe) * (k1_ratio + 6.0f * thetae) * sin_theta * sin_theta;
        rhoQ = pi * nu * (1.0f / cl) * e11;
        const float shgm = 1.0f - 0.11f * logf(1.0f + 0.035f * xx);
        const float k0_ratio = k2 > 0.0f ? kdiv(k0, k2) : 1.0f;
        const float e12 = (wp2 * isafe) * w * isafe * k0_ratio * shgm * cos_theta;
        rhoV = 2.0f * pi * nu * (1.0f / cl) * e12;
    }
    const float rscale = g * R.L;  // rho_inv * dl/dlambda = rho_phys * g * L
    c.jI = M.emission_scale * jIh;
    c.jQ = M.emission_scale * jQh;
    c.jV = M.emission_scale * jVh;
    c.aI = M.absorption_scale * aI;
    c.aQ = M.absorption_scale * aQ;
    c.aV = M.absorption_scale * aV;
    c.rQ = M.faraday_scale * rhoQ * rscale;
    c.rV = M.faraday_scale * rhoV * rscale;
}

// The frequency-independent part of the RIAF coefficients in FP32
// (riaf.hpp:410-500, 529-547; the FP64 form is riaf_local in kp_riaf.h).
KP_DEV void riaf_local_scaled(const RiafMixed& R, const St<double>& sd, RiafLocal<float>& L) {
    L.valid = 0;
    L.aligned = 0;
    L.has_chi = 0;
    const RiafDev<float>& M = R.f;
    St<float> s;
#pragma unroll
    for (int i = 0; i < 16; ++i) s.v[i] = float(sd.v[i]);
    const float* x = s.v;
    const float* k = s.v + 4;
    const float r = fmaxf(x[1], M.min_radius);
    if (r <= M.rh + M.horizon_buffer || r < M.r_min || r > M.r_max) return;
    float sn, cs;
    sincosf(x[2], &sn, &cs);
    const float zc = r * cs;
    const float rc = fmaxf(r * sn, M.min_radius);
    const float hd = fmaxf(M.disk_h, 1e-6f);
    const float ne_norm = M.nth0 * expf(kdiv(-(zc * zc), 2.0f * rc * rc * hd * hd)) * powf(r, M.pow_nth);
    if (!(ne_norm > 0.0f)) return;
    float ne = ne_norm * M.ne_unit;
    float thetae = fmaxf(M.Te0 * powf(r, M.pow_T) * M.thetae_unit, M.min_thetae);
    // 8 pi 0.1 m_p c^2 / 6 folded into one constant (in range for FP32)
    const float bk = float(8.0 * PI_D * 0.1 * 1.67262192369e-24 * KP_CL * KP_CL / 6.0);
    float b = fmaxf(sqrtf(fmaxf(kdiv(bk * ne, r), 0.0f)), 1e-12f);
    ne *= M.ne_factor;
    thetae *= M.thetae_factor;
    b *= M.b_factor;
    const BLGcov<float> g = bl_gcov(M.mass, M.spin, x[1], sn, cs);
    float u[4];
    riaf_fluid_u(M, r, sn, cs, g, u);
    float ucov[4];
    glower(g, u, ucov);
    const float uk = ucov[0] * k[0] + ucov[1] * k[1] + ucov[2] * k[2] + ucov[3] * k[3];
    const float gscale = fmaxf(fabsf(-uk), M.min_frequency_scale);
    float bc[4], bcov[4];
    riaf_b_unit(g, u, bc);
    glower(g, bc, bcov);
    const float kdotb = k[0] * bcov[0] + k[1] * bcov[1] + k[2] * bcov[2] + k[3] * bcov[3];
    const float cos_theta = clampv(kdiv(kdotb, fmaxf(gscale, 1e-30f)), -1.0f, 1.0f);
    L.valid = 1;
    L.ne = ne;
    L.thetae = thetae;
    L.b = b;
    L.nu_scale = gscale;
    L.cos_theta = cos_theta;
    L.sin_theta = fmaxf(sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta)), M.min_sin_theta);
    L.aligned = cos_theta == 1.0f || cos_theta == -1.0f;
    if (L.aligned) return;
    const float ukc = gdot(g, u, k);
    const float b1 = screen_inner(g, u, k, s.v + 8, bc, ukc);
    const float b2 = screen_inner(g, u, k, s.v + 12, bc, ukc);
    const float bp2 = b1 * b1 + b2 * b2;
    if (bp2 > 1e-30f) {
        const float ib = krcp(bp2);
        L.has_chi = 1;
        L.cos2chi = (b2 * b2 - b1 * b1) * ib;
        L.sin2chi = -2.0f * b1 * b2 * ib;
    }
}

// Rescaled FP32 coefficients at observing frequency freq; bconst = 2 h freq^3 / c^2.
KP_DEV void riaf_coeffs_scaled_at(const RiafMixed& R, const RiafLocal<float>& L, float freq, float bconst,
                                  Coeffs<float>& c) {
    coeffs_zero(c);
    if (!L.valid) return;
    const RiafDev<float>& M = R.f;
    if (L.aligned) {
        // exactly field-aligned: only rho_V survives (riaf.hpp:505-512)
        Coeffs<float> tmp;
        thermal_coeffs_scaled(R, freq, bconst, L.nu_scale, L.ne, L.thetae, L.b, M.min_sin_theta, L.cos_theta, tmp);
        c.rV = M.faraday_rotation ? tmp.rV : 0.0f;
        return;
    }
    Coeffs<float> mc;
    thermal_coeffs_scaled(R, freq, bconst, L.nu_scale, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, mc);
    c.jI = mc.jI;
    c.jV = mc.jV;
    c.aI = mc.aI;
    c.aV = mc.aV;
    c.rV = M.faraday_rotation ? mc.rV : 0.0f;
    if (L.has_chi) {
        c.jQ = mc.jQ * L.cos2chi;
        c.jU = mc.jQ * L.sin2chi;
        c.aQ = mc.aQ * L.cos2chi;
        c.aU = mc.aQ * L.sin2chi;
        c.rQ = mc.rQ * L.cos2chi;
        c.rU = mc.rQ * L.sin2chi;
    }
}

// RIAF coefficients in FP32, rescaled (riaf.hpp:410-549).
KP_DEV void riaf_coefficients_scaled(const RiafMixed& R, const St<double>& sd, Coeffs<float>& c) {
    RiafLocal<float> L;
    riaf_local_scaled(R, sd, L);
    riaf_coeffs_scaled_at(R, L, R.f.freq_cgs, R.bconst, c);
}

// FP32 Strang step with expm1-based attenuation/source factors (exact for all
// depths, no small-argument cancellation); a per-step Faraday angle above one
// radian is applied in FP64.
KP_DEV float att32(float kappa, float dl) {
    const float x = kappa * dl;
    return x > 87.0f ? 0.0f : expf(-x);
}

KP_DEV float src32(float kappa, float dl) {
    const float x = kappa * dl;
    if (fabsf(x) < 1e-4f) return dl * (1.0f - x * 0.5f + x * x * (1.0f / 6.0f));
    return kdiv(-expm1f(-x), kappa);
}

KP_DEV void faraday32(Stokes<float>& s, const Coeffs<float>& c, float dl) {
    const float rn = norm3(c.rQ, c.rU, c.rV);
    if (rn == 0.0f || dl == 0.0f) return;
    const float theta = rn * dl;
    if (fabsf(theta) > 1.0f) {
        // large Faraday angle: FP64 rotation
        Stokes<double> d{double(s.I), double(s.Q), double(s.U), double(s.V)};
        Coeffs<double> cd;
        coeffs_zero(cd);
        cd.rQ = c.rQ;
        cd.rU = c.rU;
        cd.rV = c.rV;
        faraday_rotate(d, cd, double(dl));
        s.Q = float(d.Q);
        s.U = float(d.U);
        s.V = float(d.V);
        return;
    }
    faraday_rotate(s, c, dl);
}

KP_DEV void absorb_emit32(Stokes<float>& s, const Coeffs<float>& c, float dl) {
    const float px = s.Q, py = s.U, pz = s.V;
    const float an = norm3(c.aQ, c.aU, c.aV);
    const float ei = att32(c.aI, dl);
    const float fi = src32(c.aI, dl);
    if (an < 1e-30f) {
        s.I = ei * s.I + fi * c.jI;
        s.Q = ei * s.Q + fi * c.jQ;
        s.U = ei * s.U + fi * c.jU;
        s.V = ei * s.V + fi * c.jV;
        return;
    }
    const float inv = krcp(an);
    const float nx = c.aQ * inv, ny = c.aU * inv, nz = c.aV * inv;
    const float ppar = nx * px + ny * py + nz * pz;
    const float jpar = nx * c.jQ + ny * c.jU + nz * c.jV;
    const float kp = c.aI + an, km = c.aI - an;
    const float sp = att32(kp, dl) * (s.I + ppar) + src32(kp, dl) * (c.jI + jpar);
    const float sm = att32(km, dl) * (s.I - ppar) + src32(km, dl) * (c.jI - jpar);
    const float pout = 0.5f * (sp - sm);
    s.I = 0.5f * (sp + sm);
    s.Q = nx * pout + ((px - nx * ppar) * ei + (c.jQ - nx * jpar) * fi);
    s.U = ny * pout + ((py - ny * ppar) * ei + (c.jU - ny * jpar) * fi);
    s.V = nz * pout + ((pz - nz * ppar) * ei + (c.jV - nz * jpar) * fi);
}

KP_DEV void strang_step32(Stokes<float>& s, const Coeffs<float>& c, float dl) {
    faraday32(s, c, 0.5f * dl);
    absorb_emit32(s, c, dl);
    faraday32(s, c, 0.5f * dl);
}

}  // namespace kpd
