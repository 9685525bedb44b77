// kp_grmhd.h -- GRMHD grid sampling and coefficients on the device (FMKS data).
//
// Restates GRMHDGridRadiationModel (src/model/grmhd_grid.hpp) for FMKS dumps
// with precomputed derived scalars: cell-centred trilinear interpolation in
// native (x1, x2, phi) with clamped radial ghosts, polar reflection
// (k + n3/2) and periodic phi (:271-430); u^mu rebuilt from the
// normal-observer velocity and b^mu from B^i (:980-1014); the sigma cut
// (:495-503); the thermal fit and the screen rotation of the field
// (:1091-1240).  The derived scalars follow iharm_loader.cpp:420-505.
//
// Data layout: one 72-byte kp_grmhd_cell record per cell ([n1][n2][n3],
// k fastest), so a trilinear sample gathers 8 records rather than 8 entries
// from each of 13 arrays.
#pragma once

#include "kp_device.h"
#include "kp_fmks.h"
#include "kp_riaf.h"
#include "../../include/kp_image.h"

namespace kpd {

struct GrmhdDev {
    int n1, n2, n3, fit;
    double startx1, startx2, startx3, dx1, dx2, dx3;
    double r_in, r_out, spin, gam;
    FmksMap map;
    double sigma_cut, sigma_cut_high;
    double emission_scale, absorption_scale, faraday_scale, max_pol_frac;
    double freq_cgs, length_unit, rho_unit, b_unit, ne_factor;
    double trat_small, trat_large, beta_crit;
    int faraday_rotation;
    int variable_kappa;          // nonthermal kappa from sigma and beta per sample
    double kappa, kappa_min;     // fixed kappa (floored); the floor of every kappa
    const kp_grmhd_cell* cells;
};

// native_fluid_metric (grmhd_grid.hpp:721-766): gcov and the first row of gcon.
KP_DEV void grmhd_native_metric(const GrmhdDev& G, double x1, double x2, double r, double th, Metric4<double>& g,
                                double* gcon0) {
    double sth, cth;
    ksincos(th, &sth, &cth);
    const double s2 = sth * sth;
    const double a2 = G.spin * G.spin;
    const double rho2 = r * r + a2 * cth * cth;
    const double f = 2.0 * r / rho2;
    const FmksTheta<double> t = fmks_theta(G.map, x1, x2);
    const double dth1 = t.dth1, dth2 = t.dth2;
    g.g00 = -1.0 + f;
    g.g01 = f * r;
    g.g02 = 0.0;
    g.g03 = -G.spin * f * s2;
    g.g11 = (1.0 + f) * r * r + rho2 * dth1 * dth1;
    g.g12 = rho2 * dth1 * dth2;
    g.g13 = -G.spin * s2 * (1.0 + f) * r;
    g.g22 = rho2 * dth2 * dth2;
    g.g23 = 0.0;
    g.g33 = s2 * (rho2 + a2 * s2 * (1.0 + f));
    const double inv_r = 1.0 / maxv(r, 1e-300);
    const double inv_dth2 = 1.0 / maxv(dth2, 1e-300);
    gcon0[0] = -(1.0 + f);
    gcon0[1] = f * inv_r;
    gcon0[2] = -f * dth1 * inv_r * inv_dth2;
    gcon0[3] = 0.0;
}

KP_DEV void lower4(const Metric4<double>& g, const double* v, double* o) {
    o[0] = g.g00 * v[0] + g.g01 * v[1] + g.g02 * v[2] + g.g03 * v[3];
    o[1] = g.g01 * v[0] + g.g11 * v[1] + g.g12 * v[2] + g.g13 * v[3];
    o[2] = g.g02 * v[0] + g.g12 * v[1] + g.g22 * v[2] + g.g23 * v[3];
    o[3] = g.g03 * v[0] + g.g13 * v[1] + g.g23 * v[2] + g.g33 * v[3];
}

// u^mu and b^mu from the native primitives; returns b^2 floored at 1e-40
// (grmhd_grid.hpp:980-1014, iharm_loader.cpp:453-487).
KP_DEV double grmhd_four_vectors(const Metric4<double>& g, const double* gcon0, const double* prim, double* ucon,
                                 double* bcon) {
    const double v[4] = {0.0, prim[2], prim[3], prim[4]};
    const double spatial = g.g11 * v[1] * v[1] + g.g22 * v[2] * v[2] + g.g33 * v[3] * v[3] +
                           2.0 * (g.g12 * v[1] * v[2] + g.g13 * v[1] * v[3] + g.g23 * v[2] * v[3]);
    const double vfac = sqrt(maxv(-1.0 / gcon0[0] * (1.0 + absv(spatial)), 0.0));
    ucon[0] = -vfac * gcon0[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) ucon[i] = v[i] - vfac * gcon0[i];
    double ucov[4];
    lower4(g, ucon, ucov);
    const double udotB = ucov[1] * prim[5] + ucov[2] * prim[6] + ucov[3] * prim[7];
    bcon[0] = udotB;
    const double iu0 = 1.0 / maxv(ucon[0], 1e-300);
#pragma unroll
    for (int i = 1; i < 4; ++i) bcon[i] = (prim[4 + i] + ucon[i] * udotB) * iu0;
    double bcov[4];
    lower4(g, bcon, bcov);
    const double bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
    return maxv(absv(bsq), 1e-40);
}

// KPOLARISIharmDerivedScalars for one cell (iharm_loader.cpp:420-505).
KP_DEV void grmhd_derive_cell(const GrmhdDev& G, int i, int j, kp_grmhd_cell& c) {
    const double mp = 1.67262171e-24, me = 9.1093826e-28;
    const double mp_me = mp / me;
    const double game = 4.0 / 3.0, gamp = 5.0 / 3.0;
    const double x1 = G.startx1 + (double(i) + 0.5) * G.dx1;
    const double x2 = G.startx2 + (double(j) + 0.5) * G.dx2;
    const double r = exp(x1);
    const double th = fmks_theta(G.map, x1, x2).th;
    Metric4<double> g;
    double gcon0[4];
    grmhd_native_metric(G, x1, x2, r, th, g, gcon0);
    double prim[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) prim[v] = double(c.prim[v]);
    double ucon[4], bcon[4];
    const double bsq = grmhd_four_vectors(g, gcon0, prim, ucon, bcon);
    const double rho = prim[0], uu = prim[1];
    const double beta = (G.gam - 1.0) * uu / (0.5 * bsq);
    const double sigma = bsq / maxv(rho, 1e-300);
    const double betasq = beta * beta / maxv(G.beta_crit * G.beta_crit, 1e-40);
    const double trat = (G.trat_large * betasq + G.trat_small) / (1.0 + betasq);
    const double thetae_unit = mp_me * (game - 1.0) * (gamp - 1.0) / ((gamp - 1.0) + (game - 1.0) * trat);
    c.derived[0] = rho * G.rho_unit / (mp + me) * G.ne_factor;
    c.derived[1] = thetae_unit * uu / maxv(rho, 1e-300);
    c.derived[2] = sqrt(bsq) * G.b_unit;
    c.derived[3] = sigma;
    c.derived[4] = beta;
}

// native_indices_from_native_coords (grmhd_grid.hpp:271-296).
KP_DEV bool grmhd_indices(const GrmhdDev& G, double x1, double x2, double phi, int& i0, int& j0, int& k0,
                          double& di, double& dj, double& dk) {
    if (!(x1 >= G.startx1 && x1 < G.startx1 + double(G.n1) * G.dx1 && x2 >= G.startx2 &&
          x2 < G.startx2 + double(G.n2) * G.dx2))
        return false;
    const double two_pi = 6.283185307179586476925286766559005768;
    phi = phi - floor(phi / two_pi) * two_pi;
    const double ii = (x1 - G.startx1) / G.dx1 - 0.5;
    const double jj = (x2 - G.startx2) / G.dx2 - 0.5;
    const double kk = (phi - G.startx3) / G.dx3 - 0.5;
    i0 = int(floor(ii));
    j0 = int(floor(jj));
    k0 = int(floor(kk));
    if (i0 < -1) i0 = -1;
    if (j0 < -1) j0 = -1;
    if (k0 < -1) k0 = -1;
    if (i0 >= G.n1) i0 = G.n1 - 1;
    if (j0 >= G.n2) j0 = G.n2 - 1;
    if (k0 >= G.n3) k0 = G.n3 - 1;
    di = clampv(ii - double(i0), 0.0, 1.0);
    dj = clampv(jj - double(j0), 0.0, 1.0);
    dk = clampv(kk - double(k0), 0.0, 1.0);
    return true;
}

// map_ghost_index (grmhd_grid.hpp:309-321) and the flat cell index.
KP_DEV size_t grmhd_cell(const GrmhdDev& G, int i, int j, int k) {
    if (i < 0) i = 0;
    if (i >= G.n1) i = G.n1 - 1;
    if (j < 0) {
        j = 0;
        k += G.n3 / 2;
    } else if (j >= G.n2) {
        j = G.n2 - 1;
        k += G.n3 / 2;
    }
    while (k < 0) k += G.n3;
    while (k >= G.n3) k -= G.n3;
    return (size_t(i) * size_t(G.n2) + size_t(j)) * size_t(G.n3) + size_t(k);
}

// sigma_smooth_factor (grmhd_grid.hpp:495-503).
KP_DEV double grmhd_sigma_factor(const GrmhdDev& G, double sigma) {
    double above = G.sigma_cut;
    if (G.sigma_cut_high > 0.0) above = G.sigma_cut_high;
    if (sigma < G.sigma_cut) return 1.0;
    if (sigma >= above) return 0.0;
    const double dsig = maxv(above - G.sigma_cut, 1e-300);
    return cos(PI_D / 2.0 / dsig * (sigma - G.sigma_cut));
}

// Trilinear samples of the 8 primitives and 5 derived scalars of one cell
// array (interp_prim_stencil / interp_derived_stencil, grmhd_grid.hpp:379-430;
// the reference's stencil order n = di-bit + 2 dj-bit + 4 dk-bit).
KP_DEV void grmhd_sample(const GrmhdDev& G, const kp_grmhd_cell* cells, int i0, int j0, int k0, double di,
                         double dj, double dk, double prim[8], double der[5]) {
    const double wi[2] = {1.0 - di, di}, wj[2] = {1.0 - dj, dj}, wk[2] = {1.0 - dk, dk};
#pragma unroll
    for (int v = 0; v < 8; ++v) prim[v] = 0.0;
#pragma unroll
    for (int v = 0; v < 5; ++v) der[v] = 0.0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int bi = n & 1, bj = (n >> 1) & 1, bk = (n >> 2) & 1;
        const double w = wi[bi] * wj[bj] * wk[bk];
        const kp_grmhd_cell& c = cells[grmhd_cell(G, i0 + bi, j0 + bj, k0 + bk)];
#pragma unroll
        for (int v = 0; v < 8; ++v) prim[v] += w * double(c.prim[v]);
#pragma unroll
        for (int v = 0; v < 5; ++v) der[v] += w * c.derived[v];
    }
}

// d Theta_e / d R_high of one cell's derived scalars: the iharm loader sets
// Theta_e = (m_p/m_e)(game-1)(gamp-1) / D * u/rho with
// D = (gamp-1) + (game-1) trat and trat = (R_high b2 + R_low) / (1 + b2),
// b2 = beta^2 / beta_crit^2 (src/grmhd/iharm_loader.cpp:494-504), so
// dTheta_e/dR_high = -Theta_e (game-1) b2 / ((1 + b2) D).
KP_DEV double grmhd_dthetae_drhigh(const GrmhdDev& G, double thetae, double beta) {
    const double game = 4.0 / 3.0, gamp = 5.0 / 3.0;
    const double b2 = beta * beta / maxv(G.beta_crit * G.beta_crit, 1e-40);
    const double trat = (G.trat_large * b2 + G.trat_small) / (1.0 + b2);
    const double D = (gamp - 1.0) + (game - 1.0) * trat;
    return -thetae * (game - 1.0) * (b2 / (1.0 + b2)) / D;
}

// grmhd_sample plus the trilinear sample of dTheta_e/dR_high.
KP_DEV void grmhd_sample_dth(const GrmhdDev& G, const kp_grmhd_cell* cells, int i0, int j0, int k0, double di,
                             double dj, double dk, double prim[8], double der[5], double& dth) {
    const double wi[2] = {1.0 - di, di}, wj[2] = {1.0 - dj, dj}, wk[2] = {1.0 - dk, dk};
#pragma unroll
    for (int v = 0; v < 8; ++v) prim[v] = 0.0;
#pragma unroll
    for (int v = 0; v < 5; ++v) der[v] = 0.0;
    dth = 0.0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const int bi = n & 1, bj = (n >> 1) & 1, bk = (n >> 2) & 1;
        const double w = wi[bi] * wj[bj] * wk[bk];
        const kp_grmhd_cell& c = cells[grmhd_cell(G, i0 + bi, j0 + bj, k0 + bk)];
#pragma unroll
        for (int v = 0; v < 8; ++v) prim[v] += w * double(c.prim[v]);
#pragma unroll
        for (int v = 0; v < 5; ++v) der[v] += w * c.derived[v];
        dth += w * grmhd_dthetae_drhigh(G, c.derived[1], c.derived[4]);
    }
}

// interpolate_time_value (src/model/time_interpolation.hpp:27-32).
KP_DEV double time_value(double a, double b, double f) {
    if (f <= 0.0 || a == b) return a;
    if (f >= 1.0) return b;
    return (1.0 - f) * a + f * b;
}

// The rest of the frequency-independent part of
// GRMHDGridRadiationModel::coefficients from sampled primitives and derived
// scalars (grmhd_grid.hpp:953-1080 and 1091-1150, 1210-1240), in the
// RiafLocal form shared with the per-frequency evaluation.
KP_DEV void grmhd_local_from(const GrmhdDev& G, const St<double>& s, double r, double th, const double prim[8],
                             const double der[5], RiafLocal<double>& L) {
    L.valid = 0;
    L.aligned = 0;
    L.has_chi = 0;
    if (!(prim[0] > 0.0 && prim[1] > 0.0)) return;
    const double x1 = s.v[1], x2 = s.v[2];
    Metric4<double> g;
    double gcon0[4];
    grmhd_native_metric(G, x1, x2, r, th, g, gcon0);
    double ucon[4], bcon[4];
    const double bsq = grmhd_four_vectors(g, gcon0, prim, ucon, bcon);
    const double bnorm = maxv(sqrt(bsq), 1e-150);
    const double inv_bnorm = 1.0 / bnorm;
    const double ne = der[0] * grmhd_sigma_factor(G, der[3]);
    const double* k = s.v + 4;
    const double* e1 = s.v + 8;
    const double* e2 = s.v + 12;
    double ucov[4], bcov[4], kcov[4];
    lower4(g, ucon, ucov);
    lower4(g, bcon, bcov);
    lower4(g, k, kcov);
    const double uk = ucov[0] * k[0] + ucov[1] * k[1] + ucov[2] * k[2] + ucov[3] * k[3];
    const double nu_scale = maxv(absv(-uk), 1e-8);
    double kdotb = 0.0, ub = 0.0, ue1 = 0.0, ue2 = 0.0, ke1 = 0.0, ke2 = 0.0, e1b = 0.0, e2b = 0.0;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        kdotb += k[mu] * bcov[mu];
        ub += ucov[mu] * bcon[mu];
        ue1 += ucov[mu] * e1[mu];
        ue2 += ucov[mu] * e2[mu];
        ke1 += kcov[mu] * e1[mu];
        ke2 += kcov[mu] * e2[mu];
        e1b += bcov[mu] * e1[mu];
        e2b += bcov[mu] * e2[mu];
    }
    const double cos_theta = clampv(kdotb / maxv(nu_scale * bnorm, 1e-30), -1.0, 1.0);
    L.valid = 1;
    L.ne = ne;
    L.thetae = der[1];
    L.b = der[2];
    // an invalid sigma or beta leaves the NonthermalSynchrotronParams default 3.5
    // (nonthermal_synchrotron.hpp:34, :496-498), floored like any kappa
    L.kappa = G.variable_kappa ? variable_kappa(der[3], der[4], 3.5, G.kappa_min) : G.kappa;
    L.nu_scale = nu_scale;
    L.cos_theta = cos_theta;
    L.sin_theta = maxv(sqrt(maxv(0.0, 1.0 - cos_theta * cos_theta)), 1e-4);
    const double ubunit = ub * inv_bnorm;
    const double kbunit = kdotb * inv_bnorm;
    double b1 = e1b * inv_bnorm;
    double b2 = e2b * inv_bnorm;
    if (absv(uk) > 1e-300) {
        const double minus_inv_uk = -1.0 / uk;
        const double eKb = minus_inv_uk * kbunit - ubunit;
        b1 += ue1 * ubunit - (minus_inv_uk * ke1 - ue1) * eKb;
        b2 += ue2 * ubunit - (minus_inv_uk * ke2 - ue2) * eKb;
    }
    const double bp2 = b1 * b1 + b2 * b2;
    if (bp2 > 1e-30) {
        L.has_chi = 1;
        L.cos2chi = (b2 * b2 - b1 * b1) / bp2;
        L.sin2chi = -2.0 * b1 * b2 / bp2;
    }
}

// The frequency-independent part of GRMHDGridRadiationModel::coefficients at
// one FMKS sample state (grmhd_grid.hpp:883-1080 and 1091-1150, 1210-1240).
KP_DEV void grmhd_local(const GrmhdDev& G, const St<double>& s, RiafLocal<double>& L) {
    L.valid = 0;
    L.aligned = 0;
    L.has_chi = 0;
    const double x1 = s.v[1], x2 = s.v[2], phi = s.v[3];
    const double r = exp(x1);
    const double th = fmks_theta(G.map, x1, x2).th;
    if (!(r >= G.r_in && r < G.r_out)) return;
    int i0, j0, k0;
    double di, dj, dk;
    if (!grmhd_indices(G, x1, x2, phi, i0, j0, k0, di, dj, dk)) return;
    double prim[8], der[5];
    grmhd_sample(G, G.cells, i0, j0, k0, di, dj, dk, prim, der);
    grmhd_local_from(G, s, r, th, prim, der, L);
}

// Fluid time interpolation between two snapshots of one grid: both are
// sampled at the state, every primitive and derived scalar is interpolated
// in time (fraction f), then the four-vectors are rebuilt
// (GRMHDGridRadiationModel::fluid_state with time_upper, grmhd_grid.hpp:940-962;
// SlowLightTemporalModel::coefficients, slow_light_driver.hpp:166-173).
// The upper snapshot is not read at f <= 0, nor the lower at f >= 1.
KP_DEV void grmhd_local_time(const GrmhdDev& G, const kp_grmhd_cell* lo, const kp_grmhd_cell* hi, double f,
                             const St<double>& s, RiafLocal<double>& L) {
    L.valid = 0;
    L.aligned = 0;
    L.has_chi = 0;
    const double x1 = s.v[1], x2 = s.v[2], phi = s.v[3];
    const double r = exp(x1);
    const double th = fmks_theta(G.map, x1, x2).th;
    if (!(r >= G.r_in && r < G.r_out)) return;
    int i0, j0, k0;
    double di, dj, dk;
    if (!grmhd_indices(G, x1, x2, phi, i0, j0, k0, di, dj, dk)) return;
    double prim[8], der[5];
    if (f >= 1.0) {
        grmhd_sample(G, hi, i0, j0, k0, di, dj, dk, prim, der);
    } else {
        grmhd_sample(G, lo, i0, j0, k0, di, dj, dk, prim, der);
        if (f > 0.0) {
            double p1[8], d1[5];
            grmhd_sample(G, hi, i0, j0, k0, di, dj, dk, p1, d1);
#pragma unroll
            for (int v = 0; v < 8; ++v) prim[v] = time_value(prim[v], p1[v], f);
#pragma unroll
            for (int v = 0; v < 5; ++v) der[v] = time_value(der[v], d1[v], f);
        }
    }
    grmhd_local_from(G, s, r, th, prim, der, L);
}

// grmhd_local_time with the sampled dTheta_e/dR_high (time-interpolated
// like Theta_e itself: the lower value's derivative where the values agree
// or f <= 0, the upper one's at f >= 1).
KP_DEV void grmhd_local_time_dth(const GrmhdDev& G, const kp_grmhd_cell* lo, const kp_grmhd_cell* hi, double f,
                                 const St<double>& s, RiafLocal<double>& L, double& dth) {
    L.valid = 0;
    L.aligned = 0;
    L.has_chi = 0;
    dth = 0.0;
    const double x1 = s.v[1], x2 = s.v[2], phi = s.v[3];
    const double r = exp(x1);
    const double th = fmks_theta(G.map, x1, x2).th;
    if (!(r >= G.r_in && r < G.r_out)) return;
    int i0, j0, k0;
    double di, dj, dk;
    if (!grmhd_indices(G, x1, x2, phi, i0, j0, k0, di, dj, dk)) return;
    double prim[8], der[5];
    if (f >= 1.0) {
        grmhd_sample_dth(G, hi, i0, j0, k0, di, dj, dk, prim, der, dth);
    } else {
        grmhd_sample_dth(G, lo, i0, j0, k0, di, dj, dk, prim, der, dth);
        if (f > 0.0) {
            double p1[8], d1[5], dth1;
            grmhd_sample_dth(G, hi, i0, j0, k0, di, dj, dk, p1, d1, dth1);
            if (der[1] != d1[1]) dth = (1.0 - f) * dth + f * dth1;
#pragma unroll
            for (int v = 0; v < 8; ++v) prim[v] = time_value(prim[v], p1[v], f);
#pragma unroll
            for (int v = 0; v < 5; ++v) der[v] = time_value(der[v], d1[v], f);
        }
    }
    grmhd_local_from(G, s, r, th, prim, der, L);
}

// Coefficients at observing frequency freq from the local plasma state
// (grmhd_grid.hpp:1150-1240); N carries the nonthermal distribution
// (grmhd_grid.hpp:1176-1205).
KP_DEV void grmhd_coeffs_at(const GrmhdDev& G, const RiafLocal<double>& L, double freq, Coeffs<double>& c,
                            const NtDev* N = nullptr) {
    coeffs_zero(c);
    if (!L.valid) return;
    const double nu = maxv(freq * L.nu_scale, 1.0);
    Coeffs<double> mc;
    if (N && (G.fit == KP_EMISSION_KAPPA || G.fit == KP_EMISSION_POWER_LAW)) {
        if (G.fit == KP_EMISSION_KAPPA && G.variable_kappa) {
            KappaConst K;
            kappa_constants(L.kappa, K);
            nonthermal_coeffs(*N, K, nu, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, mc);
        } else {
            nonthermal_coeffs(*N, N->kc, nu, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, mc);
        }
    } else {
        thermal_coeffs(G.fit, nu, L.ne, L.thetae, L.b, L.sin_theta, L.cos_theta, G.max_pol_frac, 1e-4,
                       G.emission_scale, G.absorption_scale, G.faraday_scale, mc);
    }
    c.jI = mc.jI;
    c.jV = mc.jV;
    c.aI = mc.aI;
    c.aV = mc.aV;
    c.rV = G.faraday_rotation ? mc.rV : 0.0;
    if (L.has_chi) {
        c.jQ = mc.jQ * L.cos2chi;
        c.jU = mc.jQ * L.sin2chi;
        c.aQ = mc.aQ * L.cos2chi;
        c.aU = mc.aQ * L.sin2chi;
        c.rQ = mc.rQ * L.cos2chi;
        c.rU = mc.rQ * L.sin2chi;
    }
}

}  // namespace kpd
