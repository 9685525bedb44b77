// kp_nonthermal.h -- kappa and power-law synchrotron coefficients (FP64).
//
// Restates nonthermal_synchrotron_magnetic_basis_coefficients
// (src/radiation/nonthermal_synchrotron.hpp:504-618): the Pandya et al. (2016)
// emission and absorption fits for the kappa and power-law distributions,
// the Marszewski et al. (2021) fast rotativity fits for kappa (anchored at
// kappa = 3.5, 4, 4.5, 5 and interpolated linearly, zero outside), the blend
// toward the thermal Pandya fit between kappa_interp_begin and
// kappa_interp_end, and the polarised-fraction caps.  The header's own CGS
// constants (:106-110, :125-127) are kept: they differ from the thermal fits'
// in the last digits.
//
// Everything that depends only on kappa or on the power-law slope (gamma
// functions, the 2F1 transformation ratios, fit exponents) is gathered in
// KappaConst / PowerLawConst.  The host fills them once per image when kappa
// is fixed; with a variable kappa (GRMHD, kappa from sigma and beta) the
// device fills them per sample with the same function.
#pragma once

#include "kp_device.h"
#include "../../include/kp_image.h"

namespace kpd {

#define KP_NT_EE 4.80320680e-10
#define KP_NT_ME 9.1093826e-28

struct KappaConst {
    double k;
    double j_low, j_high;    // I emission normalisations (:226-231)
    double a_low, a_high;    // I absorption normalisations without the (w kappa) powers and 2F1 (:327-336)
    double g1, g2;           // Gamma ratios of the 1/(1-z) transformation of 2F1(k-1/3, k+1; k+2/3; z)
    double x_j, x_jq, x_ai, x_aq, x_av;   // blend exponents of the fits
    double q_high;           // (4/5)^2 + k/50 (Q emission, :258)
    double vj_low, vj_high;  // k^(-66/125), k^(-11/25) (V emission, :286-292)
    double ai_high;          // (3/k)^(19/4) + 3/5 (I absorption, :339)
    double aq_high;          // 21^2 k^(-144/25) + 11/20 (Q absorption, :362)
    double av_low, av_nfac;  // k^(-7/10) and the V absorption polynomial in k (:387-396)
};

struct PowerLawConst {
    int valid;               // p > 2 and gamma_min >= 1 (power_law_density_from_eta, :112-123)
    double dens;             // n_pl = dens * B^2
    double jc, ac;           // emission / absorption constants (:134-141, :178-185)
    double qj, vj, qa, va;   // Q, V emission and absorption factors (:150-152, :160, :193-195, :203-204)
};

struct NtDev {
    int dist;                // 2 kappa, 3 power law
    int variable_kappa;      // GRMHD: kappa from sigma and beta per sample
    double kappa_min;        // floor of the variable kappa (grmhd_grid.hpp:1187-1189)
    double kappa_width;      // > 0: fixed width; otherwise (k - 3) / k * Theta_e
    double interp_begin, interp_end;
    double max_pol_em, max_pol_abs, min_sin;
    double es, as, fs;
    double p;
    KappaConst kc;           // fixed kappa
    PowerLawConst pl;
};

// The two large routines stay out of line: one copy serves every kernel that
// can meet a nonthermal model, the thermal kernels' register allocation and
// build time are not touched, and a call is paid only on the nonthermal path.
#define KP_DEV_OUTLINE __host__ __device__ inline __attribute__((noinline))

KP_DEV double nt_nonzero(double x) {
    if (absv(x) < 1e-300) return x >= 0.0 ? 1e-300 : -1e-300;
    return x;
}

KP_DEV_OUTLINE void kappa_constants(double k, KappaConst& c) {
    const double pi = PI_D;
    c.k = k;
    c.j_low = 4.0 * pi * tgamma(k - 4.0 / 3.0) / (pow(3.0, 7.0 / 3.0) * tgamma(k - 2.0));
    c.j_high = 0.25 * pow(3.0, (k - 1.0) / 2.0) * (k - 2.0) * (k - 1.0) * tgamma(k / 4.0 - 1.0 / 3.0) *
               tgamma(k / 4.0 + 4.0 / 3.0);
    c.a_low = pow(3.0, 1.0 / 6.0) * 10.0 / 41.0 * (4.0 * pi * pi) * (k - 2.0) * (k - 1.0) * k / (3.0 * k - 1.0) *
              tgamma(5.0 / 3.0);
    c.a_high = 2.0 * pow(pi, 2.5) / 3.0 * (k - 2.0) * (k - 1.0) * k * (2.0 * tgamma(2.0 + k / 2.0) / (2.0 + k) - 1.0);
    const double a = k - 1.0 / 3.0, b = k + 1.0, cc = k + 2.0 / 3.0;
    const double gc = tgamma(cc);
    c.g1 = gc * tgamma(b - a) / nt_nonzero(tgamma(b) * tgamma(cc - a));
    c.g2 = gc * tgamma(a - b) / nt_nonzero(tgamma(a) * tgamma(cc - b));
    c.x_j = 3.0 * pow(k, -1.5);
    c.x_jq = 3.7 * pow(k, -1.6);
    c.x_ai = pow(-7.0 / 4.0 + 8.0 * k / 5.0, -43.0 / 50.0);
    c.x_aq = 7.0 / 5.0 * pow(k, -23.0 / 20.0);
    c.x_av = 61.0 / 50.0 * pow(k, -142.0 / 125.0) + 7.0 / 1000.0;
    c.q_high = pow(4.0 / 5.0, 2.0) + k / 50.0;
    c.vj_low = pow(k, -66.0 / 125.0);
    c.vj_high = pow(k, -11.0 / 25.0);
    c.ai_high = pow(3.0 / k, 19.0 / 4.0) + 3.0 / 5.0;
    c.aq_high = 21.0 * 21.0 * pow(k, -144.0 / 25.0) + 11.0 / 20.0;
    c.av_low = pow(k, -7.0 / 10.0);
    c.av_nfac = 13.0 * 13.0 * pow(k, -8.0) + 13.0 / 2500.0 * k - 263.0 / 5000.0 + 47.0 / (200.0 * k);
}

KP_DEV void power_law_constants(double p, double eta, double gmin, double gmax, PowerLawConst& c) {
    c.valid = (p > 2.0) && (gmin >= 1.0);
    c.dens = eta / 2.0 * (p - 2.0) / (p - 1.0) / maxv(KP_NT_ME * KP_CL * KP_CL * gmin, 1e-300);
    const double span = pow(gmin, 1.0 - p) - pow(gmax, 1.0 - p);
    c.jc = pow(3.0, p / 2.0) * (p - 1.0) / maxv(2.0 * (p + 1.0) * span, 1e-300) * tgamma((3.0 * p - 1.0) / 12.0) *
           tgamma((3.0 * p + 19.0) / 12.0);
    c.ac = pow(3.0, (p + 1.0) / 2.0) * (p - 1.0) / maxv(4.0 * span, 1e-300) * tgamma((3.0 * p + 2.0) / 12.0) *
           tgamma((3.0 * p + 22.0) / 12.0);
    c.qj = (p + 1.0) / (p + 7.0 / 3.0);
    c.vj = (171.0 / 250.0) * pow(p, 49.0 / 100.0);
    c.qa = -pow((17.0 / 500.0) * p - 43.0 / 1250.0, 43.0 / 500.0);
    c.va = -pow((71.0 / 100.0) * p + 22.0 / 625.0, 197.0 / 500.0);
}

// hyp2f1_series (nonthermal_synchrotron.hpp:55-73): at most 200 terms,
// stopped once a term falls below 1e-13 of the sum.
KP_DEV double hyp2f1_series(double a, double b, double c, double z) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n <= 200; ++n) {
        const double m = double(n) - 1.0;
        term *= (a + m) * (b + m) * z / nt_nonzero((c + m) * double(n));
        sum += term;
        if (absv(term) < absv(sum) * 1e-13) break;
    }
    return sum;
}

// stable_hyp2f1 (:83-96) for 2F1(k - 1/3, k + 1; k + 2/3; z): the series
// inside (-1, 1), else the 1/(1-z) transformation with the Gamma ratios of K.
KP_DEV double kappa_hyp2f1(const KappaConst& K, double z) {
    const double a = K.k - 1.0 / 3.0, b = K.k + 1.0, c = K.k + 2.0 / 3.0;
    if (z > -1.0 && z < 1.0) return hyp2f1_series(a, b, c, z);
    const double w = 1.0 - z;
    const double zp = 1.0 / w;
    return pow(w, -a) * K.g1 * hyp2f1_series(a, c - b, a - b + 1.0, zp) +
           pow(w, -b) * K.g2 * hyp2f1_series(b, c - a, b - a + 1.0, zp);
}

// pref * n_low * X^low_power * (1 + X^(x * high_power) * (n_low / n_high)^x)^(-1/x)
KP_DEV double nt_blend(double pref, double n_low, double n_high, double X, double low_power, double high_power,
                       double x) {
    return pref * n_low * pow(X, low_power) * pow(1.0 + pow(X, x * high_power) * pow(n_low / n_high, x), -1.0 / x);
}

// kappa_rho_Q_at / kappa_rho_V_at (:393-460) at anchor a = 0..3
// (kappa 3.5, 4, 4.5, 5); kv = K0(1/w) / K2(1/w).
KP_DEV double kappa_rho_q_at(int a, double pref_q, double w, double X) {
    const double sw = sqrt(w), e5 = exp(-5.0 * w), x84 = pow(X, 0.84);
    double wterm, fx;
    if (a == 0) {
        wterm = 17.0 * w - 3.0 * sw + 7.0 * sw * e5;
        fx = 1.0 - exp(-x84 / 30.0) - sin(X / 10.0) * exp(-1.5 * pow(X, 0.471));
    } else if (a == 1) {
        wterm = 46.0 / 3.0 * w - 5.0 / 3.0 * sw + 17.0 / 3.0 * sw * e5;
        fx = 1.0 - exp(-x84 / 18.0) - sin(X / 6.0) * exp(-7.0 * sqrt(X) / 4.0);
    } else if (a == 2) {
        wterm = 14.0 * w - 13.0 / 8.0 * sw + 9.0 / 2.0 * sw * e5;
        fx = 1.0 - exp(-x84 / 12.0) - sin(X / 4.0) * exp(-2.0 * pow(X, 0.525));
    } else {
        wterm = 25.0 / 2.0 * w - sw + 5.0 * sw * e5;
        fx = 1.0 - exp(-x84 / 8.0) - sin(3.0 * X / 8.0) * exp(-9.0 * pow(X, 0.541) / 4.0);
    }
    return pref_q * fx * wterm;
}

KP_DEV double kappa_rho_v_at(int a, double pref_v, double w, double X, double kv) {
    double wterm, c;
    if (a == 0) {
        wterm = (w * w + 2.0 * w + 1.0) / (25.0 / 8.0 * w * w + 4.0 * w + 1.0);
        c = 0.447;
    } else if (a == 1) {
        wterm = (w * w + 54.0 * w + 50.0) / (30.0 / 11.0 * w * w + 134.0 * w + 50.0);
        c = 0.391;
    } else if (a == 2) {
        wterm = (w * w + 43.0 * w + 38.0) / (7.0 / 3.0 * w * w + 185.0 / 2.0 * w + 38.0);
        c = 0.348;
    } else {
        wterm = (w + 13.0 / 14.0) / (2.0 * w + 13.0 / 14.0);
        c = 0.313;
    }
    const double gx = 1.0 - 0.17 * log(1.0 + c * pow(X, -0.5));
    return pref_v * kv * gx * wterm;
}

// polarised-fraction cap of an (I, Q, V) triple (:540-545, :557-562, :595-606)
KP_DEV void nt_cap(double i, double& q, double& v, double frac) {
    const double pol = sqrt(q * q + v * v);
    if (i > 0.0 && pol > 0.0 && i < pol / maxv(frac, 1e-30)) {
        const double s = i / pol * frac;
        q *= s;
        v *= s;
    }
}

// variable_kappa_from_sigma_beta (:493-502) with the model's floor
// (grmhd_grid.hpp:1185-1189).
KP_DEV double variable_kappa(double sigma, double beta, double kappa_default, double kappa_min) {
    double k = kappa_default;
    if (sigma > 0.0 && beta >= 0.0) {
        k = 2.8 + 0.7 / sqrt(sigma) + 3.7 * pow(sigma, -0.19) * tanh(23.4 * pow(sigma, 0.26) * beta);
        if (k < 3.1) k = 3.1;
    }
    return k < kappa_min ? kappa_min : k;
}

// nonthermal_synchrotron_magnetic_basis_coefficients (:504-618) in the
// magnetic screen basis: j / nu^2, alpha nu, rho nu.  sin_theta is the
// caller's (already floored) value; K holds the constants of this kappa.
KP_DEV_OUTLINE void nonthermal_coeffs(const NtDev& N, const KappaConst& K, double nu_in, double ne, double thetae, double b,
                              double sin_theta, double cos_theta, Coeffs<double>& c) {
    coeffs_zero(c);
    if (!(nu_in > 0.0) || !(b > 0.0) || !(ne > 0.0)) return;
    const double nu = maxv(nu_in, 1.0);
    const double inu2 = 1.0 / maxv(nu * nu, 1e-300);
    const double nuc = KP_NT_EE * b / (2.0 * PI_D * KP_NT_ME * KP_CL);
    const double sint = maxv(sin_theta, N.min_sin);
    const double sgn = cos_theta >= 0.0 ? 1.0 : -1.0;
    if (N.dist == KP_EMISSION_POWER_LAW) {
        const PowerLawConst& P = N.pl;
        const double n = P.valid ? P.dens * b * b : 0.0;
        if (!(nuc > 0.0) || !(n > 0.0) || !(N.p > 1.0)) return;  // every term is zero
        const double ratio = nu_in / maxv(nuc * sint, 1e-300);
        const double jI = n * KP_NT_EE * KP_NT_EE * nuc / KP_CL * P.jc * sint * pow(ratio, -(N.p - 1.0) / 2.0);
        double cost = cos_theta;
        if (absv(cost) < 1e-12) cost = cost >= 0.0 ? 1e-12 : -1e-12;
        const double jV = P.vj * (cost / sint * pow(nu_in / maxv(3.0 * nuc * sint, 1e-300), -0.5)) * jI;
        double ji = jI * inu2, jq = P.qj * jI * inu2, jv = jV * inu2;
        nt_cap(ji, jq, jv, N.max_pol_em);
        const double aI = n * KP_NT_EE * KP_NT_EE / maxv(nu_in * KP_NT_ME * KP_CL, 1e-300) * P.ac *
                          pow(ratio, -(N.p + 2.0) / 2.0);
        const double t6 = pow(maxv((31.0 / 10.0) * pow(sint, -48.0 / 25.0) - 31.0 / 10.0, 0.0), 64.0 / 125.0);
        const double aV = -aI * P.va * t6 * pow(ratio, -0.5) * sgn;
        double ai = aI * nu, aq = -(aI * P.qa) * nu, av = aV * nu;
        nt_cap(ai, aq, av, N.max_pol_abs);
        c.jI = N.es * ji;
        c.jQ = N.es * jq;
        c.jV = N.es * jv;
        c.aI = N.as * ai;
        c.aQ = N.as * aq;
        c.aV = N.as * av;
        return;  // the fast fits carry no power-law rotativity
    }
    if (N.dist != KP_EMISSION_KAPPA) return;
    const double k = K.k;
    const double w = N.kappa_width > 0.0 ? N.kappa_width : (k - 3.0) / maxv(k, 1e-300) * thetae;
    const double X = nu_in / maxv(w * w * k * k * nuc * sint, 1e-300);
    const bool fit = k > 2.0 && w > 0.0 && X > 0.0;
    double ji = 0.0, jq = 0.0, jv = 0.0, ai = 0.0, aq = 0.0, av = 0.0, rq = 0.0, rv = 0.0;
    if (fit) {
        // emission (:217-308)
        const double pref = ne * KP_NT_EE * KP_NT_EE * nuc * sint / KP_CL;
        const double hp = (3.0 * k - 4.0) / 6.0;
        ji = nt_blend(pref, K.j_low, K.j_high, X, 1.0 / 3.0, hp, K.x_j) * inu2;
        jq = -nt_blend(pref, -0.5 * K.j_low, -K.q_high * K.j_high, X, 1.0 / 3.0, hp, K.x_jq) * inu2;
        const double ang1 = maxv(pow(sint, -12.0 / 5.0) - 1.0, 0.0);
        const double ang2 = maxv(pow(sint, -5.0 / 2.0) - 1.0, 0.0);
        const double vl = -(9.0 / 16.0) * pow(ang1, 12.0 / 25.0) * K.vj_low / w * pow(X, -7.0 / 20.0) * K.j_low;
        const double vh = -(49.0 / 64.0) * pow(ang2, 11.0 / 25.0) * K.vj_high / w * pow(X, -0.5) * K.j_high;
        jv = -nt_blend(pref, vl, vh, X, 1.0 / 3.0, hp, K.x_j) * sgn * inu2;
        // absorption (:310-391)
        const double hyp = kappa_hyp2f1(K, -k * w);
        if (absv(hyp) > 1e-300) {
            const double pa = ne * KP_NT_EE / maxv(b * sint, 1e-300);
            const double wk = w * k;
            const double lo = K.a_low / pow(wk, 16.0 / 3.0 - k) * hyp;
            const double hi = K.a_high / pow(wk, 5.0);
            const double hq = (3.0 * k - 1.0) / 6.0;
            ai = nt_blend(pa, lo, hi * K.ai_high + 1e-300, X, -5.0 / 3.0, hq, K.x_ai) * nu;
            aq = -nt_blend(pa, -25.0 / 48.0 * lo, -K.aq_high * hi + 1e-300, X, -5.0 / 3.0, hq, K.x_aq) * nu;
            const double b1 = maxv(pow(sint, -114.0 / 50.0) - 1.0, 0.0);
            const double b2 = maxv(pow(sint, -41.0 / 20.0) - 1.0, 0.0);
            const double al = -77.0 / (100.0 * w) * pow(b1, 223.0 / 500.0) * pow(X, -7.0 / 20.0) * K.av_low * lo;
            const double ah = -143.0 / 10.0 * pow(w, -116.0 / 125.0) * pow(b2, 0.5) * K.av_nfac * pow(X, -0.5) * hi +
                              1e-300;
            av = -nt_blend(pa, al, ah, X, -5.0 / 3.0, hq, K.x_av) * sgn * nu;
        }
    }
    // rotativity (:393-489): defined for 3.5 <= kappa <= 5 only
    if (X > 0.0 && k >= 3.5 && k <= 5.0) {
        const int a0 = k >= 5.0 ? 3 : (k < 4.0 ? 0 : (k < 4.5 ? 1 : 2));
        const double kl = 3.5 + 0.5 * a0;
        const double pq = -(ne * KP_NT_EE * KP_NT_EE * nuc * nuc * sint * sint) /
                          maxv(KP_NT_ME * KP_CL * nu_in * nu_in * nu_in, 1e-300);
        const double q0 = kappa_rho_q_at(a0, pq, w, X);
        rq = a0 == 3 ? q0 : ((kl + 0.5 - k) * q0 + (k - kl) * kappa_rho_q_at(a0 + 1, pq, w, X)) / 0.5;
        if (w > 0.0) {
            const double pv = 2.0 * ne * KP_NT_EE * KP_NT_EE * nuc * cos_theta /
                              maxv(KP_NT_ME * KP_CL * nu_in * nu_in, 1e-300);
            double k0, k1, k2;
            bessel_k012(1.0 / w, k0, k1, k2);
            const double kv = k0 / maxv(k2 + 1e-300, 1e-300);
            const double v0 = kappa_rho_v_at(a0, pv, w, X, kv);
            rv = a0 == 3 ? v0 : ((kl + 0.5 - k) * v0 + (k - kl) * kappa_rho_v_at(a0 + 1, pv, w, X, kv)) / 0.5;
        }
        rq *= nu;
        rv *= nu;
    }
    // blend toward the thermal Pandya fit (:522-540, :585-592)
    double mix = 0.0;
    if (k >= N.interp_begin && k < N.interp_end)
        mix = (k - N.interp_begin) / maxv(N.interp_end - N.interp_begin, 1e-300);
    else if (k >= N.interp_end)
        mix = 1.0;
    if (mix != 0.0) {
        Coeffs<double> th;
        thermal_coeffs(1, nu_in, ne, thetae, b, sin_theta, cos_theta, N.max_pol_em, N.min_sin, 1.0, 1.0, 1.0, th);
        ji = (1.0 - mix) * ji + mix * th.jI;
        jq = (1.0 - mix) * jq + mix * th.jQ;
        jv = (1.0 - mix) * jv + mix * th.jV;
        ai = (1.0 - mix) * ai + mix * th.aI;
        aq = (1.0 - mix) * aq + mix * th.aQ;
        av = (1.0 - mix) * av + mix * th.aV;
    }
    nt_cap(ji, jq, jv, N.max_pol_em);
    nt_cap(ai, aq, av, N.max_pol_abs);
    c.jI = N.es * ji;
    c.jQ = N.es * jq;
    c.jV = N.es * jv;
    c.aI = N.as * ai;
    c.aQ = N.as * aq;
    c.aV = N.as * av;
    c.rQ = N.fs * rq;
    c.rV = N.fs * rv;
}

}  // namespace kpd
