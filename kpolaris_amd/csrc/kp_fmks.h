// kp_fmks.h -- FMKS (funky modified Kerr-Schild) geometry on the device.
//
// Coordinates x = (t, x1 = ln r, x2, phi) with theta(x1, x2) blending the MKS
// map and the polynomial jet map (src/geometry/kerr_fmks.hpp:32-117).  The
// transport RHS follows the reference's route (kerr_fmks.hpp:336-479): map k,
// e1, e2 to spherical Kerr-Schild components, contract with the spherical-KS
// connection (src/geometry/kerr_schild_spherical.hpp:208-356) and map back
// with the second derivatives of theta.  The spherical-KS gcon and metric
// derivatives are stored by their nonzero entries only; the dense sums of the
// reference add exact zeros, so the sparse sums are the same numbers.
#pragma once

#include "kp_device.h"

#include <cmath>

namespace kpd {

// FMKS map parameters (KerrFMKSMetric fields, kerr_fmks.hpp:11-19) and the
// parameter-only factors of theta(x1, x2) and its derivatives, formed once on
// the host so that the transport RHS multiplies instead of dividing by them.
struct FmksMap {
    double startx1, hslope, mks_smooth, poly_alpha, poly_xt, poly_norm;
    double inv_xt;      // 1 / poly_xt
    double inv_alpha1;  // 1 / (poly_alpha + 1)
    double dthj_c;      // 2 alpha norm / ((alpha + 1) xt)
    double d2thj_c;     // 4 norm alpha / (alpha + 1)
    // the powers z^(alpha - k), k = 0, 1, 2, of the jet map: exponent and
    // the factor a negative z contributes (-1 for an odd integer exponent)
    double pw[3], pw_neg[3];
};

// fmks_power's sign rule (kerr_fmks.hpp:37-47) for exponent p and z < 0.
inline double fmks_power_sign(double p) {
    const double n = std::floor(p + 0.5);
    const bool integer = std::fabs(p - n) < 1e-10 && n >= 0.0;
    return (integer && std::fmod(n, 2.0) != 0.0) ? -1.0 : 1.0;
}

inline FmksMap fmks_map(double startx1, double hslope, double mks_smooth, double poly_alpha, double poly_xt,
                        double poly_norm) {
    FmksMap f{startx1, hslope, mks_smooth, poly_alpha, poly_xt, poly_norm, 0.0, 0.0, 0.0, 0.0, {}, {}};
    f.inv_xt = 1.0 / poly_xt;
    f.inv_alpha1 = 1.0 / (poly_alpha + 1.0);
    f.dthj_c = (2.0 * poly_alpha * poly_norm) / ((poly_alpha + 1.0) * poly_xt);
    f.d2thj_c = 4.0 * poly_norm * poly_alpha / (poly_alpha + 1.0);
    for (int k = 0; k < 3; ++k) {
        f.pw[k] = poly_alpha - double(k);
        f.pw_neg[k] = fmks_power_sign(f.pw[k]);
    }
    return f;
}

// fmks_power (kerr_fmks.hpp:37-47) for the three exponents alpha, alpha - 1
// and alpha - 2 of the jet map: |z|^p as exp(p ln|z|) with |z| floored at
// 1e-300 (so p = 0 gives 1), times the sign factor of an odd integer power
// when z < 0.  The reference forms integer powers by repeated products and
// others with pow; this form agrees with both to a few ulps (|alpha ln|z|| <= 3
// on the jet map's range) and has no branch: a branch inside the transport
// RHS, inlined into the eleven RHS evaluations of an adaptive attempt, made the
// FMKS kernels spill 0.9-2 KB per lane, so everything here is arithmetic and
// selects between computed values.
template <class T> KP_DEV void fmks_powers(const FmksMap& f, T z, T& za, T& zam1, T& zam2) {
    const T lz = log(maxv(absv(z), T(1e-300)));
    const bool neg = z < T(0);
    za = exp(T(f.pw[0]) * lz) * (neg ? T(f.pw_neg[0]) : T(1));
    zam1 = exp(T(f.pw[1]) * lz) * (neg ? T(f.pw_neg[1]) : T(1));
    zam2 = exp(T(f.pw[2]) * lz) * (neg ? T(f.pw_neg[2]) : T(1));
}

// theta_from_x2 and its first derivatives (kerr_fmks.hpp:49-80).
template <class T> struct FmksTheta {
    T th, dth1, dth2;
};

template <class T> KP_DEV FmksTheta<T> fmks_theta(const FmksMap& f, T x1, T x2) {
    const T pi = T(PI_D);
    T s2, c2;
    ksincos(T(2) * pi * x2, &s2, &c2);
    const T thG = pi * x2 + ((T(1) - T(f.hslope)) / T(2)) * s2;
    const T y = T(2) * x2 - T(1);
    const T z = y * T(f.inv_xt);
    T za, zam1, zam2;
    fmks_powers(f, z, za, zam1, zam2);
    (void)zam2;
    const T thJ = T(f.poly_norm) * y * (T(1) + za * T(f.inv_alpha1)) + T(0.5) * pi;
    const T expfac = exp(T(f.mks_smooth) * (T(f.startx1) - x1));
    const T dthG = pi + (T(1) - T(f.hslope)) * pi * c2;
    const T dthJ = T(2) * T(f.poly_norm) * (T(1) + za * T(f.inv_alpha1)) + T(f.dthj_c) * y * zam1;
    FmksTheta<T> o;
    o.th = thG + expfac * (thJ - thG);
    o.dth1 = -T(f.mks_smooth) * expfac * (thJ - thG);
    o.dth2 = dthG + expfac * (dthJ - dthG);
    return o;
}

// Spherical-KS connection work: gcon and d gcov / dr, d gcov / dtheta,
// nonzero entries (kerr_schild_spherical.hpp:208-268).
template <class T> struct SksConn {
    T g00, g01, g11, g13, g22, g33;             // gcon
    T r00, r01, r03, r11, r13, r22, r33;        // d gcov / dr
    T t00, t01, t03, t11, t13, t22, t33;        // d gcov / dtheta
};

template <class T> KP_DEV SksConn<T> sks_conn(T mass, T spin, T r_in, T s, T cc) {
    SksConn<T> c;
    const T r = maxv(r_in, tiny<T>());
    const T a = spin, a2 = a * a;
    const T s2 = s * s;
    const T safe_s2 = maxv(s2, tiny<T>());
    const T rho2 = r * r + a2 * cc * cc;
    // the reference's seven divisions by rho^2, rho^4 and rho^2 sin^2 as two
    // reciprocals (each quotient moves by a few ulps at most)
    const T irho2 = T(1) / rho2;
    const T irho4 = irho2 * irho2;
    const T f = T(2) * mass * r * irho2;
    const T delta = r * r - T(2) * mass * r + a2;
    c.g00 = -(T(1) + f);
    c.g01 = f;
    c.g11 = delta * irho2;
    c.g13 = a * irho2;
    c.g22 = irho2;
    c.g33 = irho2 / safe_s2;
    const T drho_dr = T(2) * r;
    const T drho_dth = -T(2) * a2 * s * cc;
    const T ds2_dth = T(2) * s * cc;
    const T df_dr = T(2) * mass * (rho2 - r * drho_dr) * irho4;
    const T df_dth = -T(2) * mass * r * drho_dth * irho4;
    c.r00 = df_dr;
    c.t00 = df_dth;
    c.r01 = df_dr;
    c.t01 = df_dth;
    c.r03 = -a * s2 * df_dr;
    c.t03 = -a * (s2 * df_dth + ds2_dth * f);
    c.r11 = df_dr;
    c.t11 = df_dth;
    c.r13 = -a * s2 * df_dr;
    c.t13 = -a * (ds2_dth * (T(1) + f) + s2 * df_dth);
    c.r22 = drho_dr;
    c.t22 = drho_dth;
    const T bracket = rho2 + a2 * s2 * (T(1) + f);
    const T dbr_dr = drho_dr + a2 * s2 * df_dr;
    const T dbr_dth = drho_dth + a2 * (ds2_dth * (T(1) + f) + s2 * df_dth);
    c.r33 = s2 * dbr_dr;
    c.t33 = ds2_dth * bracket + s2 * dbr_dth;
    return c;
}

// Row products sum_b D[nu][b] v[b] of a symmetric derivative matrix with the
// (0,0) (0,1) (0,3) (1,1) (1,3) (2,2) (3,3) pattern.
template <class T> KP_DEV void sks_rows(T d00, T d01, T d03, T d11, T d13, T d22, T d33, const T* v, T* o) {
    o[0] = d00 * v[0] + d01 * v[1] + d03 * v[3];
    o[1] = d01 * v[0] + d11 * v[1] + d13 * v[3];
    o[2] = d22 * v[2];
    o[3] = d03 * v[0] + d13 * v[1] + d33 * v[3];
}

// spherical_ks_connection_contract_fixed_p (kerr_schild_spherical.hpp:318-356):
// Gamma^mu_ab p^a q^b given the rows of p.
template <class T>
KP_DEV void sks_contract(const SksConn<T>& c, const T* rp_r, const T* rp_t, const T* p, const T* q, T* out) {
    T rq_r[4], rq_t[4];
    sks_rows(c.r00, c.r01, c.r03, c.r11, c.r13, c.r22, c.r33, q, rq_r);
    sks_rows(c.t00, c.t01, c.t03, c.t11, c.t13, c.t22, c.t33, q, rq_t);
    const T pD1q = p[0] * rq_r[0] + p[1] * rq_r[1] + p[2] * rq_r[2] + p[3] * rq_r[3];
    const T pD2q = p[0] * rq_t[0] + p[1] * rq_t[1] + p[2] * rq_t[2] + p[3] * rq_t[3];
    T t[4];
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) t[nu] = p[1] * rq_r[nu] + q[1] * rp_r[nu] + p[2] * rq_t[nu] + q[2] * rp_t[nu];
    t[1] -= pD1q;
    t[2] -= pD2q;
    out[0] = T(0.5) * (c.g00 * t[0] + c.g01 * t[1]);
    out[1] = T(0.5) * (c.g01 * t[0] + c.g11 * t[1] + c.g13 * t[3]);
    out[2] = T(0.5) * (c.g22 * t[2]);
    out[3] = T(0.5) * (c.g13 * t[1] + c.g33 * t[3]);
}

// FMKSConnectionWork (kerr_fmks.hpp:336-402).
template <class T> struct FmksConn {
    T r, th, dth1, dth2, d2th11, d2th12, d2th22;
    SksConn<T> sph;
};

template <class T> KP_DEV FmksConn<T> fmks_conn(T mass, T spin, const FmksMap& f, T x1, T x2) {
    const T pi = T(PI_D);
    const T y = T(2) * x2 - T(1);
    const T z = y * T(f.inv_xt);
    const T expfac = exp(T(f.mks_smooth) * (T(f.startx1) - x1));
    T za, zam1, zam2;
    fmks_powers(f, z, za, zam1, zam2);
    T s2, c2;
    ksincos(T(2) * pi * x2, &s2, &c2);
    const T thG = pi * x2 + ((T(1) - T(f.hslope)) / T(2)) * s2;
    const T thJ = T(f.poly_norm) * y * (T(1) + za * T(f.inv_alpha1)) + T(0.5) * pi;
    const T dthG = pi + (T(1) - T(f.hslope)) * pi * c2;
    const T dthJ = T(2) * T(f.poly_norm) * (T(1) + za * T(f.inv_alpha1)) + T(f.dthj_c) * y * zam1;
    const T d2thG = -T(2) * (T(1) - T(f.hslope)) * pi * pi * s2;
    const T ixt = T(f.inv_xt);
    const T d2thJ = T(f.d2thj_c) * (T(2) * zam1 * ixt + y * (T(f.poly_alpha) - T(1)) * zam2 * (ixt * ixt));
    FmksConn<T> c;
    c.r = exp(x1);
    c.th = thG + expfac * (thJ - thG);
    c.dth1 = -T(f.mks_smooth) * expfac * (thJ - thG);
    c.dth2 = dthG + expfac * (dthJ - dthG);
    c.d2th11 = -T(f.mks_smooth) * c.dth1;
    c.d2th12 = -T(f.mks_smooth) * expfac * (dthJ - dthG);
    c.d2th22 = d2thG + expfac * (d2thJ - d2thG);
    T sn, cs;
    ksincos(c.th, &sn, &cs);
    c.sph = sks_conn(mass, spin, c.r, sn, cs);
    return c;
}

// fmks_to_spherical_vector / fmks_from_spherical_connection_contract
// (kerr_fmks.hpp:404-432).
template <class T> KP_DEV void fmks_to_sph(const FmksConn<T>& c, const T* v, T* o) {
    o[0] = v[0];
    o[1] = c.r * v[1];
    o[2] = c.dth1 * v[1] + c.dth2 * v[2];
    o[3] = v[3];
}

template <class T> KP_DEV void fmks_from_sph(const FmksConn<T>& c, T* hs, const T* p, const T* q, T* o) {
    hs[1] += c.r * p[1] * q[1];
    hs[2] += c.d2th11 * p[1] * q[1] + c.d2th12 * (p[1] * q[2] + p[2] * q[1]) + c.d2th22 * p[2] * q[2];
    o[0] = hs[0];
    // divisions, not reciprocals kept in FmksConn: two more live values there
    // cost the pass-B kernels more in spills than the divisions do.  The theta
    // row reuses the quotient hs[1] / r (the reference forms (dth1 hs[1]) / r:
    // the two differ by an ulp), one FP64 division less per contraction.
    o[1] = hs[1] / c.r;
    o[2] = (hs[2] - c.dth1 * o[1]) / c.dth2;
    o[3] = hs[3];
}

// transport_rhs for KerrFMKSMetric (kerr_fmks.hpp:455-479), delivered
// component-wise like bl_rhs_apply: every input component is read before the
// sink sees any output.
template <class T, class Y, class Sink>
KP_DEV void fmks_rhs_apply(T mass, T spin, const FmksMap& f, const Y& y, Sink&& sink) {
    const FmksConn<T> c = fmks_conn(mass, spin, f, T(y[1]), T(y[2]));
    const T k[4] = {y[4], y[5], y[6], y[7]};
    const T e1[4] = {y[8], y[9], y[10], y[11]};
    const T e2[4] = {y[12], y[13], y[14], y[15]};
    T ks[4], rk_r[4], rk_t[4];
    fmks_to_sph(c, k, ks);
    sks_rows(c.sph.r00, c.sph.r01, c.sph.r03, c.sph.r11, c.sph.r13, c.sph.r22, c.sph.r33, ks, rk_r);
    sks_rows(c.sph.t00, c.sph.t01, c.sph.t03, c.sph.t11, c.sph.t13, c.sph.t22, c.sph.t33, ks, rk_t);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(mu, k[mu]);
    T hs[4], out[4], vs[4];
    sks_contract(c.sph, rk_r, rk_t, ks, ks, hs);
    fmks_from_sph(c, hs, k, k, out);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(4 + mu, -out[mu]);
    fmks_to_sph(c, e1, vs);
    sks_contract(c.sph, rk_r, rk_t, ks, vs, hs);
    fmks_from_sph(c, hs, k, e1, out);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(8 + mu, -out[mu]);
    fmks_to_sph(c, e2, vs);
    sks_contract(c.sph, rk_r, rk_t, ks, vs, hs);
    fmks_from_sph(c, hs, k, e2, out);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(12 + mu, -out[mu]);
}

// Integration context for the FMKS metric.
template <class T> struct GeoFmks {
    using real = T;
    T m, a;
    uint32_t* rhs_count;
    FmksMap f;
    template <class Y, class Sink> KP_DEV void apply(const Y& y, Sink&& sink) const {
        fmks_rhs_apply(m, a, f, y, sink);
    }
};

// FMKS covariant metric at (x1, x2): the Jacobian congruence of the KS metric
// written out (the explicit form of grmhd_grid.hpp:721-760, equal to
// KerrFMKSMetric::gcov_matrix, kerr_fmks.hpp:141-157, up to rounding).
template <class T> KP_DEV Metric4<T> fmks_gcov(T mass, T spin, const FmksMap& f, T x1, T x2) {
    const FmksTheta<T> t = fmks_theta(f, x1, x2);
    T sth, cth;
    ksincos(t.th, &sth, &cth);
    const T r = exp(x1);
    const T s2 = sth * sth;
    const T a2 = spin * spin;
    const T rho2 = r * r + a2 * cth * cth;
    const T fr = T(2) * mass * r / rho2;
    Metric4<T> g;
    g.g00 = T(-1) + fr;
    g.g01 = fr * r;
    g.g02 = T(0);
    g.g03 = -spin * fr * s2;
    g.g11 = (T(1) + fr) * r * r + rho2 * t.dth1 * t.dth1;
    g.g12 = rho2 * t.dth1 * t.dth2;
    g.g13 = -spin * s2 * (T(1) + fr) * r;
    g.g22 = rho2 * t.dth2 * t.dth2;
    g.g23 = T(0);
    g.g33 = s2 * (rho2 + a2 * s2 * (T(1) + fr));
    return g;
}

}  // namespace kpd
