/*
 * kp_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU oracle of the imaging path.
 *
 * A plain-C11 restatement, in IEEE double with the C library's math, of the
 * reference algorithm for the Boyer-Lindquist RIAF image: camera rays, Kerr BL
 * connection, fixed and step-doubling adaptive RK4, pass A segment endpoints,
 * pass B transport with camera crossing and depth control, the RIAF plasma
 * model, thermal synchrotron fits, the Strang-split Stokes step and the
 * observer-basis transform.  Each function cites the reference file:line it
 * follows.  It is the checker for the HIP kernels (tests, smoke, bench's CPU
 * baseline); the product never links or calls it.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   transfer step, observer transform   closed forms of tests/test_core.cpp:247-335
 *                                       (pure emission, absorption+emission,
 *                                       pi/2 Faraday rotation, Jones-vector oracle)
 *   thermal Pandya coefficients         Symphony fast-fit checkpoints,
 *                                       tests/test_core.cpp:2188-2237 (rel 1e-5)
 *   thermal rho_Q/rho_V, Dexter fit     no reference vectors: exact Maxwell-Juettner
 *                                       integrals, the large-X saddle point (exponent,
 *                                       amplitudes), high-frequency and cold-plasma
 *                                       limits, Kirchhoff's law a = j / B_nu
 *                                       (tests/test_oracle_fit_pins.py)
 *   Bessel K0/K1/K2                     scipy.special.kv (NR polynomial accuracy 1e-7)
 *   BL metric, connection, RHS          g g^-1 = I, finite differences of gcov,
 *                                       specialised == generic connection
 *   RK4 / adaptive RK4                  convergence order on Kerr geodesics, null
 *                                       and frame invariants, tolerance monotonicity
 *   camera                              -u.k = 1, k.k = 0, e_i.e_j = delta_ij
 *   whole RIAF image (pass A + B)       end-to-end 8x8 golden fluxes of
 *                                       tests/test_hdf5_contract.py:671-681 (rel 2e-6)
 *   FMKS map, metric, RHS               theta map and its inverse, g g^-1 = I and the
 *                                       KS congruence, specialised RHS == connection
 *                                       from finite differences of the metric
 *   GRMHD sampling, derived scalars     trilinear exactness on linear fields and
 *                                       ghosts; n_e, B, sigma, beta, Theta_e closed
 *                                       forms of iharm_loader.cpp:420-505
 *   kappa / power-law coefficients      Symphony checkpoints and the reference's
 *                                       coefficient checks; 2F1 against scipy
 *   slow-light time interpolation       interpolation rules, pair bisection by brute
 *                                       force, identical snapshots == static image
 *   direct_only, equatorial wedge       tests/test_direct_only.cpp and
 *                                       tests/test_equatorial.cpp checks (bracket truth
 *                                       table, dz/ds closed form, Kerr turn endpoints,
 *                                       analytic slab behind a screen)
 *   response accumulation               closed forms of
 *                                       tests/test_physical_response.cpp:44-90
 *   d/dR_high of slow-light images      central differences of the pinned image over
 *                                       R_high (no reference output exists)
 *   Dexter and kappa fit constants      PARITY UNPINNED to their printed digits: held
 *                                       to the fits' accuracy (~1 %) against the exact
 *                                       integrals
 */
#include "kp_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TINY 1e-300
#define LARGE 1e300
#define PI_D 3.141592653589793238462643383279502884

static inline double absv(double x) { return x < 0.0 ? -x : x; }
static inline double minv(double a, double b) { return a < b ? a : b; }
static inline double maxv(double a, double b) { return a > b ? a : b; }
static inline double clampv(double x, double lo, double hi) { return minv(maxv(x, lo), hi); }
static inline double sq(double x) { return x * x; }

/* ======================================================================
 * Boyer-Lindquist metric (src/geometry/kerr_boyer_lindquist.hpp)
 * ==================================================================== */
typedef struct bl_work {
    double r, th, a, a2, m, sinth, costh, sin2, cos2, sigma, delta;
    double dsigma_dr, dsigma_dth, ddelta_dr, dsin2_dth;
} bl_work;

/* KerrBoyerLindquistMetric::build_work (kerr_boyer_lindquist.hpp:28-45). */
static bl_work bl_build_work(double mass, double spin, const double* X) {
    bl_work w;
    w.r = maxv(X[1], TINY);
    w.th = X[2];
    w.a = spin;
    w.a2 = spin * spin;
    w.m = mass;
    w.sinth = sin(w.th);
    w.costh = cos(w.th);
    w.sin2 = maxv(w.sinth * w.sinth, TINY);
    w.cos2 = w.costh * w.costh;
    w.sigma = w.r * w.r + w.a2 * w.cos2;
    w.delta = w.r * w.r - 2.0 * w.m * w.r + w.a2;
    w.dsigma_dr = 2.0 * w.r;
    w.dsigma_dth = -2.0 * w.a2 * w.sinth * w.costh;
    w.ddelta_dr = 2.0 * (w.r - w.m);
    w.dsin2_dth = 2.0 * w.sinth * w.costh;
    return w;
}

/* gcov(mu,nu) element formulas (kerr_boyer_lindquist.hpp:47-71). */
void kpo_bl_gcov(double mass, double spin, const double x[4], double g[16]) {
    const bl_work w = bl_build_work(mass, spin, x);
    for (int i = 0; i < 16; ++i) g[i] = 0.0;
    g[0] = -(1.0 - 2.0 * w.m * w.r / w.sigma);
    g[3] = g[12] = -2.0 * w.m * w.a * w.r * w.sin2 / w.sigma;
    g[5] = w.sigma / w.delta;
    g[10] = w.sigma;
    g[15] = w.sin2 * (w.r * w.r + w.a2 + 2.0 * w.m * w.a2 * w.r * w.sin2 / w.sigma);
}

/* gcon(mu,nu) element formulas (kerr_boyer_lindquist.hpp:73-99). */
void kpo_bl_gcon(double mass, double spin, const double x[4], double g[16]) {
    const bl_work w = bl_build_work(mass, spin, x);
    for (int i = 0; i < 16; ++i) g[i] = 0.0;
    const double inv_sigma_delta = 1.0 / (w.sigma * w.delta);
    const double r2pa2 = w.r * w.r + w.a2;
    g[0] = -(r2pa2 * r2pa2 - w.a2 * w.delta * w.sin2) * inv_sigma_delta;
    g[3] = g[12] = -2.0 * w.m * w.a * w.r * inv_sigma_delta;
    g[5] = w.delta / w.sigma;
    g[10] = 1.0 / w.sigma;
    g[15] = (w.delta - w.a2 * w.sin2) / (w.sigma * w.delta * w.sin2);
}

/* KerrBoyerLindquistMetric::dot (kerr_boyer_lindquist.hpp:117-127). */
double kpo_bl_dot(double mass, double spin, const double x[4],
                  const double u[4], const double v[4]) {
    double g[16];
    kpo_bl_gcov(mass, spin, x, g);
    double out = 0.0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu)
            out += g[mu * 4 + nu] * u[mu] * v[nu];
    return out;
}

/* KerrBoyerLindquistMetric::dgcov (kerr_boyer_lindquist.hpp:129-177); dg[mu][nu][alpha]. */
void kpo_bl_dgcov(double mass, double spin, const double x[4], double dg[64]) {
    const bl_work w = bl_build_work(mass, spin, x);
    for (int i = 0; i < 64; ++i) dg[i] = 0.0;
    const double sig = w.sigma, sig2 = sig * sig, del = w.delta, del2 = del * del;
    const double r = w.r, m = w.m, a = w.a, a2 = w.a2, s2 = w.sin2, ds2 = w.dsin2_dth;
    const double sig_r = w.dsigma_dr, sig_t = w.dsigma_dth, del_r = w.ddelta_dr;
#define SETSYM(mu, nu, al, val) do { dg[((mu)*4+(nu))*4+(al)] = (val); dg[((nu)*4+(mu))*4+(al)] = (val); } while (0)
    SETSYM(0, 0, 1, 2.0 * m * (sig - r * sig_r) / sig2);
    SETSYM(0, 0, 2, -2.0 * m * r * sig_t / sig2);
    SETSYM(0, 3, 1, -2.0 * m * a * s2 * (sig - r * sig_r) / sig2);
    SETSYM(0, 3, 2, -2.0 * m * a * r * (ds2 * sig - s2 * sig_t) / sig2);
    SETSYM(1, 1, 1, (sig_r * del - sig * del_r) / del2);
    SETSYM(1, 1, 2, sig_t / del);
    SETSYM(2, 2, 1, sig_r);
    SETSYM(2, 2, 2, sig_t);
    const double extra = 2.0 * m * a2 * r * s2 / sig;
    const double aterm = r * r + a2 + extra;
    const double extra_r = 2.0 * m * a2 * s2 * (sig - r * sig_r) / sig2;
    const double extra_t = 2.0 * m * a2 * r * (ds2 * sig - s2 * sig_t) / sig2;
    SETSYM(3, 3, 1, s2 * (2.0 * r + extra_r));
    SETSYM(3, 3, 2, ds2 * aterm + s2 * extra_t);
#undef SETSYM
}

/* fill_connection_from_metric_derivs (src/geometry/connection.hpp:7-24). */
void kpo_bl_connection(double mass, double spin, const double x[4], double gamma[64]) {
    double gcon[16], dg[64];
    kpo_bl_gcon(mass, spin, x, gcon);
    kpo_bl_dgcov(mass, spin, x, dg);
    for (int mu = 0; mu < 4; ++mu)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                double comp = 0.0;
                for (int nu = 0; nu < 4; ++nu)
                    comp += 0.5 * gcon[mu * 4 + nu] *
                            (dg[(nu * 4 + b) * 4 + a] + dg[(nu * 4 + a) * 4 + b] -
                             dg[(a * 4 + b) * 4 + nu]);
                gamma[(mu * 4 + a) * 4 + b] = comp;
            }
}

/* Generic transport_rhs (src/geodesic/rk4.hpp:7-34): test oracle for the
 * specialised contraction. */
void kpo_bl_transport_rhs_generic(double mass, double spin, const kpo_state* s, kpo_state* d) {
    double gamma[64];
    kpo_bl_connection(mass, spin, s->v, gamma);
    for (int mu = 0; mu < 4; ++mu) d->v[mu] = s->v[4 + mu];
    for (int mu = 0; mu < 4; ++mu) {
        double dk = 0.0, de1 = 0.0, de2 = 0.0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                const double g = gamma[(mu * 4 + a) * 4 + b];
                dk -= g * s->v[4 + a] * s->v[4 + b];
                de1 -= g * s->v[4 + a] * s->v[8 + b];
                de2 -= g * s->v[4 + a] * s->v[12 + b];
            }
        d->v[4 + mu] = dk;
        d->v[8 + mu] = de1;
        d->v[12 + mu] = de2;
    }
}

typedef struct bl_conn {
    double dr00, dr03, dr11, dr22, dr33;
    double dt00, dt03, dt11, dt22, dt33;
    double g00, g03, g11, g22, g33;
} bl_conn;

/* build_bl_connection_work (kerr_boyer_lindquist.hpp:229-272). */
static bl_conn bl_connection_work(double mass, double spin, const double* X) {
    const bl_work w = bl_build_work(mass, spin, X);
    bl_conn c;
    const double sig = w.sigma, sig2 = sig * sig, del = w.delta, del2 = del * del;
    const double r = w.r, m = w.m, a = w.a, a2 = w.a2, s2 = w.sin2, ds2 = w.dsin2_dth;
    const double sig_r = w.dsigma_dr, sig_t = w.dsigma_dth, del_r = w.ddelta_dr;
    c.dr00 = 2.0 * m * (sig - r * sig_r) / sig2;
    c.dt00 = -2.0 * m * r * sig_t / sig2;
    c.dr03 = -2.0 * m * a * s2 * (sig - r * sig_r) / sig2;
    c.dt03 = -2.0 * m * a * r * (ds2 * sig - s2 * sig_t) / sig2;
    c.dr11 = (sig_r * del - sig * del_r) / del2;
    c.dt11 = sig_t / del;
    c.dr22 = sig_r;
    c.dt22 = sig_t;
    const double extra = 2.0 * m * a2 * r * s2 / sig;
    const double aterm = r * r + a2 + extra;
    const double extra_r = 2.0 * m * a2 * s2 * (sig - r * sig_r) / sig2;
    const double extra_t = 2.0 * m * a2 * r * (ds2 * sig - s2 * sig_t) / sig2;
    c.dr33 = s2 * (2.0 * r + extra_r);
    c.dt33 = ds2 * aterm + s2 * extra_t;
    const double inv_sigma_delta = 1.0 / (sig * del);
    const double r2pa2 = r * r + a2;
    c.g00 = -(r2pa2 * r2pa2 - a2 * del * s2) * inv_sigma_delta;
    c.g03 = -2.0 * m * a * r * inv_sigma_delta;
    c.g11 = del / sig;
    c.g22 = 1.0 / sig;
    c.g33 = (del - a2 * s2) / (sig * del * s2);
    return c;
}

/* bl_derivative_row_dot (kerr_boyer_lindquist.hpp:180-194). */
static inline double row_dot(double d00, double d03, double d11, double d22, double d33,
                             int row, const double* v) {
    if (row == 0) return d00 * v[0] + d03 * v[3];
    if (row == 1) return d11 * v[1];
    if (row == 2) return d22 * v[2];
    return d03 * v[0] + d33 * v[3];
}

/* bl_derivative_bilinear (kerr_boyer_lindquist.hpp:196-206). */
static inline double bilinear(double d00, double d03, double d11, double d22, double d33,
                              const double* p, const double* q) {
    return d00 * p[0] * q[0] + d03 * (p[0] * q[3] + p[3] * q[0]) + d11 * p[1] * q[1] +
           d22 * p[2] * q[2] + d33 * p[3] * q[3];
}

/* bl_connection_contract: Gamma^mu_ab p^a q^b (kerr_boyer_lindquist.hpp:274-295). */
static void bl_contract(const bl_conn* c, const double* p, const double* q, double* out) {
    double t[4];
    for (int nu = 0; nu < 4; ++nu) {
        t[nu] = p[1] * row_dot(c->dr00, c->dr03, c->dr11, c->dr22, c->dr33, nu, q) +
                p[2] * row_dot(c->dt00, c->dt03, c->dt11, c->dt22, c->dt33, nu, q) +
                q[1] * row_dot(c->dr00, c->dr03, c->dr11, c->dr22, c->dr33, nu, p) +
                q[2] * row_dot(c->dt00, c->dt03, c->dt11, c->dt22, c->dt33, nu, p);
    }
    t[1] -= bilinear(c->dr00, c->dr03, c->dr11, c->dr22, c->dr33, p, q);
    t[2] -= bilinear(c->dt00, c->dt03, c->dt11, c->dt22, c->dt33, p, q);
    out[0] = 0.5 * (c->g00 * t[0] + c->g03 * t[3]);
    out[1] = 0.5 * c->g11 * t[1];
    out[2] = 0.5 * c->g22 * t[2];
    out[3] = 0.5 * (c->g03 * t[0] + c->g33 * t[3]);
}

/* Specialised BL transport_rhs (kerr_boyer_lindquist.hpp:306-322). */
void kpo_bl_transport_rhs(double mass, double spin, const kpo_state* s, kpo_state* d) {
    const bl_conn c = bl_connection_work(mass, spin, s->v);
    double ck[4], ce1[4], ce2[4];
    bl_contract(&c, s->v + 4, s->v + 4, ck);
    bl_contract(&c, s->v + 4, s->v + 8, ce1);
    bl_contract(&c, s->v + 4, s->v + 12, ce2);
    for (int mu = 0; mu < 4; ++mu) {
        d->v[mu] = s->v[4 + mu];
        d->v[4 + mu] = -ck[mu];
        d->v[8 + mu] = -ce1[mu];
        d->v[12 + mu] = -ce2[mu];
    }
}

/* ======================================================================
 * FMKS geometry (src/geometry/kerr_fmks.hpp, src/geometry/kerr_schild_spherical.hpp,
 * src/grmhd/grid_mapping.hpp): x = (t, x1 = ln r, x2, phi) with
 * theta(x1, x2) blending the MKS and the polynomial jet maps.
 * ==================================================================== */
/* fmks_power / FMKSGridMapping::power (kerr_fmks.hpp:37-47, grid_mapping.hpp:45-55). */
static double fmks_power(double x, double p) {
    const int n = (int)(p + 0.5);
    if (absv(p - (double)n) < 1e-10 && n >= 0) {
        double y = 1.0;
        const double ax = absv(x);
        for (int i = 0; i < n; ++i) y *= ax;
        if (x < 0.0 && (n % 2) == 1) y = -y;
        return y;
    }
    return pow(absv(x), p);
}

/* theta_from_x2 (kerr_fmks.hpp:49-56). */
double kpo_fmks_theta(const kpo_fmks* f, double x1, double x2) {
    const double thG = PI_D * x2 + ((1.0 - f->hslope) / 2.0) * sin(2.0 * PI_D * x2);
    const double y = 2.0 * x2 - 1.0;
    const double thJ = f->poly_norm * y * (1.0 + fmks_power(y / f->poly_xt, f->poly_alpha) / (f->poly_alpha + 1.0)) +
                       0.5 * PI_D;
    return thG + exp(f->mks_smooth * (f->startx1 - x1)) * (thJ - thG);
}

/* dtheta_dx1, dtheta_dx2 (kerr_fmks.hpp:58-80). */
static double fmks_dtheta_dx1(const kpo_fmks* f, double x1, double x2) {
    const double thG = PI_D * x2 + ((1.0 - f->hslope) / 2.0) * sin(2.0 * PI_D * x2);
    const double y = 2.0 * x2 - 1.0;
    const double thJ = f->poly_norm * y * (1.0 + fmks_power(y / f->poly_xt, f->poly_alpha) / (f->poly_alpha + 1.0)) +
                       0.5 * PI_D;
    return -f->mks_smooth * exp(f->mks_smooth * (f->startx1 - x1)) * (thJ - thG);
}

static double fmks_dtheta_dx2(const kpo_fmks* f, double x1, double x2) {
    const double y = 2.0 * x2 - 1.0;
    const double expfac = exp(f->mks_smooth * (f->startx1 - x1));
    const double dthG = PI_D + (1.0 - f->hslope) * PI_D * cos(2.0 * PI_D * x2);
    const double dthJ = 2.0 * f->poly_norm * (1.0 + fmks_power(y / f->poly_xt, f->poly_alpha) / (f->poly_alpha + 1.0)) +
                        (2.0 * f->poly_alpha * f->poly_norm * y * fmks_power(y / f->poly_xt, f->poly_alpha - 1.0)) /
                            ((f->poly_alpha + 1.0) * f->poly_xt);
    return dthG + expfac * (dthJ - dthG);
}

/* native_x2_from_theta: 64 bisections (kerr_fmks.hpp:107-117). */
double kpo_fmks_x2_from_theta(const kpo_fmks* f, double x1, double theta) {
    double lo = 0.0, hi = 1.0;
    for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (kpo_fmks_theta(f, x1, mid) < theta)
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* Kerr-Schild spherical gcov at (r, theta) (kerr_schild_spherical.hpp:48-66). */
static void ks_gcov_rth(double mass, double spin, double r, double sth, double cth, double g[16]) {
    const double s2 = sth * sth;
    const double a2 = spin * spin;
    const double rho2 = r * r + a2 * cth * cth;
    for (int i = 0; i < 16; ++i) g[i] = 0.0;
    const double f = 2.0 * mass * r / rho2;
    g[0] = -1.0 + f;
    g[1] = f;
    g[3] = -spin * f * s2;
    g[4] = g[1];
    g[5] = 1.0 + f;
    g[7] = -spin * s2 * (1.0 + f);
    g[10] = rho2;
    g[12] = g[3];
    g[13] = g[7];
    g[15] = s2 * (rho2 + a2 * s2 * (1.0 + f));
}

/* KerrFMKSMetric::gcov_matrix: J^T g_KS J with J = dx_KS/dX (kerr_fmks.hpp:125-157). */
void kpo_fmks_gcov(double mass, double spin, const kpo_fmks* f, const double x[4], double g[16]) {
    const double r = exp(x[1]);
    const double th = kpo_fmks_theta(f, x[1], x[2]);
    double gks[16], jac[16];
    ks_gcov_rth(mass, spin, r, sin(th), cos(th), gks);
    for (int i = 0; i < 16; ++i) jac[i] = (i % 5 == 0) ? 1.0 : 0.0;
    jac[1 * 4 + 1] = r;
    jac[2 * 4 + 1] = fmks_dtheta_dx1(f, x[1], x[2]);
    jac[2 * 4 + 2] = fmks_dtheta_dx2(f, x[1], x[2]);
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) {
            double sum = 0.0;
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) sum += gks[a * 4 + b] * jac[a * 4 + mu] * jac[b * 4 + nu];
            g[mu * 4 + nu] = sum;
        }
}

/* invert_4x4: Gauss-Jordan with a tiny-pivot floor (kerr_fmks.hpp:159-182). */
static void invert_4x4(const double* a, double* inv) {
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            m[i][j] = a[i * 4 + j];
            m[i][j + 4] = i == j ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        double pivot = m[col][col];
        if (absv(pivot) < TINY) pivot = pivot < 0.0 ? -TINY : TINY;
        const double ip = 1.0 / pivot;
        for (int j = 0; j < 8; ++j) m[col][j] *= ip;
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            const double factor = m[row][col];
            for (int j = 0; j < 8; ++j) m[row][j] -= factor * m[col][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) inv[i * 4 + j] = m[i][j + 4];
}

void kpo_fmks_gcon(double mass, double spin, const kpo_fmks* f, const double x[4], double g[16]) {
    double cov[16];
    kpo_fmks_gcov(mass, spin, f, x, cov);
    invert_4x4(cov, g);
}

/* Spherical-KS connection work (kerr_schild_spherical.hpp:208-268):
 * gcon and the r- and theta-derivatives of gcov. */
typedef struct sks_work {
    double gcon[16], dgr[16], dgth[16];
} sks_work;

static void sks_build(double mass, double spin, double r_in, double th, sks_work* c) {
    const double r = maxv(r_in, TINY);
    const double s = sin(th), cc = cos(th);
    const double a = spin, a2 = a * a;
    const double s2 = s * s;
    const double safe_s2 = maxv(s2, TINY);
    const double rho2 = r * r + a2 * cc * cc;
    const double rho4 = rho2 * rho2;
    const double f = 2.0 * mass * r / rho2;
    const double delta = r * r - 2.0 * mass * r + a2;
    for (int i = 0; i < 16; ++i) c->gcon[i] = c->dgr[i] = c->dgth[i] = 0.0;
    c->gcon[0] = -(1.0 + f);
    c->gcon[1] = c->gcon[4] = f;
    c->gcon[5] = delta / rho2;
    c->gcon[7] = c->gcon[13] = a / rho2;
    c->gcon[10] = 1.0 / rho2;
    c->gcon[15] = 1.0 / (rho2 * safe_s2);
    const double drho_dr = 2.0 * r;
    const double drho_dth = -2.0 * a2 * s * cc;
    const double ds2_dth = 2.0 * s * cc;
    const double df_dr = 2.0 * mass * (rho2 - r * drho_dr) / rho4;
    const double df_dth = -2.0 * mass * r * drho_dth / rho4;
#define SYM(M, mu, nu, v) do { (M)[(mu) * 4 + (nu)] = (v); (M)[(nu) * 4 + (mu)] = (v); } while (0)
    SYM(c->dgr, 0, 0, df_dr);
    SYM(c->dgth, 0, 0, df_dth);
    SYM(c->dgr, 0, 1, df_dr);
    SYM(c->dgth, 0, 1, df_dth);
    SYM(c->dgr, 0, 3, -a * s2 * df_dr);
    SYM(c->dgth, 0, 3, -a * (s2 * df_dth + ds2_dth * f));
    SYM(c->dgr, 1, 1, df_dr);
    SYM(c->dgth, 1, 1, df_dth);
    SYM(c->dgr, 1, 3, -a * s2 * df_dr);
    SYM(c->dgth, 1, 3, -a * (ds2_dth * (1.0 + f) + s2 * df_dth));
    SYM(c->dgr, 2, 2, drho_dr);
    SYM(c->dgth, 2, 2, drho_dth);
    {
        const double bracket = rho2 + a2 * s2 * (1.0 + f);
        const double dbr_dr = drho_dr + a2 * s2 * df_dr;
        const double dbr_dth = drho_dth + a2 * (ds2_dth * (1.0 + f) + s2 * df_dth);
        SYM(c->dgr, 3, 3, s2 * dbr_dr);
        SYM(c->dgth, 3, 3, ds2_dth * bracket + s2 * dbr_dth);
    }
#undef SYM
}

/* spherical_ks_connection_contract_fixed_p (kerr_schild_spherical.hpp:318-356):
 * Gamma^mu_ab p^a q^b with the p row products precomputed. */
static void sks_contract(const sks_work* c, const double* p, const double* q, double* out) {
    double rp_r[4], rp_t[4], rq_r[4], rq_t[4];
    for (int nu = 0; nu < 4; ++nu) {
        double ar = 0.0, at = 0.0, br = 0.0, bt = 0.0;
        for (int b = 0; b < 4; ++b) {
            ar += c->dgr[nu * 4 + b] * p[b];
            at += c->dgth[nu * 4 + b] * p[b];
            br += c->dgr[nu * 4 + b] * q[b];
            bt += c->dgth[nu * 4 + b] * q[b];
        }
        rp_r[nu] = ar;
        rp_t[nu] = at;
        rq_r[nu] = br;
        rq_t[nu] = bt;
    }
    double pD1q = 0.0, pD2q = 0.0;
    for (int nu = 0; nu < 4; ++nu) {
        pD1q += p[nu] * rq_r[nu];
        pD2q += p[nu] * rq_t[nu];
    }
    double t[4];
    for (int nu = 0; nu < 4; ++nu) {
        double term = p[1] * rq_r[nu] + q[1] * rp_r[nu] + p[2] * rq_t[nu] + q[2] * rp_t[nu];
        if (nu == 1)
            term -= pD1q;
        else if (nu == 2)
            term -= pD2q;
        t[nu] = term;
    }
    for (int mu = 0; mu < 4; ++mu) {
        double comp = 0.0;
        for (int nu = 0; nu < 4; ++nu) comp += c->gcon[mu * 4 + nu] * t[nu];
        out[mu] = 0.5 * comp;
    }
}

/* FMKSConnectionWork (kerr_fmks.hpp:336-402): r, theta and the first and
 * second derivatives of theta(x1, x2), with the spherical-KS work at (r, theta). */
typedef struct fmks_work {
    double r, th, dth1, dth2, d2th11, d2th12, d2th22;
    sks_work sph;
} fmks_work;

static void fmks_build(double mass, double spin, const kpo_fmks* f, const double* X, fmks_work* c) {
    const double x1 = X[1], x2 = X[2];
    const double y = 2.0 * x2 - 1.0;
    const double z = y / f->poly_xt;
    const double expfac = exp(f->mks_smooth * (f->startx1 - x1));
    double za, zam1, zam2;
    if (absv(f->poly_alpha - 14.0) < 1e-12) {
        const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
        zam2 = z8 * z4;
        zam1 = zam2 * z;
        za = zam2 * z2;
    } else {
        za = fmks_power(z, f->poly_alpha);
        zam1 = fmks_power(z, f->poly_alpha - 1.0);
        zam2 = fmks_power(z, f->poly_alpha - 2.0);
    }
    const double thG = PI_D * x2 + ((1.0 - f->hslope) / 2.0) * sin(2.0 * PI_D * x2);
    const double thJ = f->poly_norm * y * (1.0 + za / (f->poly_alpha + 1.0)) + 0.5 * PI_D;
    const double dthG = PI_D + (1.0 - f->hslope) * PI_D * cos(2.0 * PI_D * x2);
    const double dthJ = 2.0 * f->poly_norm * (1.0 + za / (f->poly_alpha + 1.0)) +
                        (2.0 * f->poly_alpha * f->poly_norm * y * zam1) / ((f->poly_alpha + 1.0) * f->poly_xt);
    const double d2thG = -2.0 * (1.0 - f->hslope) * PI_D * PI_D * sin(2.0 * PI_D * x2);
    const double d2thJ = 4.0 * f->poly_norm * f->poly_alpha / (f->poly_alpha + 1.0) *
                         (2.0 * zam1 / f->poly_xt + y * (f->poly_alpha - 1.0) * zam2 / (f->poly_xt * f->poly_xt));
    c->r = exp(x1);
    c->th = thG + expfac * (thJ - thG);
    c->dth1 = -f->mks_smooth * expfac * (thJ - thG);
    c->dth2 = dthG + expfac * (dthJ - dthG);
    c->d2th11 = -f->mks_smooth * c->dth1;
    c->d2th12 = -f->mks_smooth * expfac * (dthJ - dthG);
    c->d2th22 = d2thG + expfac * (d2thJ - d2thG);
    sks_build(mass, spin, c->r, c->th, &c->sph);
}

/* fmks_to_spherical_vector (kerr_fmks.hpp:404-413). */
static void fmks_to_sph(const fmks_work* c, const double* v, double* out) {
    out[0] = v[0];
    out[1] = c->r * v[1];
    out[2] = c->dth1 * v[1] + c->dth2 * v[2];
    out[3] = v[3];
}

/* fmks_from_spherical_connection_contract (kerr_fmks.hpp:415-432). */
static void fmks_from_sph(const fmks_work* c, double* hs, const double* p, const double* q, double* out) {
    hs[1] += c->r * p[1] * q[1];
    hs[2] += c->d2th11 * p[1] * q[1] + c->d2th12 * (p[1] * q[2] + p[2] * q[1]) + c->d2th22 * p[2] * q[2];
    out[0] = hs[0];
    out[1] = hs[1] / c->r;
    out[2] = (hs[2] - c->dth1 * hs[1] / c->r) / c->dth2;
    out[3] = hs[3];
}

/* transport_rhs for KerrFMKSMetric (kerr_fmks.hpp:455-479). */
void kpo_fmks_transport_rhs(double mass, double spin, const kpo_fmks* f, const kpo_state* s, kpo_state* d) {
    fmks_work c;
    fmks_build(mass, spin, f, s->v, &c);
    const double* k = s->v + 4;
    double ks[4], e1s[4], e2s[4];
    fmks_to_sph(&c, k, ks);
    fmks_to_sph(&c, s->v + 8, e1s);
    fmks_to_sph(&c, s->v + 12, e2s);
    double hs[4], out[4];
    for (int mu = 0; mu < 4; ++mu) d->v[mu] = k[mu];
    sks_contract(&c.sph, ks, ks, hs);
    fmks_from_sph(&c, hs, k, k, out);
    for (int mu = 0; mu < 4; ++mu) d->v[4 + mu] = -out[mu];
    sks_contract(&c.sph, ks, e1s, hs);
    fmks_from_sph(&c, hs, k, s->v + 8, out);
    for (int mu = 0; mu < 4; ++mu) d->v[8 + mu] = -out[mu];
    sks_contract(&c.sph, ks, e2s, hs);
    fmks_from_sph(&c, hs, k, s->v + 12, out);
    for (int mu = 0; mu < 4; ++mu) d->v[12 + mu] = -out[mu];
}

/* Generic transport RHS from the FMKS metric by central differences of gcov
 * and the Christoffel formula: the check on the specialised contraction. */
void kpo_fmks_transport_rhs_fd(double mass, double spin, const kpo_fmks* f, const kpo_state* s, double h,
                               kpo_state* d) {
    double gcon[16], dg[64];
    kpo_fmks_gcon(mass, spin, f, s->v, gcon);
    for (int a = 0; a < 4; ++a) {
        double xp[4], xm[4], gp[16], gm[16];
        for (int i = 0; i < 4; ++i) xp[i] = xm[i] = s->v[i];
        xp[a] += h;
        xm[a] -= h;
        kpo_fmks_gcov(mass, spin, f, xp, gp);
        kpo_fmks_gcov(mass, spin, f, xm, gm);
        for (int i = 0; i < 16; ++i) dg[i * 4 + a] = (gp[i] - gm[i]) / (2.0 * h);
    }
    double gam[64];
    for (int mu = 0; mu < 4; ++mu)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                double sum = 0.0;
                for (int nu = 0; nu < 4; ++nu)
                    sum += gcon[mu * 4 + nu] *
                           (dg[(nu * 4 + a) * 4 + b] + dg[(nu * 4 + b) * 4 + a] - dg[(a * 4 + b) * 4 + nu]);
                gam[(mu * 4 + a) * 4 + b] = 0.5 * sum;
            }
    const double* k = s->v + 4;
    for (int mu = 0; mu < 4; ++mu) {
        d->v[mu] = k[mu];
        double ck = 0.0, c1 = 0.0, c2 = 0.0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                const double gm = gam[(mu * 4 + a) * 4 + b];
                ck += gm * k[a] * k[b];
                c1 += gm * k[a] * s->v[8 + b];
                c2 += gm * k[a] * s->v[12 + b];
            }
        d->v[4 + mu] = -ck;
        d->v[8 + mu] = -c1;
        d->v[12 + mu] = -c2;
    }
}

/* ======================================================================
 * Integrators (src/geodesic/rk4.hpp)
 * ==================================================================== */
typedef struct metric_ctx {
    double mass, spin;
    int64_t* rhs_counter;
    int fmks;               /* 0: Boyer-Lindquist, 1: FMKS, 2: the flat test path below */
    const kpo_fmks* fm;
} metric_ctx;

/* Kind 2 is a manufactured test path, not a spacetime: zero connection
 * (straight lines, constant wavevector and basis), radius 1 everywhere and
 * the equatorial height ratio x[3] -- the LinearHeight fixture of the
 * reference's emission-selection tests (tests/test_equatorial.cpp:14-21). */
static inline void rhs(const metric_ctx* m, const kpo_state* s, kpo_state* d) {
    if (m->rhs_counter) ++*m->rhs_counter;
    if (m->fmks == 2) {
        for (int i = 0; i < 16; ++i) d->v[i] = i < 4 ? s->v[i + 4] : 0.0;
        return;
    }
    if (m->fmks)
        kpo_fmks_transport_rhs(m->mass, m->spin, m->fm, s, d);
    else
        kpo_bl_transport_rhs(m->mass, m->spin, s, d);
}

/* metric_domain_radius (metric_domain.hpp:159-172): r = x1 (BL) or exp(x1) (FMKS). */
static inline double radius_of(const metric_ctx* m, const double* x) {
    return m->fmks == 2 ? 1.0 : (m->fmks ? exp(x[1]) : x[1]);
}

static void metric_gcov(const metric_ctx* m, const double* x, double* g) {
    if (m->fmks)
        kpo_fmks_gcov(m->mass, m->spin, m->fm, x, g);
    else
        kpo_bl_gcov(m->mass, m->spin, x, g);
}

/* axpy(y, a, x) = y + x * a (src/geodesic/state.hpp:39-44). */
static inline void axpy(const kpo_state* y, double a, const kpo_state* x, kpo_state* out) {
    for (int i = 0; i < 16; ++i) out->v[i] = y->v[i] + x->v[i] * a;
}

/* rk4_step_from_k1 (rk4.hpp:49-59): s + (k1 + 2 k2 + 2 k3 + k4) * (h / 6). */
static void rk4_from_k1(const metric_ctx* m, const kpo_state* s, double h,
                        const kpo_state* k1, kpo_state* out) {
    kpo_state tmp, k2, k3, k4;
    axpy(s, 0.5 * h, k1, &tmp);
    rhs(m, &tmp, &k2);
    axpy(s, 0.5 * h, &k2, &tmp);
    rhs(m, &tmp, &k3);
    axpy(s, h, &k3, &tmp);
    rhs(m, &tmp, &k4);
    const double h6 = h / 6.0;
    for (int i = 0; i < 16; ++i)
        out->v[i] = s->v[i] + (k1->v[i] + 2.0 * k2.v[i] + 2.0 * k3.v[i] + k4.v[i]) * h6;
}

/* rk4_step (rk4.hpp:36-46). */
static void rk4(const metric_ctx* m, const kpo_state* s, double h, kpo_state* out) {
    kpo_state k1;
    rhs(m, s, &k1);
    rk4_from_k1(m, s, h, &k1, out);
}

void kpo_bl_rk4_step(double mass, double spin, const kpo_state* s, double h, kpo_state* out) {
    metric_ctx m = {mass, spin, NULL, 0, NULL};
    rk4(&m, s, h, out);
}

/* state_error_norm (rk4.hpp:74-89). */
double kpo_state_error_norm(const kpo_state* a, const kpo_state* b) {
    double err = 0.0;
    for (int mu = 0; mu < 4; ++mu) {
        const double sx = maxv(1.0, maxv(absv(a->v[mu]), absv(b->v[mu])));
        const double sk = maxv(1.0, maxv(absv(a->v[4 + mu]), absv(b->v[4 + mu])));
        const double se1 = maxv(1.0, maxv(absv(a->v[8 + mu]), absv(b->v[8 + mu])));
        const double se2 = maxv(1.0, maxv(absv(a->v[12 + mu]), absv(b->v[12 + mu])));
        err = maxv(err, absv(a->v[mu] - b->v[mu]) / sx);
        err = maxv(err, absv(a->v[4 + mu] - b->v[4 + mu]) / sk);
        err = maxv(err, absv(a->v[8 + mu] - b->v[8 + mu]) / se1);
        err = maxv(err, absv(a->v[12 + mu] - b->v[12 + mu]) / se2);
    }
    return err;
}

typedef struct adaptive_control {
    double tolerance, min_step, max_step;
    int max_attempts;
} adaptive_control;

typedef struct adaptive_result {
    kpo_state state, mid_state;
    double used_h, next_h, error;
    int accepted;
} adaptive_result;

/* adaptive_rk4_step: step doubling (rk4.hpp:109-161); tolerance floor
 * 8 * DBL_EPSILON (src/common/math.hpp:56-60). */
static void adaptive_rk4(const metric_ctx* m, const kpo_state* state, double requested_h,
                         adaptive_control control, adaptive_result* out) {
    double h = minv(maxv(absv(requested_h), control.min_step), control.max_step);
    if (requested_h < 0.0) h = -h;
    double last_err = LARGE;
    const double tol_floor = 8.0 * 2.220446049250313080847263336181640625e-16;
    for (int attempt = 0; attempt < control.max_attempts; ++attempt) {
        kpo_state k1, full, half, two_half;
        rhs(m, state, &k1);
        rk4_from_k1(m, state, h, &k1, &full);
        rk4_from_k1(m, state, h * 0.5, &k1, &half);
        rk4(m, &half, h * 0.5, &two_half);
        const double err = kpo_state_error_norm(&full, &two_half);
        last_err = err;
        const double tol = maxv(control.tolerance, tol_floor);
        if (err <= tol) {
            double factor = 2.0;
            if (err > 0.0) {
                factor = 0.9 * pow(tol / err, 0.2);
                factor = minv(2.0, maxv(0.2, factor));
            }
            out->state = two_half;
            out->mid_state = half;
            out->used_h = h;
            out->next_h = minv(control.max_step, maxv(control.min_step, absv(h) * factor));
            if (requested_h < 0.0) out->next_h = -out->next_h;
            out->error = err;
            out->accepted = 1;
            return;
        }
        if (absv(h) <= control.min_step * 1.0001) break;
        double factor = 0.9 * pow(tol / maxv(err, TINY), 0.25);
        factor = minv(0.8, maxv(0.1, factor));
        h *= factor;
    }
    const double next_abs = maxv(control.min_step, absv(h) * 0.5);
    out->state = *state;
    out->mid_state = *state;
    out->used_h = 0.0;
    out->next_h = requested_h < 0.0 ? -next_abs : next_abs;
    out->error = last_err;
    out->accepted = 0;
}

int32_t kpo_bl_adaptive_rk4_step(double mass, double spin, const kpo_state* s,
                                 double requested_h, double tolerance,
                                 double min_step, double max_step, int32_t max_attempts,
                                 kpo_state* state, kpo_state* mid_state,
                                 double* used_h, double* next_h, double* error) {
    metric_ctx m = {mass, spin, NULL, 0, NULL};
    adaptive_control c = {tolerance, min_step, max_step, max_attempts};
    adaptive_result r;
    adaptive_rk4(&m, s, requested_h, c, &r);
    if (state) *state = r.state;
    if (mid_state) *mid_state = r.mid_state;
    if (used_h) *used_h = r.used_h;
    if (next_h) *next_h = r.next_h;
    if (error) *error = r.error;
    return r.accepted;
}

/* ======================================================================
 * Camera (src/camera/camera.hpp), BL pinhole only
 * ==================================================================== */
static void raise_cov(const double* gcon, const double* cov, double* out) {
    for (int mu = 0; mu < 4; ++mu) {
        double sum = 0.0;
        for (int nu = 0; nu < 4; ++nu) sum += gcon[mu * 4 + nu] * cov[nu];
        out[mu] = sum;
    }
}

static double dot_g(const double* g, const double* u, const double* v) {
    double out = 0.0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) out += g[mu * 4 + nu] * u[mu] * v[nu];
    return out;
}

/* project_out_camera (camera.hpp:353-363). */
static void project_out(const double* g, double* a, const double* b) {
    const double b2 = dot_g(g, b, b);
    if (absv(b2) > TINY) {
        const double f = dot_g(g, a, b) / b2;
        for (int i = 0; i < 4; ++i) a[i] = a[i] - b[i] * f;
    }
}

/* normalize_camera_basis_vector (camera.hpp:365-372). */
static void normalize_basis(const double* g, double* v) {
    const double n2 = dot_g(g, v, v);
    const double n = sqrt(maxv(absv(n2), TINY));
    const double inv = 1.0 / n;
    for (int i = 0; i < 4; ++i) v[i] = v[i] * inv;
}

/* levi_civita4 + tetrad_orientation (camera.hpp:374-407). */
static int levi_civita4(int i, int j, int k, int l) {
    if (i == j || i == k || i == l || j == k || j == l || k == l) return 0;
    int a[4] = {i, j, k, l};
    int sign = 1;
    for (int m = 0; m < 4; ++m)
        for (int n = m + 1; n < 4; ++n)
            if (a[m] > a[n]) sign = -sign;
    return sign;
}

static double tetrad_orientation(const double* e0, const double* e1, const double* e2,
                                 const double* e3) {
    double out = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k)
                for (int l = 0; l < 4; ++l)
                    out += (double)levi_civita4(i, j, k, l) * e0[i] * e1[j] * e2[k] * e3[l];
    return out;
}

/* make_ipole_camera_tetrad for BL (camera.hpp:415-453) with the BL branches
 * of ipole_camera_observer_velocity (:396-413), ipole_centering_covector
 * (:350-360) and native_vector_phi0_to_metric (:300-320).
 * tetrad[0..3]=e0, [4..7]=e1, [8..11]=e2, [12..15]=e3. */
void kpo_bl_camera_tetrad(const kp_pass_a_params* p, double tetrad[16]) {
    const double mass = p->mass, spin = p->spin;
    const double x[4] = {0.0, p->camera.radius, p->camera.inclination, 0.0};
    double gcov[16], gcon[16];
    kpo_bl_gcov(mass, spin, x, gcov);
    kpo_bl_gcon(mass, spin, x, gcon);
    double* e0 = tetrad;
    double* e1 = tetrad + 4;
    double* e2 = tetrad + 8;
    double* e3 = tetrad + 12;
    {
        const double r = maxv(x[1], TINY);
        const double delta = maxv(r * r - 2.0 * mass * r + spin * spin, TINY);
        const double cov[4] = {-1.0, -2.0 * mass * r / delta, 0.0, 0.0};
        raise_cov(gcon, cov, e0);
        const double center[4] = {1.0, 1.0 / r + 2.0 * mass * r / delta, 0.0, 0.0};
        raise_cov(gcon, center, e3);
    }
    {
        const double r = maxv(p->camera.radius, TINY);
        const double delta = maxv(r * r - 2.0 * mass * r + spin * spin, TINY);
        /* e2 trial: native (0,0,1,0) */
        double vr = r * 0.0;
        e2[0] = 0.0 - 2.0 * mass * r * vr / delta;
        e2[1] = vr;
        e2[2] = PI_D * 1.0;
        e2[3] = 0.0 - spin * vr / delta;
        /* e1 trial: native (1,1,1,1) */
        vr = r * 1.0;
        e1[0] = 1.0 - 2.0 * mass * r * vr / delta;
        e1[1] = vr;
        e1[2] = PI_D * 1.0;
        e1[3] = 1.0 - spin * vr / delta;
    }
    normalize_basis(gcov, e0);
    project_out(gcov, e3, e0);
    project_out(gcov, e3, e0);
    normalize_basis(gcov, e3);

    project_out(gcov, e2, e0);
    project_out(gcov, e2, e3);
    project_out(gcov, e2, e0);
    project_out(gcov, e2, e3);
    normalize_basis(gcov, e2);

    project_out(gcov, e1, e0);
    project_out(gcov, e1, e2);
    project_out(gcov, e1, e3);
    project_out(gcov, e1, e0);
    project_out(gcov, e1, e2);
    project_out(gcov, e1, e3);
    normalize_basis(gcov, e1);

    if (tetrad_orientation(e0, e1, e2, e3) < 0.0)
        for (int i = 0; i < 4; ++i) e1[i] = e1[i] * -1.0;
}

/* pixel_offsets (camera.hpp:96-106). */
void kpo_pixel_offsets(int32_t pixel, const kp_camera_params* cam, double* sx, double* sy) {
    const int i = pixel % cam->nx;
    const int j = pixel / cam->nx;
    const double x_extent = cam->xspan > 0.0 ? 2.0 * cam->xspan : cam->fov;
    const double y_extent = cam->yspan > 0.0 ? 2.0 * cam->yspan
                          : (cam->xspan > 0.0 ? 2.0 * cam->xspan * (double)cam->ny / (double)cam->nx
                                              : (cam->fov_y > 0.0 ? cam->fov_y : cam->fov));
    *sx = (((double)i + 0.5) / (double)cam->nx - 0.5) * x_extent + cam->x_offset;
    *sy = (((double)j + 0.5) / (double)cam->ny - 0.5) * y_extent + cam->y_offset;
}

/* initialize_pinhole_camera_ray (camera.hpp:484-513) with
 * normalize_spacelike (:304-313) and orthogonalize_spacelike (:315-325). */
static void camera_ray_from_tetrad(const kp_pass_a_params* p, const double* xcam, const double* tetrad,
                                   const double* gcov, int32_t pixel, kpo_state* out) {
    const double* e0 = tetrad;
    const double* e1 = tetrad + 4;
    const double* e2 = tetrad + 8;
    const double* e3 = tetrad + 12;
    double sx = 0.0, sy = 0.0;
    kpo_pixel_offsets(pixel, &p->camera, &sx, &sy);
    const double spatial_norm = sqrt(1.0 + sx * sx + sy * sy);
    const double kt0 = 1.0;
    const double kt1 = sx / spatial_norm;
    const double kt2 = sy / spatial_norm;
    const double kt3 = 1.0 / spatial_norm;
    double sd[4];
    for (int i = 0; i < 4; ++i) sd[i] = e1[i] * kt1 + e2[i] * kt2 + e3[i] * kt3;
    for (int i = 0; i < 4; ++i) out->v[i] = xcam[i];
    for (int i = 0; i < 4; ++i) out->v[4 + i] = e0[i] * kt0 + sd[i];
    double a[4], b[4];
    for (int i = 0; i < 4; ++i) a[i] = e1[i] - sd[i] * kt1;
    {
        const double n2 = dot_g(gcov, a, a);
        if (n2 > 0.0) {
            const double inv = 1.0 / sqrt(n2);
            for (int i = 0; i < 4; ++i) a[i] = a[i] * inv;
        }
    }
    for (int i = 0; i < 4; ++i) b[i] = e2[i] - sd[i] * kt2;
    {
        const double denom = dot_g(gcov, a, a);
        if (absv(denom) > TINY) {
            const double f = dot_g(gcov, b, a) / denom;
            for (int i = 0; i < 4; ++i) b[i] = b[i] - a[i] * f;
        }
        const double n2 = dot_g(gcov, b, b);
        if (n2 > 0.0) {
            const double inv = 1.0 / sqrt(n2);
            for (int i = 0; i < 4; ++i) b[i] = b[i] * inv;
        }
    }
    for (int i = 0; i < 4; ++i) {
        out->v[8 + i] = a[i];
        out->v[12 + i] = b[i];
    }
}

void kpo_bl_camera_ray(const kp_pass_a_params* p, int32_t pixel, kpo_state* out) {
    double tetrad[16], gcov[16];
    const double x[4] = {0.0, p->camera.radius, p->camera.inclination, 0.0};
    kpo_bl_camera_tetrad(p, tetrad);
    kpo_bl_gcov(p->mass, p->spin, x, gcov);
    camera_ray_from_tetrad(p, x, tetrad, gcov, pixel, out);
}

/* Camera position and ipole tetrad in FMKS (camera.hpp:303-315 and 415-453 with
 * the FMKS branches: position (0, ln R, x2(ln R, inclination), 0), normal
 * observer, centring covector (1, 1, 0, 0), trial vectors (0,0,1,0), (1,1,1,1)). */
void kpo_fmks_camera_tetrad(const kp_pass_a_params* p, const kpo_fmks* f, double xcam[4], double tetrad[16]) {
    const double x1 = log(maxv(p->camera.radius, TINY));
    xcam[0] = 0.0;
    xcam[1] = x1;
    xcam[2] = kpo_fmks_x2_from_theta(f, x1, p->camera.inclination);
    xcam[3] = 0.0;
    double gcov[16], gcon[16];
    kpo_fmks_gcov(p->mass, p->spin, f, xcam, gcov);
    kpo_fmks_gcon(p->mass, p->spin, f, xcam, gcon);
    double* e0 = tetrad;
    double* e1 = tetrad + 4;
    double* e2 = tetrad + 8;
    double* e3 = tetrad + 12;
    /* normal_observer_velocity (camera.hpp:141-152) */
    for (int mu = 0; mu < 4; ++mu) e0[mu] = -gcon[mu * 4 + 0];
    {
        const double n2 = dot_g(gcov, e0, e0);
        if (n2 < 0.0) {
            const double inv = 1.0 / sqrt(-n2);
            for (int mu = 0; mu < 4; ++mu) e0[mu] = e0[mu] * inv;
        }
    }
    const double center[4] = {1.0, 1.0, 0.0, 0.0};
    raise_cov(gcon, center, e3);
    e2[0] = 0.0; e2[1] = 0.0; e2[2] = 1.0; e2[3] = 0.0;
    e1[0] = 1.0; e1[1] = 1.0; e1[2] = 1.0; e1[3] = 1.0;
    normalize_basis(gcov, e0);
    project_out(gcov, e3, e0);
    project_out(gcov, e3, e0);
    normalize_basis(gcov, e3);
    project_out(gcov, e2, e0);
    project_out(gcov, e2, e3);
    project_out(gcov, e2, e0);
    project_out(gcov, e2, e3);
    normalize_basis(gcov, e2);
    project_out(gcov, e1, e0);
    project_out(gcov, e1, e2);
    project_out(gcov, e1, e3);
    project_out(gcov, e1, e0);
    project_out(gcov, e1, e2);
    project_out(gcov, e1, e3);
    normalize_basis(gcov, e1);
    if (tetrad_orientation(e0, e1, e2, e3) < 0.0)
        for (int i = 0; i < 4; ++i) e1[i] = e1[i] * -1.0;
}

void kpo_fmks_camera_ray(const kp_pass_a_params* p, const kpo_fmks* f, int32_t pixel, kpo_state* out) {
    double x[4], tetrad[16], gcov[16];
    kpo_fmks_camera_tetrad(p, f, x, tetrad);
    kpo_fmks_gcov(p->mass, p->spin, f, x, gcov);
    camera_ray_from_tetrad(p, x, tetrad, gcov, pixel, out);
}

/* ======================================================================
 * Frame diagnostics (src/geodesic/state.hpp:44-76)
 * ==================================================================== */
static double max_frame_error(const metric_ctx* m, const kpo_state* s) {
    double g[16];
    metric_gcov(m, s->v, g);
    const double* k = s->v + 4;
    const double* e1 = s->v + 8;
    const double* e2 = s->v + 12;
    double out = absv(dot_g(g, k, k));
    out = maxv(out, absv(dot_g(g, e1, k)));
    out = maxv(out, absv(dot_g(g, e2, k)));
    out = maxv(out, absv(dot_g(g, e1, e1) - 1.0));
    out = maxv(out, absv(dot_g(g, e2, e2) - 1.0));
    out = maxv(out, absv(dot_g(g, e1, e2)));
    return out;
}

/* ======================================================================
 * Thermal synchrotron (src/radiation/thermal_synchrotron.hpp)
 * ==================================================================== */
#define EE 4.803204712570263e-10
#define ME 9.1093837015e-28
#define CL 2.99792458e10
#define HPL 6.62607015e-27
#define MP 1.67262192369e-24

/* Numerical-Recipes polynomial Bessel functions (thermal_synchrotron.hpp:60-165). */
static double bessel_i0(double ax) {
    if (ax < 3.75) {
        double y = ax / 3.75;
        y *= y;
        double p = 0.0045813;
        p = 0.0360768 + y * p;
        p = 0.2659732 + y * p;
        p = 1.2067492 + y * p;
        p = 3.0899424 + y * p;
        p = 3.5156229 + y * p;
        return 1.0 + y * p;
    }
    const double y = 3.75 / ax;
    double p = 0.00392377;
    p = -0.01647633 + y * p;
    p = 0.02635537 + y * p;
    p = -0.02057706 + y * p;
    p = 0.00916281 + y * p;
    p = -0.00157565 + y * p;
    p = 0.00225319 + y * p;
    p = 0.01328592 + y * p;
    p = 0.39894228 + y * p;
    return exp(ax) / sqrt(ax) * p;
}

static double bessel_i1(double ax) {
    if (ax < 3.75) {
        double y = ax / 3.75;
        y *= y;
        double p = 0.00032411;
        p = 0.00301532 + y * p;
        p = 0.02658733 + y * p;
        p = 0.15084934 + y * p;
        p = 0.51498869 + y * p;
        p = 0.87890594 + y * p;
        p = 0.5 + y * p;
        return ax * p;
    }
    const double y = 3.75 / ax;
    double p = -0.00420059;
    p = 0.01787654 + y * p;
    p = -0.02895312 + y * p;
    p = 0.02282967 + y * p;
    p = -0.01031555 + y * p;
    p = 0.00163801 + y * p;
    p = -0.00362018 + y * p;
    p = -0.03988024 + y * p;
    p = 0.39894228 + y * p;
    return exp(ax) / sqrt(ax) * p;
}

static double bessel_k0(double x) {
    if (x <= 0.0) return LARGE;
    if (x <= 2.0) {
        const double y = x * x * 0.25;
        double p = 0.00000740;
        p = 0.00010750 + y * p;
        p = 0.00262698 + y * p;
        p = 0.03488590 + y * p;
        p = 0.23069756 + y * p;
        p = 0.42278420 + y * p;
        p = -0.57721566 + y * p;
        return -log(x * 0.5) * bessel_i0(x) + p;
    }
    const double y = 2.0 / x;
    double p = 0.00053208;
    p = -0.00251540 + y * p;
    p = 0.00587872 + y * p;
    p = -0.01062446 + y * p;
    p = 0.02189568 + y * p;
    p = -0.07832358 + y * p;
    p = 1.25331414 + y * p;
    return exp(-x) / sqrt(x) * p;
}

static double bessel_k1(double x) {
    if (x <= 0.0) return LARGE;
    if (x <= 2.0) {
        const double y = x * x * 0.25;
        double p = -0.00004686;
        p = -0.00110404 + y * p;
        p = -0.01919402 + y * p;
        p = -0.18156897 + y * p;
        p = -0.67278579 + y * p;
        p = 0.15443144 + y * p;
        p = 1.0 + y * p;
        return log(x * 0.5) * bessel_i1(x) + p / x;
    }
    const double y = 2.0 / x;
    double p = -0.00068245;
    p = 0.00325614 + y * p;
    p = -0.00780353 + y * p;
    p = 0.01504268 + y * p;
    p = -0.03655620 + y * p;
    p = 0.23498619 + y * p;
    p = 1.25331414 + y * p;
    return exp(-x) / sqrt(x) * p;
}

/* thermal_bessel_k0k1k2 (thermal_synchrotron.hpp:211-231): K2 from the recurrence. */
void kpo_bessel_k012(double x, double* k0, double* k1, double* k2) {
    const double k0c = bessel_k0(x);
    const double k1c = bessel_k1(x);
    const double k2c = maxv(k0c + 2.0 * k1c / maxv(x, TINY), TINY);
    if (k0) *k0 = k0c;
    if (k1) *k1 = k1c;
    if (k2) *k2 = k2c;
}

/* symphony_bnu_inv (thermal_synchrotron.hpp:233-251): B_nu / nu^3. */
static double bnu_inv(double nu, double thetae) {
    if (!(thetae > 0.0)) return 0.0;
    const double x = HPL * nu / (ME * CL * CL * thetae);
    if (x < 2e-3) {
        const double denom = x / 24.0 * (24.0 + x * (12.0 + x * (4.0 + x)));
        return (2.0 * HPL / maxv(denom, TINY)) / (CL * CL);
    }
    const double denom = exp(x) - 1.0;
    return (2.0 * HPL / maxv(denom, TINY)) / (CL * CL);
}

/* symphony_pandya_* (thermal_synchrotron.hpp:311-395). */
static double pandya_nu_c(double b) { return EE * b / (2.0 * PI_D * ME * CL); }

static double pandya_x(double nu, double thetae, double b, double sin_theta) {
    const double nu_c = pandya_nu_c(b);
    const double nu_s = 2.0 / 9.0 * nu_c * sin_theta * thetae * thetae;
    return maxv(nu / maxv(nu_s, TINY), TINY);
}

static double pandya_jI(double nu, double ne, double thetae, double b, double sin_theta) {
    if (!(thetae > 0.0) || !(ne > 0.0) || !(b > 0.0) || !(sin_theta > 0.0)) return 0.0;
    const double nu_c = pandya_nu_c(b);
    const double x = pandya_x(nu, thetae, b, sin_theta);
    const double prefactor = ne * EE * EE * nu_c / CL;
    const double term1 = sqrt(2.0) * PI_D / 27.0 * sin_theta;
    const double x13 = cbrt(x);
    const double shape = sqrt(x) + 1.8877486253633869933 * sqrt(x13);
    const double term2 = shape * shape;
    return prefactor * term1 * term2 * exp(-x13);
}

static double pandya_jQ(double nu, double ne, double thetae, double b, double sin_theta) {
    if (!(thetae > 0.0) || !(ne > 0.0) || !(b > 0.0) || !(sin_theta > 0.0)) return 0.0;
    const double nu_c = pandya_nu_c(b);
    const double x = pandya_x(nu, thetae, b, sin_theta);
    const double t2425 = pow(thetae, 24.0 / 25.0);
    const double prefactor = ne * EE * EE * nu_c / CL;
    const double term1 = sqrt(2.0) * PI_D / 27.0 * sin_theta;
    const double term2 = (7.0 * t2425 + 35.0) / (10.0 * t2425 + 75.0);
    const double x13 = cbrt(x);
    const double shape = sqrt(x) + term2 * 1.8877486253633869933 * sqrt(x13);
    const double term3 = shape * shape;
    return prefactor * term1 * term3 * exp(-x13);
}

static double pandya_jV(double nu, double ne, double thetae, double b, double sin_theta,
                        double cos_theta) {
    if (!(thetae > 0.0) || !(ne > 0.0) || !(b > 0.0) || !(sin_theta > 0.0)) return 0.0;
    const double nu_c = pandya_nu_c(b);
    const double x = pandya_x(nu, thetae, b, sin_theta);
    const double prefactor = ne * EE * EE * nu_c / CL;
    const double sin_shift = sin_theta * 0.4356824462767121 - cos_theta * 0.9001004421765051;
    const double term1 = (37.0 - 87.0 * sin_shift) / (100.0 * (thetae + 1.0));
    const double term2 = pow(1.0 + (pow(thetae, 3.0 / 5.0) / 25.0 + 7.0 / 10.0) *
                                       pow(x, 9.0 / 25.0),
                             5.0 / 3.0);
    return prefactor * term1 * term2 * exp(-cbrt(x));
}

/* Faraday x-variable shared by rho_Q/rho_V (thermal_synchrotron.hpp:409-411). */
static double faraday_x(double thetae, double sin_theta, double omega0, double two_pi_nu) {
    return maxv(thetae * sqrt(maxv(0.0, sqrt(2.0) * sin_theta *
                                            (1e3 * omega0 / maxv(two_pi_nu, TINY)))),
                1e-12);
}

/* symphony_rho_V (thermal_synchrotron.hpp:431-459). */
double kpo_symphony_rho_v(double nu, double ne, double thetae, double b, double theta,
                          double min_sin_theta) {
    const double sin_theta = maxv(sin(theta), min_sin_theta);
    const double omega0 = EE * b / (ME * CL);
    const double wp2 = 4.0 * PI_D * ne * EE * EE / ME;
    const double two_pi_nu = 2.0 * PI_D * nu;
    const double x = faraday_x(thetae, sin_theta, omega0, two_pi_nu);
    double k0 = 0.0, k2 = 0.0;
    kpo_bessel_k012(1.0 / thetae, &k0, NULL, &k2);
    const double shgm = 1.0 - 0.11 * log(1.0 + 0.035 * x);
    const double k_ratio = k2 > 0.0 ? k0 / k2 : 1.0;
    const double fit_factor = k_ratio * shgm;
    const double safe = maxv(two_pi_nu, TINY);
    double eps12 = wp2 * omega0;
    eps12 /= safe;
    eps12 /= safe;
    eps12 /= safe;
    eps12 *= fit_factor * cos(theta);
    return 2.0 * PI_D * nu / CL * eps12;
}

/* symphony_rho_QV (thermal_synchrotron.hpp:461-506). */
static void rho_qv(double nu, double ne, double thetae, double b, double sin_theta,
                   double cos_theta, double* rhoQ, double* rhoV) {
    const double omega0 = EE * b / (ME * CL);
    const double wp2 = 4.0 * PI_D * ne * EE * EE / ME;
    const double two_pi_nu = 2.0 * PI_D * nu;
    const double x = faraday_x(thetae, sin_theta, omega0, two_pi_nu);
    double k0 = 0.0, k1 = 0.0, k2 = 0.0;
    kpo_bessel_k012(1.0 / thetae, &k0, &k1, &k2);
    const double exp_x_47 = exp(-x / 47.2);
    const double extraterm = (0.011 * exp_x_47 - 369.69342869090815 * pow(x + 1e-16, -8.0 / 3.0)) *
                             (0.5 + 0.5 * tanh((log(x) - 4.787491742782046) / 0.1));
    const double jffunc = 2.011 * exp(-pow(x, 1.035) / 4.7) -
                          cos(x / 2.0) * exp(-pow(x, 1.2) / 2.73) - 0.011 * exp_x_47 + extraterm;
    const double k1_ratio = k2 > 0.0 ? k1 / k2 : 1.0;
    const double safe = maxv(two_pi_nu, TINY);
    double eps11m22 = jffunc * wp2 * omega0 * omega0;
    eps11m22 /= safe;
    eps11m22 /= safe;
    eps11m22 /= safe;
    eps11m22 /= safe;
    eps11m22 *= (k1_ratio + 6.0 * thetae) * sin_theta * sin_theta;
    *rhoQ = PI_D * nu / CL * eps11m22;

    const double shgm = 1.0 - 0.11 * log(1.0 + 0.035 * x);
    const double k0_ratio = k2 > 0.0 ? k0 / k2 : 1.0;
    const double fit_factor = k0_ratio * shgm;
    double eps12 = wp2 * omega0;
    eps12 /= safe;
    eps12 /= safe;
    eps12 /= safe;
    eps12 *= fit_factor * cos_theta;
    *rhoV = 2.0 * PI_D * nu / CL * eps12;
}

/* thermal_synchrotron_magnetic_basis_coefficients_fit (thermal_synchrotron.hpp:512-616):
 * invariant coefficients in the magnetic screen basis; fit 1 = Pandya, 4 = Dexter. */
void kpo_thermal_coefficients(int32_t fit, double nu_in, double ne, double thetae, double b,
                              double theta, double sin_theta_in, double cos_theta,
                              double max_pol_frac, double min_sin_theta, double emission_scale,
                              double absorption_scale, double faraday_scale, kpo_coeffs* c) {
    memset(c, 0, sizeof(*c));
    (void)theta;
    if (!(nu_in > 0.0) || !(ne > 0.0) || !(thetae > 0.0) || !(b > 0.0)) return;
    const double nu = maxv(nu_in, 1.0);
    const double sin_theta = maxv(sin_theta_in, min_sin_theta);
    double jI = 0.0, jQ = 0.0, jV = 0.0;
    if (fit == KP_EMISSION_THERMAL_DEXTER) {
        const double nus = 3.0 * EE * b * sin_theta / (4.0 * PI_D * ME * CL) * thetae * thetae + 1.0;
        const double x = maxv(nu / maxv(nus, TINY), TINY);
        const double xm13 = 1.0 / cbrt(x);
        const double exp_shape = exp(-1.8899 / xm13);
        const double pref = ne * EE * EE * nu / (2.0 * sqrt(3.0) * CL * thetae * thetae);
        const double i_val = 2.5651 * (1.0 + 1.92 * xm13 + 0.9977 * xm13 * xm13) * exp_shape;
        const double iq = 2.5651 * (1.0 + 0.93193 * xm13 + 0.499873 * xm13 * xm13) * exp_shape;
        jI = pref * i_val;
        jQ = pref * iq;
        const double nus_v = 3.0 * EE * b * sin_theta / (4.0 * PI_D * ME * CL) * thetae * thetae + 1.0;
        const double xv = maxv(nu / maxv(nus_v, TINY), TINY);
        const double xv_m13 = 1.0 / cbrt(xv);
        const double iv = (1.81384 / xv + 3.42319 * xv_m13 * xv_m13 + 0.0292545 / sqrt(xv) +
                           2.03773 * xv_m13) * exp(-1.8899 / xv_m13);
        double cos_for_tan = cos_theta;
        if (absv(cos_for_tan) < 1e-10) cos_for_tan = cos_for_tan >= 0.0 ? 1e-10 : -1e-10;
        const double tan_theta = sin_theta / cos_for_tan;
        jV = 2.0 * ne * EE * EE * nu /
             (tan_theta * 3.0 * sqrt(3.0) * CL * thetae * thetae * thetae) * iv;
    } else {
        jI = pandya_jI(nu, ne, thetae, b, sin_theta);
        jQ = pandya_jQ(nu, ne, thetae, b, sin_theta);
        jV = pandya_jV(nu, ne, thetae, b, sin_theta, cos_theta);
    }
    const double nusq = nu * nu;
    double jI_inv = jI / maxv(nusq, TINY);
    double jQ_inv = jQ / maxv(nusq, TINY);
    double jV_inv = jV / maxv(nusq, TINY);
    const double jP_scale = maxv(absv(jQ_inv), absv(jV_inv));
    const double jP = jP_scale > 0.0
                          ? jP_scale * sqrt(sq(jQ_inv / jP_scale) + sq(jV_inv / jP_scale))
                          : 0.0;
    if (jI_inv > 0.0 && jP > 0.0 && jI_inv < jP / maxv(max_pol_frac, 1e-30)) {
        const double scale = jI_inv / jP * max_pol_frac;
        jQ_inv *= scale;
        jV_inv *= scale;
    }
    double aI = 0.0, aQ = 0.0, aV = 0.0;
    const double bi = bnu_inv(nu, thetae);
    if (bi > 0.0) {
        aI = jI_inv / bi;
        aQ = jQ_inv / bi;
        aV = jV_inv / bi;
    }
    double rhoQ = 0.0, rhoV = 0.0;
    rho_qv(nu, ne, thetae, b, sin_theta, cos_theta, &rhoQ, &rhoV);
    c->jI = emission_scale * jI_inv;
    c->jQ = emission_scale * jQ_inv;
    c->jV = emission_scale * jV_inv;
    c->aI = absorption_scale * aI;
    c->aQ = absorption_scale * aQ;
    c->aV = absorption_scale * aV;
    c->rQ = faraday_scale * rhoQ * nu;
    c->rV = faraday_scale * rhoV * nu;
}

/* ======================================================================
 * Nonthermal synchrotron: kappa and power-law distributions
 * (src/radiation/nonthermal_synchrotron.hpp).  The fits are Pandya et al.
 * (2016) for j and alpha and Marszewski et al. (2021) for the kappa rho;
 * their coefficients are the published ones.  The header carries its own
 * CGS constants (:106-110, :125-127), which differ from the thermal fits'
 * in the last digits, so they are kept separately here.
 * ==================================================================== */
#define NT_EE 4.80320680e-10
#define NT_ME 9.1093826e-28

static double nt_nonzero(double x) {
    if (absv(x) < TINY) return x >= 0.0 ? TINY : -TINY;
    return x;
}

/* hyp2f1_series (nonthermal_synchrotron.hpp:55-73): the Gauss series,
 * at most 200 terms, stopped at |term| < 1e-13 |sum|. */
static double nt_hyp2f1_series(double a, double b, double c, double z) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n <= 200; ++n) {
        const double m = (double)n - 1.0;
        term *= (a + m) * (b + m) * z / nt_nonzero((c + m) * (double)n);
        sum += term;
        if (absv(term) < absv(sum) * 1e-13) break;
    }
    return sum;
}

/* stable_hyp2f1 (:83-96): the series inside (-1, 1); outside, the 1/(1-z)
 * transformation (Abramowitz & Stegun 15.3.8) with two series in 1/(1-z). */
double kpo_hyp2f1(double a, double b, double c, double z) {
    if (z > -1.0 && z < 1.0) return nt_hyp2f1_series(a, b, c, z);
    const double w = 1.0 - z;
    const double zp = 1.0 / w;
    const double gc = tgamma(c);
    const double t1 = pow(w, -a) * gc * tgamma(b - a) / nt_nonzero(tgamma(b) * tgamma(c - a)) *
                      nt_hyp2f1_series(a, c - b, a - b + 1.0, zp);
    const double t2 = pow(w, -b) * gc * tgamma(a - b) / nt_nonzero(tgamma(a) * tgamma(c - b)) *
                      nt_hyp2f1_series(b, c - a, b - a + 1.0, zp);
    return t1 + t2;
}

/* The common shape of the kappa fits (:238-250 and their siblings):
 * pref * N_low * X^low_power * (1 + X^(x * high_power) * (N_low / N_high)^x)^(-1/x). */
static double nt_blend(double pref, double n_low, double n_high, double X, double low_power,
                       double high_power, double x) {
    return pref * n_low * pow(X, low_power) * pow(1.0 + pow(X, x * high_power) * pow(n_low / n_high, x), -1.0 / x);
}

typedef struct nt_local {
    double nu, ne, b, sin_raw, cos_theta;  /* state as given (nu unclamped) */
    double nuc, sint, w, X;                /* derived */
} nt_local;

/* nu_c = e B / (2 pi m_e c) (:101-108). */
static double nt_nu_c(double b) { return NT_EE * b / (2.0 * PI_D * NT_ME * CL); }

/* ---- power law (:112-215) ---- */
static double pl_density(double b, const kpo_nonthermal* p) {
    if (!(p->power_law_p > 2.0) || !(p->power_law_gamma_min >= 1.0)) return 0.0;
    const double u_nth = p->power_law_eta * b * b / 2.0;
    return u_nth * (p->power_law_p - 2.0) / (p->power_law_p - 1.0) /
           maxv(NT_ME * CL * CL * p->power_law_gamma_min, TINY);
}

static double pl_gamma_span(const kpo_nonthermal* p) {
    return pow(p->power_law_gamma_min, 1.0 - p->power_law_p) - pow(p->power_law_gamma_max, 1.0 - p->power_law_p);
}

static double pl_j_i(const nt_local* s, const kpo_nonthermal* p, double n) {
    const double pp = p->power_law_p;
    if (!(s->nu > 0.0) || !(s->nuc > 0.0) || !(n > 0.0) || !(pp > 1.0)) return 0.0;
    const double pref = n * NT_EE * NT_EE * s->nuc / CL;
    const double t1 = pow(3.0, pp / 2.0) * (pp - 1.0) * s->sint;
    const double t2 = 2.0 * (pp + 1.0) * pl_gamma_span(p);
    const double t3 = tgamma((3.0 * pp - 1.0) / 12.0) * tgamma((3.0 * pp + 19.0) / 12.0);
    const double t4 = pow(s->nu / maxv(s->nuc * s->sint, TINY), -(pp - 1.0) / 2.0);
    return pref * t1 / maxv(t2, TINY) * t3 * t4;
}

static double pl_j_v(const nt_local* s, const kpo_nonthermal* p, double jI) {
    double cost = s->cos_theta;
    if (absv(cost) < 1e-12) cost = cost >= 0.0 ? 1e-12 : -1e-12;
    const double t1 = -(171.0 / 250.0) * pow(p->power_law_p, 49.0 / 100.0);
    const double t2 = cost / s->sint * pow(s->nu / maxv(3.0 * s->nuc * s->sint, TINY), -0.5);
    return -t1 * t2 * jI;
}

static double pl_a_i(const nt_local* s, const kpo_nonthermal* p, double n) {
    const double pp = p->power_law_p;
    if (!(s->nu > 0.0) || !(s->nuc > 0.0) || !(n > 0.0) || !(pp > 1.0)) return 0.0;
    const double pref = n * NT_EE * NT_EE / maxv(s->nu * NT_ME * CL, TINY);
    const double t1 = pow(3.0, (pp + 1.0) / 2.0) * (pp - 1.0);
    const double t2 = 4.0 * pl_gamma_span(p);
    const double t3 = tgamma((3.0 * pp + 2.0) / 12.0) * tgamma((3.0 * pp + 22.0) / 12.0);
    const double t4 = pow(s->nu / maxv(s->nuc * s->sint, TINY), -(pp + 2.0) / 2.0);
    return pref * t1 / maxv(t2, TINY) * t3 * t4;
}

static double pl_a_v(const nt_local* s, const kpo_nonthermal* p, double aI) {
    const double t5 = -pow((71.0 / 100.0) * p->power_law_p + 22.0 / 625.0, 197.0 / 500.0);
    const double t6 = pow(maxv((31.0 / 10.0) * pow(s->sint, -48.0 / 25.0) - 31.0 / 10.0, 0.0), 64.0 / 125.0);
    const double t7 = pow(s->nu / maxv(s->nuc * s->sint, TINY), -0.5);
    const double sign = s->cos_theta >= 0.0 ? 1.0 : -1.0;
    return -aI * t5 * t6 * t7 * sign;
}

/* ---- kappa (:217-489) ---- */
/* Emission normalisations of the I fit (:226-231): low- and high-frequency limits. */
static double kappa_jn_low(double k) {
    return 4.0 * PI_D * tgamma(k - 4.0 / 3.0) / (pow(3.0, 7.0 / 3.0) * tgamma(k - 2.0));
}
static double kappa_jn_high(double k) {
    return 0.25 * pow(3.0, (k - 1.0) / 2.0) * (k - 2.0) * (k - 1.0) * tgamma(k / 4.0 - 1.0 / 3.0) *
           tgamma(k / 4.0 + 4.0 / 3.0);
}
/* Absorption normalisations of the I fit (:327-336), without hyp2f1 in the low one. */
static double kappa_an_low(double k, double w) {
    return pow(3.0, 1.0 / 6.0) * 10.0 / 41.0 * pow(2.0 * PI_D, 2.0) / pow(w * k, 16.0 / 3.0 - k) *
           (k - 2.0) * (k - 1.0) * k / (3.0 * k - 1.0) * tgamma(5.0 / 3.0);
}
static double kappa_an_high(double k, double w) {
    return 2.0 * pow(PI_D, 2.5) / 3.0 * (k - 2.0) * (k - 1.0) * k / pow(w * k, 5.0) *
           (2.0 * tgamma(2.0 + k / 2.0) / (2.0 + k) - 1.0);
}

static double kappa_j(const nt_local* s, double k, int stokes) {
    if (!(k > 2.0) || !(s->w > 0.0) || !(s->X > 0.0)) return 0.0;
    const double pref = s->ne * NT_EE * NT_EE * s->nuc * s->sint / CL;
    const double lo = kappa_jn_low(k), hi = kappa_jn_high(k);
    const double high_power = (3.0 * k - 4.0) / 6.0;
    if (stokes == 0) return nt_blend(pref, lo, hi, s->X, 1.0 / 3.0, high_power, 3.0 * pow(k, -1.5));
    if (stokes == 1)
        return nt_blend(pref, -0.5 * lo, -(pow(4.0 / 5.0, 2.0) + k / 50.0) * hi, s->X, 1.0 / 3.0, high_power,
                        3.7 * pow(k, -1.6));
    const double angle1 = maxv(pow(s->sint, -12.0 / 5.0) - 1.0, 0.0);
    const double angle2 = maxv(pow(s->sint, -5.0 / 2.0) - 1.0, 0.0);
    const double n_low = -pow(3.0 / 4.0, 2.0) * pow(angle1, 12.0 / 25.0) * pow(k, -66.0 / 125.0) / s->w *
                         pow(s->X, -7.0 / 20.0) * lo;
    const double n_high = -pow(7.0 / 8.0, 2.0) * pow(angle2, 11.0 / 25.0) * pow(k, -11.0 / 25.0) / s->w *
                          pow(s->X, -0.5) * hi;
    const double v = nt_blend(pref, n_low, n_high, s->X, 1.0 / 3.0, high_power, 3.0 * pow(k, -1.5));
    return -v * (s->cos_theta >= 0.0 ? 1.0 : -1.0);
}

static double kappa_a(const nt_local* s, double k, int stokes) {
    const double pref = s->ne * NT_EE / maxv(s->b * s->sint, TINY);
    const double hyp = kpo_hyp2f1(k - 1.0 / 3.0, k + 1.0, k + 2.0 / 3.0, -k * s->w);
    if (!(k > 2.0) || !(s->w > 0.0) || !(s->X > 0.0) || !(absv(hyp) > TINY)) return 0.0;
    const double lo = kappa_an_low(k, s->w) * hyp, hi = kappa_an_high(k, s->w);
    const double high_power = (3.0 * k - 1.0) / 6.0;
    if (stokes == 0) {
        const double x = pow(-7.0 / 4.0 + 8.0 * k / 5.0, -43.0 / 50.0);
        return nt_blend(pref, lo, hi * (pow(3.0 / k, 19.0 / 4.0) + 3.0 / 5.0) + TINY, s->X, -5.0 / 3.0,
                        high_power, x);
    }
    if (stokes == 1) {
        const double x = 7.0 / 5.0 * pow(k, -23.0 / 20.0);
        return nt_blend(pref, -25.0 / 48.0 * lo, -(pow(21.0, 2.0) * pow(k, -144.0 / 25.0) + 11.0 / 20.0) * hi + TINY,
                        s->X, -5.0 / 3.0, high_power, x);
    }
    const double angle1 = maxv(pow(s->sint, -114.0 / 50.0) - 1.0, 0.0);
    const double angle2 = maxv(pow(s->sint, -41.0 / 20.0) - 1.0, 0.0);
    const double n_low = -77.0 / (100.0 * s->w) * pow(angle1, 223.0 / 500.0) * pow(s->X, -7.0 / 20.0) *
                         pow(k, -7.0 / 10.0) * lo;
    const double nfac = 13.0 * 13.0 * pow(k, -8.0) + 13.0 / 2500.0 * k - 263.0 / 5000.0 + 47.0 / (200.0 * k);
    const double n_high = -143.0 / 10.0 * pow(s->w, -116.0 / 125.0) * pow(angle2, 0.5) * nfac * pow(s->X, -0.5) *
                          hi + TINY;
    const double x = 61.0 / 50.0 * pow(k, -142.0 / 125.0) + 7.0 / 1000.0;
    const double v = nt_blend(pref, n_low, n_high, s->X, -5.0 / 3.0, high_power, x);
    return -v * (s->cos_theta >= 0.0 ? 1.0 : -1.0);
}

/* kappa_rho_Q_at / kappa_rho_V_at (:393-460): the fast rho fits at the
 * anchors kappa = 3.5, 4, 4.5, 5 (Marszewski et al. 2021, table 2). */
static double kappa_rho_q_at(const nt_local* s, int anchor) {
    if (!(s->X > 0.0)) return 0.0;
    const double w = s->w, X = s->X, sw = sqrt(w);
    const double pref = -(s->ne * NT_EE * NT_EE * s->nuc * s->nuc * s->sint * s->sint) /
                        maxv(NT_ME * CL * s->nu * s->nu * s->nu, TINY);
    static const double wa[4] = {17.0, 46.0 / 3.0, 14.0, 25.0 / 2.0};
    static const double wb[4] = {3.0, 5.0 / 3.0, 13.0 / 8.0, 1.0};
    static const double wc[4] = {7.0, 17.0 / 3.0, 9.0 / 2.0, 5.0};
    static const double fd[4] = {30.0, 18.0, 12.0, 8.0};
    const double wterm = wa[anchor] * w - wb[anchor] * sw + wc[anchor] * sw * exp(-5.0 * w);
    double wave;
    switch (anchor) {
        case 0: wave = sin(X / 10.0) * exp(-1.5 * pow(X, 0.471)); break;
        case 1: wave = sin(X / 6.0) * exp(-7.0 * sqrt(X) / 4.0); break;
        case 2: wave = sin(X / 4.0) * exp(-2.0 * pow(X, 0.525)); break;
        default: wave = sin(3.0 * X / 8.0) * exp(-9.0 * pow(X, 0.541) / 4.0); break;
    }
    const double fx = 1.0 - exp(-pow(X, 0.84) / fd[anchor]) - wave;
    return pref * fx * wterm;
}

static double kappa_rho_v_at(const nt_local* s, int anchor) {
    if (!(s->X > 0.0) || !(s->w > 0.0)) return 0.0;
    const double w = s->w;
    const double pref = 2.0 * s->ne * NT_EE * NT_EE * s->nuc * s->cos_theta / maxv(NT_ME * CL * s->nu * s->nu, TINY);
    double k0 = 0.0, k1 = 0.0, k2 = 0.0;
    kpo_bessel_k012(1.0 / w, &k0, &k1, &k2);
    const double bessel = k0 / maxv(k2 + TINY, TINY);
    static const double gc[4] = {0.447, 0.391, 0.348, 0.313};
    double wterm;
    switch (anchor) {
        case 0: wterm = (w * w + 2.0 * w + 1.0) / (25.0 / 8.0 * w * w + 4.0 * w + 1.0); break;
        case 1: wterm = (w * w + 54.0 * w + 50.0) / (30.0 / 11.0 * w * w + 134.0 * w + 50.0); break;
        case 2: wterm = (w * w + 43.0 * w + 38.0) / (7.0 / 3.0 * w * w + 185.0 / 2.0 * w + 38.0); break;
        default: wterm = (w + 13.0 / 14.0) / (2.0 * w + 13.0 / 14.0); break;
    }
    const double gx = 1.0 - 0.17 * log(1.0 + gc[anchor] * pow(s->X, -0.5));
    return pref * bessel * gx * wterm;
}

/* interpolate_kappa_rho_{Q,V} (:462-489): linear in kappa between the
 * anchors, zero outside [3.5, 5] where the fast fit is not defined. */
static double kappa_rho(const nt_local* s, double k, int stokes) {
    if (k < 3.5 || k > 5.0) return 0.0;
    double (*at)(const nt_local*, int) = stokes == 1 ? kappa_rho_q_at : kappa_rho_v_at;
    if (k >= 5.0) return at(s, 3);
    const int seg = k < 4.0 ? 0 : (k < 4.5 ? 1 : 2);
    const double k_lo = 3.5 + 0.5 * seg, k_hi = k_lo + 0.5;
    return ((k_hi - k) * at(s, seg) + (k - k_lo) * at(s, seg + 1)) / 0.5;
}

/* polarised-fraction cap of emission or absorption (:540-545, :557-562, :595-606) */
static void nt_cap(double i, double* q, double* v, double frac) {
    const double pol = sqrt(*q * *q + *v * *v);
    if (i > 0.0 && pol > 0.0 && i < pol / maxv(frac, 1e-30)) {
        const double s = i / pol * frac;
        *q *= s;
        *v *= s;
    }
}

/* variable_kappa_from_sigma_beta (:493-502). */
double kpo_variable_kappa(double sigma, double beta, double kappa_default) {
    if (!(sigma > 0.0) || !(beta >= 0.0)) return kappa_default;
    const double k = 2.8 + 0.7 / sqrt(sigma) + 3.7 * pow(sigma, -0.19) * tanh(23.4 * pow(sigma, 0.26) * beta);
    return k < 3.1 ? 3.1 : k;
}

/* nonthermal_synchrotron_magnetic_basis_coefficients (:504-618): invariant
 * coefficients (j / nu^2, alpha nu, rho nu) in the magnetic screen basis. */
void kpo_nonthermal_coefficients(const kpo_nonthermal* p, double nu_in, double ne, double thetae, double b,
                                 double sin_theta, double cos_theta, kpo_coeffs* c) {
    memset(c, 0, sizeof(*c));
    if (!(nu_in > 0.0) || !(b > 0.0) || !(ne > 0.0)) return;
    const double nu = maxv(nu_in, 1.0);
    const double inu2 = 1.0 / maxv(nu * nu, TINY);
    nt_local s;
    s.nu = nu_in;
    s.ne = ne;
    s.b = b;
    s.sin_raw = sin_theta;
    s.cos_theta = cos_theta;
    s.nuc = nt_nu_c(b);
    s.sint = maxv(sin_theta, p->min_sin_theta);
    if (p->distribution == KP_EMISSION_POWER_LAW) {
        const double n = pl_density(b, p);
        const double jI = pl_j_i(&s, p, n);
        const double jQ_phys = -(p->power_law_p + 1.0) / (p->power_law_p + 7.0 / 3.0) * jI;
        double ji = jI * inu2;
        double jq = -jQ_phys * inu2;
        double jv = pl_j_v(&s, p, jI) * inu2;
        nt_cap(ji, &jq, &jv, p->max_pol_frac_emission);
        const double aI = pl_a_i(&s, p, n);
        const double qfac = -pow((17.0 / 500.0) * p->power_law_p - 43.0 / 1250.0, 43.0 / 500.0);
        double ai = aI * nu;
        double aq = -(aI * qfac) * nu;
        double av = pl_a_v(&s, p, aI) * nu;
        nt_cap(ai, &aq, &av, p->max_pol_frac_absorption);
        c->jI = p->emission_scale * ji;
        c->jQ = p->emission_scale * jq;
        c->jV = p->emission_scale * jv;
        c->aI = p->absorption_scale * ai;
        c->aQ = p->absorption_scale * aq;
        c->aV = p->absorption_scale * av;
        return;  /* the fast fits carry no power-law rotativity */
    }
    if (p->distribution != KP_EMISSION_KAPPA) return;
    const double k = p->kappa;
    /* kappa_width <= 0 selects the width that keeps the mean energy of the
     * thermal distribution at Theta_e (:522-524, :217-222) */
    s.w = p->kappa_width > 0.0 ? p->kappa_width : (k - 3.0) / maxv(k, TINY) * thetae;
    s.X = nu_in / maxv(s.w * s.w * k * k * s.nuc * s.sint, TINY);
    kpo_coeffs kc;
    kc.jI = kappa_j(&s, k, 0) * inu2;
    kc.jQ = -kappa_j(&s, k, 1) * inu2;
    kc.jV = kappa_j(&s, k, 3) * inu2;
    kc.aI = kappa_a(&s, k, 0) * nu;
    kc.aQ = -kappa_a(&s, k, 1) * nu;
    kc.aV = kappa_a(&s, k, 3) * nu;
    kc.rQ = kappa_rho(&s, k, 1) * nu;
    kc.rV = kappa_rho(&s, k, 3) * nu;
    double mix = 0.0;
    if (k >= p->kappa_interp_begin && k < p->kappa_interp_end)
        mix = (k - p->kappa_interp_begin) / maxv(p->kappa_interp_end - p->kappa_interp_begin, TINY);
    else if (k >= p->kappa_interp_end)
        mix = 1.0;
    /* the thermal Pandya fit it blends toward, with unit scales (:525-540) */
    kpo_coeffs th;
    kpo_thermal_coefficients(KP_EMISSION_THERMAL_PANDYA, nu_in, ne, thetae, b, 0.0, sin_theta, cos_theta,
                             p->max_pol_frac_emission, p->min_sin_theta, 1.0, 1.0, 1.0, &th);
    double ji = (1.0 - mix) * kc.jI + mix * th.jI;
    double jq = (1.0 - mix) * kc.jQ + mix * th.jQ;
    double jv = (1.0 - mix) * kc.jV + mix * th.jV;
    double ai = (1.0 - mix) * kc.aI + mix * th.aI;
    double aq = (1.0 - mix) * kc.aQ + mix * th.aQ;
    double av = (1.0 - mix) * kc.aV + mix * th.aV;
    nt_cap(ji, &jq, &jv, p->max_pol_frac_emission);
    nt_cap(ai, &aq, &av, p->max_pol_frac_absorption);
    c->jI = p->emission_scale * ji;
    c->jQ = p->emission_scale * jq;
    c->jV = p->emission_scale * jv;
    c->aI = p->absorption_scale * ai;
    c->aQ = p->absorption_scale * aq;
    c->aV = p->absorption_scale * av;
    c->rQ = p->faraday_scale * kc.rQ;
    c->rV = p->faraday_scale * kc.rV;
}

/* ======================================================================
 * RIAF model (src/model/riaf.hpp), BL metric branches
 * ==================================================================== */
/* density_profile_bl (riaf.hpp:54-60). */
static double riaf_density(const kp_riaf_model* m, double r, double th) {
    const double zc = r * cos(th);
    const double rc = maxv(r * sin(th), m->min_radius);
    const double h = maxv(m->disk_h, 1e-6);
    const double vertical = exp(-(zc * zc) / (2.0 * rc * rc * h * h));
    return m->nth0 * vertical * pow(maxv(r, m->min_radius), m->pow_nth);
}

/* thetae_profile (riaf.hpp:67-74). */
static double riaf_thetae(const kp_riaf_model* m, double r) {
    const double kbol = 1.380649e-16;
    const double thetae_unit = m->te_unit * kbol / (ME * CL * CL);
    return maxv(m->Te0 * pow(maxv(r, m->min_radius), m->pow_T) * thetae_unit, m->min_thetae);
}

/* length_unit_cgs / dlambda_scale (riaf.hpp:76-85). */
static double riaf_dlambda_scale(const kp_riaf_model* m) {
    const double l = 6.67430e-8 * m->mbh_solar * 1.98847e33 / (CL * CL);
    return l / maxv(m->freq_cgs, 1.0);
}

/* magnetic_field_cgs (riaf.hpp:87-96). */
static double riaf_bfield(const kp_riaf_model* m, double ne_cgs, double r) {
    const double b = sqrt(maxv(8.0 * PI_D * 0.1 * ne_cgs * MP * CL * CL /
                                   (6.0 * maxv(r, m->min_radius)),
                               0.0));
    return maxv(b, 1e-12);
}

/* bl_metric (riaf.hpp:137-166): the model's own unit-mass BL metric. */
static void riaf_bl_metric(double r, double th, double spin, double* gcov, double* gcon) {
    const double a = spin, a2 = a * a;
    const double sinth = sin(th);
    const double sin2 = maxv(sinth * sinth, 1e-30);
    const double costh = cos(th);
    const double sigma = r * r + a2 * costh * costh;
    const double delta = r * r - 2.0 * r + a2;
    for (int i = 0; i < 16; ++i) gcov[i] = gcon[i] = 0.0;
    gcov[0] = -(1.0 - 2.0 * r / sigma);
    gcov[3] = gcov[12] = -2.0 * a * r * sin2 / sigma;
    gcov[5] = sigma / maxv(delta, TINY);
    gcov[10] = sigma;
    gcov[15] = (r * r + a2 + 2.0 * a2 * r * sin2 / sigma) * sin2;
    const double isd = 1.0 / maxv(sigma * delta, TINY);
    gcon[0] = -((r * r + a2) * (r * r + a2) - a2 * delta * sin2) * isd;
    gcon[3] = gcon[12] = -2.0 * a * r * isd;
    gcon[5] = delta / sigma;
    gcon[10] = 1.0 / sigma;
    gcon[15] = (delta - a2 * sin2) / maxv(sigma * delta * sin2, TINY);
}

/* isco_radius (riaf.hpp:168-176). */
double kpo_riaf_isco_radius(double spin) {
    const double a = spin;
    const double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    const double z2 = sqrt(3.0 * a * a + z1 * z1);
    const double sign = a < 0.0 ? -1.0 : 1.0;
    return 3.0 + z2 - sign * sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

static double signed_nonzero(double x) {
    if (absv(x) > TINY) return x;
    return x < 0.0 ? -TINY : TINY;
}

/* fluid_four_velocity_from_bl_coords, BL branch (riaf.hpp:210-273). */
static void riaf_fluid_u(const kp_riaf_model* m, double mass, double spin, double r, double th,
                         const double* metric_gcov, double* u) {
    const double rh = mass + sqrt(maxv(mass * mass - spin * spin, 0.0));
    const double risco = kpo_riaf_isco_radius(spin);
    double gcov[16], gcon[16];
    riaf_bl_metric(r, th, spin, gcov, gcon);
    double vt = 1.0, vr = 0.0, vth = 0.0, vphi = 0.0;
    if (r < rh) {
        vt = 1.0 / sqrt(maxv(-gcon[0], TINY));
    } else if (r < risco) {
        double gcov_i[16], gcon_i[16];
        riaf_bl_metric(risco, th, spin, gcov_i, gcon_i);
        const double omega_isco = 1.0 / (risco * sqrt(risco) + spin);
        const double qi = gcov_i[0] + 2.0 * omega_isco * gcov_i[3] +
                          omega_isco * omega_isco * gcov_i[15];
        const double ut_i = 1.0 / sqrt(maxv(-qi, TINY));
        const double e = (gcov_i[0] + gcov_i[3] * omega_isco) * ut_i;
        const double l = (gcov_i[12] + gcov_i[15] * omega_isco) * ut_i;
        const double k_con = gcon[0] * e * e + 2.0 * gcon[3] * e * l + gcon[15] * l * l;
        const double ur_cov = -sqrt(maxv(-(1.0 + k_con) / maxv(gcon[5], TINY), 0.0));
        const double vr_tmp = gcon[5] * ur_cov;
        const double vt_tmp = gcon[0] * e + gcon[3] * l;
        const double vphi_tmp = gcon[12] * e + gcon[15] * l;
        const double omega_k = vphi_tmp / maxv(vt_tmp, TINY);
        const double omega_ff = gcon[3] / signed_nonzero(gcon[0]);
        const double omega = omega_k + (1.0 - m->keplerian_factor) * (omega_ff - omega_k);
        const double ur_ff = -sqrt(maxv(-(1.0 + gcon[0]) * gcon[5], 0.0));
        vr = vr_tmp + m->infall_factor * (ur_ff - vr_tmp);
        const double q = gcov[0] + 2.0 * omega * gcov[3] + omega * omega * gcov[15];
        vt = sqrt(maxv(-(1.0 + vr * vr * gcov[5]) / q, 0.0));
        vphi = omega * vt;
    } else {
        const double r_eff = maxv(r, m->min_radius);
        const double omega_k = 1.0 / (r_eff * sqrt(r_eff) + spin);
        const double omega_ff = gcon[3] / signed_nonzero(gcon[0]);
        const double omega = omega_k + (1.0 - m->keplerian_factor) * (omega_ff - omega_k);
        vr = m->infall_factor * -sqrt(maxv(-(1.0 + gcon[0]) * gcon[5], 0.0));
        const double q = gcov[0] + 2.0 * omega * gcov[3] + omega * omega * gcov[15];
        vt = sqrt(maxv(-(1.0 + vr * vr * gcov[5]) / q, 0.0));
        vphi = omega * vt;
    }
    u[0] = vt;
    u[1] = vr;
    u[2] = vth;
    u[3] = vphi;
    const double norm2 = dot_g(metric_gcov, u, u);
    if (norm2 < 0.0) {
        const double f = 1.0 / sqrt(maxv(-norm2, TINY));
        for (int i = 0; i < 4; ++i) u[i] = u[i] * f;
    }
}

void kpo_riaf_fluid_velocity(const kp_pass_a_params* p, const kp_riaf_model* m,
                             const double x[4], double u[4]) {
    double g[16];
    kpo_bl_gcov(p->mass, p->spin, x, g);
    riaf_fluid_u(m, p->mass, p->spin, maxv(x[1], m->min_radius), x[2], g, u);
}

static void lower_g(const double* g, const double* v, double* out) {
    for (int mu = 0; mu < 4; ++mu) {
        double s = 0.0;
        for (int nu = 0; nu < 4; ++nu) s += g[mu * 4 + nu] * v[nu];
        out[mu] = s;
    }
}

/* magnetic_unit_four_vector_with_gcov, BL branch (riaf.hpp:350-386). */
static void riaf_b_unit(const double* gcov, const double* ucon, double* bcon) {
    double btrial[4] = {0.0, 0.0, 0.0, 1.0};
    double udotb = dot_g(gcov, ucon, btrial);
    for (int i = 0; i < 4; ++i) bcon[i] = btrial[i] + ucon[i] * udotb;
    double bsq = dot_g(gcov, bcon, bcon);
    if (!(bsq > 1e-30)) {
        const double bt[4] = {0.0, 0.0, 0.0, 1.0};
        const double udotz = dot_g(gcov, ucon, bt);
        for (int i = 0; i < 4; ++i) bcon[i] = bt[i] + ucon[i] * udotz;
        bsq = dot_g(gcov, bcon, bcon);
    }
    if (!(bsq > 1e-30)) {
        const double bt[4] = {0.0, 1.0, 0.0, 0.0};
        const double udotx = dot_g(gcov, ucon, bt);
        for (int i = 0; i < 4; ++i) bcon[i] = bt[i] + ucon[i] * udotx;
        bsq = dot_g(gcov, bcon, bcon);
    }
    if (bsq > 1e-30) {
        const double f = 1.0 / sqrt(bsq);
        for (int i = 0; i < 4; ++i) bcon[i] = bcon[i] * f;
    }
}

/* screen_inner_product_with_gcov (riaf.hpp:320-336). */
static double screen_inner(const double* g, const double* u, const double* k, const double* a,
                           const double* b) {
    const double uk = dot_g(g, u, k);
    if (absv(uk) <= TINY) return dot_g(g, a, b);
    double eK[4];
    for (int i = 0; i < 4; ++i) eK[i] = k[i] * (-1.0 / uk) - u[i];
    return dot_g(g, a, b) + dot_g(g, u, a) * dot_g(g, u, b) - dot_g(g, eK, a) * dot_g(g, eK, b);
}

/* PlasmaPerturbation::apply (plasma_perturbation.hpp:17-32). */
static void apply_perturbation(const kpo_perturbation* p, double* ne, double* thetae, double* b) {
    if (!p || p->log_scale == 0.0) return;
    const double f = exp(p->log_scale);
    if (p->parameter == 1) {
        *ne *= f;
        *b *= exp(p->log_scale * 0.5);
    } else if (p->parameter == 2) {
        *thetae *= f;
    } else if (p->parameter == 3) {
        *b *= f;
    }
}

/* RIAFAnalyticRadiationModel::coefficients, BL metric, compiled Pandya fit by
 * default and Dexter when emission_type == 4 (riaf.hpp:410-549). */
void kpo_riaf_coefficients(const kp_pass_a_params* p, const kp_riaf_model* m,
                           const kpo_state* s, const kpo_perturbation* pert, kpo_coeffs* c) {
    memset(c, 0, sizeof(*c));
    const double* x = s->v;
    const double* k = s->v + 4;
    const double r_coord = x[1], th = x[2];
    const double r_bl = maxv(r_coord, m->min_radius);
    const double rh = p->mass + sqrt(maxv(p->mass * p->mass - p->spin * p->spin, 0.0));
    if (r_bl <= rh + m->horizon_buffer || r_bl < m->r_min || r_bl > m->r_max) return;
    const double ne_norm = riaf_density(m, r_bl, th);
    if (!(ne_norm > 0.0)) return;
    double ne = ne_norm * m->ne_unit;
    double thetae = riaf_thetae(m, r_bl);
    double b = riaf_bfield(m, ne, r_bl);
    apply_perturbation(pert, &ne, &thetae, &b);

    double g[16];
    kpo_bl_gcov(p->mass, p->spin, x, g);
    double ucon[4], ucov[4];
    riaf_fluid_u(m, p->mass, p->spin, r_bl, th, g, ucon);
    lower_g(g, ucon, ucov);
    double uk = 0.0;
    for (int mu = 0; mu < 4; ++mu) uk += ucov[mu] * k[mu];
    const double nu_scale = maxv(absv(-uk), m->min_frequency_scale);

    double bcon[4], bcov[4];
    riaf_b_unit(g, ucon, bcon);
    lower_g(g, bcon, bcov);
    double kdotb = 0.0;
    for (int mu = 0; mu < 4; ++mu) kdotb += k[mu] * bcov[mu];
    const double cos_theta = clampv(kdotb / maxv(nu_scale, 1e-30), -1.0, 1.0);
    const double sin_theta = maxv(sqrt(maxv(0.0, 1.0 - cos_theta * cos_theta)), m->min_sin_theta);
    const double theta = acos(clampv(cos_theta, -1.0, 1.0));
    const double nu = maxv(m->freq_cgs * nu_scale, 1.0);

    if (theta <= 0.0 || theta >= PI_D) {
        c->rV = m->faraday_scale * kpo_symphony_rho_v(nu, ne, thetae, b, theta, m->min_sin_theta) * nu;
        return;
    }
    kpo_coeffs mc;
    if (m->emission_type == KP_EMISSION_KAPPA || m->emission_type == KP_EMISSION_POWER_LAW) {
        /* riaf.hpp:497-522: fixed kappa, the width from Theta_e */
        kpo_nonthermal nt;
        memset(&nt, 0, sizeof(nt));
        nt.distribution = m->emission_type;
        nt.max_pol_frac_emission = nt.max_pol_frac_absorption = m->max_pol_frac;
        nt.min_sin_theta = m->min_sin_theta;
        nt.emission_scale = m->emission_scale;
        nt.absorption_scale = m->absorption_scale;
        nt.faraday_scale = m->faraday_scale;
        nt.kappa = m->nonthermal_kappa;
        nt.kappa_width = -1.0;
        nt.kappa_interp_begin = m->variable_kappa_interp_start;
        nt.kappa_interp_end = m->variable_kappa_max;
        nt.power_law_p = m->powerlaw_p;
        nt.power_law_eta = m->powerlaw_eta;
        nt.power_law_gamma_min = m->powerlaw_gamma_min;
        nt.power_law_gamma_max = m->powerlaw_gamma_max;
        nt.power_law_gamma_cutoff = m->powerlaw_gamma_cutoff;
        kpo_nonthermal_coefficients(&nt, nu, ne, thetae, b, sin_theta, cos_theta, &mc);
    } else {
        const int fit = m->emission_type == KP_EMISSION_THERMAL_DEXTER ? KP_EMISSION_THERMAL_DEXTER
                                                                        : KP_EMISSION_THERMAL_PANDYA;
        kpo_thermal_coefficients(fit, nu, ne, thetae, b, theta, sin_theta, cos_theta, m->max_pol_frac,
                                 m->min_sin_theta, m->emission_scale, m->absorption_scale,
                                 m->faraday_scale, &mc);
    }
    const double b1 = screen_inner(g, ucon, k, s->v + 8, bcon);
    const double b2 = screen_inner(g, ucon, k, s->v + 12, bcon);
    const double bproj2 = b1 * b1 + b2 * b2;
    c->jI = mc.jI;
    c->jV = mc.jV;
    c->aI = mc.aI;
    c->aV = mc.aV;
    c->rV = mc.rV;
    if (bproj2 > 1e-30) {
        const double cos2chi = (b2 * b2 - b1 * b1) / bproj2;
        const double sin2chi = -2.0 * b1 * b2 / bproj2;
        c->jQ = mc.jQ * cos2chi;
        c->jU = mc.jQ * sin2chi;
        c->aQ = mc.aQ * cos2chi;
        c->aU = mc.aQ * sin2chi;
        c->rQ = mc.rQ * cos2chi;
        c->rU = mc.rQ * sin2chi;
    }
}

/* ======================================================================
 * Transfer (src/radiation/semi_analytic.hpp, src/common/math.hpp:62-86)
 * ==================================================================== */
static double stable_attenuation(double kappa, double dl) {
    const double x = kappa * dl;
    if (x > 700.0) return 0.0;
    return exp(-x);
}

static double stable_source_factor(double kappa, double dl) {
    const double x = kappa * dl;
    if (absv(x) < 1e-6) return dl * (1.0 - x / 2.0 + x * x / 6.0 - x * x * x / 24.0);
    return (1.0 - exp(-x)) / kappa;
}

/* Scaled Euclidean norm (src/common/vec.hpp:63-71). */
static double norm3(double x, double y, double z) {
    const double scale = maxv(absv(x), maxv(absv(y), absv(z)));
    if (scale == 0.0) return 0.0;
    const double a = x / scale, b = y / scale, c = z / scale;
    return scale * sqrt(a * a + b * b + c * c);
}

/* faraday_rotate (semi_analytic.hpp:9-37). */
static void faraday_rotate(kpo_stokes* s, const kpo_coeffs* c, double dl) {
    const double px = s->Q, py = s->U, pz = s->V;
    const double rx = c->rQ, ry = c->rU, rz = c->rV;
    const double rn = norm3(rx, ry, rz);
    if (rn == 0.0 || dl == 0.0) return;
    const double theta = rn * dl;
    double ox, oy, oz;
    if (absv(theta) < 1e-8) {
        ox = px + (ry * pz - rz * py) * dl;
        oy = py + (rz * px - rx * pz) * dl;
        oz = pz + (rx * py - ry * px) * dl;
    } else {
        const double inv = 1.0 / rn;
        const double nx = rx * inv, ny = ry * inv, nz = rz * inv;
        const double ct = cos(theta), st = sin(theta);
        const double nd = (nx * px + ny * py + nz * pz) * (1.0 - ct);
        ox = px * ct + nx * nd + (ny * pz - nz * py) * st;
        oy = py * ct + ny * nd + (nz * px - nx * pz) * st;
        oz = pz * ct + nz * nd + (nx * py - ny * px) * st;
    }
    s->Q = ox;
    s->U = oy;
    s->V = oz;
}

/* absorb_emit (semi_analytic.hpp:39-91). */
static void absorb_emit(kpo_stokes* s, const kpo_coeffs* c, double dl) {
    const double px = s->Q, py = s->U, pz = s->V;
    const double an = norm3(c->aQ, c->aU, c->aV);
    if (an < TINY) {
        const double e = stable_attenuation(c->aI, dl);
        const double f = stable_source_factor(c->aI, dl);
        s->I = e * s->I + f * c->jI;
        s->Q = e * s->Q + f * c->jQ;
        s->U = e * s->U + f * c->jU;
        s->V = e * s->V + f * c->jV;
        return;
    }
    const double inv = 1.0 / an;
    const double nx = c->aQ * inv, ny = c->aU * inv, nz = c->aV * inv;
    const double ppar = nx * px + ny * py + nz * pz;
    const double pperp_x = px - nx * ppar, pperp_y = py - ny * ppar, pperp_z = pz - nz * ppar;
    const double jpar = nx * c->jQ + ny * c->jU + nz * c->jV;
    const double jperp_x = c->jQ - nx * jpar, jperp_y = c->jU - ny * jpar, jperp_z = c->jV - nz * jpar;
    const double s_plus = s->I + ppar, s_minus = s->I - ppar;
    const double j_plus = c->jI + jpar, j_minus = c->jI - jpar;
    const double k_plus = c->aI + an, k_minus = c->aI - an;
    const double sp = stable_attenuation(k_plus, dl) * s_plus + stable_source_factor(k_plus, dl) * j_plus;
    const double sm = stable_attenuation(k_minus, dl) * s_minus + stable_source_factor(k_minus, dl) * j_minus;
    const double ei = stable_attenuation(c->aI, dl);
    const double fi = stable_source_factor(c->aI, dl);
    const double i_out = 0.5 * (sp + sm);
    const double ppar_out = 0.5 * (sp - sm);
    s->I = i_out;
    s->Q = nx * ppar_out + (pperp_x * ei + jperp_x * fi);
    s->U = ny * ppar_out + (pperp_y * ei + jperp_y * fi);
    s->V = nz * ppar_out + (pperp_z * ei + jperp_z * fi);
}

/* semi_analytic_stokes_step (semi_analytic.hpp:93-100). */
void kpo_semi_analytic_step(kpo_stokes* s, const kpo_coeffs* c, double dl) {
    faraday_rotate(s, c, 0.5 * dl);
    absorb_emit(s, c, dl);
    faraday_rotate(s, c, 0.5 * dl);
}

/* transform_to_observer_basis (src/frame/stokes_rotation.hpp:39-57). */
void kpo_transform_to_observer_basis(const kpo_stokes* in, double r11, double r12, double r21,
                                     double r22, kpo_stokes* out) {
    kpo_stokes o = *in;
    o.Q = 0.5 * in->Q * (r11 * r11 - r12 * r12 - r21 * r21 + r22 * r22) + in->U * (r11 * r12 - r21 * r22);
    o.U = in->Q * (r11 * r21 - r12 * r22) + in->U * (r11 * r22 + r12 * r21);
    o.V = (r11 * r22 - r12 * r21) * in->V;
    *out = o;
}

/* radiation_substep_count with its rate helpers (src/geodesic/pass_b.hpp:126-179). */
int32_t kpo_radiation_substep_count(const kpo_coeffs* c, double geometric_step,
                                    double dlambda_scale, int32_t requested,
                                    double max_radiation_step, double max_radiation_depth,
                                    double max_absorption_depth, double max_faraday_depth) {
    int nsub = requested > 0 ? requested : 1;
    const double abs_h = absv(geometric_step);
    if (max_radiation_step > 0.0)
        while (abs_h / (double)nsub > max_radiation_step && nsub < 4096) nsub *= 2;
    const double dl = abs_h * absv(dlambda_scale);
    const double alim = max_absorption_depth >= 0.0 ? max_absorption_depth : max_radiation_depth;
    if (alim > 0.0) {
        const double depth = (absv(c->aI) + sqrt(sq(c->aQ) + sq(c->aU) + sq(c->aV))) * dl;
        while (depth / (double)nsub > alim && nsub < 4096) nsub *= 2;
    }
    const double flim = max_faraday_depth >= 0.0 ? max_faraday_depth : max_radiation_depth;
    if (flim > 0.0) {
        const double depth = sqrt(sq(c->rQ) + sq(c->rU) + sq(c->rV)) * dl;
        while (depth / (double)nsub > flim && nsub < 4096) nsub *= 2;
    }
    return nsub;
}

/* ======================================================================
 * Passes (src/geodesic/pass_b.hpp, src/geodesic/pass_a_core.hpp)
 * ==================================================================== */
static int finite_state(const kpo_state* s) {
    for (int i = 0; i < 16; ++i) {
        const double v = s->v[i];
        if (!(v == v) || absv(v) >= LARGE) return 0;
    }
    return 1;
}

/* metric_surface_path_state (pass_a_core.hpp:73-84). */
static void surface_path_state(const metric_ctx* m, const kpo_state* first, double h, int composed,
                               kpo_state* out) {
    if (composed) {
        kpo_state half;
        rk4(m, first, h * 0.5, &half);
        rk4(m, &half, h * 0.5, out);
        return;
    }
    rk4(m, first, h, out);
}

/* locate_metric_surface_crossing for a radial surface (pass_a_core.hpp:94-197);
 * the surface value is the domain radius minus the surface radius. */
static void locate_crossing(const metric_ctx* m, const kpo_state* first, const kpo_state* second,
                            double full_h, double surface, int composed, kpo_state* out_state,
                            double* out_h) {
    *out_state = *first;
    *out_h = 0.0;
    double lo = 0.0, hi = 1.0;
    double flo = radius_of(m, first->v) - surface;
    if (flo == 0.0) return;
    if (!finite_state(second)) return;
    kpo_state hi_state;
    surface_path_state(m, first, full_h, composed, &hi_state);
    if (!finite_state(&hi_state)) return;
    double fhi = radius_of(m, hi_state.v) - surface;
    if (!(fhi == fhi) || absv(fhi) >= LARGE) return;
    if (fhi == 0.0) {
        *out_state = hi_state;
        *out_h = full_h;
        return;
    }
    if ((flo > 0.0) == (fhi > 0.0)) return;
    int use_secant = 1;
    for (int iter = 0; iter < 12; ++iter) {
        const double width = hi - lo;
        double trial = (lo + hi) * 0.5;
        const double denom = flo - fhi;
        if (use_secant && absv(denom) > TINY) {
            const double secant = lo + width * flo / denom;
            const double guard = width * 0.1;
            if (secant > lo + guard && secant < hi - guard) trial = secant;
        }
        kpo_state ts;
        surface_path_state(m, first, full_h * trial, composed, &ts);
        if (!finite_state(&ts)) {
            hi = trial;
            use_secant = 0;
            continue;
        }
        const double ft = radius_of(m, ts.v) - surface;
        if (!(ft == ft) || absv(ft) >= LARGE) {
            hi = trial;
            use_secant = 0;
            continue;
        }
        if (ft == 0.0) {
            *out_state = ts;
            *out_h = full_h * trial;
            return;
        }
        if ((flo > 0.0) == (ft > 0.0)) {
            lo = trial;
            flo = ft;
            *out_state = ts;
            *out_h = full_h * trial;
        } else {
            hi = trial;
            fhi = ft;
        }
    }
}

/* vertical_position_velocity (pass_a_core.hpp:202-224): the spin-axis height
 * z = r cos(theta) and dz/ds along the ray; the FMKS theta depends on both x1
 * and x2, and r = exp(x1). */
static void vertical_pv(const metric_ctx* m, const kpo_state* s, double* z, double* dz) {
    if (m->fmks) {
        const double x1 = s->v[1], x2 = s->v[2];
        const double r = exp(x1);
        const double th = kpo_fmks_theta(m->fm, x1, x2);
        const double dth = fmks_dtheta_dx1(m->fm, x1, x2) * s->v[5] + fmks_dtheta_dx2(m->fm, x1, x2) * s->v[6];
        *z = r * cos(th);
        *dz = *z * s->v[5] - r * sin(th) * dth;
    } else {
        *z = s->v[1] * cos(s->v[2]);
        *dz = cos(s->v[2]) * s->v[5] - s->v[1] * sin(s->v[2]) * s->v[6];
    }
}

/* vertical_turn_bracket (pass_a_core.hpp:226-238): an away-from-midplane to
 * toward-midplane reversal on one side of the midplane, both heights above
 * 64 ulp of the radius scale. */
static int vertical_turn_bracket(double z0, double v0, double z1, double v1, double h, double radius_scale) {
    const double floor = 64.0 * 0x1p-52 * radius_scale;
    return absv(z0) > floor && absv(z1) > floor && ((z0 > 0.0) == (z1 > 0.0)) && z0 * v0 * h > 0.0 &&
           z1 * v1 * h <= 0.0;
}

/* truncate_at_first_vertical_turn (pass_a_core.hpp:240-275): bisect the step
 * on its own RK composition for the last point whose dz/ds keeps the start's
 * sign; *last and *h become that point and its step. */
static int truncate_vertical_turn(const metric_ctx* m, const kpo_state* first, kpo_state* last, double* h,
                                  int composed) {
    double z0, v0, z1, v1;
    vertical_pv(m, first, &z0, &v0);
    vertical_pv(m, last, &z1, &v1);
    const double scale = maxv(radius_of(m, first->v), radius_of(m, last->v));
    if (!finite_state(last) || !vertical_turn_bracket(z0, v0, z1, v1, *h, scale)) return 0;
    double lo = 0.0, hi = 1.0;
    kpo_state endpoint = *first;
    for (int iter = 0; iter < 48; ++iter) {
        const double mid = (lo + hi) * 0.5;
        if (mid == lo || mid == hi) break;
        kpo_state trial;
        surface_path_state(m, first, *h * mid, composed, &trial);
        double z, v;
        vertical_pv(m, &trial, &z, &v);
        if (!finite_state(&trial)) {
            hi = mid;
            continue;
        }
        if ((v > 0.0) == (v0 > 0.0) && v != 0.0) {
            lo = mid;
            endpoint = trial;
        } else {
            hi = mid;
        }
    }
    *last = endpoint;
    *h *= lo;
    return 1;
}

/* ---- equatorial emission wedge (emission_selection.hpp:11-122) ---------- */

/* equatorial_height_ratio (:21-33): cos of the physical polar angle (FMKS:
 * mapped from x1, x2), x[3] on the flat test path. */
static double height_ratio(const metric_ctx* m, const double* x) {
    if (m->fmks == 2) return x[3];
    if (m->fmks) return cos(kpo_fmks_theta(m->fm, x[1], x[2]));
    return cos(x[2]);
}

static int narrow_wedge(const kp_emission_selection* sel) {
    return sel->equatorial_h_over_r > 0.0 && sel->equatorial_h_over_r < 1.0;
}

/* apply_emission_selection (:35-42): outside the wedge only the emissivities
 * are removed; the Faraday-rotation switch is applied where the model's
 * coefficients are evaluated. */
static void select_emission(const kp_emission_selection* sel, const metric_ctx* m, const kpo_state* s,
                            kpo_coeffs* c) {
    if (narrow_wedge(sel) && absv(height_ratio(m, s->v)) > sel->equatorial_h_over_r)
        c->jI = c->jQ = c->jU = c->jV = 0.0;
}

/* equatorial_resolution_substeps (:45-60). */
static int wedge_substeps(const kp_emission_selection* sel, const metric_ctx* m, const kpo_state* first,
                          const kpo_state* middle, const kpo_state* last, double outer) {
    if (!narrow_wedge(sel)) return 1;
    if (radius_of(m, first->v) > outer && radius_of(m, middle->v) > outer && radius_of(m, last->v) > outer)
        return 1;
    const double q0 = height_ratio(m, first->v), qm = height_ratio(m, middle->v), q1 = height_ratio(m, last->v);
    const double span = absv(qm - q0) + absv(q1 - qm);
    const double allowed = 2.0 * sel->equatorial_h_over_r / (double)sel->equatorial_samples;
    const double ratio = span / allowed;
    return ratio > 1.000001 ? (int)minv(ceil(ratio), 1000000.0) : 1;
}

/* refine_equatorial_boundary (:69-122): split a resolved step at the first
 * wedge surface it crosses (bisection on the step's own RK composition);
 * *h, *last and *middle become the split step's. */
static int refine_wedge(const kp_emission_selection* sel, const metric_ctx* m, const kpo_state* first,
                        kpo_state* middle, kpo_state* last, double* h, int composed, double outer) {
    if (!narrow_wedge(sel) || *h == 0.0) return 0;
    if (radius_of(m, first->v) > outer && radius_of(m, middle->v) > outer && radius_of(m, last->v) > outer)
        return 0;
    const double tolerance = 128.0 * 0x1p-52;
    const double hr = sel->equatorial_h_over_r;
    const double q0 = height_ratio(m, first->v), qm = height_ratio(m, middle->v), q1 = height_ratio(m, last->v);
    double edge = 0.0, lo = 0.0, hi = 1.0;
    int found = 0;
    for (int half = 0; half < 2 && !found; ++half) {
        const double a = half ? qm : q0, b = half ? q1 : qm;
        double earliest = 2.0;
        for (int sign = -1; sign <= 1; sign += 2) {
            const double surface = sign * hr;
            if (absv(q0 - surface) <= tolerance || absv(a - surface) <= tolerance) continue;
            if ((a - surface) * (b - surface) > 0.0) continue;
            const double fraction = (surface - a) / (b - a);
            if (fraction < earliest) {
                earliest = fraction;
                edge = surface;
                found = 1;
            }
        }
        if (found) {
            lo = half ? 0.5 : 0.0;
            hi = half ? 1.0 : 0.5;
        }
    }
    if (!found) return 0;
    kpo_state low;
    surface_path_state(m, first, *h * lo, composed, &low);
    const double low_value = height_ratio(m, low.v) - edge;
    for (int iter = 0; iter < 48; ++iter) {
        const double mid = (lo + hi) * 0.5;
        if (mid == lo || mid == hi) break;
        kpo_state trial;
        surface_path_state(m, first, *h * mid, composed, &trial);
        const double value = height_ratio(m, trial.v) - edge;
        if ((value > 0.0) == (low_value > 0.0))
            lo = mid;
        else
            hi = mid;
    }
    const double fraction = (lo + hi) * 0.5;
    if (fraction <= tolerance || fraction >= 1.0 - tolerance) return 0;
    *h *= fraction;
    surface_path_state(m, first, *h, composed, last);
    rk4(m, first, *h * 0.5, middle);
    return 1;
}

/* The pass-B step's wedge handling (pass_b.hpp:640-655): the fixed-step
 * midpoint sample, then the resolution substep count (returned when above
 * 1: the caller retakes the step), then the split at the first surface; 0
 * when the step goes on (*crosses recomputed after a split). */
static int wedge_step(const kp_pass_a_params* p, const metric_ctx* m, const kpo_state* old, kpo_state* sample,
                      kpo_state* next, double* h, int* crosses) {
    const kp_emission_selection* sel = &p->emission_selection;
    if (!narrow_wedge(sel)) return 0;
    if (!p->adaptive) rk4(m, old, *h * 0.5, sample);
    const int n = wedge_substeps(sel, m, old, sample, next, p->outer_radius);
    if (n > 1) return n;
    if (refine_wedge(sel, m, old, sample, next, h, p->adaptive, p->outer_radius)) {
        const double s0 = radius_of(m, old->v) - p->camera.radius;
        const double s1 = radius_of(m, next->v) - p->camera.radius;
        *crosses = (s0 == 0.0 || s1 == 0.0) ? 1 : ((s0 < 0.0 && s1 > 0.0) || (s0 > 0.0 && s1 < 0.0));
    }
    return 0;
}

/* trace_pass_a_segment_endpoint_pixel_metric (pass_b.hpp:337-486), BL hooks
 * reduced per metric_domain.hpp:159-223 (no time domain, no step limits);
 * direct_only ends the segment at the first vertical turn (:461-470). */
static int pass_a_endpoint(const kp_pass_a_params* p, const metric_ctx* m, const kpo_state* init,
                           kpo_state* state_out, int32_t* steps_out, int32_t* reason_out) {
    kpo_state state = *init;
    int reason = KP_REASON_MAX_STEPS;
    int steps = 0;
    adaptive_control ac = {p->adaptive_tolerance, p->min_step, p->max_step, 10};
    double h_current = -absv(p->step);
    int entered = 0;
    int steps_inside = 0;
    for (int n = 0; n < p->max_steps; ++n) {
        const double r_bl = radius_of(m, state.v);
        if (!entered) {
            if (r_bl <= p->outer_radius) {
                entered = 1;
                steps_inside = 0;
            }
        } else {
            if (r_bl - p->inner_radius <= 0.0) {
                reason = KP_REASON_REACHED_INNER_BOUNDARY;
                break;
            }
            if (steps_inside > 0 && r_bl >= p->outer_radius) {
                reason = KP_REASON_ESCAPED_DOMAIN;
                break;
            }
        }
        const kpo_state old = state;
        const double eff_min = absv(p->min_step);
        const double old_r = r_bl;
        const double old_inner = r_bl - p->inner_radius;
        double used_h = h_current;
        kpo_state next;
        if (p->adaptive) {
            adaptive_control tc = ac;
            tc.min_step = eff_min;
            if (absv(used_h) < tc.min_step) tc.min_step = absv(used_h);
            adaptive_result st;
            adaptive_rk4(m, &old, used_h, tc, &st);
            if (!st.accepted) {
                h_current = st.next_h;
                if (absv(h_current) <= eff_min * 1.0001) {
                    reason = KP_REASON_ADAPTIVE_STEP_UNDERFLOW;
                    break;
                }
                continue;
            }
            next = st.state;
            used_h = st.used_h;
            h_current = st.next_h;
        } else {
            rk4(m, &state, used_h, &next);
        }
        const double next_r = radius_of(m, next.v);
        const double next_inner = next_r - p->inner_radius;
        if (!entered && next_r <= p->outer_radius) {
            entered = 1;
            steps_inside = 0;
        } else if (!entered && next_r >= old_r) {
            state = next;
            reason = KP_REASON_ESCAPED_DOMAIN;
            steps += 1;
            break;
        }
        int crossing = KP_REASON_NONE;
        if (entered && old_inner > 0.0 && next_inner <= 0.0) {
            crossing = KP_REASON_REACHED_INNER_BOUNDARY;
        } else if (entered && steps_inside > 0 && old_r < p->outer_radius &&
                   next_r >= p->outer_radius) {
            crossing = KP_REASON_ESCAPED_DOMAIN;
        }
        if (crossing != KP_REASON_NONE) {
            const int inner = crossing == KP_REASON_REACHED_INNER_BOUNDARY;
            const double surface = inner ? p->inner_radius : p->outer_radius;
            kpo_state ev;
            double ev_h;
            locate_crossing(m, &old, &next, used_h, surface, p->adaptive, &ev, &ev_h);
            next = ev;
            used_h = ev_h;
        }
        if (p->direct_only && truncate_vertical_turn(m, &old, &next, &used_h, p->adaptive))
            crossing = KP_REASON_REACHED_DIRECT_TURN;
        if (crossing != KP_REASON_NONE) {
            state = next;
            reason = crossing;
            steps += 1;
            break;
        }
        state = next;
        steps += 1;
        if (entered) steps_inside += 1;
    }
    const int valid = reason == KP_REASON_REACHED_DIRECT_TURN ||
                      (entered && (reason == KP_REASON_REACHED_INNER_BOUNDARY || reason == KP_REASON_ESCAPED_DOMAIN));
    *state_out = state;
    *steps_out = steps;
    *reason_out = reason;
    return valid;
}

void kpo_vertical_pv(const kpo_fmks* fm, const kpo_state* s, double* z, double* dzds) {
    metric_ctx m = {0.0, 0.0, NULL, fm != NULL, fm};
    vertical_pv(&m, s, z, dzds);
}

int32_t kpo_vertical_turn_bracket(double z0, double v0, double z1, double v1, double h, double radius_scale) {
    return vertical_turn_bracket(z0, v0, z1, v1, h, radius_scale);
}

int32_t kpo_truncate_vertical_turn(double mass, double spin, const kpo_fmks* fm, const kpo_state* first,
                                   kpo_state* last, double* h, int32_t composed) {
    metric_ctx m = {mass, spin, NULL, fm != NULL, fm};
    return truncate_vertical_turn(&m, first, last, h, composed);
}

/* The wedge pieces on metric kind (0 BL, 1 FMKS with fm, 2 the flat test path). */
int32_t kpo_wedge_substeps(const kp_emission_selection* sel, int32_t kind, const kpo_fmks* fm,
                           const kpo_state* first, const kpo_state* middle, const kpo_state* last, double outer) {
    metric_ctx m = {1.0, 0.0, NULL, kind, fm};
    return wedge_substeps(sel, &m, first, middle, last, outer);
}

int32_t kpo_refine_wedge(const kp_emission_selection* sel, int32_t kind, const kpo_fmks* fm,
                         const kpo_state* first, kpo_state* middle, kpo_state* last, double* h, int32_t composed,
                         double outer) {
    metric_ctx m = {1.0, 0.0, NULL, kind, fm};
    return refine_wedge(sel, &m, first, middle, last, h, composed, outer);
}

void kpo_select_emission(const kp_emission_selection* sel, int32_t kind, const kpo_fmks* fm, const kpo_state* s,
                         kpo_coeffs* c) {
    metric_ctx m = {1.0, 0.0, NULL, kind, fm};
    select_emission(sel, &m, s, c);
}

double kpo_height_ratio(int32_t kind, const kpo_fmks* fm, const kpo_state* s) {
    metric_ctx m = {1.0, 0.0, NULL, kind, fm};
    return height_ratio(&m, s->v);
}

void kpo_path_state(int32_t kind, double mass, double spin, const kpo_fmks* fm, const kpo_state* first, double h,
                    int32_t composed, kpo_state* out) {
    metric_ctx m = {mass, spin, NULL, kind, fm};
    surface_path_state(&m, first, h, composed, out);
}

void kpo_rk4_step(double mass, double spin, const kpo_fmks* fm, const kpo_state* s, double h, kpo_state* out) {
    metric_ctx m = {mass, spin, NULL, fm != NULL, fm};
    rk4(&m, s, h, out);
}

int32_t kpo_bl_pass_a_endpoint(const kp_pass_a_params* p, int32_t pixel, kpo_state* endpoint,
                               int32_t* steps, int32_t* reason) {
    metric_ctx m = {p->mass, p->spin, NULL, 0, NULL};
    kpo_state init;
    kpo_bl_camera_ray(p, pixel, &init);
    return pass_a_endpoint(p, &m, &init, endpoint, steps, reason);
}

static double spatial_distance(const double* a, const double* b) {
    return sqrt(sq(a[1] - b[1]) + sq(a[2] - b[2]) + sq(a[3] - b[3]));
}

static double max_abs_diff4(const double* a, const double* b) {
    double out = absv(a[0] - b[0]);
    for (int i = 1; i < 4; ++i) out = maxv(out, absv(a[i] - b[i]));
    return out;
}

typedef struct pixel_result {
    kpo_stokes observed[KPO_MAX_FREQUENCIES];
    double r11, r12, r21, r22;
    double closure_x, closure_k, final_null, frame_error;
    int pass_a_steps, steps, reason;
    int waiting;           /* slow light, decoupled: the ray needs a time outside the sequence */
    double requested_time; /* that time */
} pixel_result;

/* Evaluate the plasma coefficients on the sample state, with the
 * emission-selection Faraday switch (emission_selection.hpp:35-42). */
static void sample_coefficients(const kp_pass_a_params* p, const kp_riaf_model* model,
                                const kpo_state* s, const kpo_perturbation* pert, kpo_coeffs* c,
                                kpo_stats* stats) {
    if (stats) stats->coefficient_evals++;
    kpo_riaf_coefficients(p, model, s, pert, c);
    if (!p->emission_selection.faraday_rotation) c->rV = 0.0;
}

/* ======================================================================
 * GRMHD grid model on FMKS data (src/model/grmhd_grid.hpp) with derived
 * scalars precomputed as the iharm loader does (src/grmhd/iharm_loader.cpp).
 * Cells are kp_grmhd_cell records [n1][n2][n3] in host memory here.
 * ==================================================================== */
static kpo_fmks grmhd_map(const kp_grmhd_model* g) {
    kpo_fmks f = {g->startx1, g->hslope, g->mks_smooth, g->poly_alpha, g->poly_xt, g->poly_norm};
    return f;
}

/* length_unit_cgs, rho_unit_cgs, b_unit_cgs, dlambda_scale (grmhd_grid.hpp:140-160). */
static double grmhd_length_unit(const kp_grmhd_model* g) {
    return 6.67430e-8 * g->mbh_solar * 1.98847e33 / (2.99792458e10 * 2.99792458e10);
}
static double grmhd_rho_unit(const kp_grmhd_model* g) {
    const double L = grmhd_length_unit(g);
    return g->M_unit / maxv(L * L * L, 1e-300);
}
static double grmhd_b_unit(const kp_grmhd_model* g) {
    return 2.99792458e10 * sqrt(4.0 * PI_D * grmhd_rho_unit(g));
}
double kpo_grmhd_dlambda_scale(const kp_grmhd_model* g) { return grmhd_length_unit(g) / maxv(g->freq_cgs, 1.0); }

/* native_fluid_metric (grmhd_grid.hpp:721-766): FMKS gcov and the first row of gcon. */
static void grmhd_native_metric(const kp_grmhd_model* g, const kpo_fmks* f, double x1, double x2, double r,
                                double th, double gcov[16], double gcon0[4]) {
    const double cth = cos(th), sth = sin(th);
    const double s2 = sth * sth;
    const double a2 = g->spin * g->spin;
    const double rho2 = r * r + a2 * cth * cth;
    const double fr = 2.0 * r / rho2;
    const double dth1 = fmks_dtheta_dx1(f, x1, x2);
    const double dth2 = fmks_dtheta_dx2(f, x1, x2);
    for (int i = 0; i < 16; ++i) gcov[i] = 0.0;
    const double g00 = -1.0 + fr, g01 = fr, g03 = -g->spin * fr * s2, g11 = 1.0 + fr;
    const double g13 = -g->spin * s2 * (1.0 + fr), g22 = rho2, g33 = s2 * (rho2 + a2 * s2 * (1.0 + fr));
    gcov[0] = g00;
    gcov[1] = gcov[4] = g01 * r;
    gcov[3] = gcov[12] = g03;
    gcov[5] = g11 * r * r + g22 * dth1 * dth1;
    gcov[6] = gcov[9] = g22 * dth1 * dth2;
    gcov[7] = gcov[13] = g13 * r;
    gcov[10] = g22 * dth2 * dth2;
    gcov[15] = g33;
    const double inv_r = 1.0 / maxv(r, 1e-300);
    const double inv_dth2 = 1.0 / maxv(dth2, 1e-300);
    gcon0[0] = -(1.0 + fr);
    gcon0[1] = fr * inv_r;
    gcon0[2] = -fr * dth1 * inv_r * inv_dth2;
    gcon0[3] = 0.0;
}

/* Fluid four-velocity and magnetic four-vector from native primitives
 * (grmhd_grid.hpp:980-1014, iharm_loader.cpp:453-487).  Returns b^2 (floored). */
static double grmhd_four_vectors(const double gcov[16], const double gcon0[4], double U1, double U2, double U3,
                                 double B1, double B2, double B3, double ucon[4], double bcon[4]) {
    const double v[4] = {0.0, U1, U2, U3};
    double spatial = 0.0;
    for (int i = 1; i < 4; ++i)
        for (int j = 1; j < 4; ++j) spatial += gcov[i * 4 + j] * v[i] * v[j];
    const double vfac = sqrt(maxv(-1.0 / gcon0[0] * (1.0 + absv(spatial)), 0.0));
    ucon[0] = -vfac * gcon0[0];
    for (int i = 1; i < 4; ++i) ucon[i] = v[i] - vfac * gcon0[i];
    double ucov[4];
    for (int mu = 0; mu < 4; ++mu) {
        double sum = 0.0;
        for (int nu = 0; nu < 4; ++nu) sum += gcov[mu * 4 + nu] * ucon[nu];
        ucov[mu] = sum;
    }
    const double B[4] = {0.0, B1, B2, B3};
    double udotB = 0.0;
    for (int i = 1; i < 4; ++i) udotB += ucov[i] * B[i];
    bcon[0] = udotB;
    for (int i = 1; i < 4; ++i) bcon[i] = (B[i] + ucon[i] * udotB) / maxv(ucon[0], 1e-300);
    double bsq = 0.0;
    for (int mu = 0; mu < 4; ++mu) {
        double bc = 0.0;
        for (int nu = 0; nu < 4; ++nu) bc += gcov[mu * 4 + nu] * bcon[nu];
        bsq += bcon[mu] * bc;
    }
    return maxv(absv(bsq), 1e-40);
}

/* KPOLARISIharmDerivedScalars (iharm_loader.cpp:420-505): per cell centre. */
void kpo_grmhd_derive_scalars(const kp_grmhd_model* g, kp_grmhd_cell* cells) {
    const double mp = 1.67262171e-24, me = 9.1093826e-28;
    const double mp_me = mp / me;
    const double game = 4.0 / 3.0, gamp = 5.0 / 3.0;
    const kpo_fmks f = grmhd_map(g);
    const double rho_unit = grmhd_rho_unit(g), b_unit = grmhd_b_unit(g);
    const int n1 = g->n1, n2 = g->n2, n3 = g->n3;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            const double x1 = g->startx1 + ((double)i + 0.5) * g->dx1;
            const double x2 = g->startx2 + ((double)j + 0.5) * g->dx2;
            const double r = exp(x1);
            const double th = kpo_fmks_theta(&f, x1, x2);
            double gcov[16], gcon0[4];
            grmhd_native_metric(g, &f, x1, x2, r, th, gcov, gcon0);
            for (int k = 0; k < n3; ++k) {
                kp_grmhd_cell* c = &cells[((size_t)i * n2 + j) * n3 + k];
                const double rho = c->prim[0], uu = c->prim[1];
                double ucon[4], bcon[4];
                const double bsq = grmhd_four_vectors(gcov, gcon0, c->prim[2], c->prim[3], c->prim[4], c->prim[5],
                                                      c->prim[6], c->prim[7], ucon, bcon);
                const double beta = (g->gam - 1.0) * uu / (0.5 * bsq);
                const double sigma = bsq / maxv(rho, 1e-300);
                const double betasq = beta * beta / maxv(g->beta_crit * g->beta_crit, 1e-40);
                const double trat = (g->trat_large * betasq + g->trat_small) / (1.0 + betasq);
                const double thetae_unit = mp_me * (game - 1.0) * (gamp - 1.0) / ((gamp - 1.0) + (game - 1.0) * trat);
                c->derived[0] = rho * rho_unit / (mp + me) * g->ne_factor;
                c->derived[1] = thetae_unit * uu / maxv(rho, 1e-300);
                c->derived[2] = sqrt(bsq) * b_unit;
                c->derived[3] = sigma;
                c->derived[4] = beta;
            }
        }
}

/* u^mu and b^mu at the centre of cell (i, j, k) from its primitives, as the
 * derived-scalar pass forms them (a test hook of grmhd_four_vectors). */
void kpo_grmhd_cell_four_vectors(const kp_grmhd_model* g, const kp_grmhd_cell* cells, int32_t i, int32_t j,
                                 int32_t k, double ucon[4], double bcon[4]) {
    const kpo_fmks f = grmhd_map(g);
    const double x1 = g->startx1 + ((double)i + 0.5) * g->dx1;
    const double x2 = g->startx2 + ((double)j + 0.5) * g->dx2;
    const double r = exp(x1);
    const double th = kpo_fmks_theta(&f, x1, x2);
    double gcov[16], gcon0[4];
    grmhd_native_metric(g, &f, x1, x2, r, th, gcov, gcon0);
    const kp_grmhd_cell* c = &cells[((size_t)i * g->n2 + j) * g->n3 + k];
    (void)grmhd_four_vectors(gcov, gcon0, c->prim[2], c->prim[3], c->prim[4], c->prim[5], c->prim[6], c->prim[7],
                             ucon, bcon);
}

/* native_indices_from_native_coords (grmhd_grid.hpp:271-296). */
static int grmhd_indices(const kp_grmhd_model* g, double x1, double x2, double phi, int* i0, int* j0, int* k0,
                         double* di, double* dj, double* dk) {
    if (!(x1 >= g->startx1 && x1 < g->startx1 + (double)g->n1 * g->dx1 && x2 >= g->startx2 &&
          x2 < g->startx2 + (double)g->n2 * g->dx2))
        return 0;
    const double two_pi = 6.283185307179586476925286766559005768;
    phi = phi - floor(phi / two_pi) * two_pi;
    const double ii = (x1 - g->startx1) / g->dx1 - 0.5;
    const double jj = (x2 - g->startx2) / g->dx2 - 0.5;
    const double kk = (phi - g->startx3) / g->dx3 - 0.5;
    *i0 = (int)floor(ii);
    *j0 = (int)floor(jj);
    *k0 = (int)floor(kk);
    if (*i0 < -1) *i0 = -1;
    if (*j0 < -1) *j0 = -1;
    if (*k0 < -1) *k0 = -1;
    if (*i0 >= g->n1) *i0 = g->n1 - 1;
    if (*j0 >= g->n2) *j0 = g->n2 - 1;
    if (*k0 >= g->n3) *k0 = g->n3 - 1;
    *di = clampv(ii - (double)*i0, 0.0, 1.0);
    *dj = clampv(jj - (double)*j0, 0.0, 1.0);
    *dk = clampv(kk - (double)*k0, 0.0, 1.0);
    return 1;
}

/* map_ghost_index (grmhd_grid.hpp:309-321): clamped radius, polar reflection
 * across the axis (k + n3/2), periodic phi. */
static size_t grmhd_cell_index(const kp_grmhd_model* g, int i, int j, int k) {
    if (i < 0) i = 0;
    if (i >= g->n1) i = g->n1 - 1;
    if (j < 0) {
        j = 0;
        k += g->n3 / 2;
    } else if (j >= g->n2) {
        j = g->n2 - 1;
        k += g->n3 / 2;
    }
    while (k < 0) k += g->n3;
    while (k >= g->n3) k -= g->n3;
    return ((size_t)i * (size_t)g->n2 + (size_t)j) * (size_t)g->n3 + (size_t)k;
}

/* make_interp_stencil (grmhd_grid.hpp:379-408). */
static void grmhd_stencil(const kp_grmhd_model* g, int i0, int j0, int k0, double di, double dj, double dk,
                          size_t cell[8], double w[8]) {
    const double wi0 = 1.0 - di, wi1 = di, wj0 = 1.0 - dj, wj1 = dj, wk0 = 1.0 - dk, wk1 = dk;
    cell[0] = grmhd_cell_index(g, i0, j0, k0);
    cell[1] = grmhd_cell_index(g, i0 + 1, j0, k0);
    cell[2] = grmhd_cell_index(g, i0, j0 + 1, k0);
    cell[3] = grmhd_cell_index(g, i0 + 1, j0 + 1, k0);
    cell[4] = grmhd_cell_index(g, i0, j0, k0 + 1);
    cell[5] = grmhd_cell_index(g, i0 + 1, j0, k0 + 1);
    cell[6] = grmhd_cell_index(g, i0, j0 + 1, k0 + 1);
    cell[7] = grmhd_cell_index(g, i0 + 1, j0 + 1, k0 + 1);
    w[0] = wi0 * wj0 * wk0;
    w[1] = wi1 * wj0 * wk0;
    w[2] = wi0 * wj1 * wk0;
    w[3] = wi1 * wj1 * wk0;
    w[4] = wi0 * wj0 * wk1;
    w[5] = wi1 * wj0 * wk1;
    w[6] = wi0 * wj1 * wk1;
    w[7] = wi1 * wj1 * wk1;
}

/* sigma_smooth_factor (grmhd_grid.hpp:495-503). */
static double sigma_smooth_factor(const kp_grmhd_model* g, double sigma) {
    double above = g->sigma_cut;
    if (g->sigma_cut_high > 0.0) above = g->sigma_cut_high;
    if (sigma < g->sigma_cut) return 1.0;
    if (sigma >= above) return 0.0;
    const double dsig = maxv(above - g->sigma_cut, 1e-300);
    return cos(PI_D / 2.0 / dsig * (sigma - g->sigma_cut));
}

/* Trilinear sample of primitive v (interp_prim_stencil, grmhd_grid.hpp:410-420). */
double kpo_grmhd_sample_prim(const kp_grmhd_model* g, const kp_grmhd_cell* cells, int v, double x1, double x2,
                             double phi) {
    int i0, j0, k0;
    double di, dj, dk;
    if (!grmhd_indices(g, x1, x2, phi, &i0, &j0, &k0, &di, &dj, &dk)) return 0.0;
    size_t cell[8];
    double w[8];
    grmhd_stencil(g, i0, j0, k0, di, dj, dk, cell, w);
    double out = 0.0;
    for (int n = 0; n < 8; ++n) out += w[n] * (double)cells[cell[n]].prim[v];
    return out;
}

/* Trilinear samples of the 8 primitives and 5 derived scalars of one cell
 * array at FMKS (x1, x2, phi): interp_prim_stencil / interp_derived_stencil
 * (grmhd_grid.hpp:379-430).  Returns 0 outside the grid. */
static int grmhd_sample_cells(const kp_grmhd_model* g, const kp_grmhd_cell* cells, double x1, double x2,
                              double phi, double prim[8], double der[5]) {
    int i0, j0, k0;
    double di, dj, dk;
    if (!grmhd_indices(g, x1, x2, phi, &i0, &j0, &k0, &di, &dj, &dk)) return 0;
    size_t cell[8];
    double w[8];
    grmhd_stencil(g, i0, j0, k0, di, dj, dk, cell, w);
    for (int v = 0; v < 8; ++v) {
        double acc = 0.0;
        for (int n = 0; n < 8; ++n) acc += w[n] * (double)cells[cell[n]].prim[v];
        prim[v] = acc;
    }
    for (int v = 0; v < 5; ++v) {
        double acc = 0.0;
        for (int n = 0; n < 8; ++n) acc += w[n] * cells[cell[n]].derived[v];
        der[v] = acc;
    }
    return 1;
}

/* The coefficients at an FMKS sample state from the sampled primitives and
 * derived scalars: the rest of fluid_state (four-vector reconstruction,
 * grmhd_grid.hpp:953-1080) and coefficients (grmhd_grid.hpp:1091-1240). */
static void grmhd_coefficients_from(const kp_grmhd_model* g, const kpo_state* s, const double prim[8],
                                    const double der[5], const kpo_perturbation* pert, int faraday_rotation,
                                    kpo_coeffs* out) {
    memset(out, 0, sizeof(*out));
    const kpo_fmks f = grmhd_map(g);
    const double x1 = s->v[1], x2 = s->v[2];
    const double r = exp(x1);
    const double th = kpo_fmks_theta(&f, x1, x2);
    if (!(prim[0] > 0.0 && prim[1] > 0.0)) return;
    double gcov[16], gcon0[4];
    grmhd_native_metric(g, &f, x1, x2, r, th, gcov, gcon0);
    double ucon[4], bcon[4];
    const double bsq = grmhd_four_vectors(gcov, gcon0, prim[2], prim[3], prim[4], prim[5], prim[6], prim[7], ucon,
                                          bcon);
    const double bnorm_geom = sqrt(bsq);
    double ne = der[0], thetae = der[1], b_cgs = der[2];
    double sigma = der[3], beta = der[4];
    ne *= sigma_smooth_factor(g, sigma);
    /* PlasmaPerturbation::apply (plasma_perturbation.hpp:17-32) */
    if (pert && pert->parameter != 0 && pert->log_scale != 0.0) {
        const double fq = exp(pert->log_scale);
        if (pert->parameter == 1) {
            ne *= fq;
            b_cgs *= exp(pert->log_scale * 0.5);
        } else if (pert->parameter == 2) {
            thetae *= fq;
        } else if (pert->parameter == 3) {
            b_cgs *= fq;
            sigma *= fq * fq;
            beta /= fq * fq;
        }
    }
    const double* k = s->v + 4;
    const double* e1 = s->v + 8;
    const double* e2 = s->v + 12;
    double ucov[4];
    for (int mu = 0; mu < 4; ++mu) {
        double sum = 0.0;
        for (int nu = 0; nu < 4; ++nu) sum += gcov[mu * 4 + nu] * ucon[nu];
        ucov[mu] = sum;
    }
    double uk = 0.0;
    for (int mu = 0; mu < 4; ++mu) uk += ucov[mu] * k[mu];
    const double nu_scale = maxv(absv(-uk), 1e-8);
    const double nu = maxv(g->freq_cgs * nu_scale, 1.0);
    const double bnorm = maxv(bnorm_geom, 1e-150);
    const double inv_bnorm = 1.0 / bnorm;
    double kdotb = 0.0, ub = 0.0, ue1 = 0.0, ue2 = 0.0, ke1 = 0.0, ke2 = 0.0, e1b = 0.0, e2b = 0.0;
    for (int mu = 0; mu < 4; ++mu) {
        double bcov = 0.0, kcov = 0.0;
        for (int n = 0; n < 4; ++n) {
            bcov += gcov[mu * 4 + n] * bcon[n];
            kcov += gcov[mu * 4 + n] * k[n];
        }
        kdotb += k[mu] * bcov;
        ub += ucov[mu] * bcon[mu];
        ue1 += ucov[mu] * e1[mu];
        ue2 += ucov[mu] * e2[mu];
        ke1 += kcov * e1[mu];
        ke2 += kcov * e2[mu];
        e1b += bcov * e1[mu];
        e2b += bcov * e2[mu];
    }
    const double cos_theta = clampv(kdotb / maxv(nu_scale * bnorm, 1e-30), -1.0, 1.0);
    const double sin_theta = maxv(sqrt(maxv(0.0, 1.0 - cos_theta * cos_theta)), 1e-4);
    kpo_coeffs mc;
    if (g->emission_type == KP_EMISSION_KAPPA || g->emission_type == KP_EMISSION_POWER_LAW) {
        /* grmhd_grid.hpp:1176-1205 */
        kpo_nonthermal nt;
        memset(&nt, 0, sizeof(nt));
        nt.distribution = g->emission_type;
        nt.max_pol_frac_emission = nt.max_pol_frac_absorption = g->max_pol_frac;
        nt.min_sin_theta = 1e-4;
        nt.emission_scale = g->emission_scale;
        nt.absorption_scale = g->absorption_scale;
        nt.faraday_scale = g->faraday_scale;
        /* variable_kappa_from_sigma_beta falls back to the params' kappa, still
         * the NonthermalSynchrotronParams default 3.5 there (:34, :496-498) */
        nt.kappa = g->variable_kappa ? kpo_variable_kappa(sigma, beta, 3.5) : g->nonthermal_kappa;
        if (nt.kappa < g->variable_kappa_min) nt.kappa = g->variable_kappa_min;
        nt.kappa_width = -1.0;
        nt.kappa_interp_begin = g->variable_kappa_interp_start < g->variable_kappa_max
                                    ? g->variable_kappa_interp_start : g->variable_kappa_max;
        nt.kappa_interp_end = g->variable_kappa_max;
        nt.power_law_p = g->powerlaw_p;
        nt.power_law_eta = g->powerlaw_eta;
        nt.power_law_gamma_min = g->powerlaw_gamma_min;
        nt.power_law_gamma_max = g->powerlaw_gamma_max;
        nt.power_law_gamma_cutoff = g->powerlaw_gamma_cutoff;
        kpo_nonthermal_coefficients(&nt, nu, ne, thetae, b_cgs, sin_theta, cos_theta, &mc);
    } else {
        kpo_thermal_coefficients(g->emission_type, nu, ne, thetae, b_cgs, 0.0, sin_theta, cos_theta,
                                 g->max_pol_frac, 1e-4, g->emission_scale, g->absorption_scale, g->faraday_scale,
                                 &mc);
    }
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
    const double bproj2 = b1 * b1 + b2 * b2;
    out->jI = mc.jI;
    out->jV = mc.jV;
    out->aI = mc.aI;
    out->aV = mc.aV;
    out->rV = faraday_rotation ? mc.rV : 0.0;
    if (bproj2 > 1e-30) {
        const double cos2chi = (b2 * b2 - b1 * b1) / bproj2;
        const double sin2chi = -2.0 * b1 * b2 / bproj2;
        out->jQ = mc.jQ * cos2chi;
        out->jU = mc.jQ * sin2chi;
        out->aQ = mc.aQ * cos2chi;
        out->aU = mc.aQ * sin2chi;
        out->rQ = mc.rQ * cos2chi;
        out->rU = mc.rQ * sin2chi;
    }
}

/* GRMHDGridRadiationModel::coefficients for FMKS data and the FMKS metric
 * (grmhd_grid.hpp:883-1080 fluid_state, 1091-1240 coefficients). */
void kpo_grmhd_coefficients(const kp_grmhd_model* g, const kpo_state* s, const kpo_perturbation* pert,
                            int faraday_rotation, kpo_coeffs* out) {
    memset(out, 0, sizeof(*out));
    const double x1 = s->v[1], x2 = s->v[2], phi = s->v[3];
    const double r = exp(x1);
    if (!(r >= g->r_in && r < g->r_out)) return;
    double prim[8], der[5];
    if (!grmhd_sample_cells(g, (const kp_grmhd_cell*)g->cells, x1, x2, phi, prim, der)) return;
    grmhd_coefficients_from(g, s, prim, der, pert, faraday_rotation, out);
}

/* interpolate_time_value (src/model/time_interpolation.hpp:27-32): exact at
 * the endpoints and for equal values. */
double kpo_interpolate_time_value(double a, double b, double fraction) {
    if (fraction <= 0.0 || a == b) return a;
    if (fraction >= 1.0) return b;
    return (1.0 - fraction) * a + fraction * b;
}

/* SlowLightTemporalModel::interpolation_fraction (slow_light_driver.hpp:62-69). */
double kpo_slow_light_fraction(double t_lo, double t_hi, double t_obs, double x0) {
    const double denom = t_hi - t_lo;
    double f = 0.0;
    if (absv(denom) > TINY) f = (t_obs + x0 - t_lo) / denom;
    return minv(1.0, maxv(0.0, f));
}

/* SlowLightResidentWindows::sample_window over the whole sequence
 * (slow_light_driver.hpp:1000-1012): the pair (k, k+1) with t_k <= t < t_k+1,
 * found by bisection; t at or past the last time gives the last pair. */
int32_t kpo_slow_light_pair(const double* times, int32_t nsnap, double t) {
    int lo = 0, hi = nsnap - 1;
    while (lo + 1 < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (t < times[mid])
            hi = mid;
        else
            lo = mid;
    }
    return lo;
}

/* SlowLightTemporalModel::coefficients (slow_light_driver.hpp:166-193) for
 * the snapshot pair (k, k+1) of a GRMHD sequence at model frequency
 * g->freq_cgs.  Fluid interpolation samples both snapshots at the state,
 * interpolates every primitive and derived scalar in time and then
 * reconstructs (grmhd_grid.hpp:940-962); coefficient interpolation mixes the
 * two snapshots' coefficients linearly. */
void kpo_slow_light_coefficients(const kp_grmhd_model* g, const kp_slow_light* seq, int32_t k,
                                 const kpo_state* s, const kpo_perturbation* pert, int faraday_rotation,
                                 kpo_coeffs* out) {
    memset(out, 0, sizeof(*out));
    const double f = kpo_slow_light_fraction(seq->times[k], seq->times[k + 1], seq->observation_time, s->v[0]);
    const kp_grmhd_cell* lo = (const kp_grmhd_cell*)seq->cells[k];
    const kp_grmhd_cell* hi = (const kp_grmhd_cell*)seq->cells[k + 1];
    const double x1 = s->v[1], x2 = s->v[2], phi = s->v[3];
    const double r = exp(x1);
    if (seq->interpolation == KP_SLOW_LIGHT_FLUID) {
        if (!(r >= g->r_in && r < g->r_out)) return;
        double p0[8], d0[5], p1[8], d1[5];
        if (!grmhd_sample_cells(g, lo, x1, x2, phi, p0, d0)) return;
        if (f > 0.0) {
            if (!grmhd_sample_cells(g, hi, x1, x2, phi, p1, d1)) return;
            for (int v = 0; v < 8; ++v) p0[v] = kpo_interpolate_time_value(p0[v], p1[v], f);
            for (int v = 0; v < 5; ++v) d0[v] = kpo_interpolate_time_value(d0[v], d1[v], f);
        }
        grmhd_coefficients_from(g, s, p0, d0, pert, faraday_rotation, out);
        return;
    }
    kp_grmhd_model ga = *g, gb = *g;
    ga.cells = lo;
    gb.cells = hi;
    kpo_coeffs a, b;
    kpo_grmhd_coefficients(&ga, s, pert, faraday_rotation, &a);
    kpo_grmhd_coefficients(&gb, s, pert, faraday_rotation, &b);
    const double w = 1.0 - f;
    out->jI = w * a.jI + f * b.jI;
    out->jQ = w * a.jQ + f * b.jQ;
    out->jU = w * a.jU + f * b.jU;
    out->jV = w * a.jV + f * b.jV;
    out->aI = w * a.aI + f * b.aI;
    out->aQ = w * a.aQ + f * b.aQ;
    out->aU = w * a.aU + f * b.aU;
    out->aV = w * a.aV + f * b.aV;
    out->rQ = w * a.rQ + f * b.rQ;
    out->rU = w * a.rU + f * b.rU;
    out->rV = w * a.rV + f * b.rV;
}

/* The radiation model of one image: the analytic RIAF, a GRMHD grid, or a
 * GRMHD snapshot sequence sampled at retarded time (slow light). */
typedef struct model_ctx {
    const kp_riaf_model* riaf;
    const kp_grmhd_model* grmhd;
    const kp_slow_light* slow;
} model_ctx;

/* The frequency lists of one pass B: output frequencies and the frequencies
 * whose coefficients drive the substep control.  A single-frequency image is
 * the list {freq_cgs} for both (pass_b.hpp:515-722 is the nfreq = 1 case of
 * pass_b.hpp:1130-1500); distinct lists follow the control variant
 * (pass_b.hpp:1505-1895). */
typedef struct freq_plan {
    const double* freqs;
    int nfreq;
    const double* ctrl;
    int nctrl;
    int same;  /* ctrl == freqs elementwise: coefficients are reused */
} freq_plan;

/* Coefficients of the model copy with freq_cgs = f, and its dlambda_scale. */
static double coeffs_at_frequency(const kp_pass_a_params* p, const model_ctx* mc, double f,
                                  const kpo_state* sample, int inside, const kpo_perturbation* pert,
                                  kpo_coeffs* c, kpo_stats* stats) {
    memset(c, 0, sizeof(*c));
    if (mc->grmhd) {
        kp_grmhd_model gf = *mc->grmhd;
        gf.freq_cgs = f;
        if (inside) {
            if (stats) stats->coefficient_evals++;
            kpo_grmhd_coefficients(&gf, sample, pert, p->emission_selection.faraday_rotation, c);
        }
        return kpo_grmhd_dlambda_scale(&gf);
    }
    kp_riaf_model mf = *mc->riaf;
    mf.freq_cgs = f;
    if (inside) sample_coefficients(p, &mf, sample, pert, c, stats);
    return riaf_dlambda_scale(&mf);
}

/* ---- parameter responses (src/diagnostics/response.hpp:25-191) ----------
 * One ray's channels (response_config.hpp:28-32): `bins` source tags, four
 * mechanism derivatives per bin (channel bins + m bins + b: emission,
 * absorption, rotation rho_V, conversion rho_Q/rho_U), four full transfer
 * reruns at q = +h, -h, +h/2, -h/2 (channel 5 bins + v); data[4 ch + k] holds
 * component k (I, Q, U, V) of channel ch. */
static kpo_stokes resp_read(const double* d, int ch) {
    kpo_stokes s = {d[4 * ch], d[4 * ch + 1], d[4 * ch + 2], d[4 * ch + 3]};
    return s;
}
static void resp_write(double* d, int ch, kpo_stokes s) {
    d[4 * ch] = s.I;
    d[4 * ch + 1] = s.Q;
    d[4 * ch + 2] = s.U;
    d[4 * ch + 3] = s.V;
}
static kpo_stokes resp_add(kpo_stokes a, kpo_stokes b) {
    kpo_stokes s = {a.I + b.I, a.Q + b.Q, a.U + b.U, a.V + b.V};
    return s;
}

/* ResponseStepMatrix (:37-53): the step's homogeneous operator, its columns
 * the unit Stokes vectors stepped without emission. */
typedef struct resp_matrix { kpo_stokes col[4]; } resp_matrix;
static void resp_matrix_make(resp_matrix* M, kpo_coeffs c, double dl) {
    c.jI = c.jQ = c.jU = c.jV = 0.0;
    memset(M, 0, sizeof(*M));
    M->col[0].I = 1.0;
    M->col[1].Q = 1.0;
    M->col[2].U = 1.0;
    M->col[3].V = 1.0;
    for (int k = 0; k < 4; ++k) kpo_semi_analytic_step(&M->col[k], &c, dl);
}
static kpo_stokes resp_apply(const resp_matrix* M, kpo_stokes s) {
    const kpo_stokes* c = M->col;
    kpo_stokes o = {c[0].I * s.I + c[1].I * s.Q + c[2].I * s.U + c[3].I * s.V,
                    c[0].Q * s.I + c[1].Q * s.Q + c[2].Q * s.U + c[3].Q * s.V,
                    c[0].U * s.I + c[1].U * s.Q + c[2].U * s.U + c[3].U * s.V,
                    c[0].V * s.I + c[1].V * s.Q + c[2].V * s.U + c[3].V * s.V};
    return o;
}

/* response_replace_block (:55-69). */
static kpo_coeffs resp_replace(kpo_coeffs base, const kpo_coeffs* v, int mechanism) {
    if (mechanism == 0) {
        base.jI = v->jI; base.jQ = v->jQ; base.jU = v->jU; base.jV = v->jV;
    } else if (mechanism == 1) {
        base.aI = v->aI; base.aQ = v->aQ; base.aU = v->aU; base.aV = v->aV;
    } else if (mechanism == 2) {
        base.rV = v->rV;
    } else {
        base.rQ = v->rQ; base.rU = v->rU;
    }
    return base;
}

/* response_scale_coefficients (:71-78). */
void kpo_response_scale_coefficients(const kpo_coeffs* c, double q, kpo_coeffs* out) {
    const double f = exp(q);
    *out = *c;
    out->jI *= f; out->jQ *= f; out->jU *= f; out->jV *= f;
    out->aI *= f; out->aQ *= f; out->aU *= f; out->aV *= f;
    out->rQ *= f; out->rU *= f; out->rV *= f;
}

/* response_transfer_step (:118-149): the source tags and the mechanism
 * derivatives propagate through the step's homogeneous operator, the step's
 * emission enters the sample's bin, each mechanism's forcing is the central
 * difference of the step with that block of coefficients replaced by the
 * variants at +h and -h; the reruns take the full variant steps. */
void kpo_response_transfer_step(int32_t bins, int32_t responses, double step, double* data, int32_t bin,
                                const kpo_stokes* incoming, const kpo_coeffs* base, const kpo_coeffs* variants,
                                double dl) {
    resp_matrix M;
    resp_matrix_make(&M, *base, dl);
    kpo_stokes emitted = {0.0, 0.0, 0.0, 0.0};
    kpo_semi_analytic_step(&emitted, base, dl);
    for (int k = 0; k < bins; ++k) {
        kpo_stokes s = resp_apply(&M, resp_read(data, k));
        if (k == bin) s = resp_add(s, emitted);
        resp_write(data, k, s);
    }
    if (!responses) return;
    for (int m = 0; m < 4; ++m) {
        kpo_stokes plus = *incoming, minus = *incoming;
        kpo_coeffs cp = resp_replace(*base, &variants[0], m), cm = resp_replace(*base, &variants[1], m);
        kpo_semi_analytic_step(&plus, &cp, dl);
        kpo_semi_analytic_step(&minus, &cm, dl);
        const double d = 2.0 * step;
        const kpo_stokes forcing = {(plus.I - minus.I) / d, (plus.Q - minus.Q) / d, (plus.U - minus.U) / d,
                                    (plus.V - minus.V) / d};
        for (int k = 0; k < bins; ++k) {
            const int ch = bins + m * bins + k;
            kpo_stokes s = resp_apply(&M, resp_read(data, ch));
            if (k == bin) s = resp_add(s, forcing);
            resp_write(data, ch, s);
        }
    }
    for (int v = 0; v < 4; ++v) {
        const int ch = 5 * bins + v;
        kpo_stokes s = resp_read(data, ch);
        kpo_semi_analytic_step(&s, &variants[v], dl);
        resp_write(data, ch, s);
    }
}

/* A ray's response request: one radial bin (every sample), the plasma
 * family (kpo_perturbation.parameter) and the difference step h; data holds
 * the ray's 9 channels. */
typedef struct resp_ctx {
    int parameter;
    double step;
    double data[4 * 9];
} resp_ctx;

/* trace_pass_b_segment_model_endpoint_pixel_metric (pass_b.hpp:515-722);
 * rc (single frequency) accumulates the ray's parameter responses as the
 * analysis driver does (analysis_driver.hpp:585-613). */
static void pass_b(const kp_pass_a_params* p, const model_ctx* model,
                   const kpo_perturbation* pert, const freq_plan* fp, const metric_ctx* m,
                   const kpo_state* camera, const kpo_state* endpoint, int pass_a_steps,
                   int pass_a_reason, int valid, pixel_result* out, kpo_stats* stats, resp_ctx* rc) {
    memset(out, 0, sizeof(*out));
    out->r11 = 1.0;
    out->r22 = 1.0;
    out->pass_a_steps = pass_a_steps;
    out->reason = pass_a_reason;
    if (rc) memset(rc->data, 0, sizeof(rc->data));
    kpo_state state = *endpoint;
    if (!valid) {
        out->closure_x = spatial_distance(state.v, camera->v);
        out->closure_k = max_abs_diff4(state.v + 4, camera->v + 4);
        double g[16];
        metric_gcov(m, state.v, g);
        out->final_null = dot_g(g, state.v + 4, state.v + 4);
        out->frame_error = max_frame_error(m, &state);
        return;
    }
    const adaptive_control tcontrol = {p->adaptive_tolerance, p->min_step, p->max_step, 10};
    kpo_stokes stokes[KPO_MAX_FREQUENCIES];
    memset(stokes, 0, sizeof(stokes));
    int reached = 0;
    int steps = 0;
    int reason = pass_a_reason;
    double h_current = p->step;
    double cap = p->max_radiation_step > 0.0 ? p->max_radiation_step : p->max_step;
    while (!reached && steps < p->max_steps) {
        double h = h_current;
        const kpo_state old = state;
        const double eff_min = absv(p->min_step);
        const double start_r = radius_of(m, old.v);
        if (start_r <= p->outer_radius && cap > 0.0) h = minv(h, cap);
        kpo_state next, sample;
        if (p->adaptive) {
            adaptive_control tc = tcontrol;
            tc.min_step = eff_min;
            adaptive_result pr;
            adaptive_rk4(m, &old, h, tc, &pr);
            if (!pr.accepted) {
                h_current = absv(pr.next_h);
                if (h_current <= eff_min * 1.0001) {
                    reason = KP_REASON_ADAPTIVE_STEP_UNDERFLOW;
                    break;
                }
                continue;
            }
            h = absv(pr.used_h);
            h_current = absv(pr.next_h);
            next = pr.state;
            sample = pr.mid_state;
        } else {
            rk4(m, &old, h, &next);
            sample = next;
        }
        const double s0 = radius_of(m, old.v) - p->camera.radius;
        const double s1 = radius_of(m, next.v) - p->camera.radius;
        int crosses = (s0 == 0.0 || s1 == 0.0) ? 1 : ((s0 < 0.0 && s1 > 0.0) || (s0 > 0.0 && s1 < 0.0));
        if (crosses) {
            const double denom = s0 - s1;
            const double frac = absv(denom) <= TINY ? 1.0 : minv(1.0, maxv(0.0, s0 / denom));
            const double h_cross = h * frac;
            if (absv(h_cross) <= 1e-14) {
                h = 0.0;
                next = old;
                sample = old;
            } else if (frac < 0.999999999999) {
                h = h_cross;
                if (p->adaptive) {
                    adaptive_control cc = tcontrol;
                    cc.min_step = minv(eff_min, absv(h));
                    cc.max_step = maxv(cc.min_step, absv(h));
                    adaptive_result cr;
                    adaptive_rk4(m, &old, h, cc, &cr);
                    if (!cr.accepted) {
                        h_current = absv(cr.next_h);
                        continue;
                    }
                    h = absv(cr.used_h);
                    next = cr.state;
                    sample = cr.mid_state;
                } else {
                    rk4(m, &old, h, &next);
                    sample = next;
                }
            }
        }
        const int wedge_n = wedge_step(p, m, &old, &sample, &next, &h, &crosses);
        if (wedge_n > 1) {
            if (h <= eff_min * 1.0001) {
                reason = KP_REASON_ADAPTIVE_STEP_UNDERFLOW;
                break;
            }
            h_current = maxv(eff_min, h / (double)wedge_n);
            cap = h_current;
            continue;
        }
        const double old_r = radius_of(m, old.v), sample_r = radius_of(m, sample.v),
                     next_r = radius_of(m, next.v);
        const int inside = sample_r - p->inner_radius > 0.0 && sample_r < p->outer_radius;
        const double local_max_rs =
            (p->max_radiation_step > 0.0 &&
             (old_r <= p->outer_radius || sample_r <= p->outer_radius || next_r <= p->outer_radius))
                ? p->max_radiation_step
                : 0.0;
        /* step control: the largest substep count over the control frequencies */
        kpo_coeffs cc[KPO_MAX_FREQUENCIES];
        double dls_c[KPO_MAX_FREQUENCIES];
        int required = 1;
        for (int f = 0; f < fp->nctrl; ++f) {
            dls_c[f] = coeffs_at_frequency(p, model, fp->ctrl[f], &sample, inside, pert, &cc[f], stats);
            select_emission(&p->emission_selection, m, &sample, &cc[f]);
            const int n = kpo_radiation_substep_count(&cc[f], h, dls_c[f], 1, local_max_rs,
                                                      p->max_radiation_depth, p->max_absorption_depth,
                                                      p->max_faraday_depth);
            required = n > required ? n : required;
        }
        if (required > 1 && h > eff_min * 1.0001) {
            h_current = maxv(eff_min, h / (double)required);
            cap = h_current;
            continue;
        }
        if (local_max_rs > 0.0 || p->max_radiation_depth > 0.0 || p->max_absorption_depth > 0.0 ||
            p->max_faraday_depth > 0.0) {
            int predictive = 1;
            for (int f = 0; f < fp->nctrl; ++f) {
                const int n = kpo_radiation_substep_count(&cc[f], h_current, dls_c[f], 1, local_max_rs,
                                                          p->max_radiation_depth,
                                                          p->max_absorption_depth,
                                                          p->max_faraday_depth);
                predictive = n > predictive ? n : predictive;
            }
            cap = predictive > 1 ? maxv(eff_min, h_current / (double)predictive) : p->max_step;
        } else {
            cap = p->max_step;
        }
        if (h > 0.0) {
            if (stats) stats->transfer_steps++;
            for (int f = 0; f < fp->nfreq; ++f) {
                kpo_coeffs c;
                double dls;
                if (fp->same) {
                    c = cc[f];
                    dls = dls_c[f];
                } else {
                    dls = coeffs_at_frequency(p, model, fp->freqs[f], &sample, inside, pert, &c, stats);
                    select_emission(&p->emission_selection, m, &sample, &c);
                }
                if (rc && f == 0) {
                    /* accumulate_response_sample (response.hpp:151-182): the
                     * coefficients at q = +h, -h, +h/2, -h/2 of the family */
                    kpo_coeffs var[4];
                    for (int v = 0; v < 4; ++v) {
                        const double q = rc->step * (v % 2 == 0 ? 1.0 : -1.0) * (v < 2 ? 1.0 : 0.5);
                        const kpo_perturbation pv = {rc->parameter, 0, q};
                        coeffs_at_frequency(p, model, fp->freqs[0], &sample, inside, &pv, &var[v], NULL);
                        select_emission(&p->emission_selection, m, &sample, &var[v]);
                    }
                    kpo_response_transfer_step(1, 1, rc->step, rc->data, 0, &stokes[0], &c, var, absv(h) * dls);
                }
                kpo_semi_analytic_step(&stokes[f], &c, absv(h) * dls);
            }
        }
        state = next;
        steps += 1;
        if (crosses) reached = 1;
    }
    /* screen_overlap (pass_b.hpp:87-98) at the final position. */
    double g[16];
    metric_gcov(m, state.v, g);
    out->r11 = dot_g(g, camera->v + 8, state.v + 8);
    out->r12 = dot_g(g, camera->v + 8, state.v + 12);
    out->r21 = dot_g(g, camera->v + 12, state.v + 8);
    out->r22 = dot_g(g, camera->v + 12, state.v + 12);
    for (int f = 0; f < fp->nfreq; ++f)
        kpo_transform_to_observer_basis(&stokes[f], out->r11, out->r12, out->r21, out->r22,
                                        &out->observed[f]);
    if (rc) {
        /* finalize_response_basis (response.hpp:184-191) */
        for (int ch = 0; ch < 9; ++ch) {
            const kpo_stokes in = resp_read(rc->data, ch);
            kpo_stokes o;
            kpo_transform_to_observer_basis(&in, out->r11, out->r12, out->r21, out->r22, &o);
            resp_write(rc->data, ch, o);
        }
    }
    out->closure_x = spatial_distance(state.v, camera->v);
    out->closure_k = max_abs_diff4(state.v + 4, camera->v + 4);
    out->final_null = dot_g(g, state.v + 4, state.v + 4);
    out->frame_error = max_frame_error(m, &state);
    out->steps = steps;
    if (reached) {
        reason = KP_REASON_REACHED_CAMERA;
    } else if (reason != KP_REASON_ADAPTIVE_STEP_UNDERFLOW &&
               reason != KP_REASON_METRIC_TIME_EXHAUSTED) {
        reason = KP_REASON_MAX_STEPS;
    }
    out->reason = reason;
}

/* Pixel outputs at the final state (the window kernel's finalisation,
 * slow_light_driver.hpp:1444-1480, and finalize_slow_light_unfinished_metric,
 * :1725-1766): screen overlap with the camera ray, observer-basis Stokes,
 * closure and frame diagnostics. */
static void slow_finish(const metric_ctx* m, const kpo_state* camera, const kpo_state* state,
                        const kpo_stokes* stokes, int nfreq, pixel_result* out) {
    double g[16];
    metric_gcov(m, state->v, g);
    out->r11 = dot_g(g, camera->v + 8, state->v + 8);
    out->r12 = dot_g(g, camera->v + 8, state->v + 12);
    out->r21 = dot_g(g, camera->v + 12, state->v + 8);
    out->r22 = dot_g(g, camera->v + 12, state->v + 12);
    for (int f = 0; f < nfreq; ++f)
        kpo_transform_to_observer_basis(&stokes[f], out->r11, out->r12, out->r21, out->r22, &out->observed[f]);
    out->closure_x = spatial_distance(state->v, camera->v);
    out->closure_k = max_abs_diff4(state->v + 4, camera->v + 4);
    out->final_null = dot_g(g, state->v + 4, state->v + 4);
    out->frame_error = max_frame_error(m, state);
}

/* advance_slow_light_vacuum_to_camera (slow_light_driver.hpp:1572-1671): a
 * ray whose fluid time ran out may finish only through outward vacuum. */
static int slow_vacuum_to_camera(const kp_pass_a_params* p, const metric_ctx* m, kpo_state* state,
                                 double* h_current, int* steps) {
    const double tol_r = 1e-8 * maxv(1.0, absv(p->outer_radius));
    if (radius_of(m, state->v) < p->outer_radius - tol_r) return KP_REASON_SLOW_LIGHT_TIME_EXHAUSTED;
    const adaptive_control control = {p->adaptive_tolerance, p->min_step, p->max_step, 10};
    while (*steps < p->max_steps) {
        double h = *h_current;
        const kpo_state old = *state;
        kpo_state next, sample;
        if (p->adaptive) {
            adaptive_result pr;
            adaptive_rk4(m, &old, h, control, &pr);
            if (!pr.accepted) {
                *h_current = absv(pr.next_h);
                if (*h_current <= p->min_step * 1.0001) return KP_REASON_MAX_STEPS;
                continue;
            }
            h = absv(pr.used_h);
            *h_current = absv(pr.next_h);
            next = pr.state;
            sample = pr.mid_state;
        } else {
            rk4(m, &old, h, &next);
            sample = next;
        }
        const double s0 = radius_of(m, old.v) - p->camera.radius;
        const double s1 = radius_of(m, next.v) - p->camera.radius;
        const int crosses = (s0 == 0.0 || s1 == 0.0) ? 1 : ((s0 < 0.0 && s1 > 0.0) || (s0 > 0.0 && s1 < 0.0));
        if (crosses) {
            const double denom = s0 - s1;
            const double frac = absv(denom) <= TINY ? 1.0 : minv(1.0, maxv(0.0, s0 / denom));
            const double h_cross = h * frac;
            if (absv(h_cross) <= 1e-14) {
                h = 0.0;
                next = old;
                sample = old;
            } else if (frac < 0.999999999999) {
                h = h_cross;
                if (p->adaptive) {
                    adaptive_control cc = control;
                    cc.min_step = minv(cc.min_step, absv(h));
                    cc.max_step = maxv(cc.min_step, absv(h));
                    adaptive_result cr;
                    adaptive_rk4(m, &old, h, cc, &cr);
                    if (!cr.accepted) {
                        *h_current = absv(cr.next_h);
                        continue;
                    }
                    next = cr.state;
                    sample = cr.mid_state;
                } else {
                    rk4(m, &old, h, &next);
                    sample = next;
                }
            }
        }
        const double old_r = radius_of(m, old.v), next_r = radius_of(m, next.v), sample_r = radius_of(m, sample.v);
        if (!isfinite(next_r) || !isfinite(sample_r)) return KP_REASON_MAX_STEPS;
        if (next_r < old_r - tol_r || sample_r < p->outer_radius - tol_r || next_r < p->outer_radius - tol_r)
            return KP_REASON_SLOW_LIGHT_TIME_EXHAUSTED;
        *state = next;
        *steps += 1;
        if (crosses) return KP_REASON_REACHED_CAMERA;
    }
    return KP_REASON_MAX_STEPS;
}

/* One pixel of a slow-light image: the window kernel
 * (launch_slow_light_rays_metric, slow_light_driver.hpp:1083-1483) applied to
 * every window in turn -- each ray's windows are independent of every other
 * ray's, so one pass over them per ray equals the per-window launches -- then
 * the vacuum finish of a ray still unfinished after the last window
 * (finalize_slow_light_unfinished_metric, :1674-1769).  Windows: snapshot
 * stepping runs one window per snapshot pair (k, k+1), cuts steps at t_k+1 and
 * samples that pair (SlowLightSingleWindow, :966-981); block stepping runs one
 * window cut at the last time and samples the pair enclosing each sample's
 * time; decoupled stepping cuts nothing and a radiating sample outside the
 * sequence leaves the ray waiting (SlowLightResidentWindows, :983-1027).
 * With windows_per_block W (0 < W < nsnap - 1) the batch scheduler
 * (slow_light_batch_driver.hpp:211-340) runs the sequence as blocks of W
 * snapshot pairs, [bW, min((b+1)W, nsnap - 1)]: block stepping then runs one
 * window per block, cut at the block's last time and sampling the pair of
 * that block enclosing each sample's time; snapshot and decoupled stepping
 * give the same image as one block (a waiting decoupled ray resumes with the
 * same call in the block holding its time). */
static void pass_b_slow(const kp_pass_a_params* p, const kp_grmhd_model* g, const kp_slow_light* seq,
                        const kpo_perturbation* pert, const freq_plan* fp, const metric_ctx* m,
                        const kpo_state* camera,
                        const kpo_state* endpoint, int pass_a_steps, int pass_a_reason, int valid,
                        pixel_result* out, kpo_stats* stats) {
    memset(out, 0, sizeof(*out));
    out->r11 = 1.0;
    out->r22 = 1.0;
    out->pass_a_steps = pass_a_steps;
    out->reason = pass_a_reason;
    kpo_state state = *endpoint;
    if (!valid) {
        /* initialize_slow_light_states_metric (:362-458) */
        out->closure_x = spatial_distance(state.v, camera->v);
        out->closure_k = max_abs_diff4(state.v + 4, camera->v + 4);
        double gg[16];
        metric_gcov(m, state.v, gg);
        out->final_null = dot_g(gg, state.v + 4, state.v + 4);
        out->frame_error = max_frame_error(m, &state);
        return;
    }
    const int mode = seq->step_mode;
    const int decoupled = mode == KP_SLOW_LIGHT_DECOUPLED;
    const double* T = seq->times;
    const int n = seq->nsnap;
    const double t_obs = seq->observation_time;
    const adaptive_control control = {p->adaptive_tolerance, p->min_step, p->max_step, 10};
    kpo_stokes stokes[KPO_MAX_FREQUENCIES];
    memset(stokes, 0, sizeof(stokes));
    double h_current = p->step;
    double cap = p->max_radiation_step > 0.0 ? p->max_radiation_step : p->max_step;
    int steps = 0, reached = 0, active = 1, reason = pass_a_reason;
    const int wpb = seq->windows_per_block;
    const int blocked = mode == KP_SLOW_LIGHT_BLOCK && wpb > 0 && wpb < n - 1;
    const int nwin = mode == KP_SLOW_LIGHT_SNAPSHOT ? n - 1 : (blocked ? (n - 1 + wpb - 1) / wpb : 1);
    for (int w = 0; w < nwin; ++w) {
        if (reached || !active) break;
        const int blo = blocked ? w * wpb : 0;
        const int bhi = blocked ? (w * wpb + wpb < n - 1 ? w * wpb + wpb : n - 1) : n - 1;
        const double upper = mode == KP_SLOW_LIGHT_SNAPSHOT ? T[w + 1] : T[bhi];
        while (!reached && steps < p->max_steps) {
            const double retry_h = h_current;
            double h = h_current;
            const kpo_state old = state;
            const double old_t = t_obs + old.v[0];
            if (!decoupled && old_t >= upper) break;
            const double start_r = radius_of(m, old.v);
            if (start_r <= p->outer_radius && cap > 0.0) h = minv(h, cap);
            kpo_state next, sample;
            if (p->adaptive) {
                adaptive_result pr;
                adaptive_rk4(m, &old, h, control, &pr);
                if (!pr.accepted) {
                    h_current = absv(pr.next_h);
                    if (h_current <= p->min_step * 1.0001) {
                        reason = KP_REASON_MAX_STEPS;
                        active = 0;
                        break;
                    }
                    continue;
                }
                h = absv(pr.used_h);
                h_current = absv(pr.next_h);
                next = pr.state;
                sample = pr.mid_state;
            } else {
                rk4(m, &old, h, &next);
                sample = next;
            }
            const double s0 = radius_of(m, old.v) - p->camera.radius;
            const double s1 = radius_of(m, next.v) - p->camera.radius;
            int crosses = (s0 == 0.0 || s1 == 0.0) ? 1 : ((s0 < 0.0 && s1 > 0.0) || (s0 > 0.0 && s1 < 0.0));
            double camera_frac = 2.0;
            if (crosses) {
                const double denom = s0 - s1;
                camera_frac = absv(denom) <= TINY ? 1.0 : minv(1.0, maxv(0.0, s0 / denom));
            }
            const double next_t = t_obs + next.v[0];
            const int crosses_time = !decoupled && next_t > upper && next_t > old_t;
            double time_frac = 2.0;
            if (crosses_time) time_frac = minv(1.0, maxv(0.0, (upper - old_t) / (next_t - old_t)));
            if (crosses_time && (!crosses || time_frac < camera_frac)) {
                crosses = 0;
                const double h_b = h * time_frac;
                if (absv(h_b) <= 1e-14) break;
                h = h_b;
                if (p->adaptive) {
                    adaptive_control bc = control;
                    bc.min_step = minv(bc.min_step, absv(h));
                    bc.max_step = maxv(bc.min_step, absv(h));
                    adaptive_result br;
                    adaptive_rk4(m, &old, h, bc, &br);
                    if (!br.accepted) {
                        h_current = absv(br.next_h);
                        continue;
                    }
                    h = absv(br.used_h);
                    next = br.state;
                    sample = br.mid_state;
                } else {
                    rk4(m, &old, h, &next);
                    sample = next;
                }
                if (next.v[0] == old.v[0]) {
                    const double time_scale = maxv(absv(old.v[0]), maxv(absv(upper), absv(old_t - old.v[0])));
                    const double rounding = 0.5 * 8.0 * 2.220446049250313080847263336181640625e-16 * time_scale;
                    if (absv(upper - old_t) > rounding) {
                        reason = KP_REASON_ADAPTIVE_STEP_UNDERFLOW;
                        active = 0;
                    }
                    break;
                }
            } else if (crosses) {
                const double h_cross = h * camera_frac;
                if (absv(h_cross) <= 1e-14) {
                    h = 0.0;
                    next = old;
                    sample = old;
                } else if (camera_frac < 0.999999999999) {
                    h = h_cross;
                    if (p->adaptive) {
                        adaptive_control cc = control;
                        cc.min_step = minv(cc.min_step, absv(h));
                        cc.max_step = maxv(cc.min_step, absv(h));
                        adaptive_result cr;
                        adaptive_rk4(m, &old, h, cc, &cr);
                        if (!cr.accepted) {
                            h_current = absv(cr.next_h);
                            continue;
                        }
                        h = absv(cr.used_h);
                        next = cr.state;
                        sample = cr.mid_state;
                    } else {
                        rk4(m, &old, h, &next);
                        sample = next;
                    }
                }
            }
            /* the wedge (slow_light_driver.hpp:1272-1291) */
            const int wedge_n = wedge_step(p, m, &old, &sample, &next, &h, &crosses);
            if (wedge_n > 1) {
                if (h <= p->min_step * 1.0001) {
                    reason = KP_REASON_ADAPTIVE_STEP_UNDERFLOW;
                    active = 0;
                    break;
                }
                h_current = maxv(p->min_step, h / (double)wedge_n);
                cap = h_current;
                continue;
            }
            const double old_r = radius_of(m, old.v), sample_r = radius_of(m, sample.v),
                         next_r = radius_of(m, next.v);
            const int radiating = sample_r >= p->inner_radius && sample_r < p->outer_radius;
            const double ct = t_obs + sample.v[0];
            if (decoupled && radiating && !(ct >= T[0] && ct <= T[n - 1])) {
                out->requested_time = ct;
                out->waiting = 1;
                h_current = retry_h;
                break;
            }
            const int pair = mode == KP_SLOW_LIGHT_SNAPSHOT ? w : blo + kpo_slow_light_pair(T + blo, bhi - blo + 1, ct);
            const double local_max_rs =
                (p->max_radiation_step > 0.0 &&
                 (old_r <= p->outer_radius || sample_r <= p->outer_radius || next_r <= p->outer_radius))
                    ? p->max_radiation_step
                    : 0.0;
            kpo_coeffs cc[KPO_MAX_FREQUENCIES];
            double dls[KPO_MAX_FREQUENCIES];
            int required = 1, predictive = 1;
            for (int f = 0; f < fp->nfreq; ++f) {
                kp_grmhd_model gf = *g;
                gf.freq_cgs = fp->freqs[f];
                dls[f] = kpo_grmhd_dlambda_scale(&gf);
                memset(&cc[f], 0, sizeof(cc[f]));
                if (radiating) {
                    if (stats) stats->coefficient_evals++;
                    kpo_slow_light_coefficients(&gf, seq, pair, &sample, pert,
                                                p->emission_selection.faraday_rotation, &cc[f]);
                    select_emission(&p->emission_selection, m, &sample, &cc[f]);
                }
                const int a = kpo_radiation_substep_count(&cc[f], h, dls[f], 1, local_max_rs, p->max_radiation_depth,
                                                          p->max_absorption_depth, p->max_faraday_depth);
                required = a > required ? a : required;
                const int b = kpo_radiation_substep_count(&cc[f], h_current, dls[f], 1, local_max_rs,
                                                          p->max_radiation_depth, p->max_absorption_depth,
                                                          p->max_faraday_depth);
                predictive = b > predictive ? b : predictive;
            }
            if (required > 1 && h > p->min_step * 1.0001) {
                h_current = maxv(p->min_step, h / (double)required);
                cap = h_current;
                continue;
            }
            cap = predictive > 1 ? maxv(p->min_step, h_current / (double)predictive) : p->max_step;
            if (h > 0.0) {
                if (stats) stats->transfer_steps++;
                for (int f = 0; f < fp->nfreq; ++f) kpo_semi_analytic_step(&stokes[f], &cc[f], absv(h) * dls[f]);
            }
            state = next;
            steps += 1;
            if (crosses) reached = 1;
        }
        if (out->waiting) break;
        if (reached || steps >= p->max_steps || !active) {
            slow_finish(m, camera, &state, stokes, fp->nfreq, out);
            out->steps = steps;
            if (reached)
                out->reason = KP_REASON_REACHED_CAMERA;
            else if (reason != KP_REASON_ADAPTIVE_STEP_UNDERFLOW)
                out->reason = KP_REASON_MAX_STEPS;
            else
                out->reason = reason;
            return;
        }
    }
    if (out->waiting) {
        /* all snapshots are resident: no later block supplies the time */
        slow_finish(m, camera, &state, stokes, fp->nfreq, out);
        out->steps = steps;
        out->reason = KP_REASON_SLOW_LIGHT_TIME_EXHAUSTED;
        return;
    }
    const int vreason = slow_vacuum_to_camera(p, m, &state, &h_current, &steps);
    slow_finish(m, camera, &state, stokes, fp->nfreq, out);
    out->steps = steps;
    out->reason = vreason;
}

/* Stokes arrays are [nfreq][nout] (frequency-major, src/image/multifrequency.hpp). */
static void store_pixel(kp_image_data* o, int p, int nout, int nfreq, const pixel_result* r) {
#define PUT(field, val) do { if (o->field) o->field[p] = (val); } while (0)
    for (int f = 0; f < nfreq; ++f) {
        const size_t q = (size_t)f * (size_t)nout + (size_t)p;
        if (o->image_i) o->image_i[q] = r->observed[f].I;
        if (o->image_q) o->image_q[q] = r->observed[f].Q;
        if (o->image_u) o->image_u[q] = r->observed[f].U;
        if (o->image_v) o->image_v[q] = r->observed[f].V;
    }
    PUT(closure_x, r->closure_x);
    PUT(closure_k, r->closure_k);
    PUT(final_null, r->final_null);
    PUT(frame_error, r->frame_error);
    PUT(det_r, r->r11 * r->r22 - r->r12 * r->r21);
    PUT(overlap_r11, r->r11);
    PUT(overlap_r12, r->r12);
    PUT(overlap_r21, r->r21);
    PUT(overlap_r22, r->r22);
    PUT(basis_identity_error, maxv(maxv(absv(r->r11 - 1.0), absv(r->r22 - 1.0)),
                                   maxv(absv(r->r12), absv(r->r21))));
    PUT(basis_rotation_angle, atan2(r->r21 - r->r12, r->r11 + r->r22));
    PUT(pass_a_steps, r->pass_a_steps);
    PUT(steps, r->steps);
    PUT(reason, r->reason);
#undef PUT
}

/* One pixel set: pixel = first + i * stride for i < count, outputs at index
 * i (compact) or at the pixel index.  Split pass A / pass B per pixel
 * (src/image/split_driver.hpp:41-104); OpenMP static schedule as the
 * reference's RangePolicy. */
static kp_status run_pixels_w(const kp_pass_a_params* p, const model_ctx* model,
                              const kpo_perturbation* pert, const freq_plan* fp, int first, int stride,
                              int count, int compact, int nout, int nthreads, kp_image_data* out,
                              kpo_stats* stats, double* waiting, const resp_ctx* resp_cfg,
                              double* resp_out) {
    double tetrad[16], gcam[16];
    double xcam[4] = {0.0, p->camera.radius, p->camera.inclination, 0.0};
    const int fmks = p->coordinate_system == KP_COORD_FMKS;
    const kpo_fmks fm = {p->fmks_startx1, p->fmks_hslope, p->fmks_mks_smooth, p->fmks_poly_alpha,
                         p->fmks_poly_xt, p->fmks_poly_norm};
    if (fmks) {
        kpo_fmks_camera_tetrad(p, &fm, xcam, tetrad);
        kpo_fmks_gcov(p->mass, p->spin, &fm, xcam, gcam);
    } else {
        kpo_bl_camera_tetrad(p, tetrad);
        kpo_bl_gcov(p->mass, p->spin, xcam, gcam);
    }
    int64_t rhs_total = 0, coeff_total = 0, transfer_total = 0, a_total = 0, b_total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(+ : rhs_total, coeff_total, transfer_total, a_total, b_total)
#endif
    for (int i = 0; i < count; ++i) {
        const int pix = first + i * stride;
        int64_t rhs_count = 0;
        metric_ctx m = {p->mass, p->spin, &rhs_count, fmks, &fm};
        kpo_stats local = {0, 0, 0, 0, 0};
        kpo_state camera, endpoint;
        camera_ray_from_tetrad(p, xcam, tetrad, gcam, pix, &camera);
        int32_t a_steps = 0, a_reason = 0;
        const int valid = pass_a_endpoint(p, &m, &camera, &endpoint, &a_steps, &a_reason);
        pixel_result r;
        resp_ctx rc;
        if (resp_cfg) rc = *resp_cfg;
        if (model->slow)
            pass_b_slow(p, model->grmhd, model->slow, pert, fp, &m, &camera, &endpoint, a_steps, a_reason, valid,
                        &r, &local);
        else
            pass_b(p, model, pert, fp, &m, &camera, &endpoint, a_steps, a_reason, valid, &r, &local,
                   resp_cfg ? &rc : NULL);
        if (resp_cfg && resp_out) {
            /* the reference's layout: component k of channel ch at (4 ch + k) npix + pixel */
            const int o = compact ? i : pix;
            for (int q = 0; q < 36; ++q) resp_out[(size_t)q * (size_t)nout + (size_t)o] = rc.data[q];
        }
        if (r.waiting) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                if (waiting) {
                    waiting[0] += 1.0;
                    waiting[1] = minv(waiting[1], r.requested_time);
                    waiting[2] = maxv(waiting[2], r.requested_time);
                }
            }
        }
        store_pixel(out, compact ? i : pix, nout, fp->nfreq, &r);
        rhs_total += rhs_count;
        coeff_total += local.coefficient_evals;
        transfer_total += local.transfer_steps;
        a_total += a_steps;
        b_total += r.steps;
    }
    if (stats) {
        stats->rhs_evals = rhs_total;
        stats->coefficient_evals = coeff_total;
        stats->transfer_steps = transfer_total;
        stats->pass_a_steps = a_total;
        stats->pass_b_steps = b_total;
    }
    return KP_OK;
}

static kp_status run_pixels(const kp_pass_a_params* p, const model_ctx* model,
                            const kpo_perturbation* pert, const freq_plan* fp, int first, int stride,
                            int count, int compact, int nout, int nthreads, kp_image_data* out,
                            kpo_stats* stats) {
    return run_pixels_w(p, model, pert, fp, first, stride, count, compact, nout, nthreads, out, stats, NULL, NULL,
                        NULL);
}

static kp_status check_params(const kp_pass_a_params* p, const kp_riaf_model* model, const void* out) {
    if (!p || !model || !out) return KP_INVALID_ARGUMENT;
    if (p->coordinate_system != KP_COORD_BOYER_LINDQUIST || p->camera.model != KP_CAMERA_PINHOLE)
        return KP_UNSUPPORTED;
    return KP_OK;
}

kp_status kpo_run_riaf_image(const kp_pass_a_params* p, const kp_riaf_model* model,
                             const kpo_perturbation* pert, int32_t pixel_begin,
                             int32_t pixel_end, int32_t nthreads, kp_image_data* out,
                             kpo_stats* stats) {
    const kp_status st = check_params(p, model, out);
    if (st != KP_OK) return st;
    const int npix = p->camera.nx * p->camera.ny;
    const int begin = pixel_begin < 0 ? 0 : pixel_begin;
    const int end = (pixel_end < 0 || pixel_end > npix) ? npix : pixel_end;
    const double f0 = model->freq_cgs;
    const freq_plan fp = {&f0, 1, &f0, 1, 1};
    const model_ctx mc = {model, NULL, NULL};
    return run_pixels(p, &mc, pert, &fp, begin, 1, end > begin ? end - begin : 0, 0, npix, nthreads,
                      out, stats);
}

/* The RIAF image with the parameter responses of one plasma family
 * (parameter 1 density, 2 temperature, 3 magnetic; PlasmaPerturbation,
 * plasma_perturbation.hpp:12-33) at difference step h, one radial bin: resp
 * [36][npix] holds the reference's channels (source, four mechanism
 * derivatives, reruns at +h, -h, +h/2, -h/2) component-major. */
kp_status kpo_run_riaf_response_image(const kp_pass_a_params* p, const kp_riaf_model* model, int32_t parameter,
                                      double step, int32_t nthreads, kp_image_data* out, double* resp,
                                      kpo_stats* stats) {
    const kp_status st = check_params(p, model, out);
    if (st != KP_OK) return st;
    if (parameter < 1 || parameter > 3 || !(step > 0.0) || !resp) return KP_INVALID_ARGUMENT;
    const int npix = p->camera.nx * p->camera.ny;
    const double f0 = model->freq_cgs;
    const freq_plan fp = {&f0, 1, &f0, 1, 1};
    const model_ctx mc = {model, NULL, NULL};
    const kpo_perturbation none = {0, 0, 0.0};
    resp_ctx cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.parameter = parameter;
    cfg.step = step;
    return run_pixels_w(p, &mc, &none, &fp, 0, 1, npix, 0, npix, nthreads, out, stats, NULL, &cfg, resp);
}

/* Strided pixel sample pixel = offset + i * stride, i < count, outputs compact
 * (index i): bounded CPU baselines and spot checks of full-size images. */
kp_status kpo_run_riaf_image_strided(const kp_pass_a_params* p, const kp_riaf_model* model,
                                     const kpo_perturbation* pert, int32_t offset, int32_t stride,
                                     int32_t count, int32_t nthreads, kp_image_data* out,
                                     kpo_stats* stats) {
    kp_status st = check_params(p, model, out);
    if (st != KP_OK) return st;
    if (stride <= 0 || count < 0 || offset < 0) return KP_INVALID_ARGUMENT;
    const int64_t npix = (int64_t)p->camera.nx * p->camera.ny;
    if ((int64_t)offset + (int64_t)(count > 0 ? count - 1 : 0) * stride >= npix && count > 0)
        return KP_INVALID_ARGUMENT;
    const double f0 = model->freq_cgs;
    const freq_plan fp = {&f0, 1, &f0, 1, 1};
    const model_ctx mc = {model, NULL, NULL};
    return run_pixels(p, &mc, pert, &fp, offset, stride, count, 1, count, nthreads, out, stats);
}

/* Multi-frequency image on shared geodesics: Stokes outputs [nfreq][npix]
 * (trace_pass_b_segment_model_multifrequency_pixel_metric, pass_b.hpp:1130-1500,
 * and the separate step-control variant, pass_b.hpp:1505-1895; driver
 * src/image/multifrequency.hpp:152-280). */
kp_status kpo_run_riaf_multifrequency_image(const kp_pass_a_params* p, const kp_riaf_model* model,
                                            const kpo_perturbation* pert, const double* freqs,
                                            int32_t nfreq, const double* ctrl, int32_t nctrl,
                                            int32_t pixel_begin, int32_t pixel_end, int32_t nthreads,
                                            kp_image_data* out, kpo_stats* stats) {
    const kp_status st = check_params(p, model, out);
    if (st != KP_OK) return st;
    if (!freqs || nfreq < 1 || nfreq > KPO_MAX_FREQUENCIES || !ctrl || nctrl < 1 ||
        nctrl > KPO_MAX_FREQUENCIES)
        return KP_INVALID_ARGUMENT;
    int same = nctrl == nfreq;
    for (int f = 0; same && f < nfreq; ++f) same = freqs[f] == ctrl[f];
    const freq_plan fp = {freqs, nfreq, ctrl, nctrl, same};
    const int npix = p->camera.nx * p->camera.ny;
    const int begin = pixel_begin < 0 ? 0 : pixel_begin;
    const int end = (pixel_end < 0 || pixel_end > npix) ? npix : pixel_end;
    const model_ctx mc = {model, NULL, NULL};
    return run_pixels(p, &mc, pert, &fp, begin, 1, end > begin ? end - begin : 0, 0, npix, nthreads,
                      out, stats);
}

/* GRMHD image on FMKS geodesics (src/image/grmhd.cpp:132-173 FMKS backend),
 * one or more frequencies; g->cells are HOST kp_grmhd_cell records with the
 * derived scalars filled (kpo_grmhd_derive_scalars). */
kp_status kpo_run_grmhd_image(const kp_pass_a_params* p, const kp_grmhd_model* g, const kpo_perturbation* pert,
                              const double* freqs, int32_t nfreq, const double* ctrl, int32_t nctrl,
                              int32_t pixel_begin, int32_t pixel_end, int32_t nthreads, kp_image_data* out,
                              kpo_stats* stats) {
    if (!p || !g || !out || !g->cells) return KP_INVALID_ARGUMENT;
    if (p->coordinate_system != KP_COORD_FMKS || p->camera.model != KP_CAMERA_PINHOLE) return KP_UNSUPPORTED;
    const double f0 = g->freq_cgs;
    if (!freqs) {
        freqs = &f0;
        nfreq = 1;
    }
    if (!ctrl) {
        ctrl = freqs;
        nctrl = nfreq;
    }
    if (nfreq < 1 || nfreq > KPO_MAX_FREQUENCIES || nctrl < 1 || nctrl > KPO_MAX_FREQUENCIES)
        return KP_INVALID_ARGUMENT;
    int same = nctrl == nfreq;
    for (int f = 0; same && f < nfreq; ++f) same = freqs[f] == ctrl[f];
    const freq_plan fp = {freqs, nfreq, ctrl, nctrl, same};
    const int npix = p->camera.nx * p->camera.ny;
    const int begin = pixel_begin < 0 ? 0 : pixel_begin;
    const int end = (pixel_end < 0 || pixel_end > npix) ? npix : pixel_end;
    const model_ctx mc = {NULL, g, NULL};
    return run_pixels(p, &mc, pert, &fp, begin, 1, end > begin ? end - begin : 0, 0, npix, nthreads, out,
                      stats);
}

kp_status kpo_run_grmhd_slow_light_image(const kp_pass_a_params* p, const kp_grmhd_model* g,
                                         const kp_slow_light* seq, const kpo_perturbation* pert,
                                         const double* freqs, int32_t nfreq,
                                         int32_t pixel_begin, int32_t pixel_end, int32_t nthreads,
                                         kp_image_data* out, kpo_stats* stats, double waiting[3]) {
    if (!p || !g || !seq || !out || !seq->times || !seq->cells || seq->nsnap < 2) return KP_INVALID_ARGUMENT;
    if (p->coordinate_system != KP_COORD_FMKS || p->camera.model != KP_CAMERA_PINHOLE) return KP_UNSUPPORTED;
    for (int k = 0; k + 1 < seq->nsnap; ++k)
        if (!(seq->times[k + 1] > seq->times[k])) return KP_INVALID_ARGUMENT;
    const double f0 = g->freq_cgs;
    if (!freqs) {
        freqs = &f0;
        nfreq = 1;
    }
    if (nfreq < 1 || nfreq > KPO_MAX_FREQUENCIES) return KP_INVALID_ARGUMENT;
    const freq_plan fp = {freqs, nfreq, freqs, nfreq, 1};
    const int npix = p->camera.nx * p->camera.ny;
    const int begin = pixel_begin < 0 ? 0 : pixel_begin;
    const int end = (pixel_end < 0 || pixel_end > npix) ? npix : pixel_end;
    const model_ctx mc = {NULL, g, seq};
    if (waiting) {
        waiting[0] = 0.0;
        waiting[1] = LARGE;
        waiting[2] = -LARGE;
    }
    const kp_status st = run_pixels_w(p, &mc, pert, &fp, begin, 1, end > begin ? end - begin : 0, 0, npix,
                                      nthreads, out, stats, waiting, NULL, NULL);
    if (st != KP_OK) return st;
    return (waiting && waiting[0] > 0.0) ? KP_TIME_SUPPORT : KP_OK;
}
