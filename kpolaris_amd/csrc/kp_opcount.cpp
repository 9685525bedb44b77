// kp_opcount.cpp -- algorithmic FLOP counts of the kernel's work units.
//
// Instantiates the same templates the HIP kernels run (kp_device.h,
// kp_riaf.h) on the host with a counting scalar type: add/sub/mul/div/sqrt
// count 1 each, a*b+c counts 2 (a mul and an add), each transcendental call
// counts 1 and is also tallied separately.  Expressions whose operands are all
// literals are compile-time constants and are not counted.  Prints JSON with
// per-unit counts: one transport_rhs, one step-doubling attempt, one RK4 step,
// one RIAF coefficient evaluation (Pandya and Dexter, averaged over sample
// states), one Strang transfer step and one substep-count evaluation; for
// GRMHD grids one FMKS transport_rhs, one FMKS step-doubling attempt and one
// thermal coefficient evaluation of a sampled plasma state at one frequency
// (Dexter and Pandya, averaged over eight frequencies 86-345 GHz and a few
// plasma states; the grid sample and four-vector reconstruction are not
// counted).
#include <cmath>
#include <cstdio>

struct OpTally {
    long add = 0, mul = 0, div = 0, sqrt = 0, trans = 0;
    long flops() const { return add + mul + div + sqrt + trans; }
};
static OpTally g_ops;

struct CD {
    double v;
    bool lit;
    CD() : v(0.0), lit(true) {}
    CD(double x) : v(x), lit(true) {}
    CD(int x) : v(double(x)), lit(true) {}
    static CD run(double x) {
        CD c(x);
        c.lit = false;
        return c;
    }
};

static inline CD mk(double v, bool lit) {
    CD c(v);
    c.lit = lit;
    return c;
}
static inline CD operator+(CD a, CD b) { if (!(a.lit && b.lit)) ++g_ops.add; return mk(a.v + b.v, a.lit && b.lit); }
static inline CD operator-(CD a, CD b) { if (!(a.lit && b.lit)) ++g_ops.add; return mk(a.v - b.v, a.lit && b.lit); }
static inline CD operator*(CD a, CD b) { if (!(a.lit && b.lit)) ++g_ops.mul; return mk(a.v * b.v, a.lit && b.lit); }
static inline CD operator/(CD a, CD b) { if (!(a.lit && b.lit)) ++g_ops.div; return mk(a.v / b.v, a.lit && b.lit); }
static inline CD operator-(CD a) { return mk(-a.v, a.lit); }
static inline CD& operator+=(CD& a, CD b) { a = a + b; return a; }
static inline CD& operator-=(CD& a, CD b) { a = a - b; return a; }
static inline CD& operator*=(CD& a, CD b) { a = a * b; return a; }
static inline CD& operator/=(CD& a, CD b) { a = a / b; return a; }
static inline bool operator<(CD a, CD b) { return a.v < b.v; }
static inline bool operator>(CD a, CD b) { return a.v > b.v; }
static inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
static inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
static inline bool operator==(CD a, CD b) { return a.v == b.v; }
static inline bool operator!=(CD a, CD b) { return a.v != b.v; }

#define KP_T1(name, fn)                                        \
    static inline CD name(CD a) {                              \
        if (!a.lit) ++g_ops.trans;                             \
        return mk(fn(a.v), a.lit);                             \
    }
KP_T1(exp, std::exp)
KP_T1(log, std::log)
KP_T1(sin, std::sin)
KP_T1(cos, std::cos)
KP_T1(cbrt, std::cbrt)
KP_T1(tanh, std::tanh)
KP_T1(tan, std::tan)
KP_T1(acos, std::acos)
#undef KP_T1
static inline CD sqrt(CD a) {
    if (!a.lit) ++g_ops.sqrt;
    return mk(std::sqrt(a.v), a.lit);
}
static inline CD pow(CD a, CD b) {
    if (!(a.lit && b.lit)) ++g_ops.trans;
    return mk(std::pow(a.v, b.v), a.lit && b.lit);
}
static inline CD atan2(CD a, CD b) {
    if (!(a.lit && b.lit)) ++g_ops.trans;
    return mk(std::atan2(a.v, b.v), a.lit && b.lit);
}
static inline void sincos(CD a, CD* s, CD* c) {
    if (!a.lit) g_ops.trans += 2;
    *s = mk(std::sin(a.v), a.lit);
    *c = mk(std::cos(a.v), a.lit);
}

#include "kp_device.h"
#include "kp_riaf.h"
#include "kp_fmks.h"

// the kernels' bounded-argument sine/cosine pair counts as two transcendentals
KP_DEV void ksincos(CD a, CD* s, CD* c) { sincos(a, s, c); }
KP_DEV void lsincos(CD a, CD* s, CD* c) { sincos(a, s, c); }
// a reciprocal or quotient counts as one division
KP_DEV CD krcp(CD x) { return CD(1.0) / x; }
KP_DEV CD kdiv(CD a, CD b) { return a / b; }

namespace kpd {
template <> KP_DEV CD tiny<CD>() { return CD(1e-300); }
template <> KP_DEV CD large<CD>() { return CD(1e300); }
template <> KP_DEV CD eps<CD>() { return CD(2.220446049250313080847263336181640625e-16); }
}  // namespace kpd

using namespace kpd;

static St<CD> ray_state(double r, double th, double a) {
    // A plausible null-ish state; the counts of these branch-free units do
    // not depend on the values.
    St<CD> s;
    const double v[16] = {0.0, r, th, 0.3, 1.2, -0.9, 0.01, 0.02, 0.0, 0.01, 0.0, 0.0, 0.0, 0.0, 0.05, 0.1};
    for (int i = 0; i < 16; ++i) s.v[i] = CD::run(v[i]);
    (void)a;
    return s;
}

static void emit(const char* name, const OpTally& t, double per, bool last = false) {
    std::printf("  \"%s\": {\"flops\": %.1f, \"add\": %.1f, \"mul\": %.1f, \"div\": %.1f, \"sqrt\": %.1f, "
                "\"transcendental\": %.1f}%s\n",
                name, t.flops() / per, t.add / per, t.mul / per, t.div / per, t.sqrt / per, t.trans / per,
                last ? "" : ",");
}

int main() {
    uint32_t rhs_count = 0;
    const CD m = CD::run(1.0), a = CD::run(0.94);
    Geo<CD> G{m, a, &rhs_count};
    std::printf("{\n");
    {
        g_ops = OpTally();
        St<CD> s = ray_state(7.0, 1.3, 0.94), d;
        bl_rhs(m, a, s, d);
        emit("rhs", g_ops, 1.0);
    }
    {
        g_ops = OpTally();
        St<CD> s = ray_state(7.0, 1.3, 0.94), out;
        rk4(G, s, CD::run(0.1), out);
        emit("rk4_step", g_ops, 1.0);
    }
    {
        g_ops = OpTally();
        St<CD> s = ray_state(7.0, 1.3, 0.94), half, two;
        CD err = doubling_attempt(G, s, CD::run(0.1), half, two);
        (void)err;
        emit("doubling_attempt", g_ops, 1.0);
    }
    for (int fit : {1, 4}) {
        RiafDev<CD> M;
        M.nth0 = CD::run(1.0); M.Te0 = CD::run(1.0); M.disk_h = CD::run(0.35); M.pow_nth = CD::run(-1.5);
        M.pow_T = CD::run(-0.84); M.ne_unit = CD::run(2e5); M.thetae_unit = CD::run(84.3);
        M.r_min = CD::run(1.0); M.r_max = CD::run(100.0); M.horizon_buffer = CD::run(0.1);
        M.emission_scale = CD::run(1.0); M.absorption_scale = CD::run(1.0); M.faraday_scale = CD::run(1.0);
        M.max_pol_frac = CD::run(0.99); M.keplerian_factor = CD::run(1.0); M.infall_factor = CD::run(0.0);
        M.min_frequency_scale = CD::run(1e-6); M.min_sin_theta = CD::run(1e-4); M.min_thetae = CD::run(1e-3);
        M.min_radius = CD::run(1e-6); M.freq_cgs = CD::run(230e9); M.dlambda_scale = CD::run(4e-8);
        M.rh = CD::run(1.341); M.risco = CD::run(2.03); M.ne_factor = CD::run(1.0);
        M.thetae_factor = CD::run(1.0); M.b_factor = CD::run(1.0); M.mass = m; M.spin = a;
        M.fit = fit;
        M.faraday_rotation = 1;
        g_ops = OpTally();
        const double radii[6] = {1.8, 3.0, 6.0, 15.0, 40.0, 90.0};
        for (double r : radii) {
            St<CD> s = ray_state(r, 1.45, 0.94);
            Coeffs<CD> c;
            riaf_coefficients(M, s, c);
        }
        emit(fit == 1 ? "coefficients_pandya" : "coefficients_dexter", g_ops, 6.0);
    }
    {
        // FMKS geometry of the synthetic GRMHD grids (poly_alpha 14, xt 0.82)
        const FmksMap f = fmks_map(std::log(1.2), 0.3, 0.5, 14.0, 0.82, 0.5 * M_PI / (1.0 + 1.0 / 15.0 / std::pow(0.82, 14.0)));
        const CD mf = CD::run(1.0), af = CD::run(0.9375);
        GeoFmks<CD> GF{mf, af, &rhs_count, f};
        const double v[16] = {0.0, std::log(7.0), 0.45, 0.3, 1.2, -0.2, 0.01, 0.02, 0.0, 0.01, 0.0, 0.0, 0.0, 0.0, 0.05, 0.1};
        St<CD> s, d, half, two;
        for (int i = 0; i < 16; ++i) s.v[i] = CD::run(v[i]);
        g_ops = OpTally();
        fmks_rhs_apply(mf, af, f, s, [&](int i, CD x) { d.v[i] = x; });
        emit("fmks_rhs", g_ops, 1.0);
        g_ops = OpTally();
        CD err = doubling_attempt(GF, s, CD::run(0.1), half, two);
        (void)err;
        emit("fmks_doubling_attempt", g_ops, 1.0);
        for (int fit : {1, 4}) {
            g_ops = OpTally();
            int n = 0;
            const double states[3][4] = {{1e5, 3.0, 20.0, 0.7}, {3e4, 12.0, 5.0, 0.9}, {2e6, 0.8, 60.0, 0.4}};
            for (const auto& st : states)
                for (int k = 0; k < 8; ++k) {
                    const double nu = 86e9 * std::pow(345.0 / 86.0, k / 7.0);
                    Coeffs<CD> c;
                    thermal_coeffs(fit, CD::run(nu), CD::run(st[0]), CD::run(st[1]), CD::run(st[2]),
                                   CD::run(st[3]), CD::run(std::sqrt(1.0 - st[3] * st[3])), CD::run(0.99),
                                   CD(1e-4), CD::run(1.0), CD::run(1.0), CD::run(1.0), c);
                    ++n;
                }
            emit(fit == 1 ? "grmhd_coefficients_pandya" : "grmhd_coefficients_dexter", g_ops, double(n));
        }
    }
    {
        Coeffs<CD> c;
        c.jI = CD::run(1e-20); c.jQ = CD::run(3e-21); c.jU = CD::run(-2e-21); c.jV = CD::run(1e-22);
        c.aI = CD::run(1e-10); c.aQ = CD::run(2e-11); c.aU = CD::run(-1e-11); c.aV = CD::run(3e-12);
        c.rQ = CD::run(1e-9); c.rU = CD::run(-4e-10); c.rV = CD::run(2e-8);
        Stokes<CD> st{CD::run(1e-5), CD::run(1e-6), CD::run(-1e-6), CD::run(1e-7)};
        g_ops = OpTally();
        strang_step(st, c, CD::run(3e7));
        emit("transfer_step", g_ops, 1.0);
        g_ops = OpTally();
        int n = substep_count(c, CD::run(0.3), CD::run(4e7), CD::run(1.0), CD::run(1.0), CD::run(1.0),
                              CD::run(4.0));
        (void)n;
        emit("substep_count", g_ops, 1.0, true);
    }
    std::printf("}\n");
    return 0;
}
