// kp_dual.h -- forward-mode (tangent) arithmetic for plasma-parameter gradients.
//
// Dual<N> carries a double value and N directional derivatives.  The thermal
// coefficient and Stokes-transfer templates of kp_device.h instantiate on it
// unchanged, so the gradient of the transfer along a fixed geodesic is exact
// to rounding (chain rule through every operation the FP64 path performs,
// including the branches it takes at the value).  Derivatives of the math
// functions are the textbook ones.
#pragma once

#include "kp_device.h"

namespace kpd {

template <int N> struct Dual {
    double v;
    double d[N];

    KP_DEV Dual() : v(0.0) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    KP_DEV Dual(double x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    KP_DEV Dual(int x) : v(double(x)) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }

    // value v with derivative a * (d of x): the chain rule for f(x), a = f'(x)
    KP_DEV static Dual chain(double value, double a, const Dual& x) {
        Dual r(value);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = a * x.d[i];
        return r;
    }

    friend KP_DEV Dual operator+(const Dual& a, const Dual& b) {
        Dual r(a.v + b.v);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
        return r;
    }
    friend KP_DEV Dual operator-(const Dual& a, const Dual& b) {
        Dual r(a.v - b.v);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
        return r;
    }
    friend KP_DEV Dual operator*(const Dual& a, const Dual& b) {
        Dual r(a.v * b.v);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
        return r;
    }
    friend KP_DEV Dual operator/(const Dual& a, const Dual& b) {
        const double q = a.v / b.v;
        Dual r(q);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) / b.v;
        return r;
    }
    friend KP_DEV Dual operator-(const Dual& a) {
        Dual r(-a.v);
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
        return r;
    }
    KP_DEV Dual& operator+=(const Dual& b) { return *this = *this + b; }
    KP_DEV Dual& operator-=(const Dual& b) { return *this = *this - b; }
    KP_DEV Dual& operator*=(const Dual& b) { return *this = *this * b; }
    KP_DEV Dual& operator/=(const Dual& b) { return *this = *this / b; }

    friend KP_DEV bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    friend KP_DEV bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    friend KP_DEV bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    friend KP_DEV bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    friend KP_DEV bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
    friend KP_DEV bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

    friend KP_DEV Dual exp(const Dual& x) {
        const double e = ::exp(x.v);
        return chain(e, e, x);
    }
    friend KP_DEV Dual log(const Dual& x) { return chain(::log(x.v), 1.0 / x.v, x); }
    friend KP_DEV Dual sqrt(const Dual& x) {
        const double s = ::sqrt(x.v);
        return chain(s, s > 0.0 ? 0.5 / s : 0.0, x);
    }
    friend KP_DEV Dual cbrt(const Dual& x) {
        const double c = ::cbrt(x.v);
        return chain(c, c != 0.0 ? 1.0 / (3.0 * c * c) : 0.0, x);
    }
    friend KP_DEV Dual tanh(const Dual& x) {
        const double t = ::tanh(x.v);
        return chain(t, 1.0 - t * t, x);
    }
    friend KP_DEV Dual cos(const Dual& x) { return chain(::cos(x.v), -::sin(x.v), x); }
    friend KP_DEV Dual sin(const Dual& x) { return chain(::sin(x.v), ::cos(x.v), x); }
    // x^y = exp(y log x) for x > 0; d = y x^(y-1) dx + x^y log(x) dy
    friend KP_DEV Dual pow(const Dual& x, const Dual& y) {
        const double p = ::pow(x.v, y.v);
        Dual r(p);
        const double ax = x.v != 0.0 ? y.v * p / x.v : 0.0;
        const double ay = x.v > 0.0 ? p * ::log(x.v) : 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = ax * x.d[i] + ay * y.d[i];
        return r;
    }
    friend KP_DEV void lsincos(const Dual& x, Dual* s, Dual* c) {
        double sv, cv;
        ::sincos(x.v, &sv, &cv);
        *s = chain(sv, cv, x);
        *c = chain(cv, -sv, x);
    }
    friend KP_DEV void sincos(const Dual& x, Dual* s, Dual* c) { lsincos(x, s, c); }
};

using D3 = Dual<3>;

template <> KP_DEV D3 tiny<D3>() { return D3(1e-300); }
template <> KP_DEV D3 large<D3>() { return D3(1e300); }
template <> KP_DEV D3 eps<D3>() { return D3(2.220446049250313080847263336181640625e-16); }

KP_DEV Coeffs<double> value_of(const Coeffs<D3>& c) {
    return {c.jI.v, c.jQ.v, c.jU.v, c.jV.v, c.aI.v, c.aQ.v, c.aU.v, c.aV.v, c.rQ.v, c.rU.v, c.rV.v};
}

}  // namespace kpd
