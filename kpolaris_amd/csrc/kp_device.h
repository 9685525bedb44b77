// kp_device.h -- device physics of the per-pixel imaging kernel (gfx950).
//
// One ray per lane.  Every function is a register-resident, fully unrolled
// restatement of the reference algorithm, templated on the arithmetic type T
// (double for the parity kernel, float for the mixed-precision bulk path).
// Reference files are cited per function; the algorithm is the reference's,
// the data layout and scheduling are this kernel's own.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Host-callable as well so the op-counting tool can instantiate the same code.
#define KP_DEV __host__ __device__ __forceinline__

namespace kpd {

template <class T> KP_DEV T tiny();
template <> KP_DEV double tiny<double>() { return 1e-300; }
template <> KP_DEV float tiny<float>() { return 1e-30f; }
template <class T> KP_DEV T large();
template <> KP_DEV double large<double>() { return 1e300; }
template <> KP_DEV float large<float>() { return 1e30f; }
template <class T> KP_DEV T eps();
template <> KP_DEV double eps<double>() { return 2.220446049250313080847263336181640625e-16; }
template <> KP_DEV float eps<float>() { return 1.1920928955078125e-7f; }

template <class T> KP_DEV T absv(T x) { return x < T(0) ? -x : x; }
template <class T> KP_DEV T minv(T a, T b) { return a < b ? a : b; }
template <class T> KP_DEV T maxv(T a, T b) { return a > b ? a : b; }
template <class T> KP_DEV T clampv(T x, T lo, T hi) { return minv(maxv(x, lo), hi); }
template <class T> KP_DEV T sq(T x) { return x * x; }

constexpr double PI_D = 3.141592653589793238462643383279502884;

// Reciprocal and quotient: exact IEEE division in double; in float on the
// device the hardware reciprocal (1 ulp), which the mixed-precision path
// accepts in place of the multi-instruction correctly rounded division.
KP_DEV double krcp(double x) { return 1.0 / x; }
KP_DEV double kdiv(double a, double b) { return a / b; }
KP_DEV float krcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
KP_DEV float kdiv(float a, float b) { return a * krcp(b); }
// other number types (the forward-mode duals): the exact quotient
template <class T> KP_DEV T krcp(const T& x) { return T(1) / x; }

// Counter increment shared by the kernels and their host build (the CPU
// emulation of the lane logic runs one lane at a time).
KP_DEV void count_add(unsigned long long* p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(p, v);
#else
    *p += v;
#endif
}

// Sine and cosine for the bounded arguments of the transport equations
// (|x| < 1e5; the Boyer-Lindquist polar angle stays within a few pi).
// Cody-Waite reduction by pi/2 with a three-part split of pi/2 (the first two
// parts carry 33 significant bits, so n * P1 and n * P2 are exact for
// |n| < 2^20), then the Taylor series of sin and cos on |r| <= pi/4 through
// r^19 and r^18 in Horner form; the truncation error there is below 1e-19.
// The library routines also handle arbitrarily large arguments, which costs
// registers this kernel needs for its state.
KP_DEV void sincos_bounded(double x, double* s_out, double* c_out) {
    const double n = rint(x * 0.6366197723675814);
    double r = fma(-n, 1.5707963267341256, x);
    r = fma(-n, 6.077100506303966e-11, r);
    r = fma(-n, 2.0222662487959506e-21, r);
    const double r2 = r * r;
    double ps = -8.22063524662433e-18;
    ps = fma(ps, r2, 2.8114572543455206e-15);
    ps = fma(ps, r2, -7.647163731819816e-13);
    ps = fma(ps, r2, 1.6059043836821613e-10);
    ps = fma(ps, r2, -2.505210838544172e-08);
    ps = fma(ps, r2, 2.7557319223985893e-06);
    ps = fma(ps, r2, -0.0001984126984126984);
    ps = fma(ps, r2, 0.008333333333333333);
    ps = fma(ps, r2, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = -1.5619206968586225e-16;
    pc = fma(pc, r2, 4.779477332387385e-14);
    pc = fma(pc, r2, -1.1470745597729725e-11);
    pc = fma(pc, r2, 2.08767569878681e-09);
    pc = fma(pc, r2, -2.755731922398589e-07);
    pc = fma(pc, r2, 2.48015873015873e-05);
    pc = fma(pc, r2, -0.001388888888888889);
    pc = fma(pc, r2, 0.041666666666666664);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int q = int(n) & 3;
    const double s_ = (q & 1) ? cr : sr;
    const double c_ = (q & 1) ? sr : cr;
    *s_out = (q & 2) ? -s_ : s_;
    *c_out = ((q + 1) & 2) ? -c_ : c_;
}

KP_DEV void ksincos(double x, double* s, double* c) { sincos_bounded(x, s, c); }
KP_DEV void ksincos(float x, float* s, float* c) { sincosf(x, s, c); }
// Library sine/cosine for unbounded arguments (Faraday rotation angles).
KP_DEV void lsincos(double x, double* s, double* c) { sincos(x, s, c); }
KP_DEV void lsincos(float x, float* s, float* c) { sincosf(x, s, c); }

// ---------------------------------------------------------------------------
// Transport state: x, k, e1, e2 (src/geodesic/state.hpp:8-16) as 16 registers.
// ---------------------------------------------------------------------------
template <class T> struct St {
    T v[16];
    KP_DEV T operator[](int i) const { return v[i]; }
};

// A 16-component vector held in LDS for one lane: component c at p[c * stride]
// (component-major, so a wavefront's 64 lanes touch 64 consecutive words).
template <class T> struct LdsVec {
    T* p;
    int stride;
    KP_DEV T operator[](int i) const { return p[i * stride]; }
    KP_DEV void set(int i, T x) const { p[i * stride] = x; }
};

template <class T> KP_DEV void st_copy(St<T>& d, const St<T>& s) {
#pragma unroll
    for (int i = 0; i < 16; ++i) d.v[i] = s.v[i];
}

template <class T> KP_DEV void st_load(St<T>& d, const LdsVec<T>& s) {
#pragma unroll
    for (int i = 0; i < 16; ++i) d.v[i] = s[i];
}

template <class T> KP_DEV void st_store(const LdsVec<T>& d, const St<T>& s) {
#pragma unroll
    for (int i = 0; i < 16; ++i) d.set(i, s.v[i]);
}

template <class T> KP_DEV void st_load(St<T>& d, const St<T>& s) { st_copy(d, s); }
template <class T> KP_DEV void st_store(St<T>& d, const St<T>& s) { st_copy(d, s); }

template <class T> KP_DEV bool st_finite(const St<T>& s) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 16; ++i) ok = ok && (s.v[i] == s.v[i]) && (absv(s.v[i]) < large<T>());
    return ok;
}

// ---------------------------------------------------------------------------
// Kerr Boyer-Lindquist connection (src/geometry/kerr_boyer_lindquist.hpp:28-45,
// 229-322).  Divisions are folded into three reciprocals (1/Sigma, 1/Delta,
// 1/sin^2), and the bilinear terms reuse the row products:
// bilinear(d, p, q) = sum_nu p_nu * rowdot(d, nu, q).
// ---------------------------------------------------------------------------
template <class T> struct BLConn {
    T dr00, dr03, dr11, dr22, dr33;
    T dt00, dt03, dt11, dt22, dt33;
    T g00, g03, g11, g22, g33;
};

template <class T> KP_DEV BLConn<T> bl_conn(T m, T a, T x1, T x2) {
    BLConn<T> c;
    const T r = maxv(x1, tiny<T>());
    T sn, cs;
    ksincos(x2, &sn, &cs);
    const T a2 = a * a;
    const T sin2 = maxv(sn * sn, tiny<T>());
    const T cos2 = cs * cs;
    const T sig = r * r + a2 * cos2;
    const T del = r * r - T(2) * m * r + a2;
    const T sig_r = T(2) * r;
    const T sig_t = -T(2) * a2 * sn * cs;
    const T del_r = T(2) * (r - m);
    const T ds2 = T(2) * sn * cs;
    const T isig = krcp(sig);
    const T idel = krcp(del);
    const T isig2 = isig * isig;
    const T A = (sig - r * sig_r) * isig2;
    const T B = (ds2 * sig - sin2 * sig_t) * isig2;
    c.dr00 = T(2) * m * A;
    c.dt00 = -T(2) * m * r * sig_t * isig2;
    c.dr03 = -T(2) * m * a * sin2 * A;
    c.dt03 = -T(2) * m * a * r * B;
    c.dr11 = (sig_r * del - sig * del_r) * (idel * idel);
    c.dt11 = sig_t * idel;
    c.dr22 = sig_r;
    c.dt22 = sig_t;
    const T extra = T(2) * m * a2 * r * sin2 * isig;
    const T aterm = r * r + a2 + extra;
    const T extra_r = T(2) * m * a2 * sin2 * A;
    const T extra_t = T(2) * m * a2 * r * B;
    c.dr33 = sin2 * (T(2) * r + extra_r);
    c.dt33 = ds2 * aterm + sin2 * extra_t;
    const T isd = isig * idel;
    const T r2pa2 = r * r + a2;
    c.g00 = -(r2pa2 * r2pa2 - a2 * del * sin2) * isd;
    c.g03 = -T(2) * m * a * r * isd;
    c.g11 = del * isig;
    c.g22 = isig;
    c.g33 = kdiv((del - a2 * sin2) * isd, sin2);
    return c;
}

// Row products rowdot(d, nu, v) for nu = 0..3 (kerr_boyer_lindquist.hpp:180-194).
template <class T>
KP_DEV void rowdot4(T d00, T d03, T d11, T d22, T d33, const T* v, T* out) {
    out[0] = d00 * v[0] + d03 * v[3];
    out[1] = d11 * v[1];
    out[2] = d22 * v[2];
    out[3] = d03 * v[0] + d33 * v[3];
}

// Gamma^mu_ab k^a q^b given the k row products (kerr_boyer_lindquist.hpp:274-295).
template <class T>
KP_DEV void bl_contract_k(const BLConn<T>& c, const T* k, const T* Ak, const T* Bk, const T* q,
                          T* out) {
    T Aq[4], Bq[4];
    rowdot4(c.dr00, c.dr03, c.dr11, c.dr22, c.dr33, q, Aq);
    rowdot4(c.dt00, c.dt03, c.dt11, c.dt22, c.dt33, q, Bq);
    T t[4];
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) t[nu] = k[1] * Aq[nu] + k[2] * Bq[nu] + q[1] * Ak[nu] + q[2] * Bk[nu];
    const T bil_r = k[0] * Aq[0] + k[1] * Aq[1] + k[2] * Aq[2] + k[3] * Aq[3];
    const T bil_t = k[0] * Bq[0] + k[1] * Bq[1] + k[2] * Bq[2] + k[3] * Bq[3];
    t[1] -= bil_r;
    t[2] -= bil_t;
    out[0] = T(0.5) * (c.g00 * t[0] + c.g03 * t[3]);
    out[1] = T(0.5) * c.g11 * t[1];
    out[2] = T(0.5) * c.g22 * t[2];
    out[3] = T(0.5) * (c.g03 * t[0] + c.g33 * t[3]);
}

// Specialised BL transport_rhs (kerr_boyer_lindquist.hpp:306-322):
// dx = k, dk = -Gamma(k,k), de_i = -Gamma(k, e_i).
template <class T> KP_DEV void bl_rhs(T m, T a, const St<T>& s, St<T>& d) {
    const BLConn<T> c = bl_conn(m, a, s.v[1], s.v[2]);
    const T* k = s.v + 4;
    T Ak[4], Bk[4];
    rowdot4(c.dr00, c.dr03, c.dr11, c.dr22, c.dr33, k, Ak);
    rowdot4(c.dt00, c.dt03, c.dt11, c.dt22, c.dt33, k, Bk);
    T ck[4], c1[4], c2[4];
    bl_contract_k(c, k, Ak, Bk, k, ck);
    bl_contract_k(c, k, Ak, Bk, s.v + 8, c1);
    bl_contract_k(c, k, Ak, Bk, s.v + 12, c2);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        d.v[mu] = k[mu];
        d.v[4 + mu] = -ck[mu];
        d.v[8 + mu] = -c1[mu];
        d.v[12 + mu] = -c2[mu];
    }
}

// Diagonal-plus-(0,3) BL covariant metric (kerr_boyer_lindquist.hpp:47-71).
template <class T> struct BLGcov {
    T g00, g03, g11, g22, g33;
};

template <class T> KP_DEV BLGcov<T> bl_gcov(T m, T a, T x1, T sn, T cs) {
    BLGcov<T> g;
    const T r = maxv(x1, tiny<T>());
    const T a2 = a * a;
    const T sin2 = maxv(sn * sn, tiny<T>());
    const T sig = r * r + a2 * cs * cs;
    const T del = r * r - T(2) * m * r + a2;
    g.g00 = -(T(1) - T(2) * m * r / sig);
    g.g03 = -T(2) * m * a * r * sin2 / sig;
    g.g11 = sig / del;
    g.g22 = sig;
    g.g33 = sin2 * (r * r + a2 + T(2) * m * a2 * r * sin2 / sig);
    return g;
}

template <class T> KP_DEV T gdot(const BLGcov<T>& g, const T* u, const T* v) {
    return g.g00 * u[0] * v[0] + g.g03 * (u[0] * v[3] + u[3] * v[0]) + g.g11 * u[1] * v[1] +
           g.g22 * u[2] * v[2] + g.g33 * u[3] * v[3];
}

template <class T> KP_DEV void glower(const BLGcov<T>& g, const T* v, T* out) {
    out[0] = g.g00 * v[0] + g.g03 * v[3];
    out[1] = g.g11 * v[1];
    out[2] = g.g22 * v[2];
    out[3] = g.g03 * v[0] + g.g33 * v[3];
}

// A general symmetric 4x4 metric (FMKS has g01, g12, g13 off-diagonals).
template <class T> struct Metric4 {
    T g00, g01, g02, g03, g11, g12, g13, g22, g23, g33;
};

template <class T> KP_DEV Metric4<T> metric4_of(const BLGcov<T>& b) {
    Metric4<T> g;
    g.g00 = b.g00;
    g.g01 = T(0);
    g.g02 = T(0);
    g.g03 = b.g03;
    g.g11 = b.g11;
    g.g12 = T(0);
    g.g13 = T(0);
    g.g22 = b.g22;
    g.g23 = T(0);
    g.g33 = b.g33;
    return g;
}

// metric.dot(x, u, v) (e.g. kerr_fmks.hpp:196-204).
template <class T> KP_DEV T dot4(const Metric4<T>& g, const T* u, const T* v) {
    return g.g00 * u[0] * v[0] + g.g11 * u[1] * v[1] + g.g22 * u[2] * v[2] + g.g33 * u[3] * v[3] +
           g.g01 * (u[0] * v[1] + u[1] * v[0]) + g.g02 * (u[0] * v[2] + u[2] * v[0]) +
           g.g03 * (u[0] * v[3] + u[3] * v[0]) + g.g12 * (u[1] * v[2] + u[2] * v[1]) +
           g.g13 * (u[1] * v[3] + u[3] * v[1]) + g.g23 * (u[2] * v[3] + u[3] * v[2]);
}

template <class T> KP_DEV T frame_error4(const Metric4<T>& g, const St<T>& s) {
    const T* k = s.v + 4;
    const T* e1 = s.v + 8;
    const T* e2 = s.v + 12;
    T out = absv(dot4(g, k, k));
    out = maxv(out, absv(dot4(g, e1, k)));
    out = maxv(out, absv(dot4(g, e2, k)));
    out = maxv(out, absv(dot4(g, e1, e1) - T(1)));
    out = maxv(out, absv(dot4(g, e2, e2) - T(1)));
    out = maxv(out, absv(dot4(g, e1, e2)));
    return out;
}

// max_frame_error(frame_errors(metric, s)) (src/geodesic/state.hpp:44-76).
template <class T> KP_DEV T frame_error(const BLGcov<T>& g, const St<T>& s) {
    const T* k = s.v + 4;
    const T* e1 = s.v + 8;
    const T* e2 = s.v + 12;
    T out = absv(gdot(g, k, k));
    out = maxv(out, absv(gdot(g, e1, k)));
    out = maxv(out, absv(gdot(g, e2, k)));
    out = maxv(out, absv(gdot(g, e1, e1) - T(1)));
    out = maxv(out, absv(gdot(g, e2, e2) - T(1)));
    out = maxv(out, absv(gdot(g, e1, e2)));
    return out;
}

// ---------------------------------------------------------------------------
// RK4 family (src/geodesic/rk4.hpp:36-59, 74-89).
// ---------------------------------------------------------------------------
template <class T> struct Geo;

template <class T> KP_DEV void rhs(const Geo<T>& g, const St<T>& s, St<T>& d) {
    ++*g.rhs_count;
    bl_rhs(g.m, g.a, s, d);
}

// BL transport RHS delivered component by component: sink(i, f_i) is called
// for i = 0..15 in order, and every component of y a later f_j depends on is
// read before sink(i, .) runs, so a sink may overwrite y[i] in place.  This
// lets an RK stage fold f into its accumulator and build the next stage state
// without holding a separate derivative vector.  Y is any indexable vector
// (registers or LDS).
template <class T, class Y, class Sink> KP_DEV void bl_rhs_apply(T m, T a, const Y& y, Sink&& sink) {
    const BLConn<T> c = bl_conn(m, a, T(y[1]), T(y[2]));
    const T k[4] = {y[4], y[5], y[6], y[7]};
    T Ak[4], Bk[4];
    rowdot4(c.dr00, c.dr03, c.dr11, c.dr22, c.dr33, k, Ak);
    rowdot4(c.dt00, c.dt03, c.dt11, c.dt22, c.dt33, k, Bk);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(mu, k[mu]);
    T out[4];
    bl_contract_k(c, k, Ak, Bk, k, out);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(4 + mu, -out[mu]);
    const T e1[4] = {y[8], y[9], y[10], y[11]};
    bl_contract_k(c, k, Ak, Bk, e1, out);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(8 + mu, -out[mu]);
    const T e2[4] = {y[12], y[13], y[14], y[15]};
    bl_contract_k(c, k, Ak, Bk, e2, out);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) sink(12 + mu, -out[mu]);
}

// Integration context for the Boyer-Lindquist metric: mass, spin and the
// lane's RHS counter.  apply() delivers the transport RHS component-wise.
template <class T> struct Geo {
    using real = T;
    T m, a;
    uint32_t* rhs_count;
    template <class Y, class Sink> KP_DEV void apply(const Y& y, Sink&& sink) const { bl_rhs_apply(m, a, y, sink); }
};

template <class T> KP_DEV void vset(St<T>& v, int i, T x) { v.v[i] = x; }
template <class T> KP_DEV void vset(const LdsVec<T>& v, int i, T x) { v.set(i, x); }

// RK4 stages (src/geodesic/rk4.hpp:36-59).  With acc = k1 + 2 k2 + 2 k3 built
// in that order, the step is base + (acc + k4) * (h / 6) exactly as the
// reference sums it.  The base vector may live in registers or in LDS.
// f = f(y); acc += 2 f; y = base + f * c
template <class G, class T, class B>
KP_DEV void stage_mid(const G& g, St<T>& y, const B& base, St<T>& acc, T c) {
    ++*g.rhs_count;
    g.apply(y, [&](int i, T f) {
        acc.v[i] = acc.v[i] + T(2) * f;
        y.v[i] = base[i] + f * c;
    });
}

// f = f(y); y = base + (acc + f) * c
template <class G, class T, class B>
KP_DEV void stage_last(const G& g, St<T>& y, const B& base, const St<T>& acc, T c) {
    ++*g.rhs_count;
    g.apply(y, [&](int i, T f) { y.v[i] = base[i] + (acc.v[i] + f) * c; });
}

// f = f(src); acc = f; y = src + f * c
template <class G, class T, class B> KP_DEV void stage_first(const G& g, const B& src, St<T>& acc, St<T>& y, T c) {
    ++*g.rhs_count;
    g.apply(src, [&](int i, T f) {
        acc.v[i] = f;
        y.v[i] = src[i] + f * c;
    });
}

// RK4 step s -> out (rk4.hpp:36-46).
template <class G, class T, class B> KP_DEV void rk4(const G& g, const B& s, T h, St<T>& out) {
    St<T> acc;
    const T hh = T(0.5) * h;
    stage_first(g, s, acc, out, hh);
    stage_mid(g, out, s, acc, hh);
    stage_mid(g, out, s, acc, h);
    stage_last(g, out, s, acc, h / T(6));
}

// state_error_norm (rk4.hpp:74-89).
template <class T> KP_DEV T err_norm(const St<T>& a, const St<T>& b) {
    T err = T(0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const T sc = maxv(T(1), maxv(absv(a.v[i]), absv(b.v[i])));
        err = maxv(err, absv(a.v[i] - b.v[i]) / sc);
    }
    return err;
}

// One step-doubling attempt of adaptive_rk4_step (rk4.hpp:121-128): the full
// step and the two half steps share k1.  The half steps are integrated first;
// the full step's last stage is compared with the two-half-step state as it is
// formed, so the full-step state is never held.  s (read only) and the k1
// scratch K may live in registers or LDS.  Returns the error norm.
// full = false computes only the two half steps (the composed path of the
// surface-crossing refinement) and returns 0.
template <class G, class T, class S, class K>
KP_DEV T doubling_attempt(const G& g, const S& s, T h, K& k1, St<T>& half, St<T>& two_half, bool full = true) {
    St<T> acc;
    ++*g.rhs_count;
    g.apply(s, [&](int i, T f) { vset(k1, i, f); });
    const T h2 = h * T(0.5);
    const T q2 = T(0.5) * h2;
    // first half step from s with k1
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const T k = k1[i];
        acc.v[i] = k;
        half.v[i] = s[i] + k * q2;
    }
    stage_mid(g, half, s, acc, q2);
    stage_mid(g, half, s, acc, h2);
    stage_last(g, half, s, acc, h2 / T(6));
    // second half step from the mid state
    stage_first(g, half, acc, two_half, q2);
    stage_mid(g, two_half, half, acc, q2);
    stage_mid(g, two_half, half, acc, h2);
    stage_last(g, two_half, half, acc, h2 / T(6));
    if (!full) return T(0);
    // full step from s with k1, compared on the fly
    St<T> y;
    const T hh = T(0.5) * h;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const T k = k1[i];
        acc.v[i] = k;
        y.v[i] = s[i] + k * hh;
    }
    stage_mid(g, y, s, acc, hh);
    stage_mid(g, y, s, acc, h);
    T err = T(0);
    const T h6 = h / T(6);
    ++*g.rhs_count;
    g.apply(y, [&](int i, T f) {
        const T full = s[i] + (acc.v[i] + f) * h6;
        const T two = two_half.v[i];
        const T sc = maxv(T(1), maxv(absv(full), absv(two)));
        err = maxv(err, absv(full - two) / sc);
    });
    return err;
}

template <class G, class T>
KP_DEV T doubling_attempt(const G& g, const St<T>& s, T h, St<T>& half, St<T>& two_half) {
    St<T> k1;
    return doubling_attempt(g, s, h, k1, half, two_half);
}

// ---------------------------------------------------------------------------
// Transfer (src/radiation/semi_analytic.hpp, src/common/math.hpp:62-86).
// ---------------------------------------------------------------------------
template <class T> struct Coeffs {
    T jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV;
};

template <class T> KP_DEV void coeffs_zero(Coeffs<T>& c) {
    c.jI = c.jQ = c.jU = c.jV = c.aI = c.aQ = c.aU = c.aV = c.rQ = c.rU = c.rV = T(0);
}

template <class T> struct Stokes {
    T I, Q, U, V;
};

// Scaled Euclidean norm (src/common/vec.hpp:63-71).  Divides by the scale: a
// reciprocal of a subnormal scale overflows to inf and turns 0 into NaN.
template <class T> KP_DEV T norm3(T x, T y, T z) {
    const T scale = maxv(absv(x), maxv(absv(y), absv(z)));
    if (scale == T(0)) return T(0);
    const T a = x / scale, b = y / scale, c = z / scale;
    return scale * sqrt(a * a + b * b + c * c);
}

template <class T> KP_DEV T stable_attenuation(T kappa, T dl) {
    const T x = kappa * dl;
    if (x > T(700)) return T(0);
    return exp(-x);
}

template <class T> KP_DEV T stable_source_factor(T kappa, T dl) {
    const T x = kappa * dl;
    if (absv(x) < T(1e-6)) return dl * (T(1) - x / T(2) + x * x / T(6) - x * x * x / T(24));
    return (T(1) - exp(-x)) / kappa;
}

// faraday_rotate (semi_analytic.hpp:9-37): Rodrigues rotation about rho.
template <class T> KP_DEV void faraday_rotate(Stokes<T>& s, const Coeffs<T>& c, T dl) {
    const T px = s.Q, py = s.U, pz = s.V;
    const T rn = norm3(c.rQ, c.rU, c.rV);
    if (rn == T(0) || dl == T(0)) return;
    const T theta = rn * dl;
    if (absv(theta) < T(1e-8)) {
        s.Q = px + (c.rU * pz - c.rV * py) * dl;
        s.U = py + (c.rV * px - c.rQ * pz) * dl;
        s.V = pz + (c.rQ * py - c.rU * px) * dl;
        return;
    }
    const T inv = krcp(rn);
    const T nx = c.rQ * inv, ny = c.rU * inv, nz = c.rV * inv;
    T st, ct;
    lsincos(theta, &st, &ct);
    const T nd = (nx * px + ny * py + nz * pz) * (T(1) - ct);
    s.Q = px * ct + nx * nd + (ny * pz - nz * py) * st;
    s.U = py * ct + ny * nd + (nz * px - nx * pz) * st;
    s.V = pz * ct + nz * nd + (nx * py - ny * px) * st;
}

// absorb_emit (semi_analytic.hpp:39-91): exact along the absorption axis.
template <class T> KP_DEV void absorb_emit(Stokes<T>& s, const Coeffs<T>& c, T dl) {
    const T px = s.Q, py = s.U, pz = s.V;
    const T an = norm3(c.aQ, c.aU, c.aV);
    const T ei = stable_attenuation(c.aI, dl);
    const T fi = stable_source_factor(c.aI, dl);
    if (an < tiny<T>()) {
        s.I = ei * s.I + fi * c.jI;
        s.Q = ei * s.Q + fi * c.jQ;
        s.U = ei * s.U + fi * c.jU;
        s.V = ei * s.V + fi * c.jV;
        return;
    }
    const T inv = T(1) / an;
    const T nx = c.aQ * inv, ny = c.aU * inv, nz = c.aV * inv;
    const T ppar = nx * px + ny * py + nz * pz;
    const T jpar = nx * c.jQ + ny * c.jU + nz * c.jV;
    const T kp = c.aI + an, km = c.aI - an;
    const T sp = stable_attenuation(kp, dl) * (s.I + ppar) + stable_source_factor(kp, dl) * (c.jI + jpar);
    const T sm = stable_attenuation(km, dl) * (s.I - ppar) + stable_source_factor(km, dl) * (c.jI - jpar);
    const T pout = T(0.5) * (sp - sm);
    s.I = T(0.5) * (sp + sm);
    s.Q = nx * pout + ((px - nx * ppar) * ei + (c.jQ - nx * jpar) * fi);
    s.U = ny * pout + ((py - ny * ppar) * ei + (c.jU - ny * jpar) * fi);
    s.V = nz * pout + ((pz - nz * ppar) * ei + (c.jV - nz * jpar) * fi);
}

// semi_analytic_stokes_step (semi_analytic.hpp:93-100).
template <class T> KP_DEV void strang_step(Stokes<T>& s, const Coeffs<T>& c, T dl) {
    faraday_rotate(s, c, T(0.5) * dl);
    absorb_emit(s, c, dl);
    faraday_rotate(s, c, T(0.5) * dl);
}

// radiation_substep_count (src/geodesic/pass_b.hpp:126-179).  The count n
// is a power of two, so x / n is formed as x * (1/n) with 1/n = 2^-k held
// exactly: both are the correctly rounded value of the same real number (the
// same bits), and the doubling loops run without a division.
template <class T>
KP_DEV int substep_count(const Coeffs<T>& c, T h, T dls, T max_rs, T max_rd, T max_ad, T max_fd) {
    int n = 1;
    T inv = T(1);
    const T ah = absv(h);
    if (max_rs > T(0))
        while (ah * inv > max_rs && n < 4096) {
            n *= 2;
            inv *= T(0.5);
        }
    const T dl = ah * absv(dls);
    const T alim = max_ad >= T(0) ? max_ad : max_rd;
    if (alim > T(0)) {
        const T depth = (absv(c.aI) + sqrt(sq(c.aQ) + sq(c.aU) + sq(c.aV))) * dl;
        while (depth * inv > alim && n < 4096) {
            n *= 2;
            inv *= T(0.5);
        }
    }
    const T flim = max_fd >= T(0) ? max_fd : max_rd;
    if (flim > T(0)) {
        const T depth = sqrt(sq(c.rQ) + sq(c.rU) + sq(c.rV)) * dl;
        while (depth * inv > flim && n < 4096) {
            n *= 2;
            inv *= T(0.5);
        }
    }
    return n;
}

// ---------------------------------------------------------------------------
// Thermal synchrotron (src/radiation/thermal_synchrotron.hpp).
// ---------------------------------------------------------------------------
#define KP_EE 4.803204712570263e-10
#define KP_ME 9.1093837015e-28
#define KP_CL 2.99792458e10
#define KP_HPL 6.62607015e-27

// Numerical-Recipes polynomial Bessel functions (thermal_synchrotron.hpp:60-165).
template <class T> KP_DEV T bessel_i0(T ax) {
    if (ax < T(3.75)) {
        T y = ax / T(3.75);
        y *= y;
        T p = T(0.0045813);
        p = T(0.0360768) + y * p;
        p = T(0.2659732) + y * p;
        p = T(1.2067492) + y * p;
        p = T(3.0899424) + y * p;
        p = T(3.5156229) + y * p;
        return T(1) + y * p;
    }
    const T y = T(3.75) / ax;
    T p = T(0.00392377);
    p = T(-0.01647633) + y * p;
    p = T(0.02635537) + y * p;
    p = T(-0.02057706) + y * p;
    p = T(0.00916281) + y * p;
    p = T(-0.00157565) + y * p;
    p = T(0.00225319) + y * p;
    p = T(0.01328592) + y * p;
    p = T(0.39894228) + y * p;
    return exp(ax) / sqrt(ax) * p;
}

template <class T> KP_DEV T bessel_i1(T ax) {
    if (ax < T(3.75)) {
        T y = ax / T(3.75);
        y *= y;
        T p = T(0.00032411);
        p = T(0.00301532) + y * p;
        p = T(0.02658733) + y * p;
        p = T(0.15084934) + y * p;
        p = T(0.51498869) + y * p;
        p = T(0.87890594) + y * p;
        p = T(0.5) + y * p;
        return ax * p;
    }
    const T y = T(3.75) / ax;
    T p = T(-0.00420059);
    p = T(0.01787654) + y * p;
    p = T(-0.02895312) + y * p;
    p = T(0.02282967) + y * p;
    p = T(-0.01031555) + y * p;
    p = T(0.00163801) + y * p;
    p = T(-0.00362018) + y * p;
    p = T(-0.03988024) + y * p;
    p = T(0.39894228) + y * p;
    return exp(ax) / sqrt(ax) * p;
}

// K0, K1 and K2 = K0 + 2 K1 / x together (thermal_synchrotron.hpp:113-231);
// the small-x and large-x branches share exp(-x)/sqrt(x) and log(x/2).
template <class T> KP_DEV void bessel_k012(T x, T& k0, T& k1, T& k2) {
    if (x <= T(0)) {
        k0 = k1 = large<T>();
    } else if (x <= T(2)) {
        const T y = x * x * T(0.25);
        T p0 = T(0.00000740);
        p0 = T(0.00010750) + y * p0;
        p0 = T(0.00262698) + y * p0;
        p0 = T(0.03488590) + y * p0;
        p0 = T(0.23069756) + y * p0;
        p0 = T(0.42278420) + y * p0;
        p0 = T(-0.57721566) + y * p0;
        T p1 = T(-0.00004686);
        p1 = T(-0.00110404) + y * p1;
        p1 = T(-0.01919402) + y * p1;
        p1 = T(-0.18156897) + y * p1;
        p1 = T(-0.67278579) + y * p1;
        p1 = T(0.15443144) + y * p1;
        p1 = T(1) + y * p1;
        const T lg = log(x * T(0.5));
        k0 = -lg * bessel_i0(x) + p0;
        k1 = lg * bessel_i1(x) + p1 / x;
    } else {
        const T y = T(2) / x;
        T p0 = T(0.00053208);
        p0 = T(-0.00251540) + y * p0;
        p0 = T(0.00587872) + y * p0;
        p0 = T(-0.01062446) + y * p0;
        p0 = T(0.02189568) + y * p0;
        p0 = T(-0.07832358) + y * p0;
        p0 = T(1.25331414) + y * p0;
        T p1 = T(-0.00068245);
        p1 = T(0.00325614) + y * p1;
        p1 = T(-0.00780353) + y * p1;
        p1 = T(0.01504268) + y * p1;
        p1 = T(-0.03655620) + y * p1;
        p1 = T(0.23498619) + y * p1;
        p1 = T(1.25331414) + y * p1;
        const T e = exp(-x) / sqrt(x);
        k0 = e * p0;
        k1 = e * p1;
    }
    k2 = maxv(k0 + T(2) * k1 / maxv(x, tiny<T>()), tiny<T>());
}

// symphony_bnu_inv (thermal_synchrotron.hpp:233-251).
template <class T> KP_DEV T bnu_inv(T nu, T thetae) {
    if (!(thetae > T(0))) return T(0);
    const T x = T(KP_HPL) * nu / (T(KP_ME) * T(KP_CL) * T(KP_CL) * thetae);
    T denom;
    if (x < T(2e-3))
        denom = x / T(24) * (T(24) + x * (T(12) + x * (T(4) + x)));
    else
        denom = exp(x) - T(1);
    return (T(2) * T(KP_HPL) / maxv(denom, tiny<T>())) / (T(KP_CL) * T(KP_CL));
}

// symphony_rho_QV (thermal_synchrotron.hpp:461-506).
template <class T>
KP_DEV void rho_qv(T nu, T ne, T thetae, T b, T sin_theta, T cos_theta, T& rhoQ, T& rhoV) {
    const T pi = T(PI_D);
    const T omega0 = T(KP_EE) * b / (T(KP_ME) * T(KP_CL));
    const T wp2 = T(4) * pi * ne * T(KP_EE) * T(KP_EE) / T(KP_ME);
    const T two_pi_nu = T(2) * pi * nu;
    const T safe = maxv(two_pi_nu, tiny<T>());
    const T x = maxv(thetae * sqrt(maxv(T(0), sqrt(T(2)) * sin_theta * (T(1e3) * omega0 / safe))), T(1e-12));
    T k0, k1, k2;
    bessel_k012(T(1) / thetae, k0, k1, k2);
    const T exp_x_47 = exp(-x / T(47.2));
    const T extraterm = (T(0.011) * exp_x_47 - T(369.69342869090815) * pow(x + T(1e-16), T(-8) / T(3))) *
                        (T(0.5) + T(0.5) * tanh((log(x) - T(4.787491742782046)) / T(0.1)));
    const T jffunc = T(2.011) * exp(-pow(x, T(1.035)) / T(4.7)) -
                     cos(x / T(2)) * exp(-pow(x, T(1.2)) / T(2.73)) - T(0.011) * exp_x_47 + extraterm;
    const T k1_ratio = k2 > T(0) ? k1 / k2 : T(1);
    T e11 = jffunc * wp2 * omega0 * omega0;
    e11 /= safe;
    e11 /= safe;
    e11 /= safe;
    e11 /= safe;
    e11 *= (k1_ratio + T(6) * thetae) * sin_theta * sin_theta;
    rhoQ = pi * nu / T(KP_CL) * e11;
    const T shgm = T(1) - T(0.11) * log(T(1) + T(0.035) * x);
    const T k0_ratio = k2 > T(0) ? k0 / k2 : T(1);
    T e12 = wp2 * omega0;
    e12 /= safe;
    e12 /= safe;
    e12 /= safe;
    e12 *= k0_ratio * shgm * cos_theta;
    rhoV = T(2) * pi * nu / T(KP_CL) * e12;
}

// thermal_synchrotron_magnetic_basis_coefficients_fit (thermal_synchrotron.hpp:512-616);
// fit 4 = Dexter, otherwise Pandya (symphony_pandya_*: :311-395).
template <class T>
KP_DEV void thermal_coeffs(int fit, T nu_in, T ne, T thetae, T b, T sin_theta_in, T cos_theta,
                           T max_pol_frac, T min_sin_theta, T emission_scale, T absorption_scale,
                           T faraday_scale, Coeffs<T>& c) {
    coeffs_zero(c);
    if (!(nu_in > T(0)) || !(ne > T(0)) || !(thetae > T(0)) || !(b > T(0))) return;
    const T pi = T(PI_D);
    const T nu = maxv(nu_in, T(1));
    const T sin_theta = maxv(sin_theta_in, min_sin_theta);
    T jI, jQ, jV;
    if (fit == 4) {
        const T nus = T(3) * T(KP_EE) * b * sin_theta / (T(4) * pi * T(KP_ME) * T(KP_CL)) * thetae * thetae + T(1);
        const T x = maxv(nu / maxv(nus, tiny<T>()), tiny<T>());
        const T xm13 = T(1) / cbrt(x);
        const T exp_shape = exp(-T(1.8899) / xm13);
        const T pref = ne * T(KP_EE) * T(KP_EE) * nu / (T(2) * sqrt(T(3)) * T(KP_CL) * thetae * thetae);
        jI = pref * (T(2.5651) * (T(1) + T(1.92) * xm13 + T(0.9977) * xm13 * xm13) * exp_shape);
        jQ = pref * (T(2.5651) * (T(1) + T(0.93193) * xm13 + T(0.499873) * xm13 * xm13) * exp_shape);
        const T iv = (T(1.81384) / x + T(3.42319) * xm13 * xm13 + T(0.0292545) / sqrt(x) +
                      T(2.03773) * xm13) * exp_shape;
        T cft = cos_theta;
        if (absv(cft) < T(1e-10)) cft = cft >= T(0) ? T(1e-10) : T(-1e-10);
        const T tan_theta = sin_theta / cft;
        jV = T(2) * ne * T(KP_EE) * T(KP_EE) * nu /
             (tan_theta * T(3) * sqrt(T(3)) * T(KP_CL) * thetae * thetae * thetae) * iv;
    } else {
        const T nu_c = T(KP_EE) * b / (T(2) * pi * T(KP_ME) * T(KP_CL));
        const T nu_s = T(2) / T(9) * nu_c * sin_theta * thetae * thetae;
        const T x = maxv(nu / maxv(nu_s, tiny<T>()), tiny<T>());
        const T pref = ne * T(KP_EE) * T(KP_EE) * nu_c / T(KP_CL);
        const T term1 = sqrt(T(2)) * pi / T(27) * sin_theta;
        const T x13 = cbrt(x);
        const T sx = sqrt(x), sx13 = sqrt(x13);
        const T ex13 = exp(-x13);
        const T shapeI = sx + T(1.8877486253633869933) * sx13;
        jI = pref * term1 * (shapeI * shapeI) * ex13;
        const T t2425 = pow(thetae, T(24) / T(25));
        const T term2 = (T(7) * t2425 + T(35)) / (T(10) * t2425 + T(75));
        const T shapeQ = sx + term2 * T(1.8877486253633869933) * sx13;
        jQ = pref * term1 * (shapeQ * shapeQ) * ex13;
        const T sin_shift = sin_theta * T(0.4356824462767121) - cos_theta * T(0.9001004421765051);
        const T v1 = (T(37) - T(87) * sin_shift) / (T(100) * (thetae + T(1)));
        const T v2 = pow(T(1) + (pow(thetae, T(3) / T(5)) / T(25) + T(7) / T(10)) * pow(x, T(9) / T(25)),
                         T(5) / T(3));
        jV = pref * v1 * v2 * ex13;
    }
    const T nusq = nu * nu;
    const T inu2 = T(1) / maxv(nusq, tiny<T>());
    const T jI_inv = jI * inu2;
    T jQ_inv = jQ * inu2;
    T jV_inv = jV * inu2;
    const T jps = maxv(absv(jQ_inv), absv(jV_inv));
    const T jP = jps > T(0) ? jps * sqrt(sq(jQ_inv / jps) + sq(jV_inv / jps)) : T(0);
    if (jI_inv > T(0) && jP > T(0) && jI_inv < jP / maxv(max_pol_frac, T(1e-30))) {
        const T scale = jI_inv / jP * max_pol_frac;
        jQ_inv *= scale;
        jV_inv *= scale;
    }
    T aI = T(0), aQ = T(0), aV = T(0);
    const T bi = bnu_inv(nu, thetae);
    if (bi > T(0)) {
        // true divisions: B_nu/nu^3 underflows to subnormals at large h nu / kT
        aI = jI_inv / bi;
        aQ = jQ_inv / bi;
        aV = jV_inv / bi;
    }
    T rhoQ, rhoV;
    rho_qv(nu, ne, thetae, b, sin_theta, cos_theta, rhoQ, rhoV);
    c.jI = emission_scale * jI_inv;
    c.jQ = emission_scale * jQ_inv;
    c.jV = emission_scale * jV_inv;
    c.aI = absorption_scale * aI;
    c.aQ = absorption_scale * aQ;
    c.aV = absorption_scale * aV;
    c.rQ = faraday_scale * rhoQ * nu;
    c.rV = faraday_scale * rhoV * nu;
}

// symphony_rho_V (thermal_synchrotron.hpp:431-459), used on the exactly
// field-aligned branch of the RIAF model.
template <class T> KP_DEV T rho_v_aligned(T nu, T ne, T thetae, T b, T theta, T min_sin_theta) {
    T sn, cs;
    sincos(theta, &sn, &cs);
    const T sin_theta = maxv(sn, min_sin_theta);
    T rq, rv;
    rho_qv(nu, ne, thetae, b, sin_theta, cs, rq, rv);
    return rv;
}

}  // namespace kpd
