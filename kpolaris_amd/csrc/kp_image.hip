// kp_image.hip -- pass A / pass B imaging kernels for gfx950 and the C ABI.
//
// Design (one ray per lane, state in VGPRs):
//  * Each kernel is a per-lane state machine whose "tick" performs exactly one
//    integration attempt (one step-doubling attempt, 11 RHS, or one RK4 step)
//    followed by cheap per-lane bookkeeping, so the 64 lanes of a wavefront
//    execute the same heavy instruction stream every tick even when their
//    rays are at different points of the reference's accept/reject control.
//  * Persistent mode: a grid sized to the resident wave count; a lane whose
//    ray terminates refills from a global pixel queue with one wave-level
//    atomic (ballot + popcount), so long photon-ring rays do not hold idle
//    lanes hostage.  This is the wavefront compaction of terminated rays.
//  * Pass A writes endpoints as structure-of-arrays [16][npix] (coalesced);
//    pass B reads them back, like the reference's split driver
//    (src/image/split_driver.hpp:33-104).
//  * Each kernel is instantiated twice: plain FP64 (the parity path) and
//    mixed (FP64 state, FP32 derivatives, coefficients and transfer; the
//    geodesic attempt falls back to FP64 inside fp64_radius, where the
//    Boyer-Lindquist Delta cancels).
#include "kp_device.h"
#include "kp_riaf.h"
#include "kp_mixed.h"
#include "kp_dual.h"
#include "kp_fmks.h"
#include "kp_grmhd.h"
#include "kp_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace kpd;

// Target occupancy of the imaging kernels (waves per SIMD): the register
// budget the compiler may use is 512 / KP_WAVES_PER_EU per lane.
#ifndef KP_WAVES_PER_EU
#define KP_WAVES_PER_EU 1
#endif
// The same for the FP32-geodesic kernels of a staged mixed-precision image.
#ifndef KP_WAVES_PER_EU_F32
#define KP_WAVES_PER_EU_F32 1
#endif
// 1: each lane's state and k1 scratch live in LDS ([32][blockDim] doubles per
// block) instead of registers, trading LDS traffic for register headroom.
#ifndef KP_LANE_LDS
#define KP_LANE_LDS 0
#endif

namespace {

#if KP_LANE_LDS
using LaneVec = LdsVec<double>;
#else
using LaneVec = St<double>;
#endif

enum : int {
    R_NONE = 0,
    R_CAMERA = 1,
    R_INNER = 2,
    R_ESCAPED = 3,
    R_MAX_STEPS = 4,
    R_SLOW_TIME = 6,
    R_UNDERFLOW = 7,
    R_TIME = 8,
};

// Pass B variants: one frequency, several frequencies on shared geodesics,
// one frequency with the tangent (plasma-parameter derivatives) of its Stokes
// vector carried alongside, a GRMHD grid, or a GRMHD snapshot sequence
// sampled at retarded time (slow light).
enum Mode : int {
    M_SINGLE = 0, M_MF = 1, M_TANGENT = 2, M_GRMHD = 3, M_SLOW = 4, M_PROBE = 5, M_STAN = 6,
    // the kappa / power-law forms of M_MF, M_GRMHD and M_SLOW: kernels of their
    // own, so the thermal kernels carry no call to the nonthermal routine
    M_MF_NT = 7, M_GRMHD_NT = 8, M_SLOW_NT = 9
};

// The mode whose lane logic a mode runs (a nonthermal mode runs its thermal one).
constexpr int base_mode(int m) {
    return m == M_MF_NT ? M_MF : (m == M_GRMHD_NT ? M_GRMHD : (m == M_SLOW_NT ? M_SLOW : m));
}
constexpr bool nt_mode(int m) { return m == M_MF_NT || m == M_GRMHD_NT || m == M_SLOW_NT; }

// Coordinates of the geodesics: Boyer-Lindquist (analytic RIAF) or FMKS
// (GRMHD grids, whose pass B carries one Stokes vector per frequency in LDS).
enum Geom : int { G_BL = 0, G_FMKS = 1 };
// Precision of the geodesic attempts (the radiation precision is separate).
enum GeoPrec : int { GP_64 = 0, GP_GUARD = 1, GP_32 = 2 };
// Parking rule of a staged mixed-precision launch (Params::park).
enum Park : int { PARK_NONE = 0, PARK_INSIDE = 1, PARK_OUTSIDE = 2 };
constexpr int KP_NQUEUE = 32;  // queue and list counters in the workspace
// A staged FP64 launch hands a ray back to FP32 at this multiple of the
// guard radius (hysteresis: no ray alternates stages on every step).
constexpr double KP_BACK_RADIUS_FACTOR = 1.25;
template <int MODE> constexpr int geom_of() {
    constexpr int B = base_mode(MODE);
    return (B == M_GRMHD || B == M_SLOW || B == M_PROBE || B == M_STAN) ? G_FMKS : G_BL;
}
template <int MODE> constexpr bool lds_stokes() {
    return base_mode(MODE) == M_MF || base_mode(MODE) == M_GRMHD || base_mode(MODE) == M_SLOW;
}

enum Counter : int {
    C_RHS_A = 0,
    C_RHS_B = 1,
    C_COEFF = 2,
    C_TRANSFER = 3,
    C_STEPS_A = 4,
    C_STEPS_B = 5,
    C_GUARD = 6,
    C_WAIT = 7,  // slow light: rays that needed a time outside the snapshots
    C_COUNT = 8
};

struct Params {
    // camera (CameraParams + the per-image tetrad built on the host)
    int32_t nx, ny, npix, adaptive;
    int32_t max_steps, persistent;
    double radius, inclination, x_extent, y_extent, x_offset, y_offset;
    double tetrad[16];
    double xcam[4];         // camera position in the geodesic coordinates
    Metric4<double> gcam;   // metric there
    FmksMap fmks;           // FMKS map of the geodesic coordinates
    GrmhdDev grmhd;         // GRMHD grid model
    // integrator (PassAParams scalars)
    double mass, spin, inner, outer, step, tol, min_step, max_step;
    double max_rs, max_rd, max_ad, max_fd;
    RiafDev<double> riaf;
    // nonthermal distributions (kappa, power law) of either model
    NtDev nt;
    // mixed precision: FP32 model constants, FP64 guard radius, 1 / nu_obs^3
    RiafMixed rm;
    float mass_f, spin_f;
    double fp64_radius, inv_nu3;
    // multi-frequency pass B (shared geodesics): output and step-control
    // frequency lists with their dl/dlambda, mixed-precision Planck constants
    // 2 h nu^3 / c^2 and 1 / nu^3
    int32_t nfreq, nctrl;
    // single_pass: the step-control list is the output list and it fits the
    // LDS budget of a second Stokes buffer, so each frequency's coefficients
    // are evaluated once per step (control, then a tentative transfer into
    // the second buffer, kept if the step is accepted)
    int32_t single_pass, reserved_mf;
    double freqs[KP_MAX_FREQUENCIES], ctrl[KP_MAX_FREQUENCIES];
    double dls_f[KP_MAX_FREQUENCIES], dls_c[KP_MAX_FREQUENCIES];
    double inv_nu3_f[KP_MAX_FREQUENCIES];
    float bconst_f[KP_MAX_FREQUENCIES], bconst_c[KP_MAX_FREQUENCIES];
    // tangent mode: d(I,Q,U,V)/d(q_density, q_temperature, q_magnetic), [3][4][npix]
    double* grad;
    // workspace
    double* ep;
    int32_t* ep_valid;
    int32_t* ep_steps;
    int32_t* ep_reason;
    uint32_t* queue;
    unsigned long long* counters;
    // persistent launches: the queue counter this launch fetches with and its
    // work source (list_in null: every pixel; else the count_in pixels a
    // previous stage parked).  A staged mixed-precision image (stage >= 0)
    // parks a lane between steps when its radius crosses the rule's radius
    // (PARK_INSIDE: r < fp64_radius; PARK_OUTSIDE: r >= back_radius), saving
    // its state to ep and pk_*, and appends its pixel to list_out.
    int32_t stage, park;
    double back_radius;
    uint32_t* fetch;
    const uint32_t* list_in;
    const uint32_t* count_in;
    uint32_t* list_out;
    uint32_t* count_out;
    double* pk_d;     // pass A: [npix] h_current; pass B: [2][npix] h_current, cap
    int32_t* pk_i;    // pass A: [4][npix] n, steps, steps_inside, entered; pass B: [2][npix] steps, reason
    double* pk_st;    // pass B: [4 nfreq][npix] Stokes
    // slow light: snapshot cell arrays and times (device), the stepping and
    // interpolation rules (kp_slow_light), the observation time, and per pixel
    // the time a waiting ray needed (NaN for the others)
    const kp_grmhd_cell* const* snaps;
    const double* snap_t;
    int32_t nsnap, slow_mode, slow_interp, reserved_slow;
    double t_obs;
    double* req_t;
    // slow-light gradients: the evaluation point's density and field factors
    // (n_e e^q_density, B e^(q_density/2 + q_magnetic))
    double tq_ne, tq_b;
    // time probe: per pixel the smallest and largest x^0 of its radiating
    // samples ([2][npix], NaN when it has none)
    double* probe;
    // outputs
    kp_image_data out;
};

// ---------------------------------------------------------------------------
// Camera ray for one pixel from the host-built ipole tetrad
// (src/camera/camera.hpp:96-106, 484-513).
// ---------------------------------------------------------------------------
KP_DEV void camera_ray(const Params& P, int pixel, St<double>& s) {
    const int i = pixel % P.nx;
    const int j = pixel / P.nx;
    const double sx = ((double(i) + 0.5) / double(P.nx) - 0.5) * P.x_extent + P.x_offset;
    const double sy = ((double(j) + 0.5) / double(P.ny) - 0.5) * P.y_extent + P.y_offset;
    const double sn = sqrt(1.0 + sx * sx + sy * sy);
    const double kt1 = sx / sn, kt2 = sy / sn, kt3 = 1.0 / sn;
    const double* e0 = P.tetrad;
    const double* e1 = P.tetrad + 4;
    const double* e2 = P.tetrad + 8;
    const double* e3 = P.tetrad + 12;
    double sd[4], a[4], b[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) sd[m] = e1[m] * kt1 + e2[m] * kt2 + e3[m] * kt3;
#pragma unroll
    for (int m = 0; m < 4; ++m) s.v[m] = P.xcam[m];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        s.v[4 + m] = e0[m] + sd[m];
        a[m] = e1[m] - sd[m] * kt1;
        b[m] = e2[m] - sd[m] * kt2;
    }
    const double na = dot4(P.gcam, a, a);
    if (na > 0.0) {
        const double f = 1.0 / sqrt(na);
#pragma unroll
        for (int m = 0; m < 4; ++m) a[m] *= f;
    }
    const double den = dot4(P.gcam, a, a);
    if (absv(den) > 1e-300) {
        const double f = dot4(P.gcam, b, a) / den;
#pragma unroll
        for (int m = 0; m < 4; ++m) b[m] = b[m] - a[m] * f;
    }
    const double nb = dot4(P.gcam, b, b);
    if (nb > 0.0) {
        const double f = 1.0 / sqrt(nb);
#pragma unroll
        for (int m = 0; m < 4; ++m) b[m] *= f;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        s.v[8 + m] = a[m];
        s.v[12 + m] = b[m];
    }
}

// ---------------------------------------------------------------------------
// Geometry of the geodesic coordinates: integration context, domain radius
// (metric_domain_radius, src/geodesic/metric_domain.hpp:159-172) and the
// metric at a point.
// ---------------------------------------------------------------------------
template <int GEOM> struct GeomT;

template <> struct GeomT<G_BL> {
    using GeoType = Geo<double>;
    static KP_DEV GeoType geo(const Params& P, uint32_t* cnt) { return GeoType{P.mass, P.spin, cnt}; }
    static KP_DEV double radius(double x1) { return x1; }
    static KP_DEV Metric4<double> metric(const Params& P, const double* x) {
        double sn, cs;
        ksincos(x[2], &sn, &cs);
        return metric4_of(bl_gcov(P.mass, P.spin, x[1], sn, cs));
    }
};

template <> struct GeomT<G_FMKS> {
    using GeoType = GeoFmks<double>;
    static KP_DEV GeoType geo(const Params& P, uint32_t* cnt) { return GeoType{P.mass, P.spin, cnt, P.fmks}; }
    static KP_DEV double radius(double x1) { return exp(x1); }
    static KP_DEV Metric4<double> metric(const Params& P, const double* x) {
        return fmks_gcov(P.mass, P.spin, P.fmks, x[1], x[2]);
    }
};

// ---------------------------------------------------------------------------
// Wave-aggregated queue fetch: one atomic per wavefront.
// ---------------------------------------------------------------------------
KP_DEV bool fetch_pixel(uint32_t* counter, bool need, uint32_t npix, int& pixel) {
    const unsigned long long mask = __ballot(need);
    if (mask == 0ull) return false;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (!need) return false;
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    const uint32_t p = base + rank;
    if (p >= npix) return false;
    pixel = (int)p;
    return true;
}

// Wave-level sum of per-lane counters, one atomic per counter per wave.
KP_DEV void flush_counters(unsigned long long* counters, int rhs_slot, uint32_t rhs, uint32_t coeff,
                           uint32_t transfer, uint32_t steps_a, uint32_t steps_b, uint32_t guard) {
    unsigned long long v[6] = {rhs, coeff, transfer, steps_a, steps_b, guard};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        unsigned long long x = v[c];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        v[c] = x;
    }
    if (__lane_id() == 0) {
        atomicAdd(&counters[rhs_slot], v[0]);
        atomicAdd(&counters[C_COEFF], v[1]);
        atomicAdd(&counters[C_TRANSFER], v[2]);
        atomicAdd(&counters[C_STEPS_A], v[3]);
        atomicAdd(&counters[C_STEPS_B], v[4]);
        atomicAdd(&counters[C_GUARD], v[5]);
    }
}

// ---------------------------------------------------------------------------
// Pass A: trace_pass_a_segment_endpoint_pixel_metric (src/geodesic/pass_b.hpp:337-486)
// with the BL metric-domain hooks reduced (src/geodesic/metric_domain.hpp:159-223).
// ---------------------------------------------------------------------------
struct LaneA {
    LaneVec s, k1;  // the lane's state and its k1 scratch
    double h_current, h_try, call_min;
    int n, steps, steps_inside, entered, att, in_call, pixel, reason;
    // surface-crossing refinement, one trial per tick (refine: 0 off,
    // 1 evaluating the full step, 2 bisecting); best_ep: the best state so
    // far is in the endpoint buffer, not the lane state
    double lo, hi, flo, fhi, trial, full_h, surface;
    int refine, iter, use_secant, crossing, best_ep;
    // staged mixed precision: the FP32 kernel hands a step that crosses a
    // surface to the FP64 kernel (want64: park now, h_req: the step's
    // requested size); hold: the FP64 kernel keeps it until it commits a step
    double h_req;
    int want64, hold;
};

KP_DEV void lane_a_init(const Params& P, int pixel, LaneA& L) {
    St<double> cam;
    camera_ray(P, pixel, cam);
    st_store(L.s, cam);
    L.h_current = -absv(P.step);
    L.h_try = 0.0;
    L.call_min = 0.0;
    L.n = 0;
    L.steps = 0;
    L.steps_inside = 0;
    L.entered = 0;
    L.att = 0;
    L.in_call = 0;
    L.pixel = pixel;
    L.reason = R_MAX_STEPS;
    L.refine = 0;
    L.best_ep = 0;
    L.want64 = 0;
    L.hold = 0;
}

// One integration attempt from the lane state.  GP_64: FP64 derivatives.
// GP_32: FP32 derivatives and stage sums, increments added to the FP64
// state.  GP_GUARD: per lane, FP64 inside the guard radius and FP32 outside
// (the lanes of a wave may take both paths).  force64: FP64 whatever GP says
// (the surface-crossing refinement of a mixed-precision lane).  guard counts
// the FP64 integration attempts of a mixed-precision image: those of
// GP_GUARD lanes inside the radius and every attempt of a staged FP64 launch
// (P.stage >= 0), not the refinement trials.
template <int GP, class GG>
KP_DEV double attempt(const Params& P, const GG& G, const LaneVec& s, LaneVec& k1, double h,
                      St<double>& half, St<double>& next, uint32_t& guard, bool full = true,
                      bool force64 = false) {
    if (GP == GP_32 || (GP == GP_GUARD && !force64 && !(s[1] < P.fp64_radius)))
        return doubling_attempt_mixed(P.mass_f, P.spin_f, s, h, half, next, G.rhs_count, full);
    if (!force64 && (GP == GP_GUARD || P.stage >= 0)) ++guard;
    return doubling_attempt(G, s, h, k1, half, next, full);
}

template <int GP, class GG>
KP_DEV void step_rk4(const Params& P, const GG& G, const LaneVec& s, double h, St<double>& next,
                     uint32_t& guard, bool force64 = false) {
    if (GP == GP_32 || (GP == GP_GUARD && !force64 && !(s[1] < P.fp64_radius))) {
        rk4_mixed(P.mass_f, P.spin_f, s, h, next, G.rhs_count);
        return;
    }
    if (!force64 && (GP == GP_GUARD || P.stage >= 0)) ++guard;
    rk4(G, s, h, next);
}

// Surface-crossing refinement as part of the tick: the lane keeps the step's
// start state (first) and evaluates one trial path per tick -- the composed
// path of the adaptive step (its two half steps) or one RK4 step, both
// through the same call as every other lane's attempt, so a refining lane
// never runs alone in its wave.  The trial sequence and the bookkeeping are
// locate_metric_surface_crossing's (src/geodesic/pass_a_core.hpp:94-197):
// the full step first, then up to 12 secant-or-bisection trials, keeping
// the last state on the starting side.  Returns true when it has finished.
template <int GEOM>
KP_DEV bool refine_update(const Params& P, LaneA& L, const St<double>& t) {
    using GT = GeomT<GEOM>;
    auto keep = [&](const St<double>& v) {
        const size_t np = (size_t)P.npix;
#pragma unroll
        for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + L.pixel] = v.v[c];
        L.best_ep = 1;
    };
    if (L.refine == 1) {
        if (!st_finite(t)) return true;
        const double fhi = GT::radius(t.v[1]) - L.surface;
        if (!(fhi == fhi) || absv(fhi) >= 1e300) return true;
        if (fhi == 0.0) {
            keep(t);
            return true;
        }
        if ((L.flo > 0.0) == (fhi > 0.0)) return true;
        L.fhi = fhi;
        L.refine = 2;
        L.iter = 0;
        L.use_secant = 1;
    } else {
        bool ok = st_finite(t);
        double ft = 0.0;
        if (ok) {
            ft = GT::radius(t.v[1]) - L.surface;
            ok = (ft == ft) && absv(ft) < 1e300;
        }
        if (!ok) {
            L.hi = L.trial;
            L.use_secant = 0;
        } else if (ft == 0.0) {
            keep(t);
            return true;
        } else if ((L.flo > 0.0) == (ft > 0.0)) {
            L.lo = L.trial;
            L.flo = ft;
            keep(t);
        } else {
            L.hi = L.trial;
            L.fhi = ft;
        }
        if (++L.iter >= 12) return true;
    }
    const double width = L.hi - L.lo;
    double trial = (L.lo + L.hi) * 0.5;
    const double denom = L.flo - L.fhi;
    if (L.use_secant && absv(denom) > 1e-300) {
        const double secant = L.lo + width * L.flo / denom;
        const double gw = width * 0.1;
        if (secant > L.lo + gw && secant < L.hi - gw) trial = secant;
    }
    L.trial = trial;
    return false;
}

// One tick of pass A; returns true when the ray has terminated (L.reason set).
template <int GP, int GEOM>
KP_DEV bool pass_a_tick(const Params& P, const typename GeomT<GEOM>::GeoType& G, LaneA& L, uint32_t& steps_acc,
                        uint32_t& guard) {
    using GT = GeomT<GEOM>;
    const double eff_min = absv(P.min_step);
    if (!L.refine && !L.in_call) {
        if (L.n >= P.max_steps) {
            L.reason = R_MAX_STEPS;
            return true;
        }
        const double r = GT::radius(L.s[1]);
        if (!L.entered) {
            if (r <= P.outer) {
                L.entered = 1;
                L.steps_inside = 0;
            }
        } else {
            if (r - P.inner <= 0.0) {
                L.reason = R_INNER;
                return true;
            }
            if (L.steps_inside > 0 && r >= P.outer) {
                L.reason = R_ESCAPED;
                return true;
            }
        }
        L.h_req = L.h_current;
        if (P.adaptive) {
            const double req = L.h_current;
            L.call_min = eff_min;
            if (absv(req) < L.call_min) L.call_min = absv(req);
            double h = minv(maxv(absv(req), L.call_min), P.max_step);
            L.h_try = req < 0.0 ? -h : h;
            L.att = 0;
            L.in_call = 1;
        }
    }
    // one integration call for every lane: a refinement trial path, an
    // adaptive attempt or a fixed RK4 step
    const double h_call = L.refine ? L.full_h * L.trial : (P.adaptive ? L.h_try : L.h_current);
    St<double> next;
    double err = 0.0;
    if (P.adaptive) {
        St<double> half;
        err = attempt<GP>(P, G, L.s, L.k1, h_call, half, next, guard, !L.refine, L.refine);
    } else {
        step_rk4<GP>(P, G, L.s, h_call, next, guard, L.refine);
    }
    if (L.refine) {
        if (!refine_update<GEOM>(P, L, next)) return false;
        L.refine = 0;
        L.reason = L.crossing;
        L.steps += 1;
        steps_acc += 1;
        return true;
    }
    double used_h;
    if (P.adaptive) {
        // adaptive_rk4_step (src/geodesic/rk4.hpp:109-161), one attempt per tick
        const double tol = maxv(P.tol, 8.0 * eps<double>());
        const double req_sign = L.h_current < 0.0 ? -1.0 : 1.0;
        if (err <= tol) {
            double factor = 2.0;
            if (err > 0.0) factor = minv(2.0, maxv(0.2, 0.9 * pow(tol / err, 0.2)));
            const double nh = minv(P.max_step, maxv(L.call_min, absv(L.h_try) * factor));
            used_h = L.h_try;
            L.h_current = req_sign * nh;
            L.in_call = 0;
        } else {
            bool fail = false;
            if (absv(L.h_try) <= L.call_min * 1.0001) {
                fail = true;
            } else {
                const double f = minv(0.8, maxv(0.1, 0.9 * pow(tol / maxv(err, 1e-300), 0.25)));
                L.h_try *= f;
                if (++L.att >= 10) fail = true;
            }
            if (fail) {
                L.h_current = req_sign * maxv(L.call_min, absv(L.h_try) * 0.5);
                L.in_call = 0;
                L.n += 1;
                if (absv(L.h_current) <= eff_min * 1.0001) {
                    L.reason = R_UNDERFLOW;
                    return true;
                }
            }
            return false;
        }
    } else {
        used_h = L.h_current;
    }
    const double old_r = GT::radius(L.s[1]);
    const double next_r = GT::radius(next.v[1]);
    const double old_inner = old_r - P.inner;
    const double next_inner = next_r - P.inner;
    if (!L.entered && next_r <= P.outer) {
        L.entered = 1;
        L.steps_inside = 0;
    } else if (!L.entered && next_r >= old_r) {
        st_store(L.s, next);
        L.reason = R_ESCAPED;
        L.steps += 1;
        steps_acc += 1;
        return true;
    }
    int crossing = R_NONE;
    if (L.entered && old_inner > 0.0 && next_inner <= 0.0)
        crossing = R_INNER;
    else if (L.entered && L.steps_inside > 0 && old_r < P.outer && next_r >= P.outer)
        crossing = R_ESCAPED;
    if (crossing != R_NONE) {
        if (GP == GP_32 && P.park == PARK_INSIDE) {
            // the crossing is located in FP64: the FP64 kernel retakes this step
            L.h_current = L.h_req;
            L.in_call = 0;
            L.want64 = 1;
            return false;
        }
        // start the refinement from the step's start state (kept in L.s)
        L.crossing = crossing;
        L.surface = crossing == R_INNER ? P.inner : P.outer;
        L.full_h = used_h;
        L.lo = 0.0;
        L.hi = 1.0;
        L.flo = old_r - L.surface;
        if (L.flo == 0.0 || !st_finite(next)) {
            L.reason = crossing;
            L.steps += 1;
            steps_acc += 1;
            return true;
        }
        L.refine = 1;
        L.trial = 1.0;
        return false;
    }
    st_store(L.s, next);
    L.steps += 1;
    steps_acc += 1;
    if (L.entered) L.steps_inside += 1;
    L.n += 1;
    L.hold = 0;
    return false;
}

KP_DEV void lane_a_store(const Params& P, const LaneA& L) {
    const size_t np = (size_t)P.npix;
    if (!L.best_ep) {
#pragma unroll
        for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + L.pixel] = L.s[c];
    }
    const int valid = L.entered && (L.reason == R_INNER || L.reason == R_ESCAPED);
    P.ep_valid[L.pixel] = valid;
    P.ep_steps[L.pixel] = L.steps;
    P.ep_reason[L.pixel] = L.reason;
}

// Parking (staged mixed precision): between steps a lane's state moves to
// the endpoint buffer and the park arrays, and a later stage resumes it.
KP_DEV bool park_due(const Params& P, double r) {
    return (P.park == PARK_INSIDE && r < P.fp64_radius) || (P.park == PARK_OUTSIDE && r >= P.back_radius);
}

// Appends the pixels of the lanes that park, one atomic per wave; called by
// every lane of the wave.
KP_DEV void append_parked(const Params& P, bool park, int pixel) {
    const unsigned long long mask = __ballot(park);
    if (mask == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(P.count_out, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (park) P.list_out[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t)pixel;
}

KP_DEV uint32_t work_count(const Params& P) { return P.list_in ? *P.count_in : (uint32_t)P.npix; }

KP_DEV void lane_a_park(const Params& P, const LaneA& L) {
    const size_t np = (size_t)P.npix, p = (size_t)L.pixel;
#pragma unroll
    for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + p] = L.s[c];
    P.pk_d[p] = L.h_current;
    P.pk_i[p] = L.n;
    P.pk_i[np + p] = L.steps;
    P.pk_i[2 * np + p] = L.steps_inside;
    P.pk_i[3 * np + p] = L.entered;
    P.pk_i[4 * np + p] = L.want64;
}

KP_DEV void lane_a_resume(const Params& P, int pixel, LaneA& L) {
    const size_t np = (size_t)P.npix, p = (size_t)pixel;
#pragma unroll
    for (int c = 0; c < 16; ++c) vset(L.s, c, P.ep[(size_t)c * np + p]);
    L.h_current = P.pk_d[p];
    L.n = P.pk_i[p];
    L.steps = P.pk_i[np + p];
    L.steps_inside = P.pk_i[2 * np + p];
    L.entered = P.pk_i[3 * np + p];
    L.h_try = 0.0;
    L.call_min = 0.0;
    L.att = 0;
    L.in_call = 0;
    L.pixel = pixel;
    L.reason = R_MAX_STEPS;
    L.refine = 0;
    L.best_ep = 0;
    L.want64 = 0;
    L.hold = P.pk_i[4 * np + p];
}

// Per-lane LDS slots: [32][blockDim] doubles, components 0-15 the state and
// 16-31 the k1 scratch of the step-doubling attempt.
KP_DEV void lane_lds(LdsVec<double>& s, LdsVec<double>& k1) {
    extern __shared__ double kp_lds[];
    s = LdsVec<double>{kp_lds + threadIdx.x, (int)blockDim.x};
    k1 = LdsVec<double>{kp_lds + 16 * blockDim.x + threadIdx.x, (int)blockDim.x};
}
KP_DEV void lane_lds(St<double>&, St<double>&) {}

template <bool PERSISTENT, int GP, int GEOM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GP == GP_32 ? KP_WAVES_PER_EU_F32 : KP_WAVES_PER_EU))) void pass_a_kernel(Params P) {
    uint32_t rhs_count = 0, steps_a = 0, guard = 0;
    const auto G = GeomT<GEOM>::geo(P, &rhs_count);
    LaneA L;
    lane_lds(L.s, L.k1);
    bool has = false;
    if (PERSISTENT) {
        const uint32_t limit = work_count(P);
        bool exhausted = false;
        while (true) {
            int idx = 0;
            const bool need = !has && !exhausted;
            if (__ballot(need)) {
                if (fetch_pixel(P.fetch, need, limit, idx)) {
                    if (P.list_in)
                        lane_a_resume(P, (int)P.list_in[idx], L);
                    else
                        lane_a_init(P, idx, L);
                    has = true;
                } else if (need) {
                    exhausted = true;
                }
            }
            if (!__ballot(has)) break;
            if (P.park != PARK_NONE) {
                const bool park = has && !L.in_call && !L.refine &&
                                  (L.want64 || (!L.hold && park_due(P, GeomT<GEOM>::radius(L.s[1]))));
                if (park) {
                    lane_a_park(P, L);
                    has = false;
                }
                append_parked(P, park, L.pixel);
            }
            if (has && pass_a_tick<GP, GEOM>(P, G, L, steps_a, guard)) {
                lane_a_store(P, L);
                has = false;
            }
        }
    } else {
        const int pix = blockIdx.x * blockDim.x + threadIdx.x;
        if (pix < P.npix) {
            lane_a_init(P, pix, L);
            while (!pass_a_tick<GP, GEOM>(P, G, L, steps_a, guard)) {
            }
            lane_a_store(P, L);
        }
    }
    flush_counters(P.counters, C_RHS_A, rhs_count, 0, 0, steps_a, 0, guard);
}

// ---------------------------------------------------------------------------
// Pass B: trace_pass_b_segment_model_endpoint_pixel_metric (src/geodesic/pass_b.hpp:515-722).
// ---------------------------------------------------------------------------
enum : int { PH_IDLE = 0, PH_MAIN = 1, PH_CROSS = 2, PH_TIME = 3 };

struct LaneB {
    LaneVec s, k1;  // the lane's state and its k1 scratch
    Stokes<double> st;
    LdsVec<double> mf;  // multi-frequency: Stokes component 4 f + c of this lane, in LDS
    LdsVec<double> mf2; // single-pass multi-frequency: the tentative Stokes of a step
    Stokes<D3> sd;      // tangent mode: Stokes vector with its three parameter derivatives
    double h_current, cap, h, h_try, call_min, call_max;
    double req;     // slow light: the time a waiting ray needed (NaN otherwise)
    double x0lo, x0hi;  // time probe: x^0 range of the radiating samples
    int steps, reason, att, phase, crosses, reached, pixel, pass_a_steps;
    int win, vac;   // slow light: snapshot window; vacuum finish (1: entering, 2: stepping)
};

// transform_to_observer_basis (src/frame/stokes_rotation.hpp:39-57).
KP_DEV Stokes<double> observer_basis(const Stokes<double>& in, double r11, double r12, double r21, double r22) {
    Stokes<double> o;
    o.I = in.I;
    o.Q = 0.5 * in.Q * (r11 * r11 - r12 * r12 - r21 * r21 + r22 * r22) + in.U * (r11 * r12 - r21 * r22);
    o.U = in.Q * (r11 * r21 - r12 * r22) + in.U * (r11 * r22 + r12 * r21);
    o.V = (r11 * r22 - r12 * r21) * in.V;
    return o;
}

// Writes the pixel's outputs.  A mixed-precision lane carries its Stokes
// vector in nu_obs^3-scaled units; SCALED converts back to invariant ones.
// MF: nfreq Stokes vectors from LDS to the [nfreq][npix] image arrays.
template <bool SCALED, int MODE> KP_DEV void lane_b_finish(const Params& P, const LaneB& L, int valid) {
    St<double> cam, fin;
    camera_ray(P, L.pixel, cam);
    st_load(fin, L.s);
    const Metric4<double> g = GeomT<geom_of<MODE>()>::metric(P, fin.v);
    double r11 = 1.0, r12 = 0.0, r21 = 0.0, r22 = 1.0;
    Stokes<double> o = {0.0, 0.0, 0.0, 0.0};
    int reason = L.reason;
    if (valid) {
        // screen_overlap (pass_b.hpp:87-98) and transform_to_observer_basis
        // (src/frame/stokes_rotation.hpp:39-57)
        r11 = dot4(g, cam.v + 8, fin.v + 8);
        r12 = dot4(g, cam.v + 8, fin.v + 12);
        r21 = dot4(g, cam.v + 12, fin.v + 8);
        r22 = dot4(g, cam.v + 12, fin.v + 12);
        Stokes<double> in = L.st;
        if (SCALED) in = {in.I * P.inv_nu3, in.Q * P.inv_nu3, in.U * P.inv_nu3, in.V * P.inv_nu3};
        o = observer_basis(in, r11, r12, r21, r22);
        if (L.reached)
            reason = R_CAMERA;
        else if (reason != R_UNDERFLOW && reason != R_TIME && reason != R_SLOW_TIME)
            reason = R_MAX_STEPS;
    }
    const kp_image_data& O = P.out;
    const int p = L.pixel;
    if (lds_stokes<MODE>()) {
        for (int f = 0; f < P.nfreq; ++f) {
            Stokes<double> of = {0.0, 0.0, 0.0, 0.0};
            if (valid) {
                const double sc = SCALED ? P.inv_nu3_f[f] : 1.0;
                const Stokes<double> in = {L.mf[4 * f] * sc, L.mf[4 * f + 1] * sc, L.mf[4 * f + 2] * sc,
                                           L.mf[4 * f + 3] * sc};
                of = observer_basis(in, r11, r12, r21, r22);
            }
            const size_t q = (size_t)f * (size_t)P.npix + (size_t)p;
            if (O.image_i) O.image_i[q] = of.I;
            if (O.image_q) O.image_q[q] = of.Q;
            if (O.image_u) O.image_u[q] = of.U;
            if (O.image_v) O.image_v[q] = of.V;
        }
    } else if (MODE == M_TANGENT || MODE == M_STAN) {
        Stokes<double> ov = {0.0, 0.0, 0.0, 0.0};
        if (valid) ov = observer_basis({L.sd.I.v, L.sd.Q.v, L.sd.U.v, L.sd.V.v}, r11, r12, r21, r22);
        if (O.image_i) O.image_i[p] = ov.I;
        if (O.image_q) O.image_q[p] = ov.Q;
        if (O.image_u) O.image_u[p] = ov.U;
        if (O.image_v) O.image_v[p] = ov.V;
        const size_t np = (size_t)P.npix;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            Stokes<double> g = {0.0, 0.0, 0.0, 0.0};
            if (valid) g = observer_basis({L.sd.I.d[k], L.sd.Q.d[k], L.sd.U.d[k], L.sd.V.d[k]}, r11, r12, r21, r22);
            double* gk = P.grad + (size_t)k * 4 * np + (size_t)p;
            gk[0] = g.I;
            gk[np] = g.Q;
            gk[2 * np] = g.U;
            gk[3 * np] = g.V;
        }
    } else {
        if (O.image_i) O.image_i[p] = o.I;
        if (O.image_q) O.image_q[p] = o.Q;
        if (O.image_u) O.image_u[p] = o.U;
        if (O.image_v) O.image_v[p] = o.V;
    }
    if (O.closure_x)
        O.closure_x[p] = sqrt(sq(fin.v[1] - cam.v[1]) + sq(fin.v[2] - cam.v[2]) + sq(fin.v[3] - cam.v[3]));
    if (O.closure_k) {
        double ck = absv(fin.v[4] - cam.v[4]);
        for (int m = 1; m < 4; ++m) ck = maxv(ck, absv(fin.v[4 + m] - cam.v[4 + m]));
        O.closure_k[p] = ck;
    }
    if (O.final_null) O.final_null[p] = dot4(g, fin.v + 4, fin.v + 4);
    if (O.frame_error) O.frame_error[p] = frame_error4(g, fin);
    if (O.det_r) O.det_r[p] = r11 * r22 - r12 * r21;
    if (O.overlap_r11) O.overlap_r11[p] = r11;
    if (O.overlap_r12) O.overlap_r12[p] = r12;
    if (O.overlap_r21) O.overlap_r21[p] = r21;
    if (O.overlap_r22) O.overlap_r22[p] = r22;
    if (O.basis_identity_error)
        O.basis_identity_error[p] = maxv(maxv(absv(r11 - 1.0), absv(r22 - 1.0)), maxv(absv(r12), absv(r21)));
    if (O.basis_rotation_angle) O.basis_rotation_angle[p] = atan2(r21 - r12, r11 + r22);
    if (O.pass_a_steps) O.pass_a_steps[p] = L.pass_a_steps;
    if (O.steps) O.steps[p] = valid ? L.steps : 0;
    if (O.reason) O.reason[p] = reason;
    if constexpr (base_mode(MODE) == M_SLOW || MODE == M_STAN) {
        P.req_t[p] = L.req;
        if (L.req == L.req) count_add(&P.counters[C_WAIT], 1ull);
    }
    if constexpr (MODE == M_PROBE) {
        P.probe[p] = L.x0lo;
        P.probe[(size_t)P.npix + p] = L.x0hi;
    }
}

// Loads a pass A endpoint; returns false (after writing the pixel's outputs)
// when the endpoint is not a valid radiating segment.
template <int MODE> KP_DEV bool lane_b_init(const Params& P, int pixel, LaneB& L) {
    const size_t np = (size_t)P.npix;
#pragma unroll
    for (int c = 0; c < 16; ++c) vset(L.s, c, P.ep[(size_t)c * np + pixel]);
    L.pixel = pixel;
    L.pass_a_steps = P.ep_steps[pixel];
    L.reason = P.ep_reason[pixel];
    L.st = {0.0, 0.0, 0.0, 0.0};
    if (lds_stokes<MODE>())
        for (int j = 0; j < 4 * P.nfreq; ++j) L.mf.set(j, 0.0);
    if (MODE == M_TANGENT || MODE == M_STAN) L.sd = {D3(0.0), D3(0.0), D3(0.0), D3(0.0)};
    L.h_current = P.step;
    L.cap = P.max_rs > 0.0 ? P.max_rs : P.max_step;
    L.h = 0.0;
    L.h_try = 0.0;
    L.call_min = 0.0;
    L.call_max = 0.0;
    L.steps = 0;
    L.att = 0;
    L.phase = PH_IDLE;
    L.crosses = 0;
    L.reached = 0;
    L.win = 0;
    L.vac = 0;
    L.req = __builtin_nan("");
    L.x0lo = __builtin_nan("");
    L.x0hi = __builtin_nan("");
    if (MODE == M_PROBE) L.cap = 0.0;
    if (!P.ep_valid[pixel]) {
        lane_b_finish<false, MODE>(P, L, 0);
        return false;
    }
    return true;
}

template <int MODE> KP_DEV void lane_b_park(const Params& P, const LaneB& L) {
    const size_t np = (size_t)P.npix, p = (size_t)L.pixel;
#pragma unroll
    for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + p] = L.s[c];
    P.pk_d[p] = L.h_current;
    P.pk_d[np + p] = L.cap;
    P.pk_i[p] = L.steps;
    P.pk_i[np + p] = L.reason;
    if (lds_stokes<MODE>()) {
        for (int j = 0; j < 4 * P.nfreq; ++j) P.pk_st[(size_t)j * np + p] = L.mf[j];
    } else {
        P.pk_st[p] = L.st.I;
        P.pk_st[np + p] = L.st.Q;
        P.pk_st[2 * np + p] = L.st.U;
        P.pk_st[3 * np + p] = L.st.V;
    }
}

template <int MODE> KP_DEV void lane_b_resume(const Params& P, int pixel, LaneB& L) {
    const size_t np = (size_t)P.npix, p = (size_t)pixel;
#pragma unroll
    for (int c = 0; c < 16; ++c) vset(L.s, c, P.ep[(size_t)c * np + p]);
    L.pixel = pixel;
    L.pass_a_steps = P.ep_steps[p];
    L.h_current = P.pk_d[p];
    L.cap = P.pk_d[np + p];
    L.steps = P.pk_i[p];
    L.reason = P.pk_i[np + p];
    if (lds_stokes<MODE>()) {
        for (int j = 0; j < 4 * P.nfreq; ++j) L.mf.set(j, P.pk_st[(size_t)j * np + p]);
    } else {
        L.st = {P.pk_st[p], P.pk_st[np + p], P.pk_st[2 * np + p], P.pk_st[3 * np + p]};
    }
    L.h = 0.0;
    L.h_try = 0.0;
    L.call_min = 0.0;
    L.call_max = 0.0;
    L.att = 0;
    L.phase = PH_IDLE;
    L.crosses = 0;
    L.reached = 0;
    L.win = 0;
    L.vac = 0;
    L.req = __builtin_nan("");
    L.x0lo = __builtin_nan("");
    L.x0hi = __builtin_nan("");
}

KP_DEV void start_call(LaneB& L, double req, double cmin, double cmax) {
    L.call_min = cmin;
    L.call_max = cmax;
    L.h_try = minv(maxv(absv(req), cmin), cmax);
    L.att = 0;
}

// Coefficients at the sample state, the substep control they drive and one
// Strang transfer step over the affine length |h|.
//  * FP64: invariant coefficients over dl = |h| dl/dlambda.
//  * Mixed: FP32 coefficients already per unit |h| and in nu_obs^3-scaled
//    intensity units (the Stokes state is kept scaled until the ray ends).
//  * Multi-frequency (MF): the frequency-independent plasma state is evaluated
//    once; each control frequency's coefficients drive the substep count (the
//    largest wins, pass_b.hpp:1416-1460) and each output frequency's the
//    transfer of its own Stokes vector, held in LDS.
template <bool MIXED, int MODE> struct Radiation;

KP_DEV void mf_strang(const Params& P, LaneB& L, int f, const Coeffs<double>& c, double dl) {
    Stokes<double> st = {L.mf[4 * f], L.mf[4 * f + 1], L.mf[4 * f + 2], L.mf[4 * f + 3]};
    strang_step(st, c, dl);
    L.mf.set(4 * f, st.I);
    L.mf.set(4 * f + 1, st.Q);
    L.mf.set(4 * f + 2, st.U);
    L.mf.set(4 * f + 3, st.V);
}

KP_DEV void mf_strang32(const Params& P, LaneB& L, int f, const Coeffs<float>& c, float dl) {
    Stokes<float> st = {float(L.mf[4 * f]), float(L.mf[4 * f + 1]), float(L.mf[4 * f + 2]), float(L.mf[4 * f + 3])};
    strang_step32(st, c, dl);
    L.mf.set(4 * f, st.I);
    L.mf.set(4 * f + 1, st.Q);
    L.mf.set(4 * f + 2, st.U);
    L.mf.set(4 * f + 3, st.V);
}

KP_DEV Coeffs<double> widen(const Coeffs<float>& c) {
    return {c.jI, c.jQ, c.jU, c.jV, c.aI, c.aQ, c.aU, c.aV, c.rQ, c.rU, c.rV};
}

template <> struct Radiation<false, M_SINGLE> {
    Coeffs<double> c;
    KP_DEV void zero() { coeffs_zero(c); }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_coefficients(P.riaf, sample, c);
        return 1;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = substep_count(c, h, P.riaf.dlambda_scale, lmrs, P.max_rd, P.max_ad, P.max_fd);
        pred = substep_count(c, h_next, P.riaf.dlambda_scale, lmrs, P.max_rd, P.max_ad, P.max_fd);
        return 0;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        strang_step(L.st, c, absv(h) * P.riaf.dlambda_scale);
        return 0;
    }
};

template <> struct Radiation<true, M_SINGLE> {
    Coeffs<float> c;
    KP_DEV void zero() { coeffs_zero(c); }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_coefficients_scaled(P.rm, sample, c);
        return 1;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        const Coeffs<double> d = widen(c);
        req = substep_count(d, h, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd);
        pred = substep_count(d, h_next, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd);
        return 0;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        (void)P;
        Stokes<float> f = {float(L.st.I), float(L.st.Q), float(L.st.U), float(L.st.V)};
        strang_step32(f, c, float(absv(h)));
        L.st = {f.I, f.Q, f.U, f.V};
        return 0;
    }
};

// Thermal (NT false) and kappa / power-law (NT true) coefficient evaluation;
// with NT false the nonthermal branch is compiled out.
template <bool NT> struct RadiationMF {
    RiafLocal<double> loc;
    KP_DEV void zero() { loc.valid = 0; }
    KP_DEV bool valid() const { return loc.valid; }
    KP_DEV void at(const Params& P, double freq, Coeffs<double>& c) const {
        riaf_coeffs_at(P.riaf, loc, freq, c, NT ? &P.nt : nullptr);
    }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_local(P.riaf, sample, loc);
        return 0;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = 1;
        pred = 1;
        for (int f = 0; f < P.nctrl; ++f) {
            Coeffs<double> c;
            riaf_coeffs_at(P.riaf, loc, P.ctrl[f], c, NT ? &P.nt : nullptr);
            req = maxv(req, substep_count(c, h, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(c, h_next, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return loc.valid ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int f = 0; f < P.nfreq; ++f) {
            Coeffs<double> c;
            riaf_coeffs_at(P.riaf, loc, P.freqs[f], c, NT ? &P.nt : nullptr);
            mf_strang(P, L, f, c, absv(h) * P.dls_f[f]);
        }
        return loc.valid ? (uint32_t)P.nfreq : 0u;
    }
};
template <> struct Radiation<false, M_MF> : RadiationMF<false> {};
template <> struct Radiation<false, M_MF_NT> : RadiationMF<true> {};

template <> struct Radiation<true, M_MF> {
    RiafLocal<float> loc;
    KP_DEV void zero() { loc.valid = 0; }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_local_scaled(P.rm, sample, loc);
        return 0;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = 1;
        pred = 1;
        for (int f = 0; f < P.nctrl; ++f) {
            Coeffs<float> c;
            riaf_coeffs_scaled_at(P.rm, loc, float(P.ctrl[f]), P.bconst_c[f], c);
            const Coeffs<double> d = widen(c);
            req = maxv(req, substep_count(d, h, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(d, h_next, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return loc.valid ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int f = 0; f < P.nfreq; ++f) {
            Coeffs<float> c;
            riaf_coeffs_scaled_at(P.rm, loc, float(P.freqs[f]), P.bconst_f[f], c);
            mf_strang32(P, L, f, c, float(absv(h)));
        }
        return loc.valid ? (uint32_t)P.nfreq : 0u;
    }
};

template <bool NT> struct RadiationGRMHD {
    RiafLocal<double> loc;
    KP_DEV void zero() { loc.valid = 0; }
    KP_DEV bool valid() const { return loc.valid; }
    KP_DEV void at(const Params& P, double freq, Coeffs<double>& c) const {
        grmhd_coeffs_at(P.grmhd, loc, freq, c, NT ? &P.nt : nullptr);
    }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        grmhd_local(P.grmhd, sample, loc);
        return 0;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = 1;
        pred = 1;
        for (int f = 0; f < P.nctrl; ++f) {
            Coeffs<double> c;
            grmhd_coeffs_at(P.grmhd, loc, P.ctrl[f], c, NT ? &P.nt : nullptr);
            req = maxv(req, substep_count(c, h, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(c, h_next, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return loc.valid ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int f = 0; f < P.nfreq; ++f) {
            Coeffs<double> c;
            grmhd_coeffs_at(P.grmhd, loc, P.freqs[f], c, NT ? &P.nt : nullptr);
            mf_strang(P, L, f, c, absv(h) * P.dls_f[f]);
        }
        return loc.valid ? (uint32_t)P.nfreq : 0u;
    }
};
template <> struct Radiation<false, M_GRMHD> : RadiationGRMHD<false> {};
template <> struct Radiation<false, M_GRMHD_NT> : RadiationGRMHD<true> {};

// Slow light: the plasma state of snapshot pair (k, k+1) at time fraction f.
// Fluid interpolation mixes the sampled fluid before the reconstruction (lo
// holds the result); coefficient interpolation evaluates both snapshots (lo,
// hi) and mixes each frequency's coefficients (slow_light_driver.hpp:166-193).
template <bool NT> struct RadiationSlow {
    RiafLocal<double> lo, hi;
    double f;
    int k;
    KP_DEV void zero() {
        lo.valid = 0;
        hi.valid = 0;
        f = 0.0;
        k = 0;
    }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        if (P.slow_interp == KP_SLOW_LIGHT_FLUID) {
            grmhd_local_time(P.grmhd, P.snaps[k], P.snaps[k + 1], f, sample, lo);
        } else {
            grmhd_local_time(P.grmhd, P.snaps[k], P.snaps[k], 0.0, sample, lo);
            grmhd_local_time(P.grmhd, P.snaps[k + 1], P.snaps[k + 1], 0.0, sample, hi);
        }
        return 0;
    }
    KP_DEV bool valid() const { return lo.valid || hi.valid; }
    KP_DEV void at(const Params& P, double freq, Coeffs<double>& c) const {
        grmhd_coeffs_at(P.grmhd, lo, freq, c, NT ? &P.nt : nullptr);
        if (P.slow_interp != KP_SLOW_LIGHT_FLUID) {
            Coeffs<double> b;
            grmhd_coeffs_at(P.grmhd, hi, freq, b, NT ? &P.nt : nullptr);
            const double w = 1.0 - f;
            c.jI = w * c.jI + f * b.jI;
            c.jQ = w * c.jQ + f * b.jQ;
            c.jU = w * c.jU + f * b.jU;
            c.jV = w * c.jV + f * b.jV;
            c.aI = w * c.aI + f * b.aI;
            c.aQ = w * c.aQ + f * b.aQ;
            c.aU = w * c.aU + f * b.aU;
            c.aV = w * c.aV + f * b.aV;
            c.rQ = w * c.rQ + f * b.rQ;
            c.rU = w * c.rU + f * b.rU;
            c.rV = w * c.rV + f * b.rV;
        }
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = 1;
        pred = 1;
        for (int q = 0; q < P.nctrl; ++q) {
            Coeffs<double> c;
            at(P, P.ctrl[q], c);
            req = maxv(req, substep_count(c, h, P.dls_c[q], lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(c, h_next, P.dls_c[q], lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return (lo.valid || hi.valid) ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int q = 0; q < P.nfreq; ++q) {
            Coeffs<double> c;
            at(P, P.freqs[q], c);
            mf_strang(P, L, q, c, absv(h) * P.dls_f[q]);
        }
        return (lo.valid || hi.valid) ? (uint32_t)P.nfreq : 0u;
    }
};
template <> struct Radiation<false, M_SLOW> : RadiationSlow<false> {};
template <> struct Radiation<false, M_SLOW_NT> : RadiationSlow<true> {};

// Time probe: geometric stepping only (no coefficients, no substep control).
template <> struct Radiation<false, M_PROBE> {
    KP_DEV void zero() {}
    KP_DEV uint32_t eval(const Params&, const St<double>&) { return 0; }
    KP_DEV uint32_t control(const Params&, double, double, double, int& req, int& pred) const {
        req = pred = 1;
        return 0;
    }
    KP_DEV uint32_t transfer(const Params&, LaneB&, double) const { return 0; }
};

// Tangent mode: the plasma scalars carry their derivatives with respect to
// the log-scale parameters of PlasmaPerturbation (plasma_perturbation.hpp:12-33):
// q_density scales n_e and B^2 together (d ln n_e = dq, d ln B = dq / 2),
// q_temperature scales Theta_e, q_magnetic scales B.  The coefficients and
// the Strang step are then evaluated in Dual<3> arithmetic; the substep
// control uses the coefficient values, so the step sequence (the "fixed
// geodesic") is the FP64 kernel's.
KP_DEV void tangent_coeffs(const RiafDev<double>& M, const RiafLocal<double>& L, Coeffs<D3>& c) {
    coeffs_zero(c);
    if (!L.valid) return;
    D3 ne(L.ne), th(L.thetae), b(L.b);
    ne.d[0] = L.ne;
    b.d[0] = 0.5 * L.b;
    th.d[1] = L.thetae;
    b.d[2] = L.b;
    const D3 nu(maxv(M.freq_cgs * L.nu_scale, 1.0));
    if (L.aligned) {
        if (M.faraday_rotation) {
            const D3 theta(L.cos_theta == 1.0 ? 0.0 : PI_D);
            c.rV = D3(M.faraday_scale) * rho_v_aligned(nu, ne, th, b, theta, D3(M.min_sin_theta)) * nu;
        }
        return;
    }
    Coeffs<D3> mc;
    thermal_coeffs(M.fit, nu, ne, th, b, D3(L.sin_theta), D3(L.cos_theta), D3(M.max_pol_frac), D3(M.min_sin_theta),
                   D3(M.emission_scale), D3(M.absorption_scale), D3(M.faraday_scale), mc);
    c.jI = mc.jI;
    c.jV = mc.jV;
    c.aI = mc.aI;
    c.aV = mc.aV;
    if (M.faraday_rotation) c.rV = mc.rV;
    if (L.has_chi) {
        const D3 c2(L.cos2chi), s2(L.sin2chi);
        c.jQ = mc.jQ * c2;
        c.jU = mc.jQ * s2;
        c.aQ = mc.aQ * c2;
        c.aU = mc.aQ * s2;
        c.rQ = mc.rQ * c2;
        c.rU = mc.rQ * s2;
    }
}

template <> struct Radiation<false, M_TANGENT> {
    Coeffs<D3> c;
    KP_DEV void zero() { coeffs_zero(c); }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        RiafLocal<double> L;
        riaf_local(P.riaf, sample, L);
        tangent_coeffs(P.riaf, L, c);
        return 1;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        const Coeffs<double> v = value_of(c);
        req = substep_count(v, h, P.riaf.dlambda_scale, lmrs, P.max_rd, P.max_ad, P.max_fd);
        pred = substep_count(v, h_next, P.riaf.dlambda_scale, lmrs, P.max_rd, P.max_ad, P.max_fd);
        return 0;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        strang_step(L.sd, c, D3(absv(h) * P.riaf.dlambda_scale));
        return 0;
    }
};

// Slow-light gradients: the slow-light sampling with dTheta_e/dR_high, the
// Dexter/Pandya coefficients in Dual<3> arithmetic with respect to
// (q_density, R_high, q_magnetic) -- n_e e^q and B e^(q/2) for the density
// family, Theta_e(R_high) through the derived scalars, B e^q for the field
// (plasma_perturbation.hpp:17-32) -- and the Strang step on the dual Stokes
// vector.  The substep control uses the values, so the geodesic and the step
// sequence are the FP64 slow-light kernel's.
KP_DEV void grmhd_tangent_coeffs(const GrmhdDev& G, const RiafLocal<double>& L, double dth, double ne_f,
                                 double b_f, double freq, Coeffs<D3>& c) {
    coeffs_zero(c);
    if (!L.valid) return;
    const double nev = L.ne * ne_f, bv = L.b * b_f;
    D3 ne(nev), th(L.thetae), b(bv);
    ne.d[0] = nev;
    b.d[0] = 0.5 * bv;
    th.d[1] = dth;
    b.d[2] = bv;
    const D3 nu(maxv(freq * L.nu_scale, 1.0));
    Coeffs<D3> mc;
    thermal_coeffs(G.fit, nu, ne, th, b, D3(L.sin_theta), D3(L.cos_theta), D3(G.max_pol_frac), D3(1e-4),
                   D3(G.emission_scale), D3(G.absorption_scale), D3(G.faraday_scale), mc);
    c.jI = mc.jI;
    c.jV = mc.jV;
    c.aI = mc.aI;
    c.aV = mc.aV;
    if (G.faraday_rotation) c.rV = mc.rV;
    if (L.has_chi) {
        const D3 c2(L.cos2chi), s2(L.sin2chi);
        c.jQ = mc.jQ * c2;
        c.jU = mc.jQ * s2;
        c.aQ = mc.aQ * c2;
        c.aU = mc.aQ * s2;
        c.rQ = mc.rQ * c2;
        c.rU = mc.rQ * s2;
    }
}

template <> struct Radiation<false, M_STAN> {
    Coeffs<D3> c;
    double f;
    int k;
    KP_DEV void zero() {
        coeffs_zero(c);
        f = 0.0;
        k = 0;
    }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        RiafLocal<double> lo;
        double dlo;
        if (P.slow_interp == KP_SLOW_LIGHT_FLUID) {
            grmhd_local_time_dth(P.grmhd, P.snaps[k], P.snaps[k + 1], f, sample, lo, dlo);
            grmhd_tangent_coeffs(P.grmhd, lo, dlo, P.tq_ne, P.tq_b, P.freqs[0], c);
        } else {
            RiafLocal<double> hi;
            double dhi;
            grmhd_local_time_dth(P.grmhd, P.snaps[k], P.snaps[k], 0.0, sample, lo, dlo);
            grmhd_local_time_dth(P.grmhd, P.snaps[k + 1], P.snaps[k + 1], 0.0, sample, hi, dhi);
            Coeffs<D3> a, b;
            grmhd_tangent_coeffs(P.grmhd, lo, dlo, P.tq_ne, P.tq_b, P.freqs[0], a);
            grmhd_tangent_coeffs(P.grmhd, hi, dhi, P.tq_ne, P.tq_b, P.freqs[0], b);
            const D3 w(1.0 - f), ff(f);
            c.jI = w * a.jI + ff * b.jI;
            c.jQ = w * a.jQ + ff * b.jQ;
            c.jU = w * a.jU + ff * b.jU;
            c.jV = w * a.jV + ff * b.jV;
            c.aI = w * a.aI + ff * b.aI;
            c.aQ = w * a.aQ + ff * b.aQ;
            c.aU = w * a.aU + ff * b.aU;
            c.aV = w * a.aV + ff * b.aV;
            c.rQ = w * a.rQ + ff * b.rQ;
            c.rU = w * a.rU + ff * b.rU;
            c.rV = w * a.rV + ff * b.rV;
        }
        return 1;
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        const Coeffs<double> v = value_of(c);
        req = substep_count(v, h, P.dls_f[0], lmrs, P.max_rd, P.max_ad, P.max_fd);
        pred = substep_count(v, h_next, P.dls_f[0], lmrs, P.max_rd, P.max_ad, P.max_fd);
        return 0;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        strang_step(L.sd, c, D3(absv(h) * P.dls_f[0]));
        return 0;
    }
};

// SlowLightTemporalModel::interpolation_fraction (slow_light_driver.hpp:62-69).
KP_DEV double slow_fraction(double t_lo, double t_hi, double t_obs, double x0) {
    const double denom = t_hi - t_lo;
    double f = 0.0;
    if (absv(denom) > 1e-300) f = (t_obs + x0 - t_lo) / denom;
    return minv(1.0, maxv(0.0, f));
}

// Slow light: the end of a ray's current window -- the next snapshot pair
// (snapshot stepping) or, after the last, the vacuum finish (vac = 1).
KP_DEV void slow_window_end(const Params& P, LaneB& L) {
    if (P.slow_mode == KP_SLOW_LIGHT_SNAPSHOT && L.win + 2 < P.nsnap)
        L.win += 1;
    else
        L.vac = 1;
}
// The time that ends the current window (snapshot: t_win+1; block: the last).
KP_DEV double slow_upper(const Params& P, const LaneB& L) {
    return P.slow_mode == KP_SLOW_LIGHT_SNAPSHOT ? P.snap_t[L.win + 1] : P.snap_t[P.nsnap - 1];
}
// SlowLightResidentWindows::sample_window (slow_light_driver.hpp:1000-1012).
KP_DEV int slow_pair(const Params& P, double t) {
    int lo = 0, hi = P.nsnap - 1;
    while (lo + 1 < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (t < P.snap_t[mid])
            hi = mid;
        else
            lo = mid;
    }
    return lo;
}

// One tick of pass B; returns true when the ray has terminated.
// Slow light (MODE == M_SLOW) follows the window kernel of
// launch_slow_light_rays_metric (slow_light_driver.hpp:1083-1483) across all
// windows of the ray in turn, then advance_slow_light_vacuum_to_camera
// (:1572-1671) for a ray still unfinished after the last window (vac).
template <bool MIXED, int GP, int MODE>
KP_DEV bool pass_b_tick(const Params& P, const typename GeomT<geom_of<MODE>()>::GeoType& G, LaneB& L,
                        uint32_t& coeff_acc, uint32_t& transfer_acc, uint32_t& steps_acc, uint32_t& guard) {
    using GT = GeomT<geom_of<MODE>()>;
    constexpr bool SLOW = base_mode(MODE) == M_SLOW || MODE == M_STAN;
    const double eff_min = absv(P.min_step);
    if (L.phase == PH_IDLE) {
        if (L.reached || L.steps >= P.max_steps) return true;
        if constexpr (SLOW) {
            if (!L.vac && P.slow_mode != KP_SLOW_LIGHT_DECOUPLED) {
                while (!L.vac && P.t_obs + L.s[0] >= slow_upper(P, L)) slow_window_end(P, L);
            }
            if (L.vac == 1) {
                const double tol_r = 1e-8 * maxv(1.0, absv(P.outer));
                if (GT::radius(L.s[1]) < P.outer - tol_r) {
                    L.reason = R_SLOW_TIME;
                    return true;
                }
                L.vac = 2;
            }
        }
        double h = L.h_current;
        if (!(SLOW && L.vac) && GT::radius(L.s[1]) <= P.outer && L.cap > 0.0) h = minv(h, L.cap);
        L.h = h;
        if (P.adaptive) start_call(L, h, eff_min, P.max_step);
        L.phase = PH_MAIN;
    }
    St<double> next, sample;
    if (P.adaptive) {
        const double err = attempt<GP>(P, G, L.s, L.k1, L.h_try, sample, next, guard);
        const double tol = maxv(P.tol, 8.0 * eps<double>());
        if (err <= tol) {
            double factor = 2.0;
            if (err > 0.0) factor = minv(2.0, maxv(0.2, 0.9 * pow(tol / err, 0.2)));
            L.h = absv(L.h_try);
            if (L.phase == PH_MAIN) L.h_current = minv(L.call_max, maxv(L.call_min, absv(L.h_try) * factor));
        } else {
            bool fail = false;
            if (absv(L.h_try) <= L.call_min * 1.0001) {
                fail = true;
            } else {
                L.h_try *= minv(0.8, maxv(0.1, 0.9 * pow(tol / maxv(err, 1e-300), 0.25)));
                if (++L.att >= 10) fail = true;
            }
            if (fail) {
                L.h_current = maxv(L.call_min, absv(L.h_try) * 0.5);
                const bool main = L.phase == PH_MAIN;
                L.phase = PH_IDLE;
                if (main && L.h_current <= eff_min * 1.0001) {
                    // the slow-light window and vacuum loops end such a ray as max_steps
                    L.reason = SLOW ? R_MAX_STEPS : R_UNDERFLOW;
                    return true;
                }
            }
            return false;
        }
    } else {
        step_rk4<GP>(P, G, L.s, L.h, next, guard);
        st_copy(sample, next);
    }
    const bool truncating = SLOW && !L.vac && P.slow_mode != KP_SLOW_LIGHT_DECOUPLED;
    if (L.phase == PH_MAIN) {
        // camera_surface_value/crossed/crossing_fraction for the pinhole sphere
        // (pass_b.hpp:54-85, 605-637)
        const double s0 = GT::radius(L.s[1]) - P.radius;
        const double s1 = GT::radius(next.v[1]) - P.radius;
        int crosses = (s0 == 0.0 || s1 == 0.0) ? 1 : ((s0 < 0.0 && s1 > 0.0) || (s0 > 0.0 && s1 < 0.0));
        double frac = 2.0;
        if (crosses) {
            const double denom = s0 - s1;
            frac = absv(denom) <= 1e-300 ? 1.0 : minv(1.0, maxv(0.0, s0 / denom));
        }
        if (truncating) {
            // a step that crosses the window's end time first is cut there
            // (slow_light_driver.hpp:1198-1247)
            const double old_t = P.t_obs + L.s[0];
            const double next_t = P.t_obs + next.v[0];
            const double upper = slow_upper(P, L);
            if (next_t > upper && next_t > old_t) {
                const double tf = minv(1.0, maxv(0.0, (upper - old_t) / (next_t - old_t)));
                if (!crosses || tf < frac) {
                    L.crosses = 0;
                    const double h_b = L.h * tf;
                    if (absv(h_b) <= 1e-14) {
                        slow_window_end(P, L);
                        L.phase = PH_IDLE;
                        return false;
                    }
                    L.h = h_b;
                    if (P.adaptive) {
                        const double cmin = minv(eff_min, absv(h_b));
                        start_call(L, h_b, cmin, maxv(cmin, absv(h_b)));
                        L.phase = PH_TIME;
                        return false;
                    }
                    step_rk4<GP>(P, G, L.s, L.h, next, guard);
                    st_copy(sample, next);
                    crosses = 0;
                    L.phase = PH_TIME;
                }
            }
        }
        if (L.phase == PH_MAIN) {
            L.crosses = crosses;
            if (crosses) {
                const double h_cross = L.h * frac;
                if (absv(h_cross) <= 1e-14) {
                    L.h = 0.0;
                    st_load(next, L.s);
                    st_load(sample, L.s);
                } else if (frac < 0.999999999999) {
                    L.h = h_cross;
                    if (P.adaptive) {
                        const double cmin = minv(eff_min, absv(h_cross));
                        start_call(L, h_cross, cmin, maxv(cmin, absv(h_cross)));
                        L.phase = PH_CROSS;
                        return false;
                    }
                    step_rk4<GP>(P, G, L.s, L.h, next, guard);
                    st_copy(sample, next);
                }
            }
        }
    }
    if constexpr (SLOW) {
        if (L.phase == PH_TIME) {
            // two half steps that round back to the old time end the window,
            // or the ray when the gap is wider than rounding (:1227-1245)
            L.crosses = 0;
            if (next.v[0] == L.s[0]) {
                const double old_t = P.t_obs + L.s[0];
                const double upper = slow_upper(P, L);
                const double scale = maxv(absv(L.s[0]), maxv(absv(upper), absv(old_t - L.s[0])));
                const double rounding = 0.5 * 8.0 * eps<double>() * scale;
                L.phase = PH_IDLE;
                if (absv(upper - old_t) > rounding) {
                    L.reason = R_UNDERFLOW;
                    return true;
                }
                slow_window_end(P, L);
                return false;
            }
        }
    }
    L.phase = PH_IDLE;
    const double h = L.h;
    const double old_r = GT::radius(L.s[1]), sample_r = GT::radius(sample.v[1]), next_r = GT::radius(next.v[1]);
    if constexpr (SLOW) {
        if (L.vac) {
            // outward vacuum only (:1647-1660)
            const double tol_r = 1e-8 * maxv(1.0, absv(P.outer));
            if (!(next_r == next_r) || absv(next_r) > 1e300 || !(sample_r == sample_r) || absv(sample_r) > 1e300) {
                L.reason = R_MAX_STEPS;
                return true;
            }
            if (next_r < old_r - tol_r || sample_r < P.outer - tol_r || next_r < P.outer - tol_r) {
                L.reason = R_SLOW_TIME;
                return true;
            }
            st_store(L.s, next);
            L.steps += 1;
            steps_acc += 1;
            if (L.crosses) L.reached = 1;
            return false;
        }
    }
    if constexpr (MODE == M_PROBE) {
        // run_slow_light_time_probe_metric (slow_light_driver.hpp:712-721)
        if (h > 0.0 && sample_r >= P.inner && sample_r < P.outer) {
            const double x0 = sample.v[0];
            L.x0lo = L.x0lo == L.x0lo ? minv(L.x0lo, x0) : x0;
            L.x0hi = L.x0hi == L.x0hi ? maxv(L.x0hi, x0) : x0;
        }
        st_store(L.s, next);
        L.steps += 1;
        steps_acc += 1;
        if (L.crosses) L.reached = 1;
        return false;
    }
    Radiation<MIXED, MODE> rad;
    rad.zero();
    if constexpr (SLOW) {
        const int radiating = sample_r >= P.inner && sample_r < P.outer;
        const double ct = P.t_obs + sample.v[0];
        if (P.slow_mode == KP_SLOW_LIGHT_DECOUPLED && radiating && !(ct >= P.snap_t[0] && ct <= P.snap_t[P.nsnap - 1])) {
            // the time is not in the sequence: with every snapshot resident no
            // later block brings it (slow_light_driver.hpp:1303-1309,
            // slow_light_batch_driver.hpp:282-290)
            L.req = ct;
            L.reason = R_SLOW_TIME;
            return true;
        }
        const int k = P.slow_mode == KP_SLOW_LIGHT_SNAPSHOT ? L.win : slow_pair(P, ct);
        rad.k = k;
        rad.f = slow_fraction(P.snap_t[k], P.snap_t[k + 1], P.t_obs, sample.v[0]);
        if (radiating) coeff_acc += rad.eval(P, sample);
    } else {
        if (sample_r - P.inner > 0.0 && sample_r < P.outer) coeff_acc += rad.eval(P, sample);
    }
    const double lmrs = (P.max_rs > 0.0 && (old_r <= P.outer || sample_r <= P.outer || next_r <= P.outer))
                            ? P.max_rs : 0.0;
    int required, pred;
    constexpr bool SHARED = !MIXED && (base_mode(MODE) == M_MF || base_mode(MODE) == M_GRMHD || base_mode(MODE) == M_SLOW);
    if constexpr (SHARED) {
        if (P.single_pass) {
            // the control frequencies are the output frequencies: one
            // coefficient evaluation per frequency serves the substep control
            // (pass_b.hpp:1416-1460) and a tentative transfer of that
            // frequency's Stokes vector into the second buffer
            required = 1;
            pred = 1;
            const bool valid = rad.valid();
            for (int f = 0; f < P.nfreq; ++f) {
                Coeffs<double> c;
                rad.at(P, P.freqs[f], c);
                required = maxv(required, substep_count(c, h, P.dls_f[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
                pred = maxv(pred, substep_count(c, L.h_current, P.dls_f[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
                if (h > 0.0) {
                    Stokes<double> st = {L.mf[4 * f], L.mf[4 * f + 1], L.mf[4 * f + 2], L.mf[4 * f + 3]};
                    strang_step(st, c, absv(h) * P.dls_f[f]);
                    L.mf2.set(4 * f, st.I);
                    L.mf2.set(4 * f + 1, st.Q);
                    L.mf2.set(4 * f + 2, st.U);
                    L.mf2.set(4 * f + 3, st.V);
                }
            }
            if (valid) coeff_acc += (uint32_t)P.nfreq;
            if (required > 1 && h > eff_min * 1.0001) {
                L.h_current = maxv(eff_min, h / double(required));
                L.cap = L.h_current;
                return false;
            }
            if (lmrs > 0.0 || P.max_rd > 0.0 || P.max_ad > 0.0 || P.max_fd > 0.0)
                L.cap = pred > 1 ? maxv(eff_min, L.h_current / double(pred)) : P.max_step;
            else
                L.cap = P.max_step;
            if (h > 0.0) {
                const LdsVec<double> t = L.mf;
                L.mf = L.mf2;
                L.mf2 = t;
                transfer_acc += 1;
            }
            st_store(L.s, next);
            L.steps += 1;
            steps_acc += 1;
            if (L.crosses) L.reached = 1;
            return false;
        }
    }
    coeff_acc += rad.control(P, h, L.h_current, lmrs, required, pred);
    if (required > 1 && h > eff_min * 1.0001) {
        L.h_current = maxv(eff_min, h / double(required));
        L.cap = L.h_current;
        return false;
    }
    if (lmrs > 0.0 || P.max_rd > 0.0 || P.max_ad > 0.0 || P.max_fd > 0.0)
        L.cap = pred > 1 ? maxv(eff_min, L.h_current / double(pred)) : P.max_step;
    else
        L.cap = P.max_step;
    if (h > 0.0) {
        coeff_acc += rad.transfer(P, L, h);
        transfer_acc += 1;
    }
    st_store(L.s, next);
    L.steps += 1;
    steps_acc += 1;
    if (L.crosses) L.reached = 1;
    return false;
}

template <bool PERSISTENT, bool MIXED, int GP, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GP == GP_32 ? KP_WAVES_PER_EU_F32 : KP_WAVES_PER_EU))) void pass_b_kernel(Params P) {
    uint32_t rhs_count = 0, coeff = 0, transfer = 0, steps_b = 0, guard = 0;
    const auto G = GeomT<geom_of<MODE>()>::geo(P, &rhs_count);
    LaneB L;
    lane_lds(L.s, L.k1);
    if (lds_stokes<MODE>()) {
        extern __shared__ double kp_lds[];
        L.mf = LdsVec<double>{kp_lds + (KP_LANE_LDS ? 32 : 0) * blockDim.x + threadIdx.x, (int)blockDim.x};
        L.mf2 = LdsVec<double>{kp_lds + ((KP_LANE_LDS ? 32 : 0) + 4 * P.nfreq) * blockDim.x + threadIdx.x,
                               (int)blockDim.x};
    }
    bool has = false;
    if (PERSISTENT) {
        const uint32_t limit = work_count(P);
        bool exhausted = false;
        while (true) {
            // refill until every lane holds a radiating ray or the queue is dry
            while (true) {
                const bool need = !has && !exhausted;
                if (!__ballot(need)) break;
                int idx = 0;
                if (fetch_pixel(P.fetch, need, limit, idx)) {
                    if (MIXED && P.list_in) {
                        lane_b_resume<MODE>(P, (int)P.list_in[idx], L);
                        has = true;
                    } else {
                        has = lane_b_init<MODE>(P, idx, L);
                    }
                } else if (need) {
                    exhausted = true;
                }
            }
            if (!__ballot(has)) break;
            if (MIXED && P.park != PARK_NONE) {
                using GT = GeomT<geom_of<MODE>()>;
                const bool park = has && L.phase == PH_IDLE && park_due(P, GT::radius(L.s[1]));
                if (park) {
                    lane_b_park<MODE>(P, L);
                    has = false;
                }
                append_parked(P, park, L.pixel);
            }
            if (has && pass_b_tick<MIXED, GP, MODE>(P, G, L, coeff, transfer, steps_b, guard)) {
                lane_b_finish<MIXED, MODE>(P, L, 1);
                has = false;
            }
        }
    } else {
        const int pix = blockIdx.x * blockDim.x + threadIdx.x;
        if (pix < P.npix && lane_b_init<MODE>(P, pix, L)) {
            while (!pass_b_tick<MIXED, GP, MODE>(P, G, L, coeff, transfer, steps_b, guard)) {
            }
            lane_b_finish<MIXED, MODE>(P, L, 1);
        }
    }
    flush_counters(P.counters, C_RHS_B, rhs_count, coeff, transfer, 0, steps_b, guard);
}

// ---------------------------------------------------------------------------
// Unsplit transport (split_transport = 0): one kernel in which each lane
// integrates pass A and then pass B of the same pixel, as the reference's
// trace_pass_b_segment_model_pixel_metric (pass_b.hpp:768-1114) does; the
// results are those of the split kernels.  The endpoint goes through the
// workspace (written and read back by the same lane).  Lanes fetch pixels
// from one queue, one wave-level atomic per refill; a wave refills only when
// all of its lanes have finished both passes.
// ---------------------------------------------------------------------------
template <bool MIXED, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP_WAVES_PER_EU))) void unsplit_kernel(Params P) {
    constexpr int GP = MIXED ? GP_GUARD : GP_64;
    uint32_t rhs_a = 0, rhs_b = 0, coeff = 0, transfer = 0, steps_a = 0, steps_b = 0, guard = 0;
    while (true) {
        int pix = 0;
        const bool has = fetch_pixel(&P.queue[0], true, (uint32_t)P.npix, pix);
        if (!__ballot(has)) break;
        if (has) {
            constexpr int GEOM = geom_of<MODE>();
            {
                const auto G = GeomT<GEOM>::geo(P, &rhs_a);
                LaneA A;
                lane_lds(A.s, A.k1);
                lane_a_init(P, pix, A);
                while (!pass_a_tick<GP, GEOM>(P, G, A, steps_a, guard)) {
                }
                lane_a_store(P, A);
            }
            const auto G = GeomT<GEOM>::geo(P, &rhs_b);
            LaneB B;
            lane_lds(B.s, B.k1);
            if (lds_stokes<MODE>()) {
                extern __shared__ double kp_lds[];
                B.mf = LdsVec<double>{kp_lds + (KP_LANE_LDS ? 32 : 0) * blockDim.x + threadIdx.x, (int)blockDim.x};
                B.mf2 = LdsVec<double>{kp_lds + ((KP_LANE_LDS ? 32 : 0) + 4 * P.nfreq) * blockDim.x + threadIdx.x,
                                       (int)blockDim.x};
            }
            if (lane_b_init<MODE>(P, pix, B)) {
                while (!pass_b_tick<MIXED, GP, MODE>(P, G, B, coeff, transfer, steps_b, guard)) {
                }
                lane_b_finish<MIXED, MODE>(P, B, 1);
            }
        }
    }
    flush_counters(P.counters, C_RHS_A, rhs_a, 0, 0, steps_a, 0, guard);
    flush_counters(P.counters, C_RHS_B, rhs_b, coeff, transfer, 0, steps_b, 0);
}

// ---------------------------------------------------------------------------
// Batched images: many jobs (each its own camera, spin, model and outputs) in
// one persistent launch per pass.  One queue runs over the concatenated rays
// of all jobs; a lane that fetches global ray index g finds its job by binary
// search over the jobs' first indices and then reads that job's parameters
// from the device job table (lanes of one wave may serve different jobs).
// The reference runs one kernel per job on partitioned execution instances
// (src/image/batch_driver.hpp:126-187).
// ---------------------------------------------------------------------------
struct BatchArgs {
    const Params* jobs;        // [njobs] device job table
    const uint32_t* start;     // [njobs + 1] first global ray index of each job
    int32_t njobs;
    uint32_t total;            // rays over all jobs
    uint32_t* queue;           // [2] pass A / pass B fetch counters
    unsigned long long* counters;
};

KP_DEV int find_job(const BatchArgs& B, uint32_t g) {
    int lo = 0, hi = B.njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.start[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP_WAVES_PER_EU))) void pass_a_batch_kernel(BatchArgs B) {
    uint32_t rhs_count = 0, steps_a = 0, guard = 0;
    LaneA L;
    lane_lds(L.s, L.k1);
    const Params* P = B.jobs;
    bool has = false, exhausted = false;
    while (true) {
        int g = 0;
        const bool need = !has && !exhausted;
        if (__ballot(need)) {
            if (fetch_pixel(&B.queue[0], need, B.total, g)) {
                const int j = find_job(B, (uint32_t)g);
                P = &B.jobs[j];
                lane_a_init(*P, g - (int)B.start[j], L);
                has = true;
            } else if (need) {
                exhausted = true;
            }
        }
        if (!__ballot(has)) break;
        if (has) {
            const Geo<double> G{P->mass, P->spin, &rhs_count};
            if (pass_a_tick<MIXED ? GP_GUARD : GP_64, G_BL>(*P, G, L, steps_a, guard)) {
                lane_a_store(*P, L);
                has = false;
            }
        }
    }
    flush_counters(B.counters, C_RHS_A, rhs_count, 0, 0, steps_a, 0, guard);
}

template <bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP_WAVES_PER_EU))) void pass_b_batch_kernel(BatchArgs B) {
    uint32_t rhs_count = 0, coeff = 0, transfer = 0, steps_b = 0, guard = 0;
    LaneB L;
    lane_lds(L.s, L.k1);
    const Params* P = B.jobs;
    bool has = false, exhausted = false;
    while (true) {
        while (true) {
            const bool need = !has && !exhausted;
            if (!__ballot(need)) break;
            int g = 0;
            if (fetch_pixel(&B.queue[1], need, B.total, g)) {
                const int j = find_job(B, (uint32_t)g);
                P = &B.jobs[j];
                has = lane_b_init<M_SINGLE>(*P, g - (int)B.start[j], L);
            } else if (need) {
                exhausted = true;
            }
        }
        if (!__ballot(has)) break;
        if (has) {
            const Geo<double> G{P->mass, P->spin, &rhs_count};
            if (pass_b_tick<MIXED, MIXED ? GP_GUARD : GP_64, M_SINGLE>(*P, G, L, coeff, transfer, steps_b, guard)) {
                lane_b_finish<MIXED, M_SINGLE>(*P, L, 1);
                has = false;
            }
        }
    }
    flush_counters(B.counters, C_RHS_B, rhs_count, coeff, transfer, 0, steps_b, guard);
}

// Derived scalars of every GRMHD cell from its primitives, one cell per
// thread: reads 32 and writes 40 bytes per cell (HBM-bound).
__global__ __launch_bounds__(256) void derive_scalars_kernel(GrmhdDev G, kp_grmhd_cell* cells, int64_t ncell) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int64_t per_i = (int64_t)G.n2 * G.n3;
    const int i = (int)(c / per_i);
    const int j = (int)((c / G.n3) % G.n2);
    grmhd_derive_cell(G, i, j, cells[c]);
}

__global__ void zero_kernel(uint32_t* q, int nq, unsigned long long* c) {
    const int t = threadIdx.x;
    if (t < nq) q[t] = 0u;
    if (t < C_COUNT) c[t] = 0ull;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct DeviceWorkspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

std::mutex g_ws_mutex;
std::vector<DeviceWorkspace> g_ws;

kp_status workspace_for(int device, size_t bytes, void** out) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if ((int)g_ws.size() <= device) g_ws.resize(device + 1);
    DeviceWorkspace& w = g_ws[device];
    if (w.bytes < bytes) {
        if (w.ptr) {
            if (hipDeviceSynchronize() != hipSuccess) return kp_internal::fail(KP_HIP_ERROR, "hipDeviceSynchronize failed");
            (void)hipFree(w.ptr);
            w.ptr = nullptr;
            w.bytes = 0;
        }
        if (hipMalloc(&w.ptr, bytes) != hipSuccess) {
            return kp_internal::fail(KP_OUT_OF_MEMORY, "hipMalloc of the image workspace failed");
        }
        w.bytes = bytes;
    }
    *out = w.ptr;
    return KP_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

#define KP_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return kp_internal::fail(KP_HIP_ERROR, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <bool PERSISTENT> const void* pass_a_fn(int gp, int geom = G_BL) {
    if (geom == G_FMKS) return (const void*)pass_a_kernel<PERSISTENT, GP_64, G_FMKS>;
    if (gp == GP_GUARD) return (const void*)pass_a_kernel<PERSISTENT, GP_GUARD, G_BL>;
    if constexpr (PERSISTENT) {
        if (gp == GP_32) return (const void*)pass_a_kernel<true, GP_32, G_BL>;
    }
    return (const void*)pass_a_kernel<PERSISTENT, GP_64, G_BL>;
}

// mixed: FP32 radiation (nu^3-scaled units); gp: geodesic precision.  The
// staged variants (GP_32, and GP_64 with FP32 radiation) exist as persistent
// kernels only.
template <bool PERSISTENT> const void* pass_b_fn(bool mixed, int gp, int mode) {
    if (mode == M_TANGENT) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_TANGENT>;
    if (mode == M_STAN) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_STAN>;
    if (mode == M_GRMHD) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_GRMHD>;
    if (mode == M_SLOW) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_SLOW>;
    if (mode == M_MF_NT) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_MF_NT>;
    if (mode == M_GRMHD_NT) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_GRMHD_NT>;
    if (mode == M_SLOW_NT) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_SLOW_NT>;
    if (mode == M_PROBE) return (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_PROBE>;
    const bool mf = mode == M_MF;
    if (!mixed)
        return mf ? (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_MF>
                  : (const void*)pass_b_kernel<PERSISTENT, false, GP_64, M_SINGLE>;
    if constexpr (PERSISTENT) {
        if (gp == GP_32)
            return mf ? (const void*)pass_b_kernel<true, true, GP_32, M_MF>
                      : (const void*)pass_b_kernel<true, true, GP_32, M_SINGLE>;
        if (gp == GP_64)
            return mf ? (const void*)pass_b_kernel<true, true, GP_64, M_MF>
                      : (const void*)pass_b_kernel<true, true, GP_64, M_SINGLE>;
    }
    return mf ? (const void*)pass_b_kernel<PERSISTENT, true, GP_GUARD, M_MF>
              : (const void*)pass_b_kernel<PERSISTENT, true, GP_GUARD, M_SINGLE>;
}

const void* unsplit_fn(bool mixed, int mode) {
    if (mode == M_TANGENT) return (const void*)unsplit_kernel<false, M_TANGENT>;
    if (mode == M_STAN) return (const void*)unsplit_kernel<false, M_STAN>;
    if (mode == M_GRMHD) return (const void*)unsplit_kernel<false, M_GRMHD>;
    if (mode == M_SLOW) return (const void*)unsplit_kernel<false, M_SLOW>;
    if (mode == M_MF_NT) return (const void*)unsplit_kernel<false, M_MF_NT>;
    if (mode == M_GRMHD_NT) return (const void*)unsplit_kernel<false, M_GRMHD_NT>;
    if (mode == M_SLOW_NT) return (const void*)unsplit_kernel<false, M_SLOW_NT>;
    if (mode == M_PROBE) return (const void*)unsplit_kernel<false, M_PROBE>;
    if (mode == M_MF)
        return mixed ? (const void*)unsplit_kernel<true, M_MF> : (const void*)unsplit_kernel<false, M_MF>;
    return mixed ? (const void*)unsplit_kernel<true, M_SINGLE> : (const void*)unsplit_kernel<false, M_SINGLE>;
}

int geom_of_mode(int mode) {
    const int b = base_mode(mode);
    return (b == M_GRMHD || b == M_SLOW || b == M_PROBE || b == M_STAN) ? G_FMKS : G_BL;
}

// LDS bytes per block: 32 doubles per lane when the lane state lives in LDS,
// plus 4 doubles per lane and output frequency for multi-frequency Stokes.
size_t lds_bytes(int block, int nfreq_mf, int stokes_buffers = 1) {
    return (size_t(KP_LANE_LDS ? 32 : 0) + size_t(4 * nfreq_mf * stokes_buffers)) * size_t(block) * sizeof(double);
}

struct OccKey {
    int device;
    const void* fn;
    int block;
    size_t lds;
};
std::mutex g_occ_mutex;
std::vector<std::pair<OccKey, int>> g_occ;

// Resident blocks per CU of one kernel at this block size and LDS (cached).
kp_status blocks_per_cu(int device, const void* fn, int block, size_t lds, int& out) {
    std::lock_guard<std::mutex> lock(g_occ_mutex);
    for (auto& e : g_occ)
        if (e.first.device == device && e.first.fn == fn && e.first.block == block && e.first.lds == lds) {
            out = e.second;
            return KP_OK;
        }
    int n = 0;
    KP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, block, lds));
    n = std::max(n, 1);
    g_occ.push_back({OccKey{device, fn, block, lds}, n});
    out = n;
    return KP_OK;
}

std::mutex g_cu_mutex;
std::vector<int> g_cus;

kp_status cu_count(int device, int& out) {
    std::lock_guard<std::mutex> lock(g_cu_mutex);
    if ((int)g_cus.size() <= device) g_cus.resize(device + 1, 0);
    if (g_cus[device] == 0) {
        hipDeviceProp_t prop;
        KP_HIP(hipGetDeviceProperties(&prop, device));
        g_cus[device] = prop.multiProcessorCount;
    }
    out = g_cus[device];
    return KP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Parameter preparation (host, double): tetrad, extents, model constants.
// ---------------------------------------------------------------------------
// Camera extents and integrator scalars shared by every model.
static kp_status prepare_common(const kp_pass_a_params* p, Params& P) {
    if (p->camera.model != KP_CAMERA_PINHOLE)
        return kp_internal::fail(KP_UNSUPPORTED, "only the pinhole camera is built");
    if (p->direct_only)
        return kp_internal::fail(KP_UNSUPPORTED, "direct_only vertical-turn truncation is not implemented");
    const double hr = p->emission_selection.equatorial_h_over_r;
    if (hr > 0.0 && hr < 1.0)
        return kp_internal::fail(KP_UNSUPPORTED, "equatorial emission wedges are not implemented");
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

// grmhd != NULL selects the GRMHD grid model on FMKS geodesics (model unused).
static kp_status run_image(const kp_pass_a_params* params, const kp_riaf_model* model, const FreqLists& fl_in,
                           const kp_run_options* options, kp_image_data* out, kp_image_stats* stats,
                           int perturbation, double log_scale, const Tangent* tangent = nullptr,
                           const kp_grmhd_model* grmhd = nullptr, const kp_slow_light* slow = nullptr,
                           kp_slow_light_probe* probe = nullptr) {
    kp_internal::clear_error();
    if (!out) return kp_internal::fail(KP_INVALID_ARGUMENT, "null output struct");
    kp_run_options opt;
    kp_run_options_default(&opt);
    if (options) opt = *options;
    if (opt.precision != KP_PRECISION_FP64 && opt.precision != KP_PRECISION_MIXED)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "precision must be KP_PRECISION_FP64 or KP_PRECISION_MIXED");
    const bool mixed = opt.precision == KP_PRECISION_MIXED;
    Params P;
    FreqLists fl = fl_in;
    kp_status st;
    if (grmhd) {
        if (mixed || perturbation || (tangent && !slow))
            return kp_internal::fail(KP_UNSUPPORTED, "GRMHD images are built for FP64; gradients for slow-light sequences");
        st = prepare_grmhd(params, grmhd, P, probe == nullptr);
        if (st != KP_OK) return st;
        if (fl.nfreq == 0) {
            fl.nfreq = fl.nctrl = 1;
            fl.freqs = fl.ctrl = &grmhd->freq_cgs;
        }
    } else {
        st = prepare(params, model, P);
        if (st != KP_OK) return st;
        if (perturbation) kp_internal::riaf_constants(*params, *model, perturbation, log_scale, P.riaf);
    }
    if (emission_nonthermal(P.nt.dist) && (mixed || tangent))
        return kp_internal::fail(KP_UNSUPPORTED,
                                 "the kappa and power-law distributions are evaluated in FP64 without tangents");
    if (mixed) {
        if (!(opt.fp64_radius >= 0.0))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "fp64_radius must be non-negative");
        if (opt.guard_mode < 0 || opt.guard_mode > 2)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "guard_mode must be 0, 1 or 2");
        prepare_mixed(P);
        P.fp64_radius = opt.fp64_radius;
        // every geodesic attempt of guard_mode 0 is FP64 and counts as a guard step
        if (opt.guard_mode == 0) P.stage = 0;
    }
    const bool nonthermal = emission_nonthermal(P.nt.dist);
    if (nonthermal && !grmhd && fl.nfreq == 0) {
        // kappa / power law run in the frequency-list kernels
        fl.nfreq = fl.nctrl = 1;
        fl.freqs = fl.ctrl = &model->freq_cgs;
    }
    const bool mf = fl.nfreq > 0;
    if (mf) {
        st = prepare_frequencies(grmhd ? grmhd->mbh_solar : model->mbh_solar, fl, P);
        if (st != KP_OK) return st;
    }
    int mode = grmhd ? (slow ? M_SLOW : (probe ? M_PROBE : M_GRMHD)) : (mf ? M_MF : M_SINGLE);
    if (nonthermal) mode = mode == M_SLOW ? M_SLOW_NT : (mode == M_GRMHD ? M_GRMHD_NT : (mode == M_MF ? M_MF_NT : mode));
    if (tangent && slow) {
        if (mixed || fl.nfreq != 1)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "slow-light gradients are built for single-frequency FP64");
        if (!tangent->log_scales || !tangent->grad)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "null log_scales or gradient array");
        const double qd = tangent->log_scales[0], qm = tangent->log_scales[1];
        if (!std::isfinite(qd) || !std::isfinite(qm))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "log_scales must be finite");
        P.tq_ne = std::exp(qd);
        P.tq_b = std::exp(0.5 * qd) * std::exp(qm);
        P.grad = tangent->grad;
        P.single_pass = 0;
        mode = M_STAN;
    } else if (tangent) {
        if (mixed || mf)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "parameter gradients are built for single-frequency FP64");
        if (!tangent->log_scales || !tangent->grad)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "null log_scales or gradient array");
        const double qd = tangent->log_scales[0], qt = tangent->log_scales[1], qm = tangent->log_scales[2];
        if (!std::isfinite(qd) || !std::isfinite(qt) || !std::isfinite(qm))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "log_scales must be finite");
        // PlasmaPerturbation::apply per family (plasma_perturbation.hpp:17-32), composed
        P.riaf.ne_factor = std::exp(qd);
        P.riaf.thetae_factor = std::exp(qt);
        P.riaf.b_factor = std::exp(0.5 * qd) * std::exp(qm);
        P.grad = tangent->grad;
        mode = M_TANGENT;
    }
    P.out = *out;
    P.persistent = opt.persistent ? 1 : 0;
    int device = 0;
    KP_HIP(hipGetDevice(&device));
    const size_t np = (size_t)P.npix;
    const bool unsplit = opt.split_transport == 0;
    // Staged mixed precision (guard_mode 2, persistent split kernels): an
    // FP32-geodesic launch parks rays that reach r < fp64_radius, an FP64-geodesic launch
    // continues them and parks those back that reach r >= back_radius, then
    // FP32 again and a final FP64 launch; no wave mixes the two paths.
    const bool staged = mixed && P.persistent && !unsplit && opt.guard_mode == 2 && P.fp64_radius > 0.0 &&
                        (mode == M_SINGLE || mode == M_MF);
    const int nst = mf ? P.nfreq : 1;
    const size_t off_ep = 0;
    const size_t off_valid = align_up(off_ep + 16 * np * sizeof(double));
    const size_t off_steps = align_up(off_valid + np * sizeof(int32_t));
    const size_t off_reason = align_up(off_steps + np * sizeof(int32_t));
    const size_t off_queue = align_up(off_reason + np * sizeof(int32_t));
    const size_t off_counters = align_up(off_queue + KP_NQUEUE * sizeof(uint32_t));
    size_t total = align_up(off_counters + C_COUNT * sizeof(unsigned long long));
    const size_t off_lists = total;
    const size_t off_pkd = align_up(off_lists + (staged ? 2 * np * sizeof(uint32_t) : 0));
    const size_t off_pki = align_up(off_pkd + (staged ? 2 * np * sizeof(double) : 0));
    const size_t off_pkst = align_up(off_pki + (staged ? 5 * np * sizeof(int32_t) : 0));
    const size_t off_slow_t = align_up(off_pkst + (staged ? 4 * (size_t)nst * np * sizeof(double) : 0));
    const size_t nsnap = slow ? (size_t)slow->nsnap : 0;
    const size_t off_slow_c = align_up(off_slow_t + nsnap * sizeof(double));
    const size_t off_req = align_up(off_slow_c + nsnap * sizeof(void*));
    const size_t off_probe = align_up(off_req + (slow ? np * sizeof(double) : 0));
    const size_t off_preason = align_up(off_probe + (probe ? 2 * np * sizeof(double) : 0));
    total = align_up(off_preason + (probe ? np * sizeof(int32_t) : 0));
    void* ws = nullptr;
    st = workspace_for(device, total, &ws);
    if (st != KP_OK) return st;
    char* base = static_cast<char*>(ws);
    P.ep = reinterpret_cast<double*>(base + off_ep);
    P.ep_valid = reinterpret_cast<int32_t*>(base + off_valid);
    P.ep_steps = reinterpret_cast<int32_t*>(base + off_steps);
    P.ep_reason = reinterpret_cast<int32_t*>(base + off_reason);
    P.queue = reinterpret_cast<uint32_t*>(base + off_queue);
    P.counters = reinterpret_cast<unsigned long long*>(base + off_counters);
    uint32_t* lists = staged ? reinterpret_cast<uint32_t*>(base + off_lists) : nullptr;
    if (staged) {
        P.pk_d = reinterpret_cast<double*>(base + off_pkd);
        P.pk_i = reinterpret_cast<int32_t*>(base + off_pki);
        P.pk_st = reinterpret_cast<double*>(base + off_pkst);
        P.back_radius = KP_BACK_RADIUS_FACTOR * P.fp64_radius;
    }

    hipStream_t stream = static_cast<hipStream_t>(opt.stream);
    if (probe) {
        P.probe = reinterpret_cast<double*>(base + off_probe);
        std::memset(&P.out, 0, sizeof(P.out));
        P.out.reason = reinterpret_cast<int32_t*>(base + off_preason);
    }
    if (slow) {
        P.snap_t = reinterpret_cast<const double*>(base + off_slow_t);
        P.snaps = reinterpret_cast<const kp_grmhd_cell* const*>(base + off_slow_c);
        P.req_t = reinterpret_cast<double*>(base + off_req);
        P.nsnap = slow->nsnap;
        P.slow_mode = slow->step_mode;
        P.slow_interp = slow->interpolation;
        P.t_obs = slow->observation_time;
        KP_HIP(hipMemcpyAsync(base + off_slow_t, slow->times, nsnap * sizeof(double), hipMemcpyHostToDevice, stream));
        KP_HIP(hipMemcpyAsync(base + off_slow_c, slow->cells, nsnap * sizeof(void*), hipMemcpyHostToDevice, stream));
    }
    const int block = opt.block_size > 0 ? opt.block_size : 128;
    if (block % 64 != 0 || block > 256)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "block_size must be a multiple of 64 and at most 256");
    int cus = 0;
    st = cu_count(device, cus);
    if (st != KP_OK) return st;
    const size_t smem_a = lds_bytes(block, 0);
    // the single-pass frequency loop serves the FP64 multi-frequency, GRMHD
    // and slow-light lanes
    if (mixed || !(base_mode(mode) == M_MF || base_mode(mode) == M_GRMHD || base_mode(mode) == M_SLOW))
        P.single_pass = 0;
    const size_t smem_b = mode == M_STAN ? lds_bytes(block, 0) : lds_bytes(block, mf ? P.nfreq : 0, P.single_pass ? 2 : 1);
    const int nblocks_flat = (int)((np + block - 1) / block);
    // grid of one launch: one block per pixel block, or (persistent) the
    // resident blocks of the kernel on every CU
    auto grid_of = [&](const void* fn, size_t smem, int& grid) -> kp_status {
        grid = nblocks_flat;
        if (!P.persistent) return KP_OK;
        int per_cu = 0;
        const kp_status s2 = blocks_per_cu(device, fn, block, smem, per_cu);
        if (s2 != KP_OK) return s2;
        if (opt.waves_per_cu > 0) per_cu = std::max(1, opt.waves_per_cu / (block / 64));
        grid = std::max(1, std::min(nblocks_flat, cus * per_cu));
        return KP_OK;
    };
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (stats) {
        for (int i = 0; i < 3; ++i) KP_HIP(hipEventCreate(&ev[i]));
    }
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, P.queue, KP_NQUEUE, P.counters);
    KP_HIP(hipGetLastError());
    if (stats) KP_HIP(hipEventRecord(ev[0], stream));
    auto launch = [&](const void* fn, size_t smem, Params& Q) -> kp_status {
        int grid = 0;
        const kp_status s2 = grid_of(fn, smem, grid);
        if (s2 != KP_OK) return s2;
        void* args[] = {&Q};
        KP_HIP(hipLaunchKernel(fn, dim3(grid), dim3(block), args, smem, stream));
        return KP_OK;
    };
    // the four stages of one pass: FP32, FP64, FP32, FP64 (final, no parking);
    // queue slots: fetch counters at first_slot + k, list counts at first_slot + 4 + k
    auto run_staged = [&](bool pass_b, int first_slot) -> kp_status {
        for (int k = 0; k < 4; ++k) {
            Params Q = P;
            const bool fp32 = (k % 2) == 0;
            Q.stage = k;
            Q.park = k == 3 ? PARK_NONE : (fp32 ? PARK_INSIDE : PARK_OUTSIDE);
            Q.fetch = P.queue + first_slot + k;
            Q.list_in = k == 0 ? nullptr : lists + (size_t)((k - 1) % 2) * np;
            Q.count_in = k == 0 ? nullptr : P.queue + first_slot + 4 + (k - 1);
            Q.list_out = lists + (size_t)(k % 2) * np;
            Q.count_out = P.queue + first_slot + 4 + k;
            const int gp = fp32 ? GP_32 : GP_64;
            const void* fn = pass_b ? pass_b_fn<true>(true, gp, mode) : pass_a_fn<true>(gp, G_BL);
            const kp_status s2 = launch(fn, pass_b ? smem_b : smem_a, Q);
            if (s2 != KP_OK) return s2;
        }
        return KP_OK;
    };
    // guard_mode 0: FP64 geodesics with the FP32 radiation path
    const int gp_plain = mixed ? (opt.guard_mode == 0 ? GP_64 : GP_GUARD) : GP_64;
    if (P.npix > 0 && unsplit) {
        Params Q = P;
        Q.fetch = P.queue;
        st = launch(unsplit_fn(mixed, mode), smem_b, Q);
        if (st != KP_OK) return st;
    }
    if (P.npix > 0 && !unsplit) {
        if (staged) {
            st = run_staged(false, 2);
        } else {
            Params Q = P;
            Q.fetch = P.queue;
            const void* fn = P.persistent ? pass_a_fn<true>(gp_plain, geom_of_mode(mode))
                                          : pass_a_fn<false>(gp_plain, geom_of_mode(mode));
            st = launch(fn, smem_a, Q);
        }
        if (st != KP_OK) return st;
    }
    if (stats) KP_HIP(hipEventRecord(ev[1], stream));
    if (P.npix > 0 && !unsplit) {
        if (staged) {
            st = run_staged(true, 12);
        } else {
            Params Q = P;
            Q.fetch = P.queue + 1;
            const void* fn = P.persistent ? pass_b_fn<true>(mixed, gp_plain, mode) : pass_b_fn<false>(mixed, gp_plain, mode);
            st = launch(fn, smem_b, Q);
        }
        if (st != KP_OK) return st;
    }
    if (stats) {
        KP_HIP(hipEventRecord(ev[2], stream));
        KP_HIP(hipEventSynchronize(ev[2]));
        float a_ms = 0.f, b_ms = 0.f;
        KP_HIP(hipEventElapsedTime(&a_ms, ev[0], ev[1]));
        KP_HIP(hipEventElapsedTime(&b_ms, ev[1], ev[2]));
        unsigned long long cnt[C_COUNT];
        KP_HIP(hipMemcpy(cnt, P.counters, sizeof(cnt), hipMemcpyDeviceToHost));
        // unsplit: the one kernel is timed as pass A; the empty second
        // interval (a few microseconds of event overhead) is not reported
        if (unsplit) b_ms = 0.f;
        std::memset(stats, 0, sizeof(*stats));
        stats->pass_a_ms = a_ms;
        stats->pass_b_ms = b_ms;
        stats->kernel_ms = double(a_ms) + double(b_ms);
        stats->rays = P.npix;
        stats->rhs_evals = (int64_t)(cnt[C_RHS_A] + cnt[C_RHS_B]);
        stats->pass_a_rhs_evals = (int64_t)cnt[C_RHS_A];
        stats->coefficient_evals = (int64_t)cnt[C_COEFF];
        stats->transfer_steps = (int64_t)cnt[C_TRANSFER];
        stats->pass_a_steps = (int64_t)cnt[C_STEPS_A];
        stats->pass_b_steps = (int64_t)cnt[C_STEPS_B];
        stats->fp64_guard_steps = (int64_t)cnt[C_GUARD];
        for (int i = 0; i < 3; ++i) (void)hipEventDestroy(ev[i]);
    }
    if (probe) {
        // SlowLightTimeProbeHostResult reduction (slow_light_driver.hpp:733-778)
        std::vector<double> x0(2 * np);
        std::vector<int32_t> rs(np);
        KP_HIP(hipStreamSynchronize(stream));
        if (np) {
            KP_HIP(hipMemcpy(x0.data(), P.probe, 2 * np * sizeof(double), hipMemcpyDeviceToHost));
            KP_HIP(hipMemcpy(rs.data(), P.out.reason, np * sizeof(int32_t), hipMemcpyDeviceToHost));
        }
        probe->min_x0 = 0.0;
        probe->max_x0 = 0.0;
        probe->radiating_pixels = 0;
        probe->returned_pixels = 0;
        double lo = 1e300, hi = -1e300;
        for (size_t i = 0; i < np; ++i) {
            if (x0[i] == x0[i]) {
                probe->radiating_pixels += 1;
                lo = std::min(lo, x0[i]);
                hi = std::max(hi, x0[np + i]);
            }
            if (rs[i] == R_CAMERA) probe->returned_pixels += 1;
        }
        if (probe->radiating_pixels) {
            probe->min_x0 = lo;
            probe->max_x0 = hi;
        }
    }
    if (slow && slow->step_mode == KP_SLOW_LIGHT_DECOUPLED && P.npix > 0) {
        // decoupled stepping: a ray that needed a time outside the sequence
        // makes the run an error, as the reference's block scheduler does
        unsigned long long waiting = 0;
        KP_HIP(hipMemcpyAsync(&waiting, P.counters + C_WAIT, sizeof(waiting), hipMemcpyDeviceToHost, stream));
        KP_HIP(hipStreamSynchronize(stream));
        if (waiting) {
            std::vector<double> req(np);
            KP_HIP(hipMemcpy(req.data(), P.req_t, np * sizeof(double), hipMemcpyDeviceToHost));
            double lo = 1e300, hi = -1e300;
            for (double t : req)
                if (t == t) {
                    lo = std::min(lo, t);
                    hi = std::max(hi, t);
                }
            char msg[256];
            std::snprintf(msg, sizeof(msg),
                          "decoupled slow-light radiation sample outside supplied physical time support: %llu rays "
                          "requested %.17g..%.17g, available %.17g..%.17g",
                          waiting, lo, hi, slow->times[0], slow->times[slow->nsnap - 1]);
            return kp_internal::fail(KP_TIME_SUPPORT, msg);
        }
    }
    return KP_OK;
}

// ---------------------------------------------------------------------------
// Batch: job table, shared endpoint workspace and a two-pass graph.
// ---------------------------------------------------------------------------
struct kp_batch {
    int device = 0;
    bool mixed = false;
    int block = 128;
    int grid_a = 1, grid_b = 1;
    int32_t njobs = 0;
    int64_t rays = 0;
    void* dmem = nullptr;
    BatchArgs args{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void batch_release(kp_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->dmem) (void)hipFree(b->dmem);
    delete b;
}

static kp_status batch_build_graph(kp_batch* b) {
    KP_HIP(hipGraphCreate(&b->graph, 0));
    int nq = 2;
    void* zargs[] = {&b->args.queue, &nq, &b->args.counters};
    void* bargs[] = {&b->args};
    hipKernelNodeParams zp{};
    zp.func = (void*)zero_kernel;
    zp.gridDim = dim3(1);
    zp.blockDim = dim3(64);
    zp.kernelParams = zargs;
    hipKernelNodeParams ap{};
    ap.func = b->mixed ? (void*)pass_a_batch_kernel<true> : (void*)pass_a_batch_kernel<false>;
    ap.gridDim = dim3(b->grid_a);
    ap.blockDim = dim3(b->block);
    ap.sharedMemBytes = (unsigned)lds_bytes(b->block, 0);
    ap.kernelParams = bargs;
    hipKernelNodeParams bp = ap;
    bp.func = b->mixed ? (void*)pass_b_batch_kernel<true> : (void*)pass_b_batch_kernel<false>;
    bp.gridDim = dim3(b->grid_b);
    hipGraphNode_t nz, na, nb;
    KP_HIP(hipGraphAddKernelNode(&nz, b->graph, nullptr, 0, &zp));
    KP_HIP(hipGraphAddKernelNode(&na, b->graph, &nz, 1, &ap));
    KP_HIP(hipGraphAddKernelNode(&nb, b->graph, &na, 1, &bp));
    KP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return KP_OK;
}

extern "C" kp_status kp_batch_create(const kp_batch_job* jobs, int32_t njobs, const kp_run_options* options,
                                     kp_batch** out) {
    kp_internal::clear_error();
    if (!out) return kp_internal::fail(KP_INVALID_ARGUMENT, "null batch handle");
    *out = nullptr;
    if (!jobs || njobs < 1) return kp_internal::fail(KP_INVALID_ARGUMENT, "a batch needs at least one job");
    kp_run_options opt;
    kp_run_options_default(&opt);
    if (options) opt = *options;
    if (opt.precision != KP_PRECISION_FP64 && opt.precision != KP_PRECISION_MIXED)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "precision must be KP_PRECISION_FP64 or KP_PRECISION_MIXED");
    const bool mixed = opt.precision == KP_PRECISION_MIXED;
    if (mixed && !(opt.fp64_radius >= 0.0))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "fp64_radius must be non-negative");
    const int block = opt.block_size > 0 ? opt.block_size : 128;
    if (block % 64 != 0 || block > 256)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "block_size must be a multiple of 64 and at most 256");
    // validate every job before any allocation (batch_driver.hpp:137-146)
    std::vector<Params> table((size_t)njobs);
    std::vector<uint32_t> start((size_t)njobs + 1, 0);
    uint64_t total = 0;
    for (int j = 0; j < njobs; ++j) {
        if (jobs[j].radiation_substeps < 1)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "batch radiation_substeps must be positive");
        kp_status st = prepare(&jobs[j].params, &jobs[j].model, table[j]);
        if (st != KP_OK) return st;
        if (emission_nonthermal(table[j].nt.dist))
            return kp_internal::fail(KP_UNSUPPORTED, "batches carry the thermal fits; image a kappa or power-law "
                                                     "model with kp_run_riaf_(multifrequency_)image");
        if (mixed) {
            prepare_mixed(table[j]);
            table[j].fp64_radius = opt.fp64_radius;
        }
        table[j].out = jobs[j].out;
        table[j].persistent = 1;
        start[j] = (uint32_t)total;
        total += (uint64_t)table[j].npix;
        if (total > 0x7fffffffull) return kp_internal::fail(KP_INVALID_ARGUMENT, "batch exceeds 2^31 rays");
    }
    start[njobs] = (uint32_t)total;
    int device = 0;
    KP_HIP(hipGetDevice(&device));
    // device layout: job table | start | queue | counters | per-job endpoints
    const size_t off_table = 0;
    const size_t off_start = align_up(off_table + table.size() * sizeof(Params));
    const size_t off_queue = align_up(off_start + start.size() * sizeof(uint32_t));
    const size_t off_counters = align_up(off_queue + 2 * sizeof(uint32_t));
    size_t off = align_up(off_counters + C_COUNT * sizeof(unsigned long long));
    std::vector<size_t> ep_off((size_t)njobs);
    for (int j = 0; j < njobs; ++j) {
        ep_off[j] = off;
        const size_t np = (size_t)table[j].npix;
        off = align_up(off + 16 * np * sizeof(double));
        off = align_up(off + 3 * np * sizeof(int32_t));
    }
    kp_batch* b = new kp_batch();
    b->device = device;
    b->mixed = mixed;
    b->block = block;
    b->njobs = njobs;
    b->rays = (int64_t)total;
    if (hipMalloc(&b->dmem, off) != hipSuccess) {
        batch_release(b);
        return kp_internal::fail(KP_OUT_OF_MEMORY, "hipMalloc of the batch workspace failed");
    }
    char* base = static_cast<char*>(b->dmem);
    for (int j = 0; j < njobs; ++j) {
        const size_t np = (size_t)table[j].npix;
        char* e = base + ep_off[j];
        table[j].ep = reinterpret_cast<double*>(e);
        char* iv = base + align_up(ep_off[j] + 16 * np * sizeof(double));
        table[j].ep_valid = reinterpret_cast<int32_t*>(iv);
        table[j].ep_steps = table[j].ep_valid + np;
        table[j].ep_reason = table[j].ep_steps + np;
        table[j].queue = nullptr;
        table[j].counters = nullptr;
    }
    b->args.jobs = reinterpret_cast<const Params*>(base + off_table);
    b->args.start = reinterpret_cast<const uint32_t*>(base + off_start);
    b->args.njobs = njobs;
    b->args.total = (uint32_t)total;
    b->args.queue = reinterpret_cast<uint32_t*>(base + off_queue);
    b->args.counters = reinterpret_cast<unsigned long long*>(base + off_counters);
    kp_status st = KP_OK;
    if (hipMemcpy(base + off_table, table.data(), table.size() * sizeof(Params), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(base + off_start, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
            hipSuccess)
        st = kp_internal::fail(KP_HIP_ERROR, "upload of the batch job table failed");
    if (st == KP_OK) {
        hipDeviceProp_t prop;
        int ba = 1, bb = 1;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &ba, mixed ? (const void*)pass_a_batch_kernel<true> : (const void*)pass_a_batch_kernel<false>, block,
                lds_bytes(block, 0)) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &bb, mixed ? (const void*)pass_b_batch_kernel<true> : (const void*)pass_b_batch_kernel<false>, block,
                lds_bytes(block, 0)) != hipSuccess) {
            st = kp_internal::fail(KP_HIP_ERROR, "occupancy query for the batch kernels failed");
        } else {
            const int wpb = block / 64;
            const int per_a = opt.waves_per_cu > 0 ? std::max(1, opt.waves_per_cu / wpb) : std::max(ba, 1);
            const int per_b = opt.waves_per_cu > 0 ? std::max(1, opt.waves_per_cu / wpb) : std::max(bb, 1);
            const int64_t flat = (int64_t)((total + block - 1) / block);
            b->grid_a = (int)std::max<int64_t>(1, std::min<int64_t>(flat, (int64_t)prop.multiProcessorCount * per_a));
            b->grid_b = (int)std::max<int64_t>(1, std::min<int64_t>(flat, (int64_t)prop.multiProcessorCount * per_b));
            st = batch_build_graph(b);
        }
    }
    if (st != KP_OK) {
        batch_release(b);
        return st;
    }
    *out = b;
    return KP_OK;
}

extern "C" kp_status kp_batch_launch(kp_batch* b, void* stream) {
    kp_internal::clear_error();
    if (!b || !b->exec) return kp_internal::fail(KP_INVALID_ARGUMENT, "invalid batch handle");
    KP_HIP(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)));
    return KP_OK;
}

extern "C" kp_status kp_batch_read_stats(kp_batch* b, kp_image_stats* stats) {
    kp_internal::clear_error();
    if (!b || !stats) return kp_internal::fail(KP_INVALID_ARGUMENT, "null argument");
    unsigned long long cnt[C_COUNT];
    KP_HIP(hipMemcpy(cnt, b->args.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    std::memset(stats, 0, sizeof(*stats));
    stats->rays = b->rays;
    stats->rhs_evals = (int64_t)(cnt[C_RHS_A] + cnt[C_RHS_B]);
    stats->pass_a_rhs_evals = (int64_t)cnt[C_RHS_A];
    stats->coefficient_evals = (int64_t)cnt[C_COEFF];
    stats->transfer_steps = (int64_t)cnt[C_TRANSFER];
    stats->pass_a_steps = (int64_t)cnt[C_STEPS_A];
    stats->pass_b_steps = (int64_t)cnt[C_STEPS_B];
    stats->fp64_guard_steps = (int64_t)cnt[C_GUARD];
    return KP_OK;
}

extern "C" void kp_batch_destroy(kp_batch* b) {
    if (!b) return;
    (void)hipDeviceSynchronize();
    batch_release(b);
}

extern "C" kp_status kp_run_riaf_batch(const kp_batch_job* jobs, int32_t njobs, const kp_run_options* options,
                                       kp_image_stats* stats) {
    kp_batch* b = nullptr;
    kp_status st = kp_batch_create(jobs, njobs, options, &b);
    if (st != KP_OK) return st;
    hipStream_t stream = options ? static_cast<hipStream_t>(options->stream) : nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (stats && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        kp_batch_destroy(b);
        return kp_internal::fail(KP_HIP_ERROR, "hipEventCreate failed");
    }
    if (stats) (void)hipEventRecord(e0, stream);
    st = kp_batch_launch(b, stream);
    if (st == KP_OK && stats) {
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) st = kp_internal::fail(KP_HIP_ERROR, "batch did not complete");
        float ms = 0.f;
        if (st == KP_OK) {
            (void)hipEventElapsedTime(&ms, e0, e1);
            st = kp_batch_read_stats(b, stats);
            stats->kernel_ms = ms;
            stats->pass_a_ms = ms;
        }
    } else if (st == KP_OK && hipStreamSynchronize(stream) != hipSuccess) {
        st = kp_internal::fail(KP_HIP_ERROR, "batch did not complete");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    kp_batch_destroy(b);
    return st;
}

extern "C" int32_t kp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" kp_status kp_run_riaf_image(const kp_pass_a_params* params, const kp_riaf_model* model,
                                       int32_t radiation_substeps, const kp_run_options* options,
                                       kp_image_data* out, kp_image_stats* stats) {
    (void)radiation_substeps;
    return run_image(params, model, FreqLists{}, options, out, stats, 0, 0.0);
}

extern "C" kp_status kp_run_riaf_multifrequency_image(const kp_pass_a_params* params, const kp_riaf_model* model,
                                                      int32_t radiation_substeps, const double* frequencies,
                                                      int32_t nfreq, const double* step_control_frequencies,
                                                      int32_t step_control_nfreq, const kp_run_options* options,
                                                      kp_image_data* out, kp_image_stats* stats) {
    (void)radiation_substeps;
    FreqLists fl;
    fl.nfreq = nfreq;
    fl.freqs = frequencies;
    fl.nctrl = step_control_nfreq;
    fl.ctrl = step_control_frequencies;
    if (nfreq < 1) return kp_internal::fail(KP_INVALID_ARGUMENT, "nfreq must be at least 1");
    return run_image(params, model, fl, options, out, stats, 0, 0.0);
}

extern "C" kp_status kp_run_riaf_image_gradient(const kp_pass_a_params* params, const kp_riaf_model* model,
                                                int32_t radiation_substeps, const double* log_scales,
                                                const kp_run_options* options, kp_image_data* out, double* grad,
                                                kp_image_stats* stats) {
    (void)radiation_substeps;
    Tangent t;
    t.log_scales = log_scales;
    t.grad = grad;
    return run_image(params, model, FreqLists{}, options, out, stats, 0, 0.0, &t);
}

extern "C" kp_status kp_run_grmhd_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                        int32_t radiation_substeps, const kp_run_options* options,
                                        kp_image_data* out, kp_image_stats* stats) {
    (void)radiation_substeps;
    if (!model) return kp_internal::fail(KP_INVALID_ARGUMENT, "null GRMHD model");
    return run_image(params, nullptr, FreqLists{}, options, out, stats, 0, 0.0, nullptr, model);
}

extern "C" kp_status kp_run_grmhd_multifrequency_image(const kp_pass_a_params* params,
                                                       const kp_grmhd_model* model, int32_t radiation_substeps,
                                                       const double* frequencies, int32_t nfreq,
                                                       const double* step_control_frequencies,
                                                       int32_t step_control_nfreq, const kp_run_options* options,
                                                       kp_image_data* out, kp_image_stats* stats) {
    (void)radiation_substeps;
    if (!model) return kp_internal::fail(KP_INVALID_ARGUMENT, "null GRMHD model");
    if (nfreq < 1) return kp_internal::fail(KP_INVALID_ARGUMENT, "nfreq must be at least 1");
    FreqLists fl;
    fl.nfreq = nfreq;
    fl.freqs = frequencies;
    fl.nctrl = step_control_nfreq;
    fl.ctrl = step_control_frequencies;
    return run_image(params, nullptr, fl, options, out, stats, 0, 0.0, nullptr, model);
}

extern "C" kp_status kp_run_grmhd_slow_light_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                    const kp_slow_light* sequence, const double* frequencies,
                                                    int32_t nfreq, const kp_run_options* options,
                                                    kp_image_data* out, kp_image_stats* stats) {
    kp_internal::clear_error();
    if (!model || !sequence) return kp_internal::fail(KP_INVALID_ARGUMENT, "null GRMHD model or sequence");
    const kp_slow_light& q = *sequence;
    if (q.nsnap < 2) return kp_internal::fail(KP_INVALID_ARGUMENT, "slow light requires at least two snapshots");
    if (!q.times || !q.cells) return kp_internal::fail(KP_INVALID_ARGUMENT, "null snapshot times or cells");
    if (q.step_mode != KP_SLOW_LIGHT_DECOUPLED && q.step_mode != KP_SLOW_LIGHT_SNAPSHOT &&
        q.step_mode != KP_SLOW_LIGHT_BLOCK)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown slow-light step mode");
    if (q.interpolation != KP_SLOW_LIGHT_FLUID && q.interpolation != KP_SLOW_LIGHT_COEFFICIENTS)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown slow-light interpolation");
    if (!std::isfinite(q.observation_time)) return kp_internal::fail(KP_INVALID_ARGUMENT, "observation time must be finite");
    for (int k = 0; k < q.nsnap; ++k) {
        if (!q.cells[k]) return kp_internal::fail(KP_INVALID_ARGUMENT, "null snapshot cells");
        if (!std::isfinite(q.times[k]) || (k > 0 && !(q.times[k] > q.times[k - 1])))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "snapshot times must be finite and strictly increasing");
    }
    if (nfreq < 0 || (nfreq > 0 && !frequencies))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "frequency list and count disagree");
    kp_grmhd_model m = *model;
    m.cells = q.cells[0];
    FreqLists fl;
    if (nfreq > 0) {
        fl.nfreq = fl.nctrl = nfreq;
        fl.freqs = fl.ctrl = frequencies;
    }
    return run_image(params, nullptr, fl, options, out, stats, 0, 0.0, nullptr, &m, sequence);
}

extern "C" kp_status kp_run_grmhd_slow_light_image_gradient(const kp_pass_a_params* params,
                                                             const kp_grmhd_model* model,
                                                             const kp_slow_light* sequence,
                                                             const double* log_scales,
                                                             const kp_run_options* options, kp_image_data* out,
                                                             double* grad, kp_image_stats* stats) {
    kp_internal::clear_error();
    if (!model || !sequence || !sequence->cells || !sequence->times || sequence->nsnap < 2)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "GRMHD model and a sequence of at least two snapshots required");
    for (int k = 0; k < sequence->nsnap; ++k) {
        if (!sequence->cells[k]) return kp_internal::fail(KP_INVALID_ARGUMENT, "null snapshot cells");
        if (!std::isfinite(sequence->times[k]) || (k > 0 && !(sequence->times[k] > sequence->times[k - 1])))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "snapshot times must be finite and strictly increasing");
    }
    if (sequence->step_mode < KP_SLOW_LIGHT_DECOUPLED || sequence->step_mode > KP_SLOW_LIGHT_BLOCK ||
        (sequence->interpolation != KP_SLOW_LIGHT_FLUID && sequence->interpolation != KP_SLOW_LIGHT_COEFFICIENTS))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown slow-light step mode or interpolation");
    kp_grmhd_model m = *model;
    m.cells = sequence->cells[0];
    Tangent t;
    t.log_scales = log_scales;
    t.grad = grad;
    FreqLists fl;
    fl.nfreq = fl.nctrl = 1;
    fl.freqs = fl.ctrl = &model->freq_cgs;
    return run_image(params, nullptr, fl, options, out, stats, 0, 0.0, &t, &m, sequence);
}

extern "C" kp_status kp_run_grmhd_image_gradient(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                 const double* log_scales, const kp_run_options* options,
                                                 kp_image_data* out, double* grad, kp_image_stats* stats) {
    kp_internal::clear_error();
    if (!model || !model->cells) return kp_internal::fail(KP_INVALID_ARGUMENT, "GRMHD model with cells required");
    // the snapshot twice, at times no sample reaches: every sample's time
    // fraction lies strictly inside the pair and interpolates equal values
    const void* cells[2] = {model->cells, model->cells};
    const double times[2] = {-1e15, 1e15};
    kp_slow_light seq;
    kp_slow_light_default(&seq);
    seq.nsnap = 2;
    seq.observation_time = 0.0;
    seq.times = times;
    seq.cells = cells;
    return kp_run_grmhd_slow_light_image_gradient(params, model, &seq, log_scales, options, out, grad, stats);
}

extern "C" kp_status kp_tune_riaf_launch(const kp_pass_a_params* params, const kp_riaf_model* model,
                                         const double* frequencies, int32_t nfreq, int32_t sample_divisor,
                                         kp_run_options* options, kp_launch_tuning* report) {
    kp_internal::clear_error();
    if (!params || !model || !options) return kp_internal::fail(KP_INVALID_ARGUMENT, "null parameters");
    if (sample_divisor < 1) return kp_internal::fail(KP_INVALID_ARGUMENT, "sample_divisor must be positive");
    if (nfreq < 0 || (nfreq > 0 && !frequencies))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "frequencies required when nfreq > 0");
    kp_pass_a_params p = *params;
    p.camera.nx = std::max(1, p.camera.nx / sample_divisor);
    p.camera.ny = std::max(1, p.camera.ny / sample_divisor);
    kp_image_data out;
    std::memset(&out, 0, sizeof(out));  // every output skipped: only the kernels run
    FreqLists fl;
    if (nfreq > 0) {
        fl.nfreq = fl.nctrl = nfreq;
        fl.freqs = fl.ctrl = frequencies;
    }
    static const int blocks[3] = {64, 128, 256};
    static const int waves[3] = {0, 2, 8};
    kp_launch_tuning r;
    std::memset(&r, 0, sizeof(r));
    r.sample_pixels = p.camera.nx * p.camera.ny;
    r.best_ms = -1.0;
    for (int b = 0; b < 3; ++b) {
        for (int w = 0; w < 3; ++w) {
            kp_run_options o = *options;
            o.block_size = blocks[b];
            o.waves_per_cu = waves[w];
            kp_image_stats st;
            double ms = 0.0;
            for (int rep = 0; rep < 2; ++rep) {  // the first run loads the kernels
                const kp_status s2 = run_image(&p, model, fl, &o, &out, &st, 0, 0.0);
                if (s2 != KP_OK) return s2;
                ms = st.kernel_ms;
            }
            ++r.ncandidates;
            if (r.best_ms < 0.0 || ms < r.best_ms) {
                r.best_ms = ms;
                r.block_size = blocks[b];
                r.waves_per_cu = waves[w];
            }
            if (b == 1 && w == 0) r.default_ms = ms;  // block 128 and the occupancy grid are the defaults
        }
    }
    options->block_size = r.block_size;
    options->waves_per_cu = r.waves_per_cu;
    if (report) *report = r;
    return KP_OK;
}

extern "C" kp_status kp_slow_light_time_probe(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                              const kp_run_options* options, kp_slow_light_probe* out) {
    kp_internal::clear_error();
    if (!model || !out) return kp_internal::fail(KP_INVALID_ARGUMENT, "null GRMHD model or probe result");
    kp_image_data dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    return run_image(params, nullptr, FreqLists{}, options, &dummy, nullptr, 0, 0.0, nullptr, model, nullptr, out);
}

extern "C" kp_status kp_grmhd_derive_scalars(const kp_grmhd_model* model, void* stream) {
    kp_internal::clear_error();
    if (!model || !model->cells || model->n1 < 1 || model->n2 < 1 || model->n3 < 1)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "GRMHD model with cells and positive dimensions required");
    kp_pass_a_params p;
    kp_pass_a_params_default(&p);
    p.coordinate_system = KP_COORD_FMKS;
    p.camera.nx = p.camera.ny = 1;
    p.camera.radius = 100.0;
    Params P;
    kp_status st = prepare_grmhd(&p, model, P);
    if (st != KP_OK) return st;
    const long long ncell = (long long)model->n1 * model->n2 * model->n3;
    if (ncell > (1ll << 31)) return kp_internal::fail(KP_INVALID_ARGUMENT, "grid too large");
    const int block = 256;
    const int grid = (int)((ncell + block - 1) / block);
    hipLaunchKernelGGL(derive_scalars_kernel, dim3(grid), dim3(block), 0, static_cast<hipStream_t>(stream), P.grmhd,
                       const_cast<kp_grmhd_cell*>(P.grmhd.cells), (int64_t)ncell);
    KP_HIP(hipGetLastError());
    return KP_OK;
}

// Host-output wrapper shared by the single- and multi-frequency entries:
// device buffers for the selected outputs, the run, copies back.
static kp_status run_image_host(const kp_pass_a_params* params, const kp_riaf_model* model, const FreqLists& fl,
                                const kp_run_options* options, kp_image_data* out, kp_image_stats* stats) {
    kp_internal::clear_error();
    if (!params || !out) return kp_internal::fail(KP_INVALID_ARGUMENT, "null argument");
    const size_t np = (size_t)std::max(0, params->camera.nx) * (size_t)std::max(0, params->camera.ny);
    const size_t nf = fl.nfreq > 0 ? (size_t)std::min(fl.nfreq, KP_MAX_FREQUENCIES) : 1;
    double* const* host_real[15] = {&out->image_i, &out->image_q, &out->image_u, &out->image_v,
                                    &out->closure_x, &out->closure_k, &out->final_null, &out->frame_error,
                                    &out->det_r, &out->overlap_r11, &out->overlap_r12, &out->overlap_r21,
                                    &out->overlap_r22, &out->basis_identity_error, &out->basis_rotation_angle};
    int32_t* const* host_int[3] = {&out->pass_a_steps, &out->steps, &out->reason};
    auto real_len = [&](int i) { return i < 4 ? nf * np : np; };
    size_t bytes = 0;
    for (int i = 0; i < 15; ++i) bytes += real_len(i) * sizeof(double);
    bytes += 3 * np * sizeof(int32_t);
    void* dbuf = nullptr;
    if (np > 0 && hipMalloc(&dbuf, bytes) != hipSuccess)
        return kp_internal::fail(KP_OUT_OF_MEMORY, "hipMalloc of device outputs failed");
    kp_image_data dev;
    std::memset(&dev, 0, sizeof(dev));
    double** dev_real[15] = {&dev.image_i, &dev.image_q, &dev.image_u, &dev.image_v, &dev.closure_x,
                             &dev.closure_k, &dev.final_null, &dev.frame_error, &dev.det_r,
                             &dev.overlap_r11, &dev.overlap_r12, &dev.overlap_r21, &dev.overlap_r22,
                             &dev.basis_identity_error, &dev.basis_rotation_angle};
    int32_t** dev_int[3] = {&dev.pass_a_steps, &dev.steps, &dev.reason};
    char* b = static_cast<char*>(dbuf);
    for (int i = 0; i < 15; ++i) {
        if (*host_real[i]) *dev_real[i] = reinterpret_cast<double*>(b);
        b += real_len(i) * sizeof(double);
    }
    for (int i = 0; i < 3; ++i) {
        if (*host_int[i]) *dev_int[i] = reinterpret_cast<int32_t*>(b);
        b += np * sizeof(int32_t);
    }
    kp_run_options opt;
    kp_run_options_default(&opt);
    if (options) opt = *options;
    kp_status st = run_image(params, model, fl, &opt, &dev, stats, 0, 0.0);
    if (st == KP_OK) {
        if (hipStreamSynchronize(static_cast<hipStream_t>(opt.stream)) != hipSuccess) {
            st = kp_internal::fail(KP_HIP_ERROR, "stream synchronisation failed");
        } else {
            for (int i = 0; i < 15 && st == KP_OK; ++i)
                if (*host_real[i] && hipMemcpy(*host_real[i], *dev_real[i], real_len(i) * sizeof(double),
                                               hipMemcpyDeviceToHost) != hipSuccess)
                    st = kp_internal::fail(KP_HIP_ERROR, "device to host copy failed");
            for (int i = 0; i < 3 && st == KP_OK; ++i)
                if (*host_int[i] && hipMemcpy(*host_int[i], *dev_int[i], np * sizeof(int32_t),
                                              hipMemcpyDeviceToHost) != hipSuccess)
                    st = kp_internal::fail(KP_HIP_ERROR, "device to host copy failed");
        }
    }
    if (dbuf) (void)hipFree(dbuf);
    return st;
}

extern "C" kp_status kp_run_riaf_image_host(const kp_pass_a_params* params, const kp_riaf_model* model,
                                            int32_t radiation_substeps, const kp_run_options* options,
                                            kp_image_data* out, kp_image_stats* stats) {
    (void)radiation_substeps;
    return run_image_host(params, model, FreqLists{}, options, out, stats);
}

extern "C" kp_status kp_run_riaf_multifrequency_image_host(
    const kp_pass_a_params* params, const kp_riaf_model* model, int32_t radiation_substeps,
    const double* frequencies, int32_t nfreq, const double* step_control_frequencies, int32_t step_control_nfreq,
    const kp_run_options* options, kp_image_data* out, kp_image_stats* stats) {
    (void)radiation_substeps;
    if (nfreq < 1) return kp_internal::fail(KP_INVALID_ARGUMENT, "nfreq must be at least 1");
    FreqLists fl;
    fl.nfreq = nfreq;
    fl.freqs = frequencies;
    fl.nctrl = step_control_nfreq;
    fl.ctrl = step_control_frequencies;
    return run_image_host(params, model, fl, options, out, stats);
}
