// kp_kernels.h -- pass A / pass B imaging kernels for gfx950 (device side).
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
//  * Each kernel is instantiated as plain FP64 (the parity path) and, for
//    the Boyer-Lindquist images, mixed: FP32 coefficients and transfer with
//    FP64 geodesics (GP_64), or FP32 geodesic derivatives outside
//    fp64_radius, where the Boyer-Lindquist Delta cancels (GP_GUARD), or the
//    staged FP32-geodesic kernels (GP_32).
//  * The kernel templates are instantiated in kp_kernels.hip, one family per
//    part (the file is compiled once per part, in parallel); the host code
//    takes their addresses through the kpsel selectors declared below.
#pragma once

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
    R_DIRECT = 9,
};

// Pass B variants: one frequency, several frequencies on shared geodesics,
// one frequency with the tangent (plasma-parameter derivatives) of its Stokes
// vector carried alongside, a GRMHD grid, or a GRMHD snapshot sequence
// sampled at retarded time (slow light).
enum Mode : int {
    M_SINGLE = 0, M_MF = 1, M_TANGENT = 2, M_GRMHD = 3, M_SLOW = 4, M_PROBE = 5, M_STAN = 6,
    // the kappa / power-law forms of M_MF, M_GRMHD and M_SLOW: kernels of their
    // own, so the thermal kernels carry no call to the nonthermal routine
    M_MF_NT = 7, M_GRMHD_NT = 8, M_SLOW_NT = 9,
    // windowed slow light (thermal, nonthermal): pass B over one block of the
    // sequence, its rays parking at the block's end or while they wait for a
    // time another block holds; instantiations of their own, so the kernels
    // of whole sequences carry none of the parking code
    M_SLOW_W = 10, M_SLOW_NT_W = 11,
    // the single-frequency Boyer-Lindquist pass B in two legs: M_SINGLE_V is
    // M_SINGLE until the ray leaves the emitting region outward, where it
    // parks; M_VAC continues it to the camera with no radiation code at all
    // (outside the region every coefficient is zero and the transfer step is
    // the identity), so no wave of either kernel carries lanes of the other leg
    M_SINGLE_V = 12, M_VAC = 13
};

// The mode whose lane logic a mode runs (a nonthermal mode runs its thermal
// one, a windowed slow-light mode the slow-light one).
constexpr int base_mode(int m) {
    return m == M_MF_NT ? M_MF
           : (m == M_GRMHD_NT ? M_GRMHD
                              : ((m == M_SLOW_NT || m == M_SLOW_W || m == M_SLOW_NT_W) ? M_SLOW
                                                                                         : (m == M_SINGLE_V ? M_SINGLE : m)));
}
constexpr bool nt_mode(int m) { return m == M_MF_NT || m == M_GRMHD_NT || m == M_SLOW_NT || m == M_SLOW_NT_W; }
constexpr bool win_mode(int m) { return m == M_SLOW_W || m == M_SLOW_NT_W; }
// the radiating leg that parks at the vacuum boundary, and the vacuum leg
constexpr bool vpark_mode(int m) { return m == M_SINGLE_V; }
constexpr bool vac_mode(int m) { return m == M_VAC; }

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
    C_VACUUM = 8,  // pass-B vacuum leg: steps whose transfer is the identity
    C_RHS_V = 9,   // pass-B vacuum leg: RHS evaluations
    C_COUNT = 10
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
    // direct_only: pass A ends at the first vertical turn of the ray
    // (pass_a_core.hpp:240-275)
    int32_t direct_only, wedge_samples;
    // emission wedge: 0 < wedge_hr < 1 keeps the emissivities only where
    // |cos theta| <= wedge_hr, resolved by wedge_samples samples across it
    // (emission_selection.hpp:11-122); 0 otherwise
    double wedge_hr;
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
    // slow light: the resident snapshot cell arrays and times (device), the
    // stepping and interpolation rules (kp_slow_light), the observation time,
    // and per pixel the time a waiting ray needed (NaN for the others).
    // Windowed modes (M_SLOW_W, M_SLOW_NT_W): the resident snapshots are one block of
    // the sequence; slow_final says whether its last snapshot is the
    // sequence's last (else a ray reaching that time parks for the next
    // block), and [sup_lo, sup_hi] is the whole sequence's time support (a
    // decoupled ray needing a time inside it but outside the block parks
    // waiting, its pk_d holding the step size its current call started with).
    const kp_grmhd_cell* const* snaps;
    const double* snap_t;
    int32_t nsnap, slow_mode, slow_interp, slow_final;
    double t_obs;
    double sup_lo, sup_hi;
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
    // vertical_position_velocity (pass_a_core.hpp:202-224): the spin-axis
    // height z = r cos(theta) and dz/ds from the position (x1, x2) and the
    // wavevector components (k1, k2)
    static KP_DEV void vertical(const Params&, double x1, double x2, double k1, double k2, double& z, double& v) {
        double sn, cs;
        ksincos(x2, &sn, &cs);
        z = x1 * cs;
        v = cs * k1 - x1 * sn * k2;
    }
    // equatorial_height_ratio (emission_selection.hpp:21-33): cos(theta)
    static KP_DEV double height_ratio(const Params&, double, double x2) {
        double sn, cs;
        ksincos(x2, &sn, &cs);
        return cs;
    }
};

template <> struct GeomT<G_FMKS> {
    using GeoType = GeoFmks<double>;
    static KP_DEV GeoType geo(const Params& P, uint32_t* cnt) { return GeoType{P.mass, P.spin, cnt, P.fmks}; }
    static KP_DEV double radius(double x1) { return exp(x1); }
    static KP_DEV Metric4<double> metric(const Params& P, const double* x) {
        return fmks_gcov(P.mass, P.spin, P.fmks, x[1], x[2]);
    }
    // the FMKS theta depends on x1 and x2, and r = exp(x1)
    static KP_DEV void vertical(const Params& P, double x1, double x2, double k1, double k2, double& z, double& v) {
        const double r = exp(x1);
        const FmksTheta<double> t = fmks_theta(P.fmks, x1, x2);
        double sn, cs;
        ksincos(t.th, &sn, &cs);
        z = r * cs;
        v = z * k1 - r * sn * (t.dth1 * k1 + t.dth2 * k2);
    }
    // the physical theta of the native (x1, x2)
    static KP_DEV double height_ratio(const Params& P, double x1, double x2) {
        double sn, cs;
        ksincos(fmks_theta(P.fmks, x1, x2).th, &sn, &cs);
        return cs;
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
                           uint32_t transfer, uint32_t steps_a, uint32_t steps_b, uint32_t guard,
                           int transfer_slot = C_TRANSFER) {
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
        atomicAdd(&counters[transfer_slot], v[2]);
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
    // 1 evaluating the full step, 2 bisecting; 3 bisecting a vertical turn
    // of a direct_only image, flo holding dz/ds at the step's start);
    // best_ep: the best state so far is in the endpoint buffer, not the lane
    // state
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
// (the lanes of a wave may take both paths); an FP32 attempt across a
// turning point of theta (k^theta = dtheta/dlambda changes sign between its
// start and its end) is taken again in FP64 while the guard is on
// (fp64_radius > 0).  force64: FP64 whatever GP says (the surface-crossing
// refinement of a mixed-precision lane).  guard counts the FP64 integration
// attempts of a mixed-precision image: those of GP_GUARD lanes inside the
// radius or across a turning point and every attempt of a staged FP64 launch
// (P.stage >= 0), not the refinement trials.
KP_DEV bool theta_turns(const Params& P, double kth0, double kth1) {
    return P.fp64_radius > 0.0 && (kth0 < 0.0) != (kth1 < 0.0);
}

template <int GP, class GG>
KP_DEV double attempt(const Params& P, const GG& G, const LaneVec& s, LaneVec& k1, double h,
                      St<double>& half, St<double>& next, uint32_t& guard, bool full = true,
                      bool force64 = false) {
    if (GP == GP_32 || (GP == GP_GUARD && !force64 && !(s[1] < P.fp64_radius))) {
        const double err = doubling_attempt_mixed(P.mass_f, P.spin_f, s, h, half, next, G.rhs_count, full);
        if (GP == GP_32 || !theta_turns(P, s[6], next.v[6])) return err;
    }
    if (!force64 && (GP == GP_GUARD || P.stage >= 0)) ++guard;
    return doubling_attempt(G, s, h, k1, half, next, full);
}

template <int GP, class GG>
KP_DEV void step_rk4(const Params& P, const GG& G, const LaneVec& s, double h, St<double>& next,
                     uint32_t& guard, bool force64 = false) {
    if (GP == GP_32 || (GP == GP_GUARD && !force64 && !(s[1] < P.fp64_radius))) {
        rk4_mixed(P.mass_f, P.spin_f, s, h, next, G.rhs_count);
        if (GP == GP_32 || !theta_turns(P, s[6], next.v[6])) return;
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
    // the kept state's step fraction ends in L.lo
    auto keep = [&](const St<double>& v) {
        const size_t np = (size_t)P.npix;
#pragma unroll
        for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + L.pixel] = v.v[c];
        L.best_ep = 1;
        L.lo = L.trial;
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

// vertical_turn_bracket (pass_a_core.hpp:226-238): an away-from-midplane to
// toward-midplane reversal on one side of the midplane, both heights above
// 64 ulp of the radius scale.
KP_DEV bool vertical_turn_bracket(double z0, double v0, double z1, double v1, double h, double radius_scale) {
    const double floor = 64.0 * eps<double>() * radius_scale;
    return absv(z0) > floor && absv(z1) > floor && ((z0 > 0.0) == (z1 > 0.0)) && z0 * v0 * h > 0.0 &&
           z1 * v1 * h <= 0.0;
}

// truncate_at_first_vertical_turn (pass_a_core.hpp:240-275), started when the
// step from the lane state to `last` (step h) brackets a vertical turn: the
// lane then bisects the step, one trial path per tick on the step's own RK
// composition (as the crossing refinement does), for the last point whose
// dz/ds keeps the start's sign; that point is the endpoint (the start state
// when no trial keeps it).  Returns true when the bisection has started.
template <int GEOM>
KP_DEV bool turn_start(const Params& P, LaneA& L, const St<double>& last, double h) {
    using GT = GeomT<GEOM>;
    double z0, v0, z1, v1;
    GT::vertical(P, L.s[1], L.s[2], L.s[5], L.s[6], z0, v0);
    GT::vertical(P, last.v[1], last.v[2], last.v[5], last.v[6], z1, v1);
    const double scale = maxv(GT::radius(L.s[1]), GT::radius(last.v[1]));
    if (!st_finite(last) || !vertical_turn_bracket(z0, v0, z1, v1, h, scale)) return false;
    L.refine = 3;
    L.full_h = h;
    L.lo = 0.0;
    L.hi = 1.0;
    L.flo = v0;
    L.iter = 0;
    L.best_ep = 0;
    L.trial = 0.5;
    return true;
}

// One bisection trial t of the vertical turn; returns true when the
// bisection has finished.
template <int GEOM>
KP_DEV bool turn_update(const Params& P, LaneA& L, const St<double>& t) {
    using GT = GeomT<GEOM>;
    double z, v;
    GT::vertical(P, t.v[1], t.v[2], t.v[5], t.v[6], z, v);
    if (!st_finite(t)) {
        L.hi = L.trial;
    } else if ((v > 0.0) == (L.flo > 0.0) && v != 0.0) {
        L.lo = L.trial;
        const size_t np = (size_t)P.npix;
#pragma unroll
        for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + L.pixel] = t.v[c];
        L.best_ep = 1;
    } else {
        L.hi = L.trial;
    }
    if (++L.iter >= 48) return true;
    const double mid = (L.lo + L.hi) * 0.5;
    if (mid == L.lo || mid == L.hi) return true;
    L.trial = mid;
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
    if (L.refine == 3) {
        if (!turn_update<GEOM>(P, L, next)) return false;
        L.refine = 0;
        L.reason = R_DIRECT;
        L.steps += 1;
        steps_acc += 1;
        return true;
    }
    if (L.refine) {
        if (!refine_update<GEOM>(P, L, next)) return false;
        if (P.direct_only && L.best_ep) {
            // the refined crossing state (kept in the endpoint buffer, at
            // step fraction lo) may itself lie past a vertical turn
            const size_t np = (size_t)P.npix;
            St<double> kept;
#pragma unroll
            for (int c = 0; c < 16; ++c) kept.v[c] = P.ep[(size_t)c * np + L.pixel];
            if (turn_start<GEOM>(P, L, kept, L.full_h * L.lo)) return false;
        }
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
    if (P.direct_only && turn_start<GEOM>(P, L, next, used_h)) return false;
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
    const int valid = L.reason == R_DIRECT || (L.entered && (L.reason == R_INNER || L.reason == R_ESCAPED));
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
// PH_PARK: a windowed slow-light ray leaves the launch (block end, or a
// decoupled sample time in a later or earlier block) and resumes in the next.
enum : int { PH_IDLE = 0, PH_MAIN = 1, PH_CROSS = 2, PH_TIME = 3, PH_PARK = 4 };

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

// A waiting slow-light ray (req set) keeps the step size its call started
// with, already in pk_d, and records the time it needs.
template <int MODE> KP_DEV void lane_b_park(const Params& P, const LaneB& L) {
    const size_t np = (size_t)P.npix, p = (size_t)L.pixel;
#pragma unroll
    for (int c = 0; c < 16; ++c) P.ep[(size_t)c * np + p] = L.s[c];
    if constexpr (win_mode(MODE)) {
        if (!(L.req == L.req)) P.pk_d[p] = L.h_current;
        P.req_t[p] = L.req;
    } else {
        P.pk_d[p] = L.h_current;
    }
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

// Emission selection outside the equatorial wedge (apply_emission_selection,
// emission_selection.hpp:35-42): every radiation form carries `dark`, set
// per sample, and removes the emissivities of each coefficient set it makes
// (absorption and Faraday terms stay).
template <class C> KP_DEV void darken(C& c) {
    c.jI = c.jQ = c.jU = c.jV = decltype(c.jI)(0);
}

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
    bool dark = false;
    KP_DEV void zero() { coeffs_zero(c); }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_coefficients(P.riaf, sample, c);
        if (dark) darken(c);
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
    bool dark = false;
    KP_DEV void zero() { coeffs_zero(c); }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_coefficients_scaled(P.rm, sample, c);
        if (dark) darken(c);
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

template <> struct Radiation<false, M_SINGLE_V> : Radiation<false, M_SINGLE> {};
template <> struct Radiation<true, M_SINGLE_V> : Radiation<true, M_SINGLE> {};
// The vacuum leg returns before any radiation code; this empty form only
// completes the tick's template.
struct RadiationNone {
    bool dark = false;
    KP_DEV void zero() {}
    KP_DEV bool valid() const { return false; }
    KP_DEV uint32_t eval(const Params&, const St<double>&) { return 0; }
    KP_DEV void at(const Params&, double, Coeffs<double>& c) const { coeffs_zero(c); }
    KP_DEV uint32_t control(const Params&, double, double, double, int& req, int& pred) const {
        req = 1;
        pred = 1;
        return 0;
    }
    KP_DEV uint32_t transfer(const Params&, LaneB&, double) const { return 0; }
};
template <> struct Radiation<false, M_VAC> : RadiationNone {};
template <> struct Radiation<true, M_VAC> : RadiationNone {};

// Thermal (NT false) and kappa / power-law (NT true) coefficient evaluation;
// with NT false the nonthermal branch is compiled out.
template <bool NT> struct RadiationMF {
    RiafLocal<double> loc;
    bool dark = false;
    KP_DEV void zero() { loc.valid = 0; }
    KP_DEV bool valid() const { return loc.valid; }
    KP_DEV void at(const Params& P, double freq, Coeffs<double>& c) const {
        riaf_coeffs_at(P.riaf, loc, freq, c, NT ? &P.nt : nullptr);
        if (dark) darken(c);
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
            at(P, P.ctrl[f], c);
            req = maxv(req, substep_count(c, h, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(c, h_next, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return loc.valid ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int f = 0; f < P.nfreq; ++f) {
            Coeffs<double> c;
            at(P, P.freqs[f], c);
            mf_strang(P, L, f, c, absv(h) * P.dls_f[f]);
        }
        return loc.valid ? (uint32_t)P.nfreq : 0u;
    }
};
template <> struct Radiation<false, M_MF> : RadiationMF<false> {};
template <> struct Radiation<false, M_MF_NT> : RadiationMF<true> {};

// Mixed-precision frequency lists: FP32 coefficients of each frequency from
// one FP32 local plasma state, in nu^3-scaled units per unit |h|.
struct RadiationScaledMF {
    RiafLocal<float> loc;
    bool dark = false;
    KP_DEV void zero() { loc.valid = 0; }
    KP_DEV bool valid() const { return loc.valid; }
    KP_DEV void at_freq(const Params& P, float freq, float bconst, Coeffs<float>& c) const {
        riaf_coeffs_scaled_at(P.rm, loc, freq, bconst, c);
        if (dark) darken(c);
    }
    // output frequency f (single pass: also the control frequency f)
    KP_DEV void at32(const Params& P, int f, Coeffs<float>& c) const {
        at_freq(P, float(P.freqs[f]), P.bconst_f[f], c);
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = 1;
        pred = 1;
        for (int f = 0; f < P.nctrl; ++f) {
            Coeffs<float> c;
            at_freq(P, float(P.ctrl[f]), P.bconst_c[f], c);
            const Coeffs<double> d = widen(c);
            req = maxv(req, substep_count(d, h, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(d, h_next, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return loc.valid ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int f = 0; f < P.nfreq; ++f) {
            Coeffs<float> c;
            at32(P, f, c);
            mf_strang32(P, L, f, c, float(absv(h)));
        }
        return loc.valid ? (uint32_t)P.nfreq : 0u;
    }
};

template <> struct Radiation<true, M_MF> : RadiationScaledMF {
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        riaf_local_scaled(P.rm, sample, loc);
        return 0;
    }
};

template <bool NT> struct RadiationGRMHD {
    RiafLocal<double> loc;
    bool dark = false;
    KP_DEV void zero() { loc.valid = 0; }
    KP_DEV bool valid() const { return loc.valid; }
    KP_DEV void at(const Params& P, double freq, Coeffs<double>& c) const {
        grmhd_coeffs_at(P.grmhd, loc, freq, c, NT ? &P.nt : nullptr);
        if (dark) darken(c);
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
            at(P, P.ctrl[f], c);
            req = maxv(req, substep_count(c, h, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(c, h_next, P.dls_c[f], lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return loc.valid ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int f = 0; f < P.nfreq; ++f) {
            Coeffs<double> c;
            at(P, P.freqs[f], c);
            mf_strang(P, L, f, c, absv(h) * P.dls_f[f]);
        }
        return loc.valid ? (uint32_t)P.nfreq : 0u;
    }
};
template <> struct Radiation<false, M_GRMHD> : RadiationGRMHD<false> {};
template <> struct Radiation<false, M_GRMHD_NT> : RadiationGRMHD<true> {};

// The FP64 local plasma state in FP32 (the GRMHD form is never field-aligned).
KP_DEV void narrow_local(const RiafLocal<double>& d, RiafLocal<float>& o) {
    o.valid = d.valid;
    o.aligned = 0;
    o.has_chi = d.has_chi;
    o.ne = float(d.ne);
    o.thetae = float(d.thetae);
    o.b = float(d.b);
    o.nu_scale = float(d.nu_scale);
    o.cos_theta = float(d.cos_theta);
    o.sin_theta = float(d.sin_theta);
    o.cos2chi = float(d.cos2chi);
    o.sin2chi = float(d.sin2chi);
}

// Mixed-precision GRMHD (thermal fits): the grid gather, the native metric
// and the four-vectors stay FP64 (grmhd_local); the local plasma state is
// narrowed to FP32 and every frequency's coefficients follow the FP32
// thermal fit in nu^3-scaled units, with the GRMHD model's constants in
// P.rm (the sin(theta) floor 1e-4 of grmhd_grid.hpp:1210-1240).  The state
// is never field-aligned here: the GRMHD form has no aligned special case.
template <> struct Radiation<true, M_GRMHD> : RadiationScaledMF {
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        RiafLocal<double> d;
        grmhd_local(P.grmhd, sample, d);
        narrow_local(d, loc);
        return 0;
    }
};

// Slow light: the plasma state of snapshot pair (k, k+1) at time fraction f.
// Fluid interpolation mixes the sampled fluid before the reconstruction (lo
// holds the result); coefficient interpolation evaluates both snapshots (lo,
// hi) and mixes each frequency's coefficients (slow_light_driver.hpp:166-193).
template <bool NT> struct RadiationSlow {
    RiafLocal<double> lo, hi;
    double f;
    int k;
    bool dark = false;
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
        if (dark) darken(c);
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
template <> struct Radiation<false, M_SLOW_W> : RadiationSlow<false> {};
template <> struct Radiation<false, M_SLOW_NT_W> : RadiationSlow<true> {};

// Mixed-precision slow light (thermal fits): the time-interpolated plasma
// state is reconstructed in FP64 as above and narrowed to FP32; each
// frequency's coefficients follow the FP32 thermal fit in nu^3-scaled units,
// mixed in FP32 between the two snapshots under coefficient interpolation.
struct RadiationSlowMixed {
    RiafLocal<float> lo, hi;
    double f;
    int k;
    bool dark = false;
    KP_DEV void zero() {
        lo.valid = 0;
        hi.valid = 0;
        f = 0.0;
        k = 0;
    }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        RiafLocal<double> d;
        if (P.slow_interp == KP_SLOW_LIGHT_FLUID) {
            grmhd_local_time(P.grmhd, P.snaps[k], P.snaps[k + 1], f, sample, d);
            narrow_local(d, lo);
        } else {
            grmhd_local_time(P.grmhd, P.snaps[k], P.snaps[k], 0.0, sample, d);
            narrow_local(d, lo);
            grmhd_local_time(P.grmhd, P.snaps[k + 1], P.snaps[k + 1], 0.0, sample, d);
            narrow_local(d, hi);
        }
        return 0;
    }
    KP_DEV bool valid() const { return lo.valid || hi.valid; }
    KP_DEV void at_freq(const Params& P, float freq, float bconst, Coeffs<float>& c) const {
        riaf_coeffs_scaled_at(P.rm, lo, freq, bconst, c);
        if (P.slow_interp != KP_SLOW_LIGHT_FLUID) {
            Coeffs<float> b;
            riaf_coeffs_scaled_at(P.rm, hi, freq, bconst, b);
            const float ff = float(f), w = 1.0f - ff;
            c.jI = w * c.jI + ff * b.jI;
            c.jQ = w * c.jQ + ff * b.jQ;
            c.jU = w * c.jU + ff * b.jU;
            c.jV = w * c.jV + ff * b.jV;
            c.aI = w * c.aI + ff * b.aI;
            c.aQ = w * c.aQ + ff * b.aQ;
            c.aU = w * c.aU + ff * b.aU;
            c.aV = w * c.aV + ff * b.aV;
            c.rQ = w * c.rQ + ff * b.rQ;
            c.rU = w * c.rU + ff * b.rU;
            c.rV = w * c.rV + ff * b.rV;
        }
        if (dark) darken(c);
    }
    // output frequency q (single pass: also the control frequency q)
    KP_DEV void at32(const Params& P, int q, Coeffs<float>& c) const {
        at_freq(P, float(P.freqs[q]), P.bconst_f[q], c);
    }
    KP_DEV uint32_t control(const Params& P, double h, double h_next, double lmrs, int& req, int& pred) const {
        req = 1;
        pred = 1;
        for (int q = 0; q < P.nctrl; ++q) {
            Coeffs<float> c;
            at_freq(P, float(P.ctrl[q]), P.bconst_c[q], c);
            const Coeffs<double> d = widen(c);
            req = maxv(req, substep_count(d, h, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
            pred = maxv(pred, substep_count(d, h_next, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
        }
        return (lo.valid || hi.valid) ? (uint32_t)P.nctrl : 0u;
    }
    KP_DEV uint32_t transfer(const Params& P, LaneB& L, double h) const {
        for (int q = 0; q < P.nfreq; ++q) {
            Coeffs<float> c;
            at32(P, q, c);
            mf_strang32(P, L, q, c, float(absv(h)));
        }
        return (lo.valid || hi.valid) ? (uint32_t)P.nfreq : 0u;
    }
};
template <> struct Radiation<true, M_SLOW> : RadiationSlowMixed {};
template <> struct Radiation<true, M_SLOW_W> : RadiationSlowMixed {};

// Time probe: geometric stepping only (no coefficients, no substep control).
template <> struct Radiation<false, M_PROBE> {
    bool dark = false;
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
    bool dark = false;
    KP_DEV void zero() { coeffs_zero(c); }
    KP_DEV uint32_t eval(const Params& P, const St<double>& sample) {
        RiafLocal<double> L;
        riaf_local(P.riaf, sample, L);
        tangent_coeffs(P.riaf, L, c);
        if (dark) darken(c);
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
    bool dark = false;
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
        if (dark) darken(c);
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
// The equatorial wedge on an accepted pass-B step (pass_b.hpp:640-655):
// the step from the lane state to next (midpoint sample, step h > 0).  A
// fixed-step image samples the step's midpoint (one RK4 step of h/2); the
// step then needs wedge_samples samples across the wedge's 2 wedge_hr of
// cos(theta) (equatorial_resolution_substeps, emission_selection.hpp:45-60:
// the count is returned and the caller retakes the step), or is split at the
// first wedge surface it crosses (refine_equatorial_boundary, :69-122: 48
// bisection trials on the step's own RK composition, run in this tick; h,
// sample, next and the camera crossing become the split step's).  Returns 1
// when the step goes on.  The paths are FP64 in every precision.
template <int GP, int GEOM>
KP_DEV int wedge_step(const Params& P, const typename GeomT<GEOM>::GeoType& G, LaneB& L, double& h,
                      St<double>& sample, St<double>& next, uint32_t& guard) {
    using GT = GeomT<GEOM>;
    if (!P.adaptive) step_rk4<GP>(P, G, L.s, h * 0.5, sample, guard, true);
    const double outer = P.outer;
    if (GT::radius(L.s[1]) > outer && GT::radius(sample.v[1]) > outer && GT::radius(next.v[1]) > outer) return 1;
    const double hr = P.wedge_hr;
    const double q0 = GT::height_ratio(P, L.s[1], L.s[2]);
    const double qm = GT::height_ratio(P, sample.v[1], sample.v[2]);
    const double q1 = GT::height_ratio(P, next.v[1], next.v[2]);
    const double ratio = (absv(qm - q0) + absv(q1 - qm)) / (2.0 * hr / double(P.wedge_samples));
    if (ratio > 1.000001) return int(minv(ceil(ratio), 1e6));
    if (h == 0.0) return 1;
    const double tol = 128.0 * eps<double>();
    double edge = 0.0, lo = 0.0, hi = 1.0;
    bool found = false;
    for (int half = 0; half < 2 && !found; ++half) {
        const double a = half ? qm : q0, b = half ? q1 : qm;
        double earliest = 2.0;
        for (int sign = -1; sign <= 1; sign += 2) {
            const double surface = double(sign) * hr;
            if (absv(q0 - surface) <= tol || absv(a - surface) <= tol) continue;
            if ((a - surface) * (b - surface) > 0.0) continue;
            const double fraction = (surface - a) / (b - a);
            if (fraction < earliest) {
                earliest = fraction;
                edge = surface;
                found = true;
            }
        }
        if (found) {
            lo = half ? 0.5 : 0.0;
            hi = half ? 1.0 : 0.5;
        }
    }
    if (!found) return 1;
    // emission_selection_path_state: two RK4 half steps (adaptive) or one
    // RK4 step; with two half steps `mid` is the first one's end
    auto path = [&](double hh, St<double>& out, St<double>& mid) {
        if (P.adaptive)
            attempt<GP>(P, G, L.s, L.k1, hh, mid, out, guard, false, true);
        else
            step_rk4<GP>(P, G, L.s, hh, out, guard, true);
    };
    St<double> trial, mid;
    double low_q = q0;
    if (lo > 0.0) {
        path(h * lo, trial, mid);
        low_q = GT::height_ratio(P, trial.v[1], trial.v[2]);
    }
    const double low_value = low_q - edge;
    for (int iter = 0; iter < 48; ++iter) {
        const double m = (lo + hi) * 0.5;
        if (m == lo || m == hi) break;
        path(h * m, trial, mid);
        const double value = GT::height_ratio(P, trial.v[1], trial.v[2]) - edge;
        if ((value > 0.0) == (low_value > 0.0))
            lo = m;
        else
            hi = m;
    }
    const double fraction = (lo + hi) * 0.5;
    if (fraction <= tol || fraction >= 1.0 - tol) return 1;
    h *= fraction;
    path(h, next, sample);
    if (!P.adaptive) step_rk4<GP>(P, G, L.s, h * 0.5, sample, guard, true);
    const double s0 = GT::radius(L.s[1]) - P.radius;
    const double s1 = GT::radius(next.v[1]) - P.radius;
    L.crosses = (s0 == 0.0 || s1 == 0.0) ? 1 : ((s0 < 0.0 && s1 > 0.0) || (s0 > 0.0 && s1 < 0.0));
    return 1;
}

// SEL: the tick carries the equatorial wedge (wedge_step, the radiation
// forms' dark flag); the kernels of images without a wedge are built
// without that code.
template <bool MIXED, int GP, int MODE, bool SEL = false>
KP_DEV bool pass_b_tick(const Params& P, const typename GeomT<geom_of<MODE>()>::GeoType& G, LaneB& L,
                        uint32_t& coeff_acc, uint32_t& transfer_acc, uint32_t& steps_acc, uint32_t& guard) {
    using GT = GeomT<geom_of<MODE>()>;
    constexpr bool SLOW = base_mode(MODE) == M_SLOW || MODE == M_STAN;
    const double eff_min = absv(P.min_step);
    if (L.phase == PH_IDLE) {
        if (L.reached || L.steps >= P.max_steps) return true;
        if constexpr (vpark_mode(MODE)) {
            // outside the emitting region and moving outward (k^r > 0) at
            // r >= outer, which the host keeps beyond every photon orbit: no
            // later sample radiates, the vacuum leg takes the ray over
            if (GT::radius(L.s[1]) >= P.outer && L.s[5] > 0.0) {
                L.phase = PH_PARK;
                return false;
            }
        }
        if constexpr (SLOW) {
            if (!L.vac && P.slow_mode != KP_SLOW_LIGHT_DECOUPLED) {
                while (!L.vac && P.t_obs + L.s[0] >= slow_upper(P, L)) slow_window_end(P, L);
            }
            if (L.vac == 1) {
                if constexpr (win_mode(MODE)) {
                    if (!P.slow_final) {
                        // the block's last window ended: the ray waits for the next block
                        // (slow_light_driver.hpp:1115-1116, slow_light_batch_driver.hpp:300)
                        L.phase = PH_PARK;
                        return false;
                    }
                }
                const double tol_r = 1e-8 * maxv(1.0, absv(P.outer));
                if (GT::radius(L.s[1]) < P.outer - tol_r) {
                    L.reason = R_SLOW_TIME;
                    return true;
                }
                L.vac = 2;
            }
        }
        double h = L.h_current;
        if constexpr (win_mode(MODE)) {
            // retry_h of this call (slow_light_driver.hpp:1156, 1302)
            if (P.slow_mode == KP_SLOW_LIGHT_DECOUPLED) P.pk_d[L.pixel] = h;
        }
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
    if constexpr (vac_mode(MODE)) {
        // no sample radiates: the substep counts are 1, the cap is max_step
        // and the transfer step leaves the Stokes vector as it is (counted,
        // not computed)
        L.cap = P.max_step;
        if (L.h > 0.0) transfer_acc += 1;
        st_store(L.s, next);
        L.steps += 1;
        steps_acc += 1;
        if (L.crosses) L.reached = 1;
        return false;
    }
    double h = L.h;
    const double old_r = GT::radius(L.s[1]);
    double sample_r = GT::radius(sample.v[1]), next_r = GT::radius(next.v[1]);
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
    if constexpr (SEL) {
        if (P.wedge_hr > 0.0) {
            const int n = wedge_step<GP, geom_of<MODE>()>(P, G, L, h, sample, next, guard);
            if (n > 1) {
                if (h <= eff_min * 1.0001) {
                    L.reason = R_UNDERFLOW;
                    return true;
                }
                L.h_current = maxv(eff_min, h / double(n));
                L.cap = L.h_current;
                return false;
            }
            L.h = h;
            sample_r = GT::radius(sample.v[1]);
            next_r = GT::radius(next.v[1]);
        }
    }
    Radiation<MIXED, MODE> rad;
    rad.zero();
    if constexpr (SEL) rad.dark = P.wedge_hr > 0.0 && absv(GT::height_ratio(P, sample.v[1], sample.v[2])) > P.wedge_hr;
    if constexpr (SLOW) {
        const int radiating = sample_r >= P.inner && sample_r < P.outer;
        const double ct = P.t_obs + sample.v[0];
        if (P.slow_mode == KP_SLOW_LIGHT_DECOUPLED && radiating && !(ct >= P.snap_t[0] && ct <= P.snap_t[P.nsnap - 1])) {
            // the time is not resident: the ray waits, unrefined, for the
            // block that holds it (slow_light_driver.hpp:1299-1304); a time
            // outside the whole sequence no block brings
            // (slow_light_batch_driver.hpp:290-297)
            L.req = ct;
            if constexpr (win_mode(MODE)) {
                if (ct >= P.sup_lo && ct <= P.sup_hi) {
                    L.phase = PH_PARK;
                    return false;
                }
            }
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
    constexpr bool SHARED32 = MIXED && (MODE == M_GRMHD || MODE == M_SLOW || MODE == M_SLOW_W);
    if constexpr (SHARED || SHARED32) {
        if (P.single_pass) {
            // the control frequencies are the output frequencies: one
            // coefficient evaluation per frequency serves the substep control
            // (pass_b.hpp:1416-1460) and a tentative transfer of that
            // frequency's Stokes vector into the second buffer
            required = 1;
            pred = 1;
            const bool valid = rad.valid();
            for (int f = 0; f < P.nfreq; ++f) {
                if constexpr (SHARED32) {
                    // FP32 coefficients per unit |h| (the scaled units of
                    // Radiation<true, M_GRMHD>::control and transfer)
                    Coeffs<float> c;
                    rad.at32(P, f, c);
                    const Coeffs<double> d = widen(c);
                    required = maxv(required, substep_count(d, h, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
                    pred = maxv(pred, substep_count(d, L.h_current, 1.0, lmrs, P.max_rd, P.max_ad, P.max_fd));
                    if (h > 0.0) {
                        Stokes<float> st = {float(L.mf[4 * f]), float(L.mf[4 * f + 1]), float(L.mf[4 * f + 2]),
                                            float(L.mf[4 * f + 3])};
                        strang_step32(st, c, float(absv(h)));
                        L.mf2.set(4 * f, st.I);
                        L.mf2.set(4 * f + 1, st.Q);
                        L.mf2.set(4 * f + 2, st.U);
                        L.mf2.set(4 * f + 3, st.V);
                    }
                } else {
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

template <bool PERSISTENT, bool MIXED, int GP, int MODE, bool SEL = false>
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
                    if ((MIXED || win_mode(MODE) || vac_mode(MODE)) && P.list_in) {
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
            if constexpr (win_mode(MODE) || vpark_mode(MODE)) {
                // windowed slow light, radiating leg: a ray that parks leaves the launch
                bool park = false;
                if (has) {
                    if (pass_b_tick<MIXED, GP, MODE, SEL>(P, G, L, coeff, transfer, steps_b, guard)) {
                        lane_b_finish<MIXED, MODE>(P, L, 1);
                        has = false;
                    } else if (L.phase == PH_PARK) {
                        lane_b_park<MODE>(P, L);
                        has = false;
                        park = true;
                    }
                }
                append_parked(P, park, L.pixel);
            } else if (has && pass_b_tick<MIXED, GP, MODE, SEL>(P, G, L, coeff, transfer, steps_b, guard)) {
                lane_b_finish<MIXED, MODE>(P, L, 1);
                has = false;
            }
        }
    } else {
        // Pass B is built persistent only: its one-ray-per-thread form (each
        // lane ticking its own ray to the end) computed wrong multi-frequency
        // FP64 images on the device while every other form of the same tick,
        // and the tick's host build, agreed with the oracle; the cause was
        // not found in the source.
        static_assert(PERSISTENT, "pass B kernels are persistent");
    }
    flush_counters(P.counters, vac_mode(MODE) ? C_RHS_V : C_RHS_B, rhs_count, coeff, transfer, 0, steps_b, guard,
                   vac_mode(MODE) ? C_VACUUM : C_TRANSFER);
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
template <bool MIXED, int GP, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP_WAVES_PER_EU))) void unsplit_kernel(Params P) {
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

template <int GP>
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
            if (pass_a_tick<GP, G_BL>(*P, G, L, steps_a, guard)) {
                lane_a_store(*P, L);
                has = false;
            }
        }
    }
    flush_counters(B.counters, C_RHS_A, rhs_count, 0, 0, steps_a, 0, guard);
}

template <bool MIXED, int GP, bool SEL = false>
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
            if (pass_b_tick<MIXED, GP, M_SINGLE, SEL>(*P, G, L, coeff, transfer, steps_b, guard)) {
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

}  // namespace

// ---------------------------------------------------------------------------
// Kernel selectors (defined in kp_kernels.hip).  gp: geodesic precision
// (GeoPrec), mode: pass-B variant (Mode), geom: Geom.  Each returns the
// kernel's host-side handle for hipLaunchKernel / graph kernel nodes, or
// null for a combination that is not instantiated.
// ---------------------------------------------------------------------------
namespace kpsel {
const void* pass_a(bool persistent, int gp, int geom);
const void* pass_a_batch(int gp);
const void* zero_counters();
const void* derive_scalars();
const void* pass_b_single(bool mixed, int gp);
const void* pass_b_mf(bool mixed, int gp);
const void* pass_b_fmks(int mode);
const void* pass_b_slow(int mode);
const void* pass_b_other(int mode);
const void* pass_b_batch(bool mixed, int gp);
const void* unsplit_bl(bool mixed, int gp, int mode);
const void* unsplit_fmks(int mode);
const void* pass_b_stan();
const void* pass_b_slow_windowed(int mode);
const void* pass_b_legs(bool mixed, int gp, bool vacuum);
const void* unsplit_slow(int mode);
const void* pass_b_grmhd_mixed(int mode);
const void* unsplit_grmhd_mixed(int mode);
const void* pass_b_slow_mixed(int mode);
const void* unsplit_slow_mixed(int mode);
// persistent split pass B with the equatorial wedge (images with a wedge run
// these whatever the launch options ask: the results are the same)
const void* pass_b_sel(bool mixed, int gp, int mode);
const void* pass_b_sel_fmks(bool mixed, int mode);
const void* pass_b_sel_fmks2(bool mixed, int mode);
const void* pass_b_batch_sel(bool mixed, int gp);
}  // namespace kpsel
