// kp_kernels.hip -- instantiation of the imaging kernels, one family per part.
//
// The build compiles this file once per part (-DKP_PART=0..KP_NPARTS-1, in
// parallel); each part instantiates its kernels and defines the selectors
// that hand their addresses to the host code (kp_kernels.h, namespace
// kpsel).  Every kernel is instantiated in exactly one part.
#include "kp_kernels.h"

#ifndef KP_PART
#error "compile kp_kernels.hip with -DKP_PART=<part>"
#endif

namespace kpsel {

#if KP_PART == 0
// Pass A (every geometry and geodesic precision), the batched pass A and the
// small utility kernels.
const void* pass_a(bool persistent, int gp, int geom) {
    if (geom == G_FMKS)
        return persistent ? (const void*)pass_a_kernel<true, GP_64, G_FMKS>
                          : (const void*)pass_a_kernel<false, GP_64, G_FMKS>;
    if (gp == GP_GUARD)
        return persistent ? (const void*)pass_a_kernel<true, GP_GUARD, G_BL>
                          : (const void*)pass_a_kernel<false, GP_GUARD, G_BL>;
    if (gp == GP_32) return persistent ? (const void*)pass_a_kernel<true, GP_32, G_BL> : nullptr;
    return persistent ? (const void*)pass_a_kernel<true, GP_64, G_BL> : (const void*)pass_a_kernel<false, GP_64, G_BL>;
}

const void* pass_a_batch(int gp) {
    return gp == GP_GUARD ? (const void*)pass_a_batch_kernel<GP_GUARD> : (const void*)pass_a_batch_kernel<GP_64>;
}

const void* zero_counters() { return (const void*)zero_kernel; }
const void* derive_scalars() { return (const void*)derive_scalars_kernel; }
#endif

#if KP_PART == 1
// Pass B runs persistent kernels only (a lane takes the next pixel from the
// queue when its ray ends).
//
// Single-frequency Boyer-Lindquist pass B: FP64, and mixed with every
// geodesic precision.
const void* pass_b_single(bool mixed, int gp) {
    if (!mixed)
        return (const void*)pass_b_kernel<true, false, GP_64, M_SINGLE>;
    if (gp == GP_32) return (const void*)pass_b_kernel<true, true, GP_32, M_SINGLE>;
    if (gp == GP_64)
        return (const void*)pass_b_kernel<true, true, GP_64, M_SINGLE>;
    return (const void*)pass_b_kernel<true, true, GP_GUARD, M_SINGLE>;
}
#endif

#if KP_PART == 2
// Multi-frequency Boyer-Lindquist pass B (thermal fits).
const void* pass_b_mf(bool mixed, int gp) {
    if (!mixed)
        return (const void*)pass_b_kernel<true, false, GP_64, M_MF>;
    if (gp == GP_32) return (const void*)pass_b_kernel<true, true, GP_32, M_MF>;
    if (gp == GP_64)
        return (const void*)pass_b_kernel<true, true, GP_64, M_MF>;
    return (const void*)pass_b_kernel<true, true, GP_GUARD, M_MF>;
}
#endif

#if KP_PART == 3
// GRMHD pass B on FMKS geodesics: thermal, nonthermal, and the slow-light
// time probe.
const void* pass_b_fmks(int mode) {
    switch (mode) {
    case M_GRMHD:
        return (const void*)pass_b_kernel<true, false, GP_64, M_GRMHD>;
    case M_GRMHD_NT:
        return (const void*)pass_b_kernel<true, false, GP_64, M_GRMHD_NT>;
    case M_PROBE:
        return (const void*)pass_b_kernel<true, false, GP_64, M_PROBE>;
    default:
        return nullptr;
    }
}
#endif

#if KP_PART == 4
// Slow-light pass B: thermal and nonthermal (part 4), and the tangent of the
// slow-light sequence (part 8).
const void* pass_b_slow(int mode) {
    switch (mode) {
    case M_SLOW:
        return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW>;
    case M_SLOW_NT:
        return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW_NT>;
    case M_STAN:
        return pass_b_stan();
    default:
        return nullptr;
    }
}
#endif

#if KP_PART == 8
const void* pass_b_stan() {
    return (const void*)pass_b_kernel<true, false, GP_64, M_STAN>;
}
#endif

#if KP_PART == 5
// RIAF tangent and nonthermal multi-frequency pass B, and the batched pass B.
const void* pass_b_other(int mode) {
    switch (mode) {
    case M_TANGENT:
        return (const void*)pass_b_kernel<true, false, GP_64, M_TANGENT>;
    case M_MF_NT:
        return (const void*)pass_b_kernel<true, false, GP_64, M_MF_NT>;
    default:
        return nullptr;
    }
}

const void* pass_b_batch(bool mixed, int gp) {
    if (!mixed) return (const void*)pass_b_batch_kernel<false, GP_64>;
    return gp == GP_GUARD ? (const void*)pass_b_batch_kernel<true, GP_GUARD>
                          : (const void*)pass_b_batch_kernel<true, GP_64>;
}
#endif

#if KP_PART == 6
// Unsplit kernels on Boyer-Lindquist geodesics.
const void* unsplit_bl(bool mixed, int gp, int mode) {
    switch (mode) {
    case M_SINGLE:
        if (!mixed) return (const void*)unsplit_kernel<false, GP_64, M_SINGLE>;
        return gp == GP_GUARD ? (const void*)unsplit_kernel<true, GP_GUARD, M_SINGLE>
                              : (const void*)unsplit_kernel<true, GP_64, M_SINGLE>;
    case M_MF:
        if (!mixed) return (const void*)unsplit_kernel<false, GP_64, M_MF>;
        return gp == GP_GUARD ? (const void*)unsplit_kernel<true, GP_GUARD, M_MF>
                              : (const void*)unsplit_kernel<true, GP_64, M_MF>;
    case M_TANGENT:
        return mixed ? nullptr : (const void*)unsplit_kernel<false, GP_64, M_TANGENT>;
    case M_MF_NT:
        return mixed ? nullptr : (const void*)unsplit_kernel<false, GP_64, M_MF_NT>;
    default:
        return nullptr;
    }
}
#endif

#if KP_PART == 7
// Unsplit kernels on FMKS geodesics: static grids here, slow light in part 9.
const void* unsplit_fmks(int mode) {
    switch (mode) {
    case M_GRMHD: return (const void*)unsplit_kernel<false, GP_64, M_GRMHD>;
    case M_GRMHD_NT: return (const void*)unsplit_kernel<false, GP_64, M_GRMHD_NT>;
    case M_PROBE: return (const void*)unsplit_kernel<false, GP_64, M_PROBE>;
    default: return unsplit_slow(mode);
    }
}
#endif

#if KP_PART == 10
// Windowed slow light: persistent pass B over one block of the sequence.
const void* pass_b_slow_windowed(int mode) {
    switch (mode) {
    case M_SLOW: return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW_W>;
    case M_SLOW_NT: return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW_NT_W>;
    default: return nullptr;
    }
}
#endif

#if KP_PART == 11
// Single-frequency Boyer-Lindquist pass B in two legs (persistent): the
// radiating leg parks rays at the vacuum boundary, the vacuum leg ends them.
const void* pass_b_legs(bool mixed, int gp, bool vacuum) {
    if (!mixed)
        return vacuum ? (const void*)pass_b_kernel<true, false, GP_64, M_VAC>
                      : (const void*)pass_b_kernel<true, false, GP_64, M_SINGLE_V>;
    if (gp == GP_64)
        return vacuum ? (const void*)pass_b_kernel<true, true, GP_64, M_VAC>
                      : (const void*)pass_b_kernel<true, true, GP_64, M_SINGLE_V>;
    if (gp == GP_GUARD)
        return vacuum ? (const void*)pass_b_kernel<true, true, GP_GUARD, M_VAC>
                      : (const void*)pass_b_kernel<true, true, GP_GUARD, M_SINGLE_V>;
    return nullptr;
}
#endif

#if KP_PART == 12
// Mixed-precision GRMHD pass B (FP64 FMKS geodesics, FP32 thermal
// coefficients), split and unsplit; slow light in part 13.
const void* pass_b_grmhd_mixed(int mode) {
    if (mode != M_GRMHD) return pass_b_slow_mixed(mode);
    return (const void*)pass_b_kernel<true, true, GP_64, M_GRMHD>;
}

const void* unsplit_grmhd_mixed(int mode) {
    return mode == M_GRMHD ? (const void*)unsplit_kernel<true, GP_64, M_GRMHD> : unsplit_slow_mixed(mode);
}
#endif

#if KP_PART == 13
// Mixed-precision slow light: resident sequences (split and unsplit) and the
// windowed block launches (persistent).
const void* pass_b_slow_mixed(int mode) {
    switch (mode) {
    case M_SLOW:
        return (const void*)pass_b_kernel<true, true, GP_64, M_SLOW>;
    case M_SLOW_W: return (const void*)pass_b_kernel<true, true, GP_64, M_SLOW_W>;
    default: return nullptr;
    }
}

const void* unsplit_slow_mixed(int mode) {
    return mode == M_SLOW ? (const void*)unsplit_kernel<true, GP_64, M_SLOW> : nullptr;
}
#endif

#if KP_PART == 14
// Images with an equatorial wedge: persistent split pass B on
// Boyer-Lindquist geodesics (FP64 and mixed) and the batched pass B.
const void* pass_b_sel(bool mixed, int gp, int mode) {
    if (!mixed) {
        switch (mode) {
        case M_SINGLE: return (const void*)pass_b_kernel<true, false, GP_64, M_SINGLE, true>;
        case M_MF: return (const void*)pass_b_kernel<true, false, GP_64, M_MF, true>;
        case M_MF_NT: return (const void*)pass_b_kernel<true, false, GP_64, M_MF_NT, true>;
        case M_TANGENT: return (const void*)pass_b_kernel<true, false, GP_64, M_TANGENT, true>;
        default: return pass_b_sel_fmks(false, mode);
        }
    }
    switch (mode) {
    case M_SINGLE:
        return gp == GP_GUARD ? (const void*)pass_b_kernel<true, true, GP_GUARD, M_SINGLE, true>
                              : (const void*)pass_b_kernel<true, true, GP_64, M_SINGLE, true>;
    case M_MF:
        return gp == GP_GUARD ? (const void*)pass_b_kernel<true, true, GP_GUARD, M_MF, true>
                              : (const void*)pass_b_kernel<true, true, GP_64, M_MF, true>;
    default: return pass_b_sel_fmks(true, mode);
    }
}

const void* pass_b_batch_sel(bool mixed, int gp) {
    if (!mixed) return (const void*)pass_b_batch_kernel<false, GP_64, true>;
    return gp == GP_GUARD ? (const void*)pass_b_batch_kernel<true, GP_GUARD, true>
                          : (const void*)pass_b_batch_kernel<true, GP_64, true>;
}
#endif

#if KP_PART == 15
// Images with an equatorial wedge on FMKS geodesics: GRMHD and resident
// slow light in FP64 here; the windowed and mixed kernels in part 16.
const void* pass_b_sel_fmks(bool mixed, int mode) {
    if (mixed || mode == M_SLOW_W || mode == M_SLOW_NT_W) return pass_b_sel_fmks2(mixed, mode);
    switch (mode) {
    case M_GRMHD: return (const void*)pass_b_kernel<true, false, GP_64, M_GRMHD, true>;
    case M_GRMHD_NT: return (const void*)pass_b_kernel<true, false, GP_64, M_GRMHD_NT, true>;
    case M_SLOW: return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW, true>;
    case M_SLOW_NT: return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW_NT, true>;
    case M_STAN: return (const void*)pass_b_kernel<true, false, GP_64, M_STAN, true>;
    default: return nullptr;
    }
}
#endif

#if KP_PART == 16
const void* pass_b_sel_fmks2(bool mixed, int mode) {
    if (mixed) {
        switch (mode) {
        case M_GRMHD: return (const void*)pass_b_kernel<true, true, GP_64, M_GRMHD, true>;
        case M_SLOW: return (const void*)pass_b_kernel<true, true, GP_64, M_SLOW, true>;
        case M_SLOW_W: return (const void*)pass_b_kernel<true, true, GP_64, M_SLOW_W, true>;
        default: return nullptr;
        }
    }
    switch (mode) {
    case M_SLOW_W: return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW_W, true>;
    case M_SLOW_NT_W: return (const void*)pass_b_kernel<true, false, GP_64, M_SLOW_NT_W, true>;
    default: return nullptr;
    }
}
#endif

#if KP_PART == 9
const void* unsplit_slow(int mode) {
    switch (mode) {
    case M_SLOW: return (const void*)unsplit_kernel<false, GP_64, M_SLOW>;
    case M_SLOW_NT: return (const void*)unsplit_kernel<false, GP_64, M_SLOW_NT>;
    case M_STAN: return (const void*)unsplit_kernel<false, GP_64, M_STAN>;
    default: return nullptr;
    }
}
#endif

}  // namespace kpsel
