// CPU build of the imaging kernel's per-lane state machines (test harness).
//
// Includes the product's kernel header and parameter preparation, built for
// the host only, and drives the pass A / pass B lane functions one pixel at a
// time, so the kernel's control logic can be checked against the oracle
// without a GPU.  Never part of the product.
#include "../../kpolaris_amd/csrc/kp_prepare.h"
#include "../../kpolaris_amd/csrc/kp_slow_blocks.h"

#include <vector>

template <bool MIXED, int MODE>
static kp_status emul_run(const kp_pass_a_params* pa, const kp_riaf_model* m, double fp64_radius,
                          const FreqLists& fl, kp_image_data* out, int64_t* rhs_evals, int64_t* guard_steps,
                          const double* log_scales = nullptr, double* grad = nullptr,
                          const kp_grmhd_model* gm = nullptr, const kp_slow_light* slow = nullptr,
                          double* req_t = nullptr, unsigned long long* waiting = nullptr) {
    constexpr bool FM = base_mode(MODE) == M_GRMHD || base_mode(MODE) == M_SLOW || MODE == M_STAN;
    constexpr int GP = MIXED && !FM ? GP_GUARD : GP_64;
    Params P;
    kp_status st = FM ? prepare_grmhd(pa, gm, P) : prepare(pa, m, P);
    if (st != KP_OK) return st;
    if (MIXED) {
        if (FM)
            prepare_mixed_grmhd(P, gm->mbh_solar);
        else
            prepare_mixed(P);
        P.fp64_radius = fp64_radius;
    }
    if (lds_stokes<MODE>() || MODE == M_STAN) {
        st = prepare_frequencies(FM ? gm->mbh_solar : m->mbh_solar, fl, P);
        if (st != KP_OK) return st;
    }
    unsigned long long counters[C_COUNT] = {};
    P.counters = counters;
    if (base_mode(MODE) == M_SLOW || MODE == M_STAN) {
        P.snaps = reinterpret_cast<const kp_grmhd_cell* const*>(slow->cells);
        P.snap_t = slow->times;
        P.nsnap = slow->nsnap;
        P.slow_mode = slow->step_mode;
        P.slow_interp = slow->interpolation;
        P.t_obs = slow->observation_time;
        P.req_t = req_t;
    }
    // the library's single-pass rule: FP64 frequency lists, and mixed GRMHD
    if (MIXED && MODE != M_GRMHD && MODE != M_SLOW) P.single_pass = 0;
    if (MODE == M_STAN) {
        P.tq_ne = std::exp(log_scales[0]);
        P.tq_b = std::exp(0.5 * log_scales[0]) * std::exp(log_scales[1]);
        P.grad = grad;
        P.single_pass = 0;
    }
    if (MODE == M_TANGENT) {
        P.riaf.ne_factor = std::exp(log_scales[0]);
        P.riaf.thetae_factor = std::exp(log_scales[1]);
        P.riaf.b_factor = std::exp(0.5 * log_scales[0]) * std::exp(log_scales[2]);
        P.grad = grad;
    }
    double mfbuf[4 * KP_MAX_FREQUENCIES], mfbuf2[4 * KP_MAX_FREQUENCIES];
    const int np = P.npix;
    std::vector<double> ep(16 * (size_t)np);
    std::vector<int32_t> valid(np), steps(np), reason(np);
    P.ep = ep.data();
    P.ep_valid = valid.data();
    P.ep_steps = steps.data();
    P.ep_reason = reason.data();
    P.out = *out;
    int64_t total = 0, guard_total = 0;
    for (int pix = 0; pix < np; ++pix) {
        uint32_t rc = 0, sa = 0, c1 = 0, c2 = 0, sb = 0, guard = 0;
        const auto G = GeomT<geom_of<MODE>()>::geo(P, &rc);
        LaneA A;
        lane_a_init(P, pix, A);
        while (!pass_a_tick<GP, geom_of<MODE>()>(P, G, A, sa, guard)) {
        }
        lane_a_store(P, A);
        LaneB B;
        B.mf = LdsVec<double>{mfbuf, 1};
        B.mf2 = LdsVec<double>{mfbuf2, 1};
        if (lane_b_init<MODE>(P, pix, B)) {
            // the lanes with the equatorial-wedge code (the library's wedge
            // kernels); without a wedge it is skipped at run time
            while (!pass_b_tick<MIXED, GP, MODE, true>(P, G, B, c1, c2, sb, guard)) {
            }
            lane_b_finish<MIXED, MODE>(P, B, 1);
        }
        total += rc;
        guard_total += guard;
    }
    if (rhs_evals) *rhs_evals = total;
    if (guard_steps) *guard_steps = guard_total;
    if (waiting) *waiting = counters[C_WAIT];
    return KP_OK;
}

static bool nonthermal(int emission_type) { return emission_nonthermal(emission_type); }

extern "C" kp_status kp_emul_run_riaf_image(const kp_pass_a_params* pa, const kp_riaf_model* m,
                                            kp_image_data* out, int64_t* rhs_evals) {
    if (m && nonthermal(m->emission_type)) {
        // kappa / power law run in the frequency-list lanes (as the library does)
        FreqLists fl;
        fl.nfreq = fl.nctrl = 1;
        fl.freqs = fl.ctrl = &m->freq_cgs;
        return emul_run<false, M_MF_NT>(pa, m, 0.0, fl, out, rhs_evals, nullptr);
    }
    return emul_run<false, false>(pa, m, 0.0, FreqLists{}, out, rhs_evals, nullptr);
}

// The mixed-precision lanes (FP32 arithmetic on the host's IEEE float; the
// device's 1-ulp hardware reciprocal is a correctly rounded division here).
extern "C" kp_status kp_emul_run_riaf_image_mixed(const kp_pass_a_params* pa, const kp_riaf_model* m,
                                                  double fp64_radius, kp_image_data* out, int64_t* rhs_evals,
                                                  int64_t* guard_steps) {
    return emul_run<true, false>(pa, m, fp64_radius, FreqLists{}, out, rhs_evals, guard_steps);
}

// Multi-frequency lanes (Stokes arrays [nfreq][npix]); mixed != 0 selects the
// mixed-precision lanes.
extern "C" kp_status kp_emul_run_riaf_multifrequency_image(const kp_pass_a_params* pa, const kp_riaf_model* m,
                                                           int32_t mixed, double fp64_radius, const double* freqs,
                                                           int32_t nfreq, const double* ctrl, int32_t nctrl,
                                                           kp_image_data* out, int64_t* rhs_evals) {
    FreqLists fl;
    fl.nfreq = nfreq;
    fl.freqs = freqs;
    fl.nctrl = nctrl;
    fl.ctrl = ctrl;
    if (mixed) return emul_run<true, true>(pa, m, fp64_radius, fl, out, rhs_evals, nullptr);
    if (m && nonthermal(m->emission_type)) return emul_run<false, M_MF_NT>(pa, m, 0.0, fl, out, rhs_evals, nullptr);
    return emul_run<false, true>(pa, m, 0.0, fl, out, rhs_evals, nullptr);
}

// The kernels' bounded-argument sine/cosine, for its accuracy test.
extern "C" void kp_emul_sincos(double x, double* s, double* c) { kpd::sincos_bounded(x, s, c); }

// Tangent (plasma-parameter derivative) lanes; grad is [3][4][npix].
extern "C" kp_status kp_emul_run_riaf_image_gradient(const kp_pass_a_params* pa, const kp_riaf_model* m,
                                                     const double* log_scales, kp_image_data* out, double* grad) {
    return emul_run<false, M_TANGENT>(pa, m, 0.0, FreqLists{}, out, nullptr, nullptr, log_scales, grad);
}

// GRMHD lanes on FMKS geodesics; model->cells are HOST records here.
extern "C" kp_status kp_emul_run_grmhd_image(const kp_pass_a_params* pa, const kp_grmhd_model* gm,
                                             const double* freqs, int32_t nfreq, const double* ctrl, int32_t nctrl,
                                             kp_image_data* out, int64_t* rhs_evals) {
    FreqLists fl;
    fl.nfreq = nfreq;
    fl.freqs = freqs;
    fl.nctrl = nctrl;
    fl.ctrl = ctrl;
    if (nonthermal(gm->emission_type))
        return emul_run<false, M_GRMHD_NT>(pa, nullptr, 0.0, fl, out, rhs_evals, nullptr, nullptr, nullptr, gm);
    return emul_run<false, M_GRMHD>(pa, nullptr, 0.0, fl, out, rhs_evals, nullptr, nullptr, nullptr, gm);
}

// Mixed-precision GRMHD lanes (thermal fits): FP32 coefficients and
// transfer on FP64 FMKS geodesics; model->cells are HOST records here.
extern "C" kp_status kp_emul_run_grmhd_image_mixed(const kp_pass_a_params* pa, const kp_grmhd_model* gm,
                                                   const double* freqs, int32_t nfreq, const double* ctrl,
                                                   int32_t nctrl, kp_image_data* out, int64_t* rhs_evals) {
    FreqLists fl;
    fl.nfreq = nfreq;
    fl.freqs = freqs;
    fl.nctrl = nctrl;
    fl.ctrl = ctrl;
    if (nonthermal(gm->emission_type)) return KP_UNSUPPORTED;
    return emul_run<true, M_GRMHD>(pa, nullptr, 0.0, fl, out, rhs_evals, nullptr, nullptr, nullptr, gm);
}

// The device derived-scalar formulas run on host cells.
extern "C" kp_status kp_emul_grmhd_derive_scalars(const kp_grmhd_model* gm, kp_grmhd_cell* cells) {
    kp_pass_a_params p;
    kp_pass_a_params_default(&p);
    p.coordinate_system = KP_COORD_FMKS;
    p.camera.nx = p.camera.ny = 1;
    p.camera.radius = 100.0;
    kp_grmhd_model g = *gm;
    g.cells = cells;
    Params P;
    const kp_status st = prepare_grmhd(&p, &g, P);
    if (st != KP_OK) return st;
    for (int i = 0; i < g.n1; ++i)
        for (int j = 0; j < g.n2; ++j)
            for (int k = 0; k < g.n3; ++k) grmhd_derive_cell(P.grmhd, i, j, cells[((size_t)i * g.n2 + j) * g.n3 + k]);
    return KP_OK;
}

// FMKS transport RHS of the device code, for its parity check.
extern "C" void kp_emul_fmks_rhs(double mass, double spin, const double fm[6], const double* s, double* d) {
    const FmksMap f = fmks_map(fm[0], fm[1], fm[2], fm[3], fm[4], fm[5]);
    St<double> y;
    for (int i = 0; i < 16; ++i) y.v[i] = s[i];
    fmks_rhs_apply(mass, spin, f, y, [&](int i, double v) { d[i] = v; });
}

// Slow-light lanes over HOST snapshot records; req_t [npix] receives the
// time each waiting ray needed (NaN for the others), waiting their count.
extern "C" kp_status kp_emul_run_grmhd_slow_light_image(const kp_pass_a_params* pa, const kp_grmhd_model* gm,
                                                        const kp_slow_light* seq, const double* freqs,
                                                        int32_t nfreq, kp_image_data* out, double* req_t,
                                                        unsigned long long* waiting) {
    FreqLists fl;
    fl.nfreq = fl.nctrl = nfreq;
    fl.freqs = fl.ctrl = freqs;
    kp_grmhd_model g = *gm;
    g.cells = seq->cells[0];
    if (nonthermal(gm->emission_type))
        return emul_run<false, M_SLOW_NT>(pa, nullptr, 0.0, fl, out, nullptr, nullptr, nullptr, nullptr, &g, seq,
                                          req_t, waiting);
    return emul_run<false, M_SLOW>(pa, nullptr, 0.0, fl, out, nullptr, nullptr, nullptr, nullptr, &g, seq, req_t,
                                   waiting);
}

// Mixed-precision slow-light lanes (thermal fits) over HOST snapshot records.
extern "C" kp_status kp_emul_run_grmhd_slow_light_image_mixed(const kp_pass_a_params* pa, const kp_grmhd_model* gm,
                                                              const kp_slow_light* seq, const double* freqs,
                                                              int32_t nfreq, kp_image_data* out, double* req_t,
                                                              unsigned long long* waiting) {
    FreqLists fl;
    fl.nfreq = fl.nctrl = nfreq;
    fl.freqs = fl.ctrl = freqs;
    kp_grmhd_model g = *gm;
    g.cells = seq->cells[0];
    if (nonthermal(gm->emission_type)) return KP_UNSUPPORTED;
    return emul_run<true, M_SLOW>(pa, nullptr, 0.0, fl, out, nullptr, nullptr, nullptr, nullptr, &g, seq, req_t,
                                  waiting);
}

// Slow-light gradient lanes (one frequency, the model's); grad is [3][4][npix].
extern "C" kp_status kp_emul_run_grmhd_slow_light_gradient(const kp_pass_a_params* pa, const kp_grmhd_model* gm,
                                                           const kp_slow_light* seq, const double* log_scales,
                                                           kp_image_data* out, double* grad, double* req_t) {
    FreqLists fl;
    fl.nfreq = fl.nctrl = 1;
    fl.freqs = fl.ctrl = &gm->freq_cgs;
    kp_grmhd_model g = *gm;
    g.cells = seq->cells[0];
    return emul_run<false, M_STAN>(pa, nullptr, 0.0, fl, out, nullptr, nullptr, log_scales, grad, &g, seq, req_t,
                                   nullptr);
}

// The FP32 rescaled thermal coefficients of the mixed-precision lanes at one
// plasma state (model constants from kp_riaf_model m); out: jI jQ jU jV aI aQ
// aU aV rQ rU rV.
extern "C" kp_status kp_emul_thermal_coeffs_scaled(const kp_pass_a_params* pa, const kp_riaf_model* m, float nu,
                                                    float g, float ne, float thetae, float b, float sin_theta,
                                                    float cos_theta, float* out) {
    Params P;
    const kp_status st = prepare(pa, m, P);
    if (st != KP_OK) return st;
    prepare_mixed(P);
    Coeffs<float> c;
    thermal_coeffs_scaled(P.rm, nu, P.rm.bconst, g, ne, thetae, b, sin_theta, cos_theta, c);
    const float v[11] = {c.jI, c.jQ, c.jU, c.jV, c.aI, c.aQ, c.aU, c.aV, c.rQ, c.rU, c.rV};
    for (int i = 0; i < 11; ++i) out[i] = v[i];
    return KP_OK;
}

// Windowed slow-light lanes: pass A for every pixel, then the library's block
// loop (kp_slow_blocks.h plan) over HOST snapshots, each block running its
// work list (every pixel first, then the rays the block before it parked)
// one lane at a time through the kernel's tick, park and resume code.
template <int MODE>
static kp_status emul_slow_blocks(const kp_pass_a_params* pa, const kp_grmhd_model* gm, const kp_slow_light* slow,
                                  const FreqLists& fl, kp_image_data* out, double* req_t,
                                  unsigned long long* waiting, int32_t* nblocks, int32_t* nrewinds) {
    kp_grmhd_model g = *gm;
    g.cells = slow->cells[0];
    Params P;
    kp_status st = prepare_grmhd(pa, &g, P);
    if (st != KP_OK) return st;
    st = prepare_frequencies(gm->mbh_solar, fl, P);
    if (st != KP_OK) return st;
    unsigned long long counters[C_COUNT] = {};
    P.counters = counters;
    const int n = slow->nsnap, end = n - 1;
    const kp_grmhd_cell* const* cells = reinterpret_cast<const kp_grmhd_cell* const*>(slow->cells);
    P.slow_mode = slow->step_mode;
    P.slow_interp = slow->interpolation;
    P.t_obs = slow->observation_time;
    P.sup_lo = slow->times[0];
    P.sup_hi = slow->times[end];
    P.req_t = req_t;
    const int np = P.npix;
    const int nst = lds_stokes<MODE>() ? P.nfreq : 1;
    std::vector<double> ep(16 * (size_t)np), pkd(2 * (size_t)np), pkst(4 * (size_t)nst * np);
    std::vector<int32_t> valid(np), steps(np), reason(np), pki(5 * (size_t)np);
    std::vector<uint32_t> lists(2 * (size_t)np);
    uint32_t counts[2] = {0, 0};
    P.ep = ep.data();
    P.ep_valid = valid.data();
    P.ep_steps = steps.data();
    P.ep_reason = reason.data();
    P.pk_d = pkd.data();
    P.pk_i = pki.data();
    P.pk_st = pkst.data();
    P.out = *out;
    for (int pix = 0; pix < np; ++pix) {
        uint32_t rc = 0, sa = 0, guard = 0;
        const auto G = GeomT<geom_of<MODE>()>::geo(P, &rc);
        LaneA A;
        lane_a_init(P, pix, A);
        while (!pass_a_tick<GP_64, geom_of<MODE>()>(P, G, A, sa, guard)) {
        }
        lane_a_store(P, A);
    }
    double mfbuf[4 * KP_MAX_FREQUENCIES], mfbuf2[4 * KP_MAX_FREQUENCIES];
    const int wpb = std::max(1, slow->windows_per_block);
    const bool decoupled = slow->step_mode == KP_SLOW_LIGHT_DECOUPLED;
    kp_slow::BlockPlan plan;
    int parity = 0, blocks = 0, rewinds = 0;
    bool first = true;
    for (;;) {
        const int stop = kp_slow::block_stop(plan, wpb, n);
        Params Q = P;
        Q.snaps = cells + plan.res_lo;
        Q.snap_t = slow->times + plan.res_lo;
        Q.nsnap = stop - plan.res_lo + 1;
        Q.slow_final = stop == end ? 1 : 0;
        Q.list_in = first ? nullptr : lists.data() + (size_t)(1 - parity) * np;
        Q.count_in = first ? nullptr : &counts[1 - parity];
        Q.list_out = lists.data() + (size_t)parity * np;
        Q.count_out = &counts[parity];
        counts[parity] = 0;
        const uint32_t work = first ? (uint32_t)np : counts[1 - parity];
        for (uint32_t w = 0; w < work; ++w) {
            uint32_t rc = 0, c1 = 0, c2 = 0, sb = 0, guard = 0;
            const auto G = GeomT<geom_of<MODE>()>::geo(Q, &rc);
            LaneB B;
            B.mf = LdsVec<double>{mfbuf, 1};
            B.mf2 = LdsVec<double>{mfbuf2, 1};
            bool has = true;
            if (first) {
                has = lane_b_init<MODE>(Q, (int)w, B);
            } else {
                lane_b_resume<MODE>(Q, (int)Q.list_in[w], B);
            }
            while (has) {
                if (pass_b_tick<false, GP_64, MODE, true>(Q, G, B, c1, c2, sb, guard)) {
                    lane_b_finish<false, MODE>(Q, B, 1);
                    break;
                }
                if (B.phase == PH_PARK) {
                    lane_b_park<MODE>(Q, B);
                    Q.list_out[counts[parity]++] = (uint32_t)B.pixel;
                    break;
                }
            }
        }
        ++blocks;
        const uint32_t parked = counts[parity];
        if (parked == 0) break;
        if (stop == end && !decoupled) return KP_INVALID_ARGUMENT;
        if (blocks > 100000) return KP_INVALID_ARGUMENT;
        first = false;
        parity ^= 1;
        double earliest = slow->times[end];
        if (decoupled)
            for (uint32_t i = 0; i < parked; ++i)
                earliest = std::min(earliest, req_t[lists[(size_t)(1 - parity) * np + i]]);
        if (kp_slow::advance_plan(plan, stop, decoupled, earliest, slow->times, n)) ++rewinds;
    }
    if (waiting) *waiting = counters[C_WAIT];
    if (nblocks) *nblocks = blocks;
    if (nrewinds) *nrewinds = rewinds;
    return KP_OK;
}

// Slow-light lanes in blocks of seq->windows_per_block snapshot pairs.
extern "C" kp_status kp_emul_run_grmhd_slow_light_blocks(const kp_pass_a_params* pa, const kp_grmhd_model* gm,
                                                         const kp_slow_light* seq, const double* freqs,
                                                         int32_t nfreq, kp_image_data* out, double* req_t,
                                                         unsigned long long* waiting, int32_t* blocks,
                                                         int32_t* rewinds) {
    FreqLists fl;
    fl.nfreq = fl.nctrl = nfreq;
    fl.freqs = fl.ctrl = freqs;
    if (nonthermal(gm->emission_type))
        return emul_slow_blocks<M_SLOW_NT_W>(pa, gm, seq, fl, out, req_t, waiting, blocks, rewinds);
    return emul_slow_blocks<M_SLOW_W>(pa, gm, seq, fl, out, req_t, waiting, blocks, rewinds);
}

// The single-frequency pass B in two legs: each pixel's ray runs the
// radiating leg (M_SINGLE_V) until it parks at the vacuum boundary, then the
// vacuum leg (M_VAC) from its parked state, through the kernels' park and
// resume code.  mixed != 0 selects the mixed-precision lanes (FP64
// geodesics); legs = 0 runs the one-leg lanes (M_SINGLE) of the same
// precision instead.  *parked counts the rays that took the vacuum leg.
template <bool MIXED>
static kp_status emul_legs(const kp_pass_a_params* pa, const kp_riaf_model* m, bool legs, kp_image_data* out,
                           int64_t* parked) {
    constexpr int GP = GP_64;
    Params P;
    kp_status st = prepare(pa, m, P);
    if (st != KP_OK) return st;
    if (MIXED) prepare_mixed(P);
    unsigned long long counters[C_COUNT] = {};
    P.counters = counters;
    const int np = P.npix;
    std::vector<double> ep(16 * (size_t)np), pkd(2 * (size_t)np), pkst(4 * (size_t)np);
    std::vector<int32_t> valid(np), steps(np), reason(np), pki(5 * (size_t)np);
    P.ep = ep.data();
    P.ep_valid = valid.data();
    P.ep_steps = steps.data();
    P.ep_reason = reason.data();
    P.pk_d = pkd.data();
    P.pk_i = pki.data();
    P.pk_st = pkst.data();
    P.out = *out;
    int64_t nparked = 0;
    for (int pix = 0; pix < np; ++pix) {
        uint32_t rc = 0, sa = 0, c1 = 0, c2 = 0, sb = 0, guard = 0;
        const auto G = GeomT<G_BL>::geo(P, &rc);
        LaneA A;
        lane_a_init(P, pix, A);
        while (!pass_a_tick<GP, G_BL>(P, G, A, sa, guard)) {
        }
        lane_a_store(P, A);
        LaneB B;
        if (!legs) {
            if (lane_b_init<M_SINGLE>(P, pix, B)) {
                while (!pass_b_tick<MIXED, GP, M_SINGLE>(P, G, B, c1, c2, sb, guard)) {
                }
                lane_b_finish<MIXED, M_SINGLE>(P, B, 1);
            }
            continue;
        }
        if (!lane_b_init<M_SINGLE_V>(P, pix, B)) continue;
        bool vacuum = false;
        while (true) {
            if (pass_b_tick<MIXED, GP, M_SINGLE_V>(P, G, B, c1, c2, sb, guard)) {
                lane_b_finish<MIXED, M_SINGLE_V>(P, B, 1);
                break;
            }
            if (B.phase == PH_PARK) {
                lane_b_park<M_SINGLE_V>(P, B);
                vacuum = true;
                break;
            }
        }
        if (!vacuum) continue;
        ++nparked;
        LaneB V;
        lane_b_resume<M_VAC>(P, pix, V);
        while (!pass_b_tick<MIXED, GP, M_VAC>(P, G, V, c1, c2, sb, guard)) {
        }
        lane_b_finish<MIXED, M_VAC>(P, V, 1);
    }
    if (parked) *parked = nparked;
    return KP_OK;
}

extern "C" kp_status kp_emul_run_riaf_image_legs(const kp_pass_a_params* pa, const kp_riaf_model* m, int32_t mixed,
                                                 int32_t legs, kp_image_data* out, int64_t* parked) {
    return mixed ? emul_legs<true>(pa, m, legs != 0, out, parked) : emul_legs<false>(pa, m, legs != 0, out, parked);
}
