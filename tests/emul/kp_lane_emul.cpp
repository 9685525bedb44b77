// CPU build of the imaging kernel's per-lane state machines (test harness).
//
// Includes the product translation unit and drives its pass A / pass B lane
// functions one pixel at a time on the host, so the kernel's control logic can
// be checked against the oracle without a GPU.  Never part of the product.
#include "../../kpolaris_amd/csrc/kp_image.hip"

#include <vector>

extern "C" kp_status kp_emul_run_riaf_image(const kp_pass_a_params* pa, const kp_riaf_model* m,
                                            kp_image_data* out, int64_t* rhs_evals) {
    Params P;
    kp_status st = prepare(pa, m, P);
    if (st != KP_OK) return st;
    const int np = P.npix;
    std::vector<double> ep(16 * (size_t)np);
    std::vector<int32_t> valid(np), steps(np), reason(np);
    P.ep = ep.data();
    P.ep_valid = valid.data();
    P.ep_steps = steps.data();
    P.ep_reason = reason.data();
    P.out = *out;
    int64_t total = 0;
    for (int pix = 0; pix < np; ++pix) {
        uint32_t rc = 0, sa = 0, c1 = 0, c2 = 0, sb = 0;
        Geo<double> G{P.mass, P.spin, &rc};
        LaneA A;
        lane_a_init(P, pix, A);
        while (!pass_a_tick(P, G, A, sa)) {
        }
        lane_a_store(P, A);
        LaneB B;
        if (lane_b_init(P, pix, B)) {
            while (!pass_b_tick(P, G, B, c1, c2, sb)) {
            }
            lane_b_finish(P, B, 1);
        }
        total += rc;
    }
    if (rhs_evals) *rhs_evals = total;
    return KP_OK;
}

// The kernels' bounded-argument sine/cosine, for its accuracy test.
extern "C" void kp_emul_sincos(double x, double* s, double* c) { kpd::sincos_bounded(x, s, c); }
