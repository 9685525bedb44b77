// kp_image.hip -- host side of the imaging library: workspace, launch
// shapes, staged mixed-precision launches, batches and the C ABI
// (include/kp_image.h).  The kernels are in kp_kernels.h / kp_kernels.hip.
#include "kp_kernels.h"
#include "kp_prepare.h"
#include "kp_internal.h"
#include "kp_slow_blocks.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace kpd;

namespace {

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Library-kept workspaces, one per (device, stream): runs on different
// streams never share buffers, runs on one stream are ordered by it.
struct DeviceWorkspace {
    int device = -1;
    void* stream = nullptr;
    void* ptr = nullptr;
    size_t bytes = 0;
};

std::mutex g_ws_mutex;
std::vector<DeviceWorkspace> g_ws;

kp_status workspace_for(int device, void* stream, size_t bytes, void** out) {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    DeviceWorkspace* w = nullptr;
    for (auto& e : g_ws)
        if (e.device == device && e.stream == stream) w = &e;
    if (!w) {
        g_ws.push_back(DeviceWorkspace{device, stream, nullptr, 0});
        w = &g_ws.back();
    }
    if (w->bytes < bytes) {
        if (w->ptr) {
            // the previous run on this stream may still use the old buffer
            if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
                return kp_internal::fail(KP_HIP_ERROR, "hipStreamSynchronize failed");
            (void)hipFree(w->ptr);
            w->ptr = nullptr;
            w->bytes = 0;
        }
        if (hipMalloc(&w->ptr, bytes) != hipSuccess) {
            return kp_internal::fail(KP_OUT_OF_MEMORY, "hipMalloc of the image workspace failed");
        }
        w->bytes = bytes;
    }
    *out = w->ptr;
    return KP_OK;
}

constexpr size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Workspace layout of one image: pass-A endpoints [16][npix] doubles and
// valid/steps/reason [npix] int32, the queue and counter block, then the
// staged mixed-precision lists and parking buffers, the slow-light snapshot
// tables and the time-probe outputs when the run needs them.
struct WsLayout {
    size_t ep, valid, steps, reason, queue, counters, lists, pkd, pki, pkst, slow_t, slow_c, req, probe, preason;
    size_t total;
};

constexpr WsLayout workspace_layout(size_t np, size_t nst, bool staged, size_t nsnap, bool probe) {
    WsLayout L{};
    L.ep = 0;
    L.valid = align_up(L.ep + 16 * np * sizeof(double));
    L.steps = align_up(L.valid + np * sizeof(int32_t));
    L.reason = align_up(L.steps + np * sizeof(int32_t));
    L.queue = align_up(L.reason + np * sizeof(int32_t));
    L.counters = align_up(L.queue + KP_NQUEUE * sizeof(uint32_t));
    L.lists = align_up(L.counters + C_COUNT * sizeof(unsigned long long));
    L.pkd = align_up(L.lists + (staged ? 2 * np * sizeof(uint32_t) : 0));
    L.pki = align_up(L.pkd + (staged ? 2 * np * sizeof(double) : 0));
    L.pkst = align_up(L.pki + (staged ? 5 * np * sizeof(int32_t) : 0));
    L.slow_t = align_up(L.pkst + (staged ? 4 * nst * np * sizeof(double) : 0));
    L.slow_c = align_up(L.slow_t + nsnap * sizeof(double));
    L.req = align_up(L.slow_c + nsnap * sizeof(void*));
    L.probe = align_up(L.req + (nsnap ? np * sizeof(double) : 0));
    L.preason = align_up(L.probe + (probe ? 2 * np * sizeof(double) : 0));
    L.total = align_up(L.preason + (probe ? np * sizeof(int32_t) : 0));
    return L;
}

// ---------------------------------------------------------------------------
// Allocation-free images (kp_run_riaf_image_noalloc): the workspace of a
// single-frequency image of at most KP_NOALLOC_MAX_PIXELS pixels lives in this
// static device array, one copy per GPU, held until the process ends.  Host
// code never uses the array's name as an address: the address of each
// device's copy comes from hipGetSymbolAddress (scratch_for).
// ---------------------------------------------------------------------------
constexpr size_t KP_NOALLOC_BYTES = workspace_layout(KP_NOALLOC_MAX_PIXELS, 1, false, 0, false).total;
static __device__ __align__(256) unsigned char scratch[KP_NOALLOC_BYTES];

constexpr int KP_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch[KP_MAX_DEVICES] = {};

// The address of the current device's copy of `scratch` (device `device`),
// looked up once per device.
kp_status scratch_for(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= KP_MAX_DEVICES)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "device index beyond the allocation-free table");
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return kp_internal::fail(KP_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = ptr;
    }
    *out = g_scratch[device];
    return KP_OK;
}

// Staged mixed precision: guard_mode 2 on persistent split launches of the
// single- and multi-frequency RIAF images.
bool staged_run(const kp_run_options& opt, bool riaf) {
    return opt.precision == KP_PRECISION_MIXED && opt.persistent && opt.split_transport != 0 &&
           opt.guard_mode == 2 && opt.fp64_radius > 0.0 && riaf;
}

// Single-frequency Boyer-Lindquist pass B in two legs (radiating leg, then
// vacuum leg from the emitting region's edge to the camera): persistent split
// launches with an outer radius beyond every photon orbit (r > 4 M), where an
// outgoing ray never turns back.
bool vacuum_legs(const kp_run_options& opt, bool single_bl, double outer, double mass) {
    return single_bl && opt.persistent && opt.split_transport != 0 && !staged_run(opt, true) &&
           outer >= 5.0 * mass;
}

#define KP_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return kp_internal::fail(KP_HIP_ERROR, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// Kernel of a launch (kernels are instantiated in kp_kernels.hip); null
// when the combination is not built.
const void* pass_a_fn(bool persistent, int gp, int geom) { return kpsel::pass_a(persistent, gp, geom); }

// mixed: FP32 radiation (nu^3-scaled units); gp: geodesic precision.  Pass
// B kernels are persistent (a lane takes the next pixel from the queue).
const void* pass_b_fn(bool mixed, int gp, int mode) {
    switch (mode) {
    case M_SINGLE: return kpsel::pass_b_single(mixed, gp);
    case M_MF: return kpsel::pass_b_mf(mixed, gp);
    case M_GRMHD: return mixed ? kpsel::pass_b_grmhd_mixed(mode) : kpsel::pass_b_fmks(mode);
    case M_GRMHD_NT:
    case M_PROBE: return mixed ? nullptr : kpsel::pass_b_fmks(mode);
    case M_SLOW: return mixed ? kpsel::pass_b_grmhd_mixed(mode) : kpsel::pass_b_slow(mode);
    case M_SLOW_NT:
    case M_STAN: return mixed ? nullptr : kpsel::pass_b_slow(mode);
    default: return mixed ? nullptr : kpsel::pass_b_other(mode);
    }
}

const void* unsplit_fn(bool mixed, int gp, int mode) {
    const int b = base_mode(mode);
    if (mixed && (mode == M_GRMHD || mode == M_SLOW)) return kpsel::unsplit_grmhd_mixed(mode);
    if (b == M_GRMHD || b == M_SLOW || b == M_PROBE || mode == M_STAN) return mixed ? nullptr : kpsel::unsplit_fmks(mode);
    return kpsel::unsplit_bl(mixed, gp, mode);
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

// Observation times of one windowed slow-light run (outs[j] is job j's image).
struct SlowJobs {
    const kp_slow_light_job* jobs = nullptr;
    int32_t njobs = 0;
    kp_image_data* outs = nullptr;
};

}  // namespace

// grmhd != NULL selects the GRMHD grid model on FMKS geodesics (model unused).
static kp_status run_image(const kp_pass_a_params* params, const kp_riaf_model* model, const FreqLists& fl_in,
                           const kp_run_options* options, kp_image_data* out, kp_image_stats* stats,
                           int perturbation, double log_scale, const Tangent* tangent = nullptr,
                           const kp_grmhd_model* grmhd = nullptr, const kp_slow_light* slow = nullptr,
                           kp_slow_light_probe* probe = nullptr, bool noalloc = false,
                           kp_slow::Source* source = nullptr, kp_slow_light_schedule* schedule = nullptr,
                           const SlowJobs* jobs = nullptr) {
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
        if (mixed && probe) return kp_internal::fail(KP_UNSUPPORTED, "the slow-light time probe is built for FP64");
        if (perturbation || (tangent && !slow))
            return kp_internal::fail(KP_UNSUPPORTED, "GRMHD gradients are built for slow-light sequences");
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
        if (grmhd)
            prepare_mixed_grmhd(P, grmhd->mbh_solar);
        else
            prepare_mixed(P);
        P.fp64_radius = opt.fp64_radius;
        // every geodesic attempt of guard_mode 0 is FP64 and counts as a guard
        // step; FMKS geodesics are FP64 under every guard_mode
        if (opt.guard_mode == 0 || grmhd) P.stage = 0;
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
    // an image with an equatorial wedge runs the wedge kernels: persistent
    // split pass B (one ray per thread and unsplit launches give the same
    // results); the slow-light time probe is geometric, without emission
    // selection (run_slow_light_time_probe_metric, slow_light_driver.hpp:632-780)
    const bool sel = P.wedge_hr > 0.0 && !probe;
    P.persistent = opt.persistent || sel ? 1 : 0;
    int device = 0;
    KP_HIP(hipGetDevice(&device));
    const size_t np = (size_t)P.npix;
    // windowed slow light: pass B over blocks of snapshots from `source`
    // (always split and persistent: rays park between blocks)
    const bool windowed = slow && source && base_mode(mode) == M_SLOW;
    if (slow && source && !windowed)
        return kp_internal::fail(KP_UNSUPPORTED, "slow-light blocks are built for images, not gradients");
    const bool unsplit = opt.split_transport == 0 && !windowed && !sel;
    // single-frequency images: pass B in a radiating and a vacuum leg
    const bool legs = !noalloc && !sel && mode == M_SINGLE && vacuum_legs(opt, true, P.outer, P.mass);
    // Staged mixed precision (guard_mode 2, persistent split kernels): an
    // FP32-geodesic launch parks rays that reach r < fp64_radius, an FP64-geodesic launch
    // continues them and parks those back that reach r >= back_radius, then
    // FP32 again and a final FP64 launch; no wave mixes the two paths.
    // (a direct_only image locates its vertical turns in FP64: guard_mode 1 kernels)
    const bool staged = staged_run(opt, mode == M_SINGLE || mode == M_MF) && !P.direct_only && !sel;
    const int nst = mf ? P.nfreq : 1;
    const size_t nsnap = slow ? (size_t)slow->nsnap : 0;
    const WsLayout W = workspace_layout(np, (size_t)nst, staged || slow != nullptr || legs, nsnap, probe != nullptr);
    void* ws = nullptr;
    if (noalloc) {
        // the static scratch array: no staged lists, no timing events
        if (staged)
            return kp_internal::fail(KP_UNSUPPORTED, "the allocation-free entry point has no staged (guard_mode 2) path");
        if (stats || grmhd || slow || probe || tangent || fl.nfreq > 1)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "the allocation-free entry point images one RIAF frequency");
        if (np > KP_NOALLOC_MAX_PIXELS || W.total > KP_NOALLOC_BYTES) {
            char msg[160];
            std::snprintf(msg, sizeof(msg), "image of %llu pixels exceeds the allocation-free maximum of %d pixels",
                          (unsigned long long)np, (int)KP_NOALLOC_MAX_PIXELS);
            return kp_internal::fail(KP_INVALID_ARGUMENT, msg);
        }
        st = scratch_for(device, &ws);
        if (st != KP_OK) return st;
    } else if (opt.workspace) {
        if (opt.workspace_bytes < W.total) {
            char msg[160];
            std::snprintf(msg, sizeof(msg), "workspace of %llu bytes is smaller than the %llu bytes this run needs",
                          (unsigned long long)opt.workspace_bytes, (unsigned long long)W.total);
            return kp_internal::fail(KP_INVALID_ARGUMENT, msg);
        }
        ws = opt.workspace;
    } else {
        st = workspace_for(device, opt.stream, W.total, &ws);
        if (st != KP_OK) return st;
    }
    char* base = static_cast<char*>(ws);
    P.ep = reinterpret_cast<double*>(base + W.ep);
    P.ep_valid = reinterpret_cast<int32_t*>(base + W.valid);
    P.ep_steps = reinterpret_cast<int32_t*>(base + W.steps);
    P.ep_reason = reinterpret_cast<int32_t*>(base + W.reason);
    P.queue = reinterpret_cast<uint32_t*>(base + W.queue);
    P.counters = reinterpret_cast<unsigned long long*>(base + W.counters);
    uint32_t* lists = (staged || windowed || legs) ? reinterpret_cast<uint32_t*>(base + W.lists) : nullptr;
    if (staged || windowed || legs) {
        P.pk_d = reinterpret_cast<double*>(base + W.pkd);
        P.pk_i = reinterpret_cast<int32_t*>(base + W.pki);
        P.pk_st = reinterpret_cast<double*>(base + W.pkst);
    }
    if (staged) P.back_radius = KP_BACK_RADIUS_FACTOR * P.fp64_radius;

    hipStream_t stream = static_cast<hipStream_t>(opt.stream);
    if (probe) {
        P.probe = reinterpret_cast<double*>(base + W.probe);
        std::memset(&P.out, 0, sizeof(P.out));
        P.out.reason = reinterpret_cast<int32_t*>(base + W.preason);
    }
    if (slow) {
        P.snap_t = reinterpret_cast<const double*>(base + W.slow_t);
        P.snaps = reinterpret_cast<const kp_grmhd_cell* const*>(base + W.slow_c);
        P.req_t = reinterpret_cast<double*>(base + W.req);
        P.nsnap = slow->nsnap;
        P.slow_mode = slow->step_mode;
        P.slow_interp = slow->interpolation;
        P.t_obs = slow->observation_time;
        P.slow_final = 1;
        P.sup_lo = slow->times[0];
        P.sup_hi = slow->times[slow->nsnap - 1];
        if (!windowed) {
            KP_HIP(hipMemcpyAsync(base + W.slow_t, slow->times, nsnap * sizeof(double), hipMemcpyHostToDevice, stream));
            KP_HIP(hipMemcpyAsync(base + W.slow_c, slow->cells, nsnap * sizeof(void*), hipMemcpyHostToDevice, stream));
        }
    }
    // Observation times of the windowed block loop, each with its own image
    // and its own ray state between blocks: the pass-A endpoints it starts
    // from (pass B parks rays there), parked rays, their lists and queue
    // counters.  One observation time uses the workspace; several get a
    // workspace-sized region each, and share the workspace's pass-A results.
    struct SlowJob {
        double t_obs = 0.0;
        int first = 0, last = 0;  // inclusive snapshot support
        kp_image_data out{};
        double* ep = nullptr;
        double* pk_d = nullptr;
        int32_t* pk_i = nullptr;
        double* pk_st = nullptr;
        double* req_t = nullptr;
        uint32_t* lists = nullptr;  // [2][npix]
        uint32_t* queue = nullptr;  // fetch counter, list counts [2]
        bool started = false, done = false;
        int parity = 0;
        uint32_t parked = 0;
    };
    std::vector<SlowJob> slow_jobs;
    struct JobMemory {
        void* p = nullptr;
        ~JobMemory() {
            if (p) (void)hipFree(p);
        }
    } job_mem;
    if (jobs && !windowed) return kp_internal::fail(KP_INVALID_ARGUMENT, "observation-time jobs need the block loop");
    if (windowed && !jobs) {
        SlowJob J;
        J.t_obs = slow->observation_time;
        J.first = 0;
        J.last = slow->nsnap - 1;
        J.out = P.out;
        J.ep = P.ep;
        J.pk_d = P.pk_d;
        J.pk_i = P.pk_i;
        J.pk_st = P.pk_st;
        J.req_t = P.req_t;
        J.lists = lists;
        J.queue = P.queue + 2;
        slow_jobs.push_back(J);
    } else if (windowed) {
        const size_t per = W.total;
        if (hipMalloc(&job_mem.p, per * (size_t)jobs->njobs) != hipSuccess) {
            job_mem.p = nullptr;
            (void)hipGetLastError();
            char msg[160];
            std::snprintf(msg, sizeof(msg), "hipMalloc of the state of %d observation times (%llu bytes) failed",
                          (int)jobs->njobs, (unsigned long long)(per * (size_t)jobs->njobs));
            return kp_internal::fail(KP_OUT_OF_MEMORY, msg);
        }
        for (int j = 0; j < jobs->njobs; ++j) {
            char* b = static_cast<char*>(job_mem.p) + per * (size_t)j;
            SlowJob J;
            J.t_obs = jobs->jobs[j].observation_time;
            J.first = jobs->jobs[j].first;
            J.last = jobs->jobs[j].last;
            J.out = jobs->outs[j];
            J.ep = reinterpret_cast<double*>(b + W.ep);
            J.pk_d = reinterpret_cast<double*>(b + W.pkd);
            J.pk_i = reinterpret_cast<int32_t*>(b + W.pki);
            J.pk_st = reinterpret_cast<double*>(b + W.pkst);
            J.req_t = reinterpret_cast<double*>(b + W.req);
            J.lists = reinterpret_cast<uint32_t*>(b + W.lists);
            J.queue = reinterpret_cast<uint32_t*>(b + W.queue);
            slow_jobs.push_back(J);
        }
    }
    const int block = opt.block_size > 0 ? opt.block_size : 128;
    if (block % 64 != 0 || block > 256)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "block_size must be a multiple of 64 and at most 256");
    int cus = 0;
    st = cu_count(device, cus);
    if (st != KP_OK) return st;
    const size_t smem_a = lds_bytes(block, 0);
    // the single-pass frequency loop serves the FP64 multi-frequency, GRMHD
    // and slow-light lanes and the mixed-precision GRMHD and slow-light lanes
    if ((mixed && mode != M_GRMHD && mode != M_SLOW) ||
        !(base_mode(mode) == M_MF || base_mode(mode) == M_GRMHD || base_mode(mode) == M_SLOW))
        P.single_pass = 0;
    const size_t smem_b = mode == M_STAN ? lds_bytes(block, 0) : lds_bytes(block, mf ? P.nfreq : 0, P.single_pass ? 2 : 1);
    const int nblocks_flat = (int)((np + block - 1) / block);
    // grid of one launch: one block per pixel block, or (persistent) the
    // resident blocks of the kernel on every CU
    auto grid_of = [&](const void* fn, size_t smem, int& grid, bool persistent) -> kp_status {
        grid = nblocks_flat;
        if (!persistent) return KP_OK;
        int per_cu = 0;
        const kp_status s2 = blocks_per_cu(device, fn, block, smem, per_cu);
        if (s2 != KP_OK) return s2;
        if (opt.waves_per_cu > 0) per_cu = std::max(1, opt.waves_per_cu / (block / 64));
        grid = std::max(1, std::min(nblocks_flat, cus * per_cu));
        return KP_OK;
    };
    // timing events, destroyed on every return path
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } events;
    hipEvent_t* ev = events.e;
    if (stats) {
        for (int i = 0; i < 4; ++i) KP_HIP(hipEventCreate(&ev[i]));
    }
    {
        int nq = KP_NQUEUE;
        void* zargs[] = {&P.queue, &nq, &P.counters};
        KP_HIP(hipLaunchKernel(kpsel::zero_counters(), dim3(1), dim3(64), zargs, 0, stream));
    }
    KP_HIP(hipGetLastError());
    if (stats) KP_HIP(hipEventRecord(ev[0], stream));
    auto launch = [&](const void* fn, size_t smem, Params& Q) -> kp_status {
        if (!fn) return kp_internal::fail(KP_UNSUPPORTED, "no kernel is built for this combination of options");
        int grid = 0;
        const kp_status s2 = grid_of(fn, smem, grid, Q.persistent != 0);
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
            const void* fn = pass_b ? pass_b_fn(true, gp, mode) : pass_a_fn(true, gp, G_BL);
            const kp_status s2 = launch(fn, pass_b ? smem_b : smem_a, Q);
            if (s2 != KP_OK) return s2;
        }
        return KP_OK;
    };
    // guard_mode 0: FP64 geodesics with the FP32 radiation path
    const int gp_plain = mixed && !grmhd ? (opt.guard_mode == 0 ? GP_64 : GP_GUARD) : GP_64;
    if (P.npix > 0 && unsplit) {
        Params Q = P;
        Q.fetch = P.queue;
        st = launch(unsplit_fn(mixed, gp_plain, mode), smem_b, Q);
        if (st != KP_OK) return st;
    }
    if (P.npix > 0 && !unsplit) {
        if (staged) {
            st = run_staged(false, 2);
        } else {
            Params Q = P;
            Q.fetch = P.queue;
            const void* fn = pass_a_fn(P.persistent != 0, gp_plain, geom_of_mode(mode));
            st = launch(fn, smem_a, Q);
        }
        if (st != KP_OK) return st;
    }
    if (stats) KP_HIP(hipEventRecord(ev[1], stream));
    // Windowed slow light: the block loop of run_slow_light_batch_metric_impl
    // (slow_light_batch_driver.hpp:211-340).  Block b holds snapshots
    // res_lo..stop (stop = start + windows_per_block, the halo kept from the
    // block before it for decoupled stepping) and runs, for every job whose
    // support overlaps it (:237-251), the rays that job's last launch parked
    // (every pixel at its first launch) on the snapshots of the block inside
    // its support; a job whose launch parks nothing is done.  A decoupled
    // ray that needs an earlier time than its block holds makes the next
    // block start at the snapshot of the earliest such time over the jobs
    // (:286-330).
    double blocks_ms = 0.0;
    auto run_blocks = [&]() -> kp_status {
        const int n = slow->nsnap;
        int begin = n - 1, end = 0;
        for (const SlowJob& J : slow_jobs) {
            begin = std::min(begin, J.first);
            end = std::max(end, J.last);
        }
        const int wpb = std::max(1, slow->windows_per_block);
        const bool decoupled = slow->step_mode == KP_SLOW_LIGHT_DECOUPLED;
        struct BlockEvents {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~BlockEvents() {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } bev;
        for (int i = 0; i < 2; ++i) KP_HIP(hipEventCreate(&bev.e[i]));
        // every job starts from the pass-A endpoints (its own copy: pass B
        // parks rays over them)
        for (SlowJob& J : slow_jobs)
            if (J.ep != P.ep)
                KP_HIP(hipMemcpyAsync(J.ep, P.ep, 16 * np * sizeof(double), hipMemcpyDeviceToDevice, stream));
        std::vector<const void*> ptrs;
        std::vector<uint32_t> parked_list;
        std::vector<double> req;
        std::vector<SlowJob*> ran;
        kp_slow::BlockPlan plan;
        plan.start = plan.res_lo = begin;
        int launches = 0, rewinds = 0;
        size_t remaining = slow_jobs.size();
        const double* const times_dev = reinterpret_cast<const double*>(base + W.slow_t);
        const kp_grmhd_cell* const* const cells_dev = reinterpret_cast<const kp_grmhd_cell* const*>(base + W.slow_c);
        const void* wfn = sel ? kpsel::pass_b_sel(mixed, GP_64, mode == M_SLOW_NT ? M_SLOW_NT_W : M_SLOW_W)
                              : (mixed ? kpsel::pass_b_grmhd_mixed(M_SLOW_W) : kpsel::pass_b_slow_windowed(mode));
        // every block advances some ray by a step or finishes it, so the
        // loop ends; the bound only turns a logic error into an error
        const long long max_blocks =
            (long long)(n + 1) * (long long)std::max(P.max_steps, 1) * (long long)slow_jobs.size() + 16;
        for (long long iter = 0;; ++iter) {
            if (iter > max_blocks) return kp_internal::fail(KP_INVALID_ARGUMENT, "slow-light block loop did not end");
            const int stop = kp_slow::block_stop(plan, wpb, end + 1);
            const int res_lo = plan.res_lo;
            kp_status s2 = source->acquire(res_lo, stop, stream, ptrs);
            if (s2 != KP_OK) return s2;
            const int cnt = stop - res_lo + 1;
            KP_HIP(hipMemcpyAsync(base + W.slow_t, slow->times + res_lo, (size_t)cnt * sizeof(double),
                                  hipMemcpyHostToDevice, stream));
            KP_HIP(hipMemcpyAsync(base + W.slow_c, ptrs.data(), (size_t)cnt * sizeof(void*), hipMemcpyHostToDevice,
                                  stream));
            KP_HIP(hipEventRecord(bev.e[0], stream));
            ran.clear();
            for (SlowJob& J : slow_jobs) {
                if (J.done || stop <= J.first || res_lo >= J.last) continue;
                // the block's snapshots inside the job's support: lo < hi
                const int lo = std::max(res_lo, J.first), hi = std::min(stop, J.last);
                Params Q = P;
                Q.out = J.out;
                Q.ep = J.ep;
                Q.pk_d = J.pk_d;
                Q.pk_i = J.pk_i;
                Q.pk_st = J.pk_st;
                Q.req_t = J.req_t;
                Q.t_obs = J.t_obs;
                Q.sup_lo = slow->times[J.first];
                Q.sup_hi = slow->times[J.last];
                Q.snap_t = times_dev + (lo - res_lo);
                Q.snaps = cells_dev + (lo - res_lo);
                Q.nsnap = hi - lo + 1;
                Q.slow_final = hi == J.last ? 1 : 0;
                Q.fetch = J.queue;
                Q.list_in = J.started ? J.lists + (size_t)(1 - J.parity) * np : nullptr;
                Q.count_in = J.started ? J.queue + 1 + (1 - J.parity) : nullptr;
                Q.list_out = J.lists + (size_t)J.parity * np;
                Q.count_out = J.queue + 1 + J.parity;
                KP_HIP(hipMemsetAsync(Q.fetch, 0, sizeof(uint32_t), stream));
                KP_HIP(hipMemsetAsync(Q.count_out, 0, sizeof(uint32_t), stream));
                s2 = launch(wfn, smem_b, Q);
                if (s2 != KP_OK) return s2;
                KP_HIP(hipMemcpyAsync(&J.parked, Q.count_out, sizeof(J.parked), hipMemcpyDeviceToHost, stream));
                ran.push_back(&J);
                ++launches;
            }
            KP_HIP(hipEventRecord(bev.e[1], stream));
            // the next block's snapshots upload while this one runs
            if (stop < end) {
                s2 = source->prefetch(stop + 1, stop + std::min(wpb, end - stop));
                if (s2 != KP_OK) return s2;
            }
            KP_HIP(hipStreamSynchronize(stream));
            float ms = 0.f;
            KP_HIP(hipEventElapsedTime(&ms, bev.e[0], bev.e[1]));
            blocks_ms += ms;
            double earliest = slow->times[end];
            for (SlowJob* J : ran) {
                J->started = true;
                if (J->parked == 0) {
                    J->done = true;
                    --remaining;
                    continue;
                }
                if (stop >= J->last && !decoupled)
                    return kp_internal::fail(KP_INVALID_ARGUMENT, "slow-light rays parked after the last block");
                J->parity ^= 1;
                if (decoupled) {
                    // slow_light_pending_range (slow_light_batch_driver.hpp:81-103)
                    parked_list.resize(J->parked);
                    req.resize(np);
                    KP_HIP(hipMemcpy(parked_list.data(), J->lists + (size_t)(1 - J->parity) * np,
                                     J->parked * sizeof(uint32_t), hipMemcpyDeviceToHost));
                    KP_HIP(hipMemcpy(req.data(), J->req_t, np * sizeof(double), hipMemcpyDeviceToHost));
                    for (uint32_t pix : parked_list) earliest = std::min(earliest, req[pix]);
                }
            }
            if (remaining == 0) break;
            if (kp_slow::advance_plan(plan, stop, decoupled, earliest, slow->times, end + 1)) ++rewinds;
        }
        if (schedule) {
            schedule->blocks = launches;
            schedule->rewinds = rewinds;
            schedule->pass_b_ms = blocks_ms;
        }
        return KP_OK;
    };
    if (P.npix > 0 && windowed) {
        st = run_blocks();
        if (st != KP_OK) return st;
    } else if (P.npix > 0 && !unsplit) {
        if (staged) {
            st = run_staged(true, 12);
        } else if (legs) {
            // radiating leg: rays that leave the emitting region outward park
            // (queue slot 2 counts them); the vacuum leg takes that list
            Params Q = P;
            Q.fetch = P.queue + 1;
            Q.list_out = lists;
            Q.count_out = P.queue + 2;
            st = launch(kpsel::pass_b_legs(mixed, gp_plain, false), smem_b, Q);
            if (st == KP_OK && stats) KP_HIP(hipEventRecord(ev[3], stream));
            if (st == KP_OK) {
                Params V = P;
                V.fetch = P.queue + 3;
                V.list_in = lists;
                V.count_in = P.queue + 2;
                st = launch(kpsel::pass_b_legs(mixed, gp_plain, true), smem_b, V);
            }
        } else {
            // pass B always runs the persistent kernels (one ray per thread
            // gives the same results; its multi-frequency FP64 instantiation
            // computed wrong images, a fault not found in the source)
            Params Q = P;
            Q.fetch = P.queue + 1;
            Q.persistent = 1;
            const void* fn = sel ? kpsel::pass_b_sel(mixed, gp_plain, mode) : pass_b_fn(mixed, gp_plain, mode);
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
        // windowed: the block launches without the host's waits between them
        if (windowed) b_ms = (float)blocks_ms;
        float v_ms = 0.f;
        if (legs && P.npix > 0) {
            float r_ms = 0.f;
            KP_HIP(hipEventElapsedTime(&r_ms, ev[1], ev[3]));
            v_ms = b_ms - r_ms;
        }
        std::memset(stats, 0, sizeof(*stats));
        stats->pass_a_ms = a_ms;
        stats->pass_b_ms = b_ms;
        stats->pass_b_vacuum_ms = v_ms;
        stats->kernel_ms = double(a_ms) + double(b_ms);
        stats->rays = P.npix * (int64_t)std::max<size_t>(1, slow_jobs.size());
        stats->rhs_evals = (int64_t)(cnt[C_RHS_A] + cnt[C_RHS_B] + cnt[C_RHS_V]);
        stats->vacuum_rhs_evals = (int64_t)cnt[C_RHS_V];
        stats->pass_a_rhs_evals = (int64_t)cnt[C_RHS_A];
        stats->coefficient_evals = (int64_t)cnt[C_COEFF];
        stats->transfer_steps = (int64_t)cnt[C_TRANSFER];
        stats->pass_a_steps = (int64_t)cnt[C_STEPS_A];
        stats->pass_b_steps = (int64_t)cnt[C_STEPS_B];
        stats->fp64_guard_steps = (int64_t)cnt[C_GUARD];
        stats->vacuum_steps = (int64_t)cnt[C_VACUUM];
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
            // the first job (observation time) with such rays names its support
            std::vector<double> req(np);
            const int njob = std::max<int>(1, (int)slow_jobs.size());
            for (int j = 0; j < njob; ++j) {
                const bool one = slow_jobs.empty();
                KP_HIP(hipMemcpy(req.data(), one ? P.req_t : slow_jobs[j].req_t, np * sizeof(double),
                                 hipMemcpyDeviceToHost));
                double lo = 1e300, hi = -1e300;
                for (double t : req)
                    if (t == t) {
                        lo = std::min(lo, t);
                        hi = std::max(hi, t);
                    }
                if (!(lo <= hi)) continue;
                const int first = one ? 0 : slow_jobs[j].first;
                const int last = one ? slow->nsnap - 1 : slow_jobs[j].last;
                char msg[300];
                std::snprintf(msg, sizeof(msg),
                              "decoupled slow-light radiation sample outside supplied physical time support: %llu rays "
                              "requested %.17g..%.17g, available %.17g..%.17g (observation time %.17g)",
                              waiting, lo, hi, slow->times[first], slow->times[last],
                              one ? slow->observation_time : slow_jobs[j].t_obs);
                return kp_internal::fail(KP_TIME_SUPPORT, msg);
            }
            return kp_internal::fail(KP_TIME_SUPPORT,
                                     "decoupled slow-light radiation sample outside supplied physical time support");
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
    bool sel = false;  // some job has an equatorial wedge: the batch runs the wedge kernel
    int gp = GP_64;  // geodesic precision of the jobs (guard_mode 0: GP_64, 1: GP_GUARD)
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
    zp.func = const_cast<void*>(kpsel::zero_counters());
    zp.gridDim = dim3(1);
    zp.blockDim = dim3(64);
    zp.kernelParams = zargs;
    hipKernelNodeParams ap{};
    ap.func = const_cast<void*>(kpsel::pass_a_batch(b->gp));
    ap.gridDim = dim3(b->grid_a);
    ap.blockDim = dim3(b->block);
    ap.sharedMemBytes = (unsigned)lds_bytes(b->block, 0);
    ap.kernelParams = bargs;
    hipKernelNodeParams bp = ap;
    bp.func = const_cast<void*>(b->sel ? kpsel::pass_b_batch_sel(b->mixed, b->gp) : kpsel::pass_b_batch(b->mixed, b->gp));
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
    if (opt.guard_mode < 0 || opt.guard_mode > 2)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "guard_mode must be 0, 1 or 2");
    if (mixed && opt.guard_mode == 2)
        return kp_internal::fail(KP_UNSUPPORTED, "staged mixed precision (guard_mode 2) runs single images only");
    // the geodesic precision of a single image with the same options
    const int gp = mixed && opt.guard_mode == 1 ? GP_GUARD : GP_64;
    const int block = opt.block_size > 0 ? opt.block_size : 128;
    if (block % 64 != 0 || block > 256)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "block_size must be a multiple of 64 and at most 256");
    // validate every job before any allocation (batch_driver.hpp:137-146)
    std::vector<Params> table((size_t)njobs);
    std::vector<uint32_t> start((size_t)njobs + 1, 0);
    uint64_t total = 0;
    bool sel = false;
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
            // every geodesic attempt of guard_mode 0 is FP64 and counts as a guard step
            if (opt.guard_mode == 0) table[j].stage = 0;
        }
        table[j].out = jobs[j].out;
        table[j].persistent = 1;
        sel = sel || table[j].wedge_hr > 0.0;
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
    b->sel = sel;
    b->gp = gp;
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
                &ba, kpsel::pass_a_batch(b->gp), block,
                lds_bytes(block, 0)) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &bb, sel ? kpsel::pass_b_batch_sel(mixed, b->gp) : kpsel::pass_b_batch(mixed, b->gp), block,
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
    stats->vacuum_steps = (int64_t)cnt[C_VACUUM];
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

extern "C" kp_status kp_image_workspace_size(int64_t npix, int32_t nfreq, int32_t nsnap,
                                             const kp_run_options* options, uint64_t* bytes) {
    kp_internal::clear_error();
    if (!bytes) return kp_internal::fail(KP_INVALID_ARGUMENT, "null result");
    if (npix < 0 || npix > (int64_t(1) << 30) || nfreq < 0 || nfreq > KP_MAX_FREQUENCIES || nsnap < 0)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "npix, nfreq or nsnap out of range");
    kp_run_options opt;
    kp_run_options_default(&opt);
    if (options) opt = *options;
    // park buffers: staged mixed precision, slow-light blocks, and the two
    // pass-B legs of single-frequency images (whose outer radius the query
    // does not know)
    const bool parks = staged_run(opt, nsnap == 0) || nsnap > 0 ||
                       (nsnap == 0 && nfreq <= 1 && vacuum_legs(opt, true, 1.0, 0.0));
    const WsLayout W = workspace_layout((size_t)npix, (size_t)std::max(nfreq, 1), parks, (size_t)nsnap, false);
    *bytes = (uint64_t)W.total;
    return KP_OK;
}

extern "C" void kp_batch_limits_default(kp_batch_limits* limits) {
    if (!limits) return;
    limits->max_in_flight = 4;
    limits->max_device_output_bytes = (int64_t)256 * 1024 * 1024;
}

// The 15 real and 3 integer per-pixel output fields of an image, in the
// order of kp_image_data.
static double** real_field(kp_image_data& d, int i) {
    double** f[15] = {&d.image_i, &d.image_q, &d.image_u, &d.image_v, &d.closure_x, &d.closure_k, &d.final_null,
                      &d.frame_error, &d.det_r, &d.overlap_r11, &d.overlap_r12, &d.overlap_r21, &d.overlap_r22,
                      &d.basis_identity_error, &d.basis_rotation_angle};
    return f[i];
}
static int32_t** int_field(kp_image_data& d, int i) {
    int32_t** f[3] = {&d.pass_a_steps, &d.steps, &d.reason};
    return f[i];
}

extern "C" kp_status kp_run_riaf_batch_host(const kp_batch_job* jobs, int32_t njobs, const kp_batch_limits* limits,
                                            const kp_run_options* options, kp_batch_timing* timing) {
    using clock = std::chrono::steady_clock;
    kp_internal::clear_error();
    kp_batch_limits lim;
    kp_batch_limits_default(&lim);
    if (limits) lim = *limits;
    if (lim.max_in_flight <= 0 || lim.max_device_output_bytes <= 0)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "batch limits must be positive");
    if (njobs < 0 || (njobs > 0 && !jobs)) return kp_internal::fail(KP_INVALID_ARGUMENT, "invalid job list");
    if (timing) std::memset(timing, 0, sizeof(*timing));
    // validate every job before launching any work (batch_driver.hpp:137-146)
    const size_t budget = (size_t)lim.max_device_output_bytes;
    std::vector<size_t> pixels((size_t)njobs);
    for (int j = 0; j < njobs; ++j) {
        const int nx = jobs[j].params.camera.nx, ny = jobs[j].params.camera.ny;
        if (nx <= 0 || ny <= 0 || nx > INT_MAX / ny)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "batch image dimensions must have a positive int-sized product");
        pixels[j] = (size_t)nx * (size_t)ny;
        if (pixels[j] > budget / KP_BATCH_BYTES_PER_PIXEL)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "one batch image exceeds the device output budget");
        if (jobs[j].radiation_substeps < 1)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "batch radiation_substeps must be positive");
    }
    if (njobs == 0) return KP_OK;
    kp_run_options opt;
    kp_run_options_default(&opt);
    if (options) opt = *options;
    hipStream_t stream = static_cast<hipStream_t>(opt.stream);
    for (int begin = 0; begin < njobs;) {
        auto t0 = clock::now();
        int count = 0;
        size_t bytes = 0;
        while (count < lim.max_in_flight && begin + count < njobs) {
            const size_t next = pixels[begin + count] * KP_BATCH_BYTES_PER_PIXEL;
            if (next > budget - bytes) break;
            bytes += next;
            ++count;
        }
        void* dmem = nullptr;
        if (hipMalloc(&dmem, bytes) != hipSuccess)
            return kp_internal::fail(KP_OUT_OF_MEMORY, "hipMalloc of the batch window outputs failed");
        std::vector<kp_batch_job> w(jobs + begin, jobs + begin + count);
        char* cur = static_cast<char*>(dmem);
        for (int i = 0; i < count; ++i) {
            const size_t np = pixels[begin + i];
            kp_image_data dev;
            std::memset(&dev, 0, sizeof(dev));
            for (int f = 0; f < 15; ++f) {
                *real_field(dev, f) = reinterpret_cast<double*>(cur);
                cur += np * sizeof(double);
            }
            for (int f = 0; f < 3; ++f) {
                *int_field(dev, f) = reinterpret_cast<int32_t*>(cur);
                cur += np * sizeof(int32_t);
            }
            w[i].out = dev;
        }
        auto t1 = clock::now();
        if (timing) {
            timing->allocate_seconds += std::chrono::duration<double>(t1 - t0).count();
            timing->windows += 1;
            timing->peak_device_output_bytes = std::max<int64_t>(timing->peak_device_output_bytes, (int64_t)bytes);
        }
        kp_batch* b = nullptr;
        kp_status st = kp_batch_create(w.data(), count, &opt, &b);
        if (st == KP_OK) st = kp_batch_launch(b, stream);
        if (st == KP_OK && hipStreamSynchronize(stream) != hipSuccess)
            st = kp_internal::fail(KP_HIP_ERROR, "batch window did not complete");
        auto t2 = clock::now();
        for (int i = 0; i < count && st == KP_OK; ++i) {
            kp_image_data host = jobs[begin + i].out;
            const size_t np = pixels[begin + i];
            for (int f = 0; f < 15 && st == KP_OK; ++f)
                if (*real_field(host, f) &&
                    hipMemcpy(*real_field(host, f), *real_field(w[i].out, f), np * sizeof(double),
                              hipMemcpyDeviceToHost) != hipSuccess)
                    st = kp_internal::fail(KP_HIP_ERROR, "device to host copy failed");
            for (int f = 0; f < 3 && st == KP_OK; ++f)
                if (*int_field(host, f) &&
                    hipMemcpy(*int_field(host, f), *int_field(w[i].out, f), np * sizeof(int32_t),
                              hipMemcpyDeviceToHost) != hipSuccess)
                    st = kp_internal::fail(KP_HIP_ERROR, "device to host copy failed");
        }
        if (b) kp_batch_destroy(b);
        (void)hipFree(dmem);
        if (st != KP_OK) return st;
        if (timing) {
            timing->kernel_seconds += std::chrono::duration<double>(t2 - t1).count();
            timing->copy_seconds += std::chrono::duration<double>(clock::now() - t2).count();
        }
        begin += count;
    }
    return KP_OK;
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

extern "C" kp_status kp_run_riaf_image_noalloc(const kp_pass_a_params* params, const kp_riaf_model* model,
                                               int32_t radiation_substeps, const kp_run_options* options,
                                               kp_image_data* out) {
    (void)radiation_substeps;
    if (options && options->workspace)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "the allocation-free entry point uses its own scratch array");
    return run_image(params, model, FreqLists{}, options, out, nullptr, 0, 0.0, nullptr, nullptr, nullptr, nullptr,
                     true);
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

// The resident runs: one observation time (jobs null) or several.
static kp_status run_resident(const kp_pass_a_params* params, const kp_grmhd_model* model,
                              const kp_slow_light* sequence, const SlowJobs* jobs, const double* frequencies,
                              int32_t nfreq, const kp_run_options* options, kp_image_data* out,
                              kp_image_stats* stats) {
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
    if (!jobs && !std::isfinite(q.observation_time))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "observation time must be finite");
    if (jobs) {
        // slow_light_batch_driver.hpp:116, 144-146
        if (jobs->njobs < 1 || !jobs->jobs || !jobs->outs)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "no slow-light observation-time jobs");
        for (int j = 0; j < jobs->njobs; ++j) {
            const kp_slow_light_job& J = jobs->jobs[j];
            if (!std::isfinite(J.observation_time) || J.first < 0 || J.first >= J.last || J.last >= q.nsnap)
                return kp_internal::fail(KP_INVALID_ARGUMENT, "invalid slow-light batch observer time or snapshot support");
        }
    }
    for (int k = 0; k < q.nsnap; ++k) {
        if (!q.cells[k]) return kp_internal::fail(KP_INVALID_ARGUMENT, "null snapshot cells");
        if (!std::isfinite(q.times[k]) || (k > 0 && !(q.times[k] > q.times[k - 1])))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "snapshot times must be finite and strictly increasing");
    }
    if (nfreq < 0 || (nfreq > 0 && !frequencies))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "frequency list and count disagree");
    if (q.windows_per_block < 0) return kp_internal::fail(KP_INVALID_ARGUMENT, "windows_per_block must be non-negative");
    kp_grmhd_model m = *model;
    m.cells = q.cells[0];
    FreqLists fl;
    if (nfreq > 0) {
        fl.nfreq = fl.nctrl = nfreq;
        fl.freqs = fl.ctrl = frequencies;
    }
    // blocks shorter than the sequence run the block scheduler over the
    // resident snapshots
    kp_slow::Resident resident(q.cells);
    if (jobs) {
        // jobs run in the block loop: windows_per_block 0 is one block
        kp_slow_light seq = q;
        if (seq.windows_per_block == 0 || seq.windows_per_block > q.nsnap - 1) seq.windows_per_block = q.nsnap - 1;
        return run_image(params, nullptr, fl, options, jobs->outs, stats, 0, 0.0, nullptr, &m, &seq, nullptr, false,
                         &resident, nullptr, jobs);
    }
    kp_slow::Source* source = (q.windows_per_block > 0 && q.windows_per_block < q.nsnap - 1) ? &resident : nullptr;
    return run_image(params, nullptr, fl, options, out, stats, 0, 0.0, nullptr, &m, sequence, nullptr, false,
                     source);
}

extern "C" kp_status kp_run_grmhd_slow_light_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                    const kp_slow_light* sequence, const double* frequencies,
                                                    int32_t nfreq, const kp_run_options* options,
                                                    kp_image_data* out, kp_image_stats* stats) {
    return run_resident(params, model, sequence, nullptr, frequencies, nfreq, options, out, stats);
}

extern "C" kp_status kp_run_grmhd_slow_light_image_jobs(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                        const kp_slow_light* sequence, const kp_slow_light_job* jobs,
                                                        int32_t njobs, const double* frequencies, int32_t nfreq,
                                                        const kp_run_options* options, kp_image_data* outs,
                                                        kp_image_stats* stats) {
    kp_internal::clear_error();
    if (njobs < 1 || !jobs || !outs)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "no slow-light observation-time jobs");
    SlowJobs J;
    J.jobs = jobs;
    J.njobs = njobs;
    J.outs = outs;
    return run_resident(params, model, sequence, &J, frequencies, nfreq, options, nullptr, stats);
}

// The streamed runs: one observation time (jobs null) or several.
static kp_status run_stream(const kp_pass_a_params* params, const kp_grmhd_model* model,
                            const kp_slow_light_stream* sequence, const SlowJobs* jobs, const double* frequencies,
                            int32_t nfreq, const kp_run_options* options, kp_image_data* out, kp_image_stats* stats,
                            kp_slow_light_schedule* schedule) {
    kp_internal::clear_error();
    if (!model || !sequence) return kp_internal::fail(KP_INVALID_ARGUMENT, "null GRMHD model or sequence");
    const kp_slow_light_stream& q = *sequence;
    // the batch scheduler's argument checks (slow_light_batch_driver.hpp:133-148)
    if (q.nsnap < 2) return kp_internal::fail(KP_INVALID_ARGUMENT, "slow light requires at least two snapshots");
    if (!q.times || !q.load) return kp_internal::fail(KP_INVALID_ARGUMENT, "null snapshot times or loader");
    if (q.windows_per_block < 1 || q.prefetch_snapshots < 0)
        return kp_internal::fail(KP_INVALID_ARGUMENT,
                                 "slow-light windows_per_block must be positive and prefetch_snapshots nonnegative");
    if (q.step_mode != KP_SLOW_LIGHT_DECOUPLED && q.step_mode != KP_SLOW_LIGHT_SNAPSHOT &&
        q.step_mode != KP_SLOW_LIGHT_BLOCK)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown slow-light step mode");
    if (q.interpolation != KP_SLOW_LIGHT_FLUID && q.interpolation != KP_SLOW_LIGHT_COEFFICIENTS)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "unknown slow-light interpolation");
    if (!jobs && !std::isfinite(q.observation_time))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "observation time must be finite");
    for (int k = 0; k < q.nsnap; ++k)
        if (!std::isfinite(q.times[k]) || (k > 0 && !(q.times[k] > q.times[k - 1])))
            return kp_internal::fail(KP_INVALID_ARGUMENT, "snapshot times must be finite and strictly increasing");
    if (jobs) {
        // slow_light_batch_driver.hpp:116, 144-146
        if (jobs->njobs < 1 || !jobs->jobs || !jobs->outs)
            return kp_internal::fail(KP_INVALID_ARGUMENT, "no slow-light observation-time jobs");
        for (int j = 0; j < jobs->njobs; ++j) {
            const kp_slow_light_job& J = jobs->jobs[j];
            if (!std::isfinite(J.observation_time) || J.first < 0 || J.first >= J.last || J.last >= q.nsnap)
                return kp_internal::fail(KP_INVALID_ARGUMENT, "invalid slow-light batch observer time or snapshot support");
        }
    }
    if (nfreq < 0 || (nfreq > 0 && !frequencies))
        return kp_internal::fail(KP_INVALID_ARGUMENT, "frequency list and count disagree");
    if (model->n1 < 1 || model->n2 < 1 || model->n3 < 1)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "GRMHD grid dimensions must be positive");
    const size_t ncells = (size_t)model->n1 * (size_t)model->n2 * (size_t)model->n3;
    int device = 0;
    KP_HIP(hipGetDevice(&device));
    const int halo = q.step_mode == KP_SLOW_LIGHT_DECOUPLED ? kp_slow::DECOUPLED_HALO : 0;
    const long long want = 2ll * q.windows_per_block + halo + 1;
    const int slots = (int)std::min<long long>(want, q.nsnap);
    kp_slow::Streamed source(q, ncells, slots, device);
    kp_grmhd_model m = *model;
    // the grid model of the derived-scalar kernel (units, electron model)
    kp_pass_a_params probe_params;
    kp_pass_a_params_default(&probe_params);
    probe_params.coordinate_system = KP_COORD_FMKS;
    probe_params.camera.nx = probe_params.camera.ny = 1;
    probe_params.camera.radius = 100.0;
    Params G;
    kp_status st = prepare_grmhd(&probe_params, &m, G, false);
    if (st != KP_OK) return st;
    st = source.init(q.derive_scalars ? kpsel::derive_scalars() : nullptr, G.grmhd);
    if (st != KP_OK) return st;
    m.cells = source.slot(0);
    kp_slow_light seq;
    kp_slow_light_default(&seq);
    seq.nsnap = q.nsnap;
    seq.step_mode = q.step_mode;
    seq.interpolation = q.interpolation;
    seq.windows_per_block = q.windows_per_block;
    seq.observation_time = q.observation_time;
    seq.times = q.times;
    seq.cells = nullptr;
    FreqLists fl;
    if (nfreq > 0) {
        fl.nfreq = fl.nctrl = nfreq;
        fl.freqs = fl.ctrl = frequencies;
    }
    kp_slow_light_schedule rep;
    std::memset(&rep, 0, sizeof(rep));
    st = run_image(params, nullptr, fl, options, jobs ? jobs->outs : out, stats, 0, 0.0, nullptr, &m, &seq, nullptr,
                   false, &source, &rep, jobs);
    if (st != KP_OK && st != KP_TIME_SUPPORT) return st;
    source.report(rep);
    if (schedule) *schedule = rep;
    return st;
}

extern "C" kp_status kp_run_grmhd_slow_light_stream(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                    const kp_slow_light_stream* sequence, const double* frequencies,
                                                    int32_t nfreq, const kp_run_options* options, kp_image_data* out,
                                                    kp_image_stats* stats, kp_slow_light_schedule* schedule) {
    return run_stream(params, model, sequence, nullptr, frequencies, nfreq, options, out, stats, schedule);
}

extern "C" kp_status kp_run_grmhd_slow_light_stream_jobs(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                         const kp_slow_light_stream* sequence,
                                                         const kp_slow_light_job* jobs, int32_t njobs,
                                                         const double* frequencies, int32_t nfreq,
                                                         const kp_run_options* options, kp_image_data* outs,
                                                         kp_image_stats* stats, kp_slow_light_schedule* schedule) {
    kp_internal::clear_error();
    if (njobs < 1 || !jobs || !outs)
        return kp_internal::fail(KP_INVALID_ARGUMENT, "no slow-light observation-time jobs");
    SlowJobs J;
    J.jobs = jobs;
    J.njobs = njobs;
    J.outs = outs;
    return run_stream(params, model, sequence, &J, frequencies, nfreq, options, nullptr, stats, schedule);
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
    // the gradient kernel runs one block; the images of decoupled and
    // snapshot stepping do not depend on the block length, block stepping's do
    if (sequence->step_mode == KP_SLOW_LIGHT_BLOCK && sequence->windows_per_block > 0 &&
        sequence->windows_per_block < sequence->nsnap - 1)
        return kp_internal::fail(KP_UNSUPPORTED, "slow-light gradients with block stepping run one block of the whole sequence");
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
    kp_grmhd_cell* cells = const_cast<kp_grmhd_cell*>(P.grmhd.cells);
    int64_t n = (int64_t)ncell;
    void* args[] = {&P.grmhd, &cells, &n};
    KP_HIP(hipLaunchKernel(kpsel::derive_scalars(), dim3(grid), dim3(block), args, 0, static_cast<hipStream_t>(stream)));
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
