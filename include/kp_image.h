/*
 * kp_image.h -- C ABI of the MI355X-native KPolaris per-pixel imaging kernel.
 *
 * Every entry point is plain C: structs of scalars, caller-owned pointers,
 * a status code, and no C++ types or exceptions across the boundary.
 * Per-pixel arrays use the reference's pixel order p = iy * nx + ix
 * (reference: src/camera/camera.hpp:96-99).
 *
 * Struct layouts mirror the reference's by-value parameter objects so that a
 * reference-side binding is a field-by-field copy:
 *   kp_camera_params      <- kpolaris::CameraParams      (src/camera/camera.hpp:12-25)
 *   kp_emission_selection <- kpolaris::EmissionSelection (src/radiation/emission_selection.hpp:11-19)
 *   kp_pass_a_params      <- kpolaris::PassAParams       (src/geodesic/pass_a_core.hpp:10-36)
 *   kp_riaf_model         <- kpolaris::RIAFAnalyticRadiationModel (src/model/riaf.hpp:17-52)
 *   kp_image_data         <- ImageHostData arrays         (src/image/result.hpp:9-31)
 */
#ifndef KP_IMAGE_H
#define KP_IMAGE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KP_ABI_VERSION 7

/* Largest frequency list of one shared-geodesic image (the reference's
 * KPOLARIS_MAX_FREQUENCIES build option, CMakeLists.txt:14). */
#define KP_MAX_FREQUENCIES 16

/* Status codes.  The reference throws std::runtime_error / std::invalid_argument
 * (src/image/riaf.cpp:15-18, src/image/batch_driver.hpp:40-41); here no
 * exception crosses the ABI and kp_last_error() holds the message. */
typedef enum kp_status {
    KP_OK = 0,
    KP_INVALID_ARGUMENT = 1,
    KP_UNSUPPORTED = 2,
    KP_HIP_ERROR = 3,
    KP_OUT_OF_MEMORY = 4,
    /* slow light, decoupled stepping: a radiating sample lies outside the
     * snapshot times (the reference's "radiation sample outside supplied
     * physical time support", src/image/slow_light_batch_driver.hpp:286-289);
     * the image is written, its waiting rays with reason 6 */
    KP_TIME_SUPPORT = 5
} kp_status;

/* Per-pixel termination codes, same values and meanings as
 * kpolaris::TerminationReason (src/common/types.hpp:37-48). */
typedef enum kp_termination_reason {
    KP_REASON_NONE = 0,
    KP_REASON_REACHED_CAMERA = 1,
    KP_REASON_REACHED_INNER_BOUNDARY = 2,
    KP_REASON_ESCAPED_DOMAIN = 3,
    KP_REASON_MAX_STEPS = 4,
    KP_REASON_INVALID_STATE = 5,
    KP_REASON_SLOW_LIGHT_TIME_EXHAUSTED = 6,
    KP_REASON_ADAPTIVE_STEP_UNDERFLOW = 7,
    KP_REASON_METRIC_TIME_EXHAUSTED = 8,
    KP_REASON_REACHED_DIRECT_TURN = 9
} kp_termination_reason;

/* kpolaris::CoordinateSystem (src/common/types.hpp:50-56). */
typedef enum kp_coordinate_system {
    KP_COORD_CARTESIAN_KS = 0,
    KP_COORD_BOYER_LINDQUIST = 1,
    KP_COORD_SPHERICAL_KS = 2,
    KP_COORD_FMKS = 3,
    KP_COORD_MKS = 4
} kp_coordinate_system;

/* kpolaris::CameraModel (src/camera/camera.hpp:7-10). */
typedef enum kp_camera_model {
    KP_CAMERA_PINHOLE = 0,
    KP_CAMERA_PARALLEL_PLANE = 1
} kp_camera_model;

/* Thermal fits (src/radiation/thermal_synchrotron.hpp:9-17) and the
 * nonthermal distributions (src/radiation/nonthermal_synchrotron.hpp). */
typedef enum kp_emission_type {
    KP_EMISSION_THERMAL_PANDYA = 1,
    KP_EMISSION_KAPPA = 2,
    KP_EMISSION_POWER_LAW = 3,
    KP_EMISSION_THERMAL_DEXTER = 4
} kp_emission_type;

/* Arithmetic of the per-pixel kernel. */
typedef enum kp_precision {
    /* FP64 throughout: the parity kernel. */
    KP_PRECISION_FP64 = 0,
    /* FP64 ray state; FP32 plasma coefficients and Stokes transfer (in
     * nu_obs^3-scaled units), Faraday rotation angles above one radian in
     * FP64.  Geodesic stepping per kp_run_options.guard_mode: FP64 (default),
     * or FP32 derivatives and RK stage sums with FP64 attempts inside
     * fp64_radius and at turning points of theta, FP64 step-error comparison
     * and surface-crossing refinement.  The GRMHD and slow-light images take
     * the thermal fits only in this precision. */
    KP_PRECISION_MIXED = 1
} kp_precision;

typedef struct kp_camera_params {
    int32_t nx;
    int32_t ny;
    double radius;
    double inclination;
    double fov;
    double fov_y;
    double xspan;
    double yspan;
    double x_offset;
    double y_offset;
    int32_t model;      /* kp_camera_model */
    int32_t reserved0;
} kp_camera_params;

typedef struct kp_emission_selection {
    double equatorial_h_over_r;
    int32_t equatorial_samples;
    int32_t faraday_rotation;
} kp_emission_selection;

typedef struct kp_pass_a_params {
    kp_camera_params camera;
    kp_emission_selection emission_selection;
    double mass;
    double spin;
    double inner_radius;
    double outer_radius;
    double step;
    int32_t max_steps;
    int32_t adaptive;
    int32_t direct_only;
    int32_t coordinate_system; /* kp_coordinate_system */
    double adaptive_tolerance;
    double min_step;
    double max_step;
    double max_radiation_step;
    double max_radiation_depth;
    double max_absorption_depth;
    double max_faraday_depth;
    double fmks_startx1;
    double fmks_hslope;
    double fmks_mks_smooth;
    double fmks_poly_alpha;
    double fmks_poly_xt;
    double fmks_poly_norm;
} kp_pass_a_params;

typedef struct kp_riaf_model {
    double nth0;
    double Te0;
    double disk_h;
    double pow_nth;
    double pow_T;
    double ne_unit;
    double te_unit;
    double b_unit;
    double freq_cgs;
    double mbh_solar;
    double r_min;
    double r_max;
    double horizon_buffer;
    double emission_scale;
    double absorption_scale;
    double faraday_scale;
    double max_pol_frac;
    double nonthermal_kappa;
    int32_t variable_kappa;
    int32_t emission_type;    /* kp_emission_type */
    double variable_kappa_min;
    double variable_kappa_interp_start;
    double variable_kappa_max;
    double powerlaw_p;
    double powerlaw_eta;
    double powerlaw_gamma_min;
    double powerlaw_gamma_max;
    double powerlaw_gamma_cutoff;
    double keplerian_factor;
    double infall_factor;
    double min_frequency_scale;
    double min_sin_theta;
    double min_thetae;
    double min_radius;
} kp_riaf_model;

/* Per-pixel outputs.  Every pointer may be NULL (that output is skipped).
 * Stokes values are invariant (I/nu^3), in the camera basis, not EVPA-flipped
 * and not masked by reason: the host packing rules apply afterwards, exactly
 * as for ImageHostData. */
typedef struct kp_image_data {
    double* image_i;
    double* image_q;
    double* image_u;
    double* image_v;
    double* closure_x;
    double* closure_k;
    double* final_null;
    double* frame_error;
    double* det_r;
    double* overlap_r11;
    double* overlap_r12;
    double* overlap_r21;
    double* overlap_r22;
    double* basis_identity_error;
    double* basis_rotation_angle;
    int32_t* pass_a_steps;
    int32_t* steps;
    int32_t* reason;
} kp_image_data;

/*
 * GRMHD grid model on FMKS data: GRMHDGridRadiationModel
 * (src/model/grmhd_grid.hpp:31-99) with primitives and derived scalars
 * (has_derived_scalars = 1, the iharm loader's default).  Field defaults are
 * the reference's (kp_grmhd_model_default).  `cells` is a DEVICE array of
 * n1*n2*n3 kp_grmhd_cell records in (i, j, k) row-major order, k fastest.
 */
typedef struct kp_grmhd_model {
    int32_t n1, n2, n3;
    int32_t emission_type;      /* kp_emission_type: 4 thermal Dexter (default), 1 thermal
                                   Pandya, 2 kappa, 3 power law */
    double spin, gam;
    double startx1, startx2, startx3, dx1, dx2, dx3;
    double r_in, r_out;
    double hslope, mks_smooth, poly_alpha, poly_xt, poly_norm;
    double freq_cgs, mbh_solar, M_unit;
    double trat_small, trat_large, beta_crit;   /* R_low, R_high, beta_crit */
    double sigma_cut, sigma_cut_high;
    double ne_factor;
    double emission_scale, absorption_scale, faraday_scale, max_pol_frac;
    /* nonthermal distributions (grmhd_grid.hpp:79-89, :1176-1205): kappa,
     * or with variable_kappa != 0 the kappa of sigma and beta, floored at
     * variable_kappa_min; blended toward thermal from
     * min(variable_kappa_interp_start, variable_kappa_max) to
     * variable_kappa_max.  Power law: p, eta (fraction of the magnetic
     * energy density), Lorentz-factor range. */
    double nonthermal_kappa;
    int32_t variable_kappa;
    int32_t reserved_nt;
    double variable_kappa_min, variable_kappa_interp_start, variable_kappa_max;
    double powerlaw_p, powerlaw_eta, powerlaw_gamma_min, powerlaw_gamma_max, powerlaw_gamma_cutoff;
    const void* cells;          /* kp_grmhd_cell[n1 * n2 * n3], device */
} kp_grmhd_model;

/*
 * One cell of the device grid: the 8 primitives as read from the dump
 * (rho, u, U1, U2, U3, B1, B2, B3 in native FMKS components, single
 * precision as the iharm reader reads them, src/grmhd/iharm_loader.cpp:262-396)
 * and the 5 derived scalars (n_e [cm^-3], Theta_e, B [G], sigma, beta;
 * iharm_loader.cpp:420-505) in double.  A trilinear sample reads 8 whole
 * records (72 bytes each) instead of 8 entries of 13 separate arrays.
 */
typedef struct kp_grmhd_cell {
    float prim[8];
    double derived[5];
} kp_grmhd_cell;

/* Launch configuration; no reference counterpart (Kokkos chose it there). */
typedef struct kp_run_options {
    int32_t precision;        /* kp_precision */
    int32_t split_transport;  /* 1: pass A and pass B kernels (reference default);
                                 0: one kernel, pass A then pass B per lane
                                 (identical results; stats report the whole
                                 time as pass_a_ms) */
    int32_t block_size;       /* rays per workgroup; 0 = tuned default */
    int32_t persistent;       /* 1: lane-refill work queue; 0: one ray per thread in
                                 pass A (pass B always refills lanes from the queue;
                                 the results are the same) */
    int32_t waves_per_cu;     /* persistent grid size; 0 = occupancy-derived */
    int32_t guard_mode;       /* mixed: geodesic precision.  0 = FP64 geodesics,
                                 FP32 coefficients and transfer (default; FP32
                                 arithmetic issues at the FP64 rate on this
                                 hardware without packed FP32, so FP32 stepping
                                 buys nothing);  1 = FP32 stepping with FP64
                                 attempts inside fp64_radius and across turning
                                 points of theta (an FP32 attempt whose
                                 dtheta/dlambda changes sign is taken again in
                                 FP64), chosen per lane in one kernel;  2 = FP32
                                 stepping staged: FP32-geodesic and
                                 FP64-geodesic kernels, rays moved between them
                                 at fp64_radius and 1.25 fp64_radius, without
                                 the turning-point retake (persistent split
                                 launches; others use 1).  GRMHD and slow-light
                                 images step FMKS geodesics in FP64 in every
                                 mode. */
    double fp64_radius;       /* mixed: attempts starting at r < fp64_radius run in FP64;
                                 0 turns every FP64 guard off */
    void* stream;             /* hipStream_t, NULL = default stream */
    /* Device workspace of the run (pass-A endpoints, queue counters, staged
     * and slow-light tables), owned by the caller, at least the size
     * kp_image_workspace_size reports, and not used by other work while the
     * run is in flight on `stream`.  NULL: the library keeps one workspace
     * per (device, stream), grown on demand (it synchronises that stream
     * and reallocates when a larger image arrives), so runs on different
     * streams never share buffers. */
    void* workspace;
    uint64_t workspace_bytes;
} kp_run_options;

/* Device workspace bytes of an image of npix pixels with nfreq output
 * frequencies (0 or 1: one), nsnap slow-light snapshots (0: none) and these
 * options (precision, guard_mode and split_transport select the staged
 * mixed-precision buffers).  The kernels never allocate: a caller that
 * passes options->workspace of this size makes the run allocation-free. */
kp_status kp_image_workspace_size(int64_t npix, int32_t nfreq, int32_t nsnap, const kp_run_options* options,
                                  uint64_t* bytes);

/* Work counters and device timings of one run.  The unit counts are what the
 * roofline's algorithmic FLOPs are priced on. */
typedef struct kp_image_stats {
    double pass_a_ms;
    double pass_b_ms;
    double kernel_ms;
    int64_t rays;
    int64_t pass_a_steps;       /* accepted steps, pass A */
    int64_t pass_b_steps;       /* accepted steps, pass B */
    int64_t rhs_evals;          /* transport_rhs evaluations, both passes */
    int64_t pass_a_rhs_evals;   /* transport_rhs evaluations in pass A */
    int64_t coefficient_evals;  /* plasma coefficient evaluations */
    int64_t transfer_steps;     /* Strang transfer steps computed */
    int64_t fp64_guard_steps;   /* mixed precision: steps taken in FP64 */
    int64_t vacuum_steps;       /* pass-B steps of the vacuum leg (outside the
                                 * emitting region, moving outward): their
                                 * transfer step is the identity and is not
                                 * computed; transfer_steps + vacuum_steps is
                                 * the reference's transfer-step count */
    int64_t vacuum_rhs_evals;   /* transport_rhs evaluations of the vacuum leg
                                 * (included in rhs_evals) */
    double pass_b_vacuum_ms;    /* the vacuum leg's share of pass_b_ms */
} kp_image_stats;

/* ABI version and last error message of the calling thread. */
int32_t kp_abi_version(void);
const char* kp_last_error(void);

/* Number of visible HIP devices; 0 when none (never an error). */
int32_t kp_device_count(void);

/*
 * RIAF image in Boyer-Lindquist coordinates: pass A + pass B per pixel.
 * Replaces run_riaf_image(const PassAParams&, RIAFAnalyticRadiationModel,
 * int radiation_substeps, int timing, int split_transport)
 * (src/image/riaf.hpp:12-17, dispatch src/image/riaf.cpp:21-45,
 * kernels src/image/split_driver.hpp:41-104).
 * `out` holds DEVICE pointers; work is stream-ordered on options->stream and
 * the call returns after the kernels are queued unless `stats` is non-NULL,
 * in which case it synchronises to read counters and event times.
 * radiation_substeps is accepted and ignored, as on the reference's segment
 * path (src/geodesic/pass_b.hpp:526).
 */
kp_status kp_run_riaf_image(const kp_pass_a_params* params,
                            const kp_riaf_model* model,
                            int32_t radiation_substeps,
                            const kp_run_options* options,
                            kp_image_data* out,
                            kp_image_stats* stats);

/* Allocation-free variant of kp_run_riaf_image.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The pass-A endpoints and queue counters of a single-frequency RIAF image of
 * at most KP_NOALLOC_MAX_PIXELS pixels live in a static device array of the
 * library (140 bytes per pixel, about 37 MB, held on every GPU the library
 * runs on until the process ends); the caller passes inputs and device
 * outputs only.  Every call on one GPU shares that array and it keeps the
 * last call's contents (the queue and counters are cleared by a kernel at the
 * start of each call): calls that overlap on one GPU (on two streams, or
 * from two threads) race and corrupt each other; calls ordered on one stream
 * do not.  Larger images, staged mixed precision (guard_mode 2) and a caller
 * workspace are rejected (KP_INVALID_ARGUMENT / KP_UNSUPPORTED).  No stats.
 * Stream-ordered on options->stream; the results equal kp_run_riaf_image's
 * bit for bit (same kernels). */
#define KP_NOALLOC_MAX_PIXELS (512 * 512)
kp_status kp_run_riaf_image_noalloc(const kp_pass_a_params* params,
                                    const kp_riaf_model* model,
                                    int32_t radiation_substeps,
                                    const kp_run_options* options,
                                    kp_image_data* out);

/* Same computation with HOST output pointers; synchronous. */
kp_status kp_run_riaf_image_host(const kp_pass_a_params* params,
                                 const kp_riaf_model* model,
                                 int32_t radiation_substeps,
                                 const kp_run_options* options,
                                 kp_image_data* out,
                                 kp_image_stats* stats);

/*
 * Multi-frequency RIAF image on shared geodesics: the geodesic of each pixel
 * is integrated once and nfreq Stokes vectors are transported along it.
 * Replaces run_riaf_multifrequency_image(const PassAParams&,
 * RIAFAnalyticRadiationModel, int radiation_substeps,
 * const std::vector<double>& frequencies,
 * const std::vector<double>& step_control_frequencies, int timing,
 * int split_transport) (src/image/riaf.hpp:31-39, src/image/riaf.cpp:80-105,
 * src/image/multifrequency.hpp:152-280; per-pixel algorithm
 * src/geodesic/pass_b.hpp:1130-1500 and, for distinct control lists,
 * :1505-1895).  frequencies and step_control_frequencies are HOST arrays of
 * 1..KP_MAX_FREQUENCIES values (Hz); the substep control uses the largest
 * count over the control frequencies.  out->image_i/q/u/v point to
 * [nfreq][npix] arrays (frequency-major); the other outputs are per pixel.
 * model->freq_cgs is not used.
 */
kp_status kp_run_riaf_multifrequency_image(const kp_pass_a_params* params,
                                           const kp_riaf_model* model,
                                           int32_t radiation_substeps,
                                           const double* frequencies,
                                           int32_t nfreq,
                                           const double* step_control_frequencies,
                                           int32_t step_control_nfreq,
                                           const kp_run_options* options,
                                           kp_image_data* out,
                                           kp_image_stats* stats);

/* Multi-frequency image with HOST output pointers; synchronous. */
kp_status kp_run_riaf_multifrequency_image_host(const kp_pass_a_params* params,
                                                const kp_riaf_model* model,
                                                int32_t radiation_substeps,
                                                const double* frequencies,
                                                int32_t nfreq,
                                                const double* step_control_frequencies,
                                                int32_t step_control_nfreq,
                                                const kp_run_options* options,
                                                kp_image_data* out,
                                                kp_image_stats* stats);

/*
 * RIAF image together with its derivatives with respect to the plasma
 * log-scale parameters of PlasmaPerturbation (src/radiation/plasma_perturbation.hpp:10-33):
 * q_density (n_e x e^q, B x e^(q/2)), q_temperature (Theta_e x e^q) and
 * q_magnetic (B x e^q), evaluated at log_scales = {q_density, q_temperature,
 * q_magnetic} (HOST array of 3; zeros = the unperturbed model) along the
 * image's geodesics.  The reference estimates these by central differences
 * of the coefficient blocks (src/diagnostics/response.hpp:118-182); here the
 * transfer is differentiated exactly in forward (tangent) mode.  grad is a
 * DEVICE array [3][4][npix]: grad[(k * 4 + s) * npix + p] = dS_s/dq_k for
 * Stokes s = I, Q, U, V in the observer basis (invariant units, like
 * out->image_*).  FP64 single-frequency only (options->precision must be
 * KP_PRECISION_FP64); out receives the image at the same point.
 */
kp_status kp_run_riaf_image_gradient(const kp_pass_a_params* params,
                                     const kp_riaf_model* model,
                                     int32_t radiation_substeps,
                                     const double* log_scales,
                                     const kp_run_options* options,
                                     kp_image_data* out,
                                     double* grad,
                                     kp_image_stats* stats);

/*
 * Batched images.  Replaces run_image_batch_metric(jobs, ImageBatchOptions,
 * ImageBatchTiming*) and ImageBatchJob (src/image/batch_driver.hpp:15-21,
 * 126-187), which run one unsplit kernel per job on partitioned execution
 * instances.  Here every job's rays share one persistent pass A launch and
 * one pass B launch with a single work queue, recorded once as a HIP graph
 * (kp_batch_create) and replayed by kp_batch_launch.  Jobs may differ in
 * camera, resolution, spin, inclination and model; all share one precision
 * and launch configuration (kp_run_options).  Each job's `out` holds DEVICE
 * pointers owned by the caller, which must stay valid while the batch exists.
 */
typedef struct kp_batch_job {
    kp_pass_a_params params;
    kp_riaf_model model;
    int32_t radiation_substeps; /* must be positive, otherwise unused (pass_b.hpp:526) */
    int32_t reserved0;
    kp_image_data out;
} kp_batch_job;

typedef struct kp_batch kp_batch;

/* Validates every job, uploads the job table, allocates the endpoint
 * workspace and instantiates the graph. */
kp_status kp_batch_create(const kp_batch_job* jobs, int32_t njobs, const kp_run_options* options,
                          kp_batch** batch);
/* Replays the graph on `stream` (hipStream_t; NULL = default stream). */
kp_status kp_batch_launch(kp_batch* batch, void* stream);
/* Work counters of the last completed launch (the caller synchronises first). */
kp_status kp_batch_read_stats(kp_batch* batch, kp_image_stats* stats);
void kp_batch_destroy(kp_batch* batch);
/* create + launch + wait + destroy; stats->kernel_ms is the graph's device time. */
kp_status kp_run_riaf_batch(const kp_batch_job* jobs, int32_t njobs, const kp_run_options* options,
                            kp_image_stats* stats);

/* ImageBatchOptions (src/image/batch_driver.hpp:23-27): at most
 * max_in_flight jobs per window, and a window's device outputs
 * (KP_BATCH_BYTES_PER_PIXEL per pixel) within max_device_output_bytes. */
typedef struct kp_batch_limits {
    int64_t max_in_flight;
    int64_t max_device_output_bytes;
} kp_batch_limits;
/* BatchImageDeviceData::bytes_per_pixel: 15 doubles and 3 int32 per pixel. */
#define KP_BATCH_BYTES_PER_PIXEL (15 * 8 + 3 * 4)
void kp_batch_limits_default(kp_batch_limits* limits); /* 4 jobs, 256 MiB */
/* ImageBatchTiming (src/image/batch_driver.hpp:29-35). */
typedef struct kp_batch_timing {
    double allocate_seconds;
    double kernel_seconds;
    double copy_seconds;
    int64_t windows;
    int64_t peak_device_output_bytes;
} kp_batch_timing;
/* run_image_batch_metric with host outputs (batch_driver.hpp:126-187): each
 * job's `out` holds HOST pointers (null fields are skipped).  Every job is
 * validated before any work: limits must be positive, each job's nx * ny a
 * positive int-sized product whose outputs fit the budget, radiation_substeps
 * positive (KP_INVALID_ARGUMENT otherwise).  Jobs then run in windows, in
 * order: a window takes jobs while both limits hold, allocates their device
 * outputs, runs them as one batch graph, copies back and frees.  njobs == 0
 * is a valid empty batch.  Synchronous on options->stream. */
kp_status kp_run_riaf_batch_host(const kp_batch_job* jobs, int32_t njobs, const kp_batch_limits* limits,
                                 const kp_run_options* options, kp_batch_timing* timing);

/*
 * GRMHD grid image on FMKS geodesics (params->coordinate_system must be
 * KP_COORD_FMKS with params->fmks_* describing the geodesic coordinates).
 * Replaces run_grmhd_image (src/image/grmhd.hpp:12-17, FMKS backend
 * src/image/grmhd.cpp) for FMKS dumps with derived scalars; FP64 only.
 * `out` holds DEVICE pointers, as for kp_run_riaf_image.
 */
kp_status kp_run_grmhd_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                             int32_t radiation_substeps, const kp_run_options* options,
                             kp_image_data* out, kp_image_stats* stats);

/* Slow light: a time sequence of GRMHD snapshots on one grid, sampled at
 * retarded time t = observation_time + x^0 (x^0 = 0 at the camera,
 * docs/slow_light.md:9-15).  Mirrors the arguments of
 * run_slow_light_image_metric (src/image/slow_light_driver.hpp:1940-1954) and
 * the residency options of run_slow_light_batch (slow_light_batch_driver.hpp:
 * 13-24): pass B runs over blocks of windows_per_block snapshot pairs, as the
 * reference's block scheduler does, or over the whole sequence at once. */
typedef enum kp_slow_light_step_mode {
    /* kpolaris::SlowLightStepMode (src/model/time_interpolation.hpp:9-12) */
    KP_SLOW_LIGHT_DECOUPLED = 0,  /* steps never cut at snapshot times; the default */
    KP_SLOW_LIGHT_SNAPSHOT = 1,   /* pair-by-pair windows, steps cut at each snapshot time */
    KP_SLOW_LIGHT_BLOCK = 2       /* steps cut at the last snapshot time only */
} kp_slow_light_step_mode;

typedef enum kp_slow_light_interpolation {
    /* kpolaris::SlowLightInterpolation (src/model/time_interpolation.hpp:19) */
    KP_SLOW_LIGHT_COEFFICIENTS = 0,
    KP_SLOW_LIGHT_FLUID = 1       /* the default */
} kp_slow_light_interpolation;

typedef struct kp_slow_light {
    int32_t nsnap;                /* snapshots, at least 2 */
    int32_t step_mode;            /* kp_slow_light_step_mode */
    int32_t interpolation;        /* kp_slow_light_interpolation */
    int32_t windows_per_block;    /* SlowLightBatchOptions::windows_per_block
                                   * (slow_light_batch_driver.hpp:20): snapshot
                                   * pairs per pass-B block, 0 = one block over
                                   * the whole sequence.  Block stepping cuts
                                   * steps at each block's last time, so its
                                   * image depends on this value; decoupled and
                                   * snapshot images do not. */
    double observation_time;      /* coordinate arrival time at the camera */
    const double* times;          /* [nsnap] strictly increasing, HOST */
    const void* const* cells;     /* [nsnap] DEVICE kp_grmhd_cell arrays (HOST array of pointers) */
} kp_slow_light;

/* Slow-light GRMHD image (nfreq frequencies on shared geodesics, as
 * kp_run_grmhd_multifrequency_image).  `model` gives the grid, units and
 * emission settings (its cells pointer is not used).  Replaces
 * run_slow_light_image_metric (src/image/slow_light_driver.hpp:1940-2028:
 * pass A endpoints, the window kernel over every snapshot pair, the vacuum
 * finish of unfinished rays) and, with decoupled stepping, the
 * all-resident case of run_slow_light_batch (slow_light_batch_driver.hpp).
 * Returns KP_TIME_SUPPORT (image written) when decoupled stepping needs a
 * time outside [times[0], times[nsnap-1]]. */
kp_status kp_run_grmhd_slow_light_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                        const kp_slow_light* sequence, const double* frequencies,
                                        int32_t nfreq, const kp_run_options* options, kp_image_data* out,
                                        kp_image_stats* stats);
void kp_slow_light_default(kp_slow_light* sequence);

/* Streamed slow light: a sequence read snapshot by snapshot from host storage
 * instead of resident in device memory, for sequences larger than the device
 * holds.  Replaces run_slow_light_batch_metric_impl's block loop for one
 * observation time (src/image/slow_light_batch_driver.hpp:105-340) with its
 * host read-ahead (SlowLightHostStream, src/image/slow_light_prefetch.hpp:
 * 15-88): a library thread calls `load` for snapshots in time order, up to
 * prefetch_snapshots ahead, into pinned host buffers; a copy stream uploads
 * each into a device ring of 2 windows_per_block + halo + 1 snapshot slots
 * (halo = 4 for decoupled stepping, decoupled_cache_halo_snapshots,
 * src/model/time_interpolation.hpp:11; 0 otherwise) while pass B runs the
 * current block.  A decoupled ray that needs an earlier time than its block
 * holds makes the scheduler rewind to that snapshot (:317-330).  The image
 * equals kp_run_grmhd_slow_light_image's over the same snapshots with the
 * same windows_per_block.
 *
 * `load` fills n1*n2*n3 kp_grmhd_cell records of snapshot `index` at
 * `cells` (host memory); with derive_scalars = 1 it fills the primitives
 * only and the library computes the derived scalars on the device
 * (kp_grmhd_derive_scalars).  It runs on the read-ahead thread (prefetch =
 * 1) or the calling thread (prefetch = 0), one call at a time, and returns
 * KP_OK or an error that ends the run.  kp_iharm_load_snapshot (kp_io.h) is
 * such a loader for iHARM dump files. */
typedef kp_status (*kp_snapshot_loader)(void* user, int32_t index, kp_grmhd_cell* cells, uint64_t ncells);

typedef struct kp_slow_light_stream {
    int32_t nsnap;                /* snapshots, at least 2 */
    int32_t step_mode;            /* kp_slow_light_step_mode */
    int32_t interpolation;        /* kp_slow_light_interpolation */
    int32_t windows_per_block;    /* snapshot pairs per block, at least 1 */
    int32_t prefetch;             /* 1: read ahead on a library thread (default) */
    int32_t prefetch_snapshots;   /* read-ahead queue length, 0 = windows_per_block */
    int32_t derive_scalars;       /* 1: loader fills primitives, derived on the device */
    int32_t reserved0;
    double observation_time;
    const double* times;          /* [nsnap] strictly increasing, HOST */
    kp_snapshot_loader load;
    void* user;                   /* passed to load */
} kp_slow_light_stream;

/* What the block scheduler did (the reference's timing lines
 * slow_light_block_*, slow_light_prefetch_*, slow_light_cache_rewind). */
typedef struct kp_slow_light_schedule {
    int32_t blocks;               /* pass-B launches */
    int32_t rewinds;              /* decoupled rewinds to an earlier snapshot */
    int32_t loads;                /* loader calls */
    int32_t slots;                /* device snapshot slots */
    int64_t slot_bytes;           /* device bytes of the slots */
    double load_ms;               /* time inside the loader */
    double wait_ms;               /* host time waiting for a snapshot to be read */
    double pass_b_ms;             /* device time of the pass-B launches */
} kp_slow_light_schedule;

void kp_slow_light_stream_default(kp_slow_light_stream* sequence);
kp_status kp_run_grmhd_slow_light_stream(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                         const kp_slow_light_stream* sequence, const double* frequencies,
                                         int32_t nfreq, const kp_run_options* options, kp_image_data* out,
                                         kp_image_stats* stats, kp_slow_light_schedule* schedule);

/* One observation time of a multi-observer run (SlowLightBatchJob,
 * src/image/slow_light_batch_driver.hpp:13-17): the image at
 * observation_time from snapshots first..last (inclusive support,
 * 0 <= first < last < nsnap). */
typedef struct kp_slow_light_job {
    double observation_time;
    int32_t first;
    int32_t last;
} kp_slow_light_job;

/* Several observation times from one pass over a streamed sequence
 * (run_slow_light_batch_metric_impl with several observer jobs,
 * src/image/slow_light_batch_driver.hpp:105-340): pass A runs once (the
 * geodesics do not depend on the time), every job starts its own rays from
 * those endpoints, and each block of snapshots runs the pass B of every job
 * whose support overlaps it, on the part of the block inside that support.
 * The blocks start at the earliest job's first snapshot and end at the
 * latest job's last; a decoupled rewind goes to the earliest time any job's
 * waiting ray needs.  outs[j] is job j's image (nfreq planes, like `out` of
 * kp_run_grmhd_slow_light_stream); sequence->observation_time is not read.
 * Decoupled and snapshot stepping give each job the image of
 * kp_run_grmhd_slow_light_stream over its own snapshots first..last; block
 * stepping's image depends on where the blocks fall, as the reference's
 * does, and equals that run's when the job's first snapshot is the earliest
 * job's.  The job state (about the workspace of one image per job) is
 * allocated by the call.  stats sums over the jobs (rays = npix * njobs);
 * schedule->blocks counts the pass-B launches.  Errors as the reference
 * (:133-148): no jobs, a non-finite time, first >= last or last >= nsnap ->
 * KP_INVALID_ARGUMENT; a decoupled sample outside a job's support ->
 * KP_TIME_SUPPORT. */
kp_status kp_run_grmhd_slow_light_stream_jobs(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                              const kp_slow_light_stream* sequence, const kp_slow_light_job* jobs,
                                              int32_t njobs, const double* frequencies, int32_t nfreq,
                                              const kp_run_options* options, kp_image_data* outs,
                                              kp_image_stats* stats, kp_slow_light_schedule* schedule);

/* The same observation-time jobs over resident device snapshots
 * (kp_slow_light, as kp_run_grmhd_slow_light_image): blocks of
 * sequence->windows_per_block pairs, 0 = one block over the sequence;
 * sequence->observation_time is not read.  Each job equals
 * kp_run_grmhd_slow_light_image's image over its own snapshots as the
 * streamed form does, and the streamed form's image of the same jobs. */
kp_status kp_run_grmhd_slow_light_image_jobs(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                             const kp_slow_light* sequence, const kp_slow_light_job* jobs,
                                             int32_t njobs, const double* frequencies, int32_t nfreq,
                                             const kp_run_options* options, kp_image_data* outs,
                                             kp_image_stats* stats);

/* Slow-light image and its derivatives along the fixed geodesics with
 * respect to (q_density, R_high, q_magnetic): q_density scales n_e by e^q and
 * B by e^(q/2), q_magnetic scales B by e^q (PlasmaPerturbation,
 * src/radiation/plasma_perturbation.hpp:12-33), R_high is the model's
 * trat_large entering Theta_e through the derived scalars
 * (src/grmhd/iharm_loader.cpp:494-504).  log_scales = {q_density,
 * q_magnetic} is the evaluation point; grad is a DEVICE [3][4][npix] array
 * (parameter-major, then I, Q, U, V).  One frequency (model->freq_cgs).
 * The reference computes these responses by finite differences and
 * perturbation reruns (src/diagnostics/response.hpp:118-191, its slow-light
 * form slow_light_driver.hpp:1376); here they are exact forward-mode
 * tangents of the same transfer. */
kp_status kp_run_grmhd_slow_light_image_gradient(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                                 const kp_slow_light* sequence, const double* log_scales,
                                                 const kp_run_options* options, kp_image_data* out,
                                                 double* grad, kp_image_stats* stats);

/* Static GRMHD image and its derivatives along the fixed geodesics with
 * respect to (q_density, R_high, q_magnetic), as
 * kp_run_grmhd_slow_light_image_gradient defines them, for one snapshot:
 * the plasma responses of a GRMHD image (PlasmaPerturbation,
 * src/radiation/plasma_perturbation.hpp:12-33; response runs,
 * src/diagnostics/response.hpp:118-191).  The snapshot is imaged as the
 * two-snapshot sequence of itself, whose fluid interpolation returns the
 * snapshot's values unchanged (src/model/time_interpolation.hpp:27-32).
 * log_scales = {q_density, q_magnetic}; grad is a DEVICE [3][4][npix] array. */
kp_status kp_run_grmhd_image_gradient(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                      const double* log_scales, const kp_run_options* options,
                                      kp_image_data* out, double* grad, kp_image_stats* stats);

/* Launch-configuration tuning (no reference counterpart: Kokkos chooses the
 * reference's launch shapes).  Times the image kernels of a reduced copy of
 * the RIAF image (camera nx / sample_divisor by ny / sample_divisor over the
 * same field of view, outputs discarded) for block_size in {64, 128, 256}
 * and waves_per_cu in {0, 2, 8}, each configuration twice, and writes the
 * fastest block_size and waves_per_cu into *options (its other fields,
 * precision included, select the kernels that are timed).  frequencies may
 * be NULL (nfreq 0): the single-frequency image.  The results of an image do
 * not depend on these two fields. */
typedef struct kp_launch_tuning {
    int32_t ncandidates;     /* configurations timed */
    int32_t block_size;      /* the chosen configuration */
    int32_t waves_per_cu;
    int32_t sample_pixels;   /* pixels of the reduced image */
    double best_ms;          /* kernel time of the reduced image, chosen configuration */
    double default_ms;       /* the same with block_size = waves_per_cu = 0 */
} kp_launch_tuning;
kp_status kp_tune_riaf_launch(const kp_pass_a_params* params, const kp_riaf_model* model, const double* frequencies,
                              int32_t nfreq, int32_t sample_divisor, kp_run_options* options,
                              kp_launch_tuning* report);

/* Time support of a slow-light image: pass A endpoints, then geometric
 * stepping to the camera without radiation, recording the x^0 range of the
 * samples inside [inner_radius, outer_radius); the fluid times needed are
 * observation_time + [min_x0, max_x0].  Replaces
 * run_slow_light_time_probe_metric (src/image/slow_light_driver.hpp:632-780;
 * the tool's slow_light_time_probe=1).  model->cells is not read. */
typedef struct kp_slow_light_probe {
    double min_x0, max_x0;        /* 0, 0 when no pixel radiates */
    int64_t radiating_pixels;     /* pixels with at least one radiating sample */
    int64_t returned_pixels;      /* pixels whose ray reached the camera */
} kp_slow_light_probe;

kp_status kp_slow_light_time_probe(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                   const kp_run_options* options, kp_slow_light_probe* out);

/* Multi-frequency GRMHD image on shared geodesics; replaces
 * run_grmhd_multifrequency_image (src/image/grmhd.hpp:31-38).  Arrays and
 * frequency lists as for kp_run_riaf_multifrequency_image. */
kp_status kp_run_grmhd_multifrequency_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                            int32_t radiation_substeps, const double* frequencies,
                                            int32_t nfreq, const double* step_control_frequencies,
                                            int32_t step_control_nfreq, const kp_run_options* options,
                                            kp_image_data* out, kp_image_stats* stats);

/* Fills the derived scalars of every cell of model->cells (DEVICE records
 * whose primitives are set) on `stream`: the iharm loader's derived-scalar
 * pass (src/grmhd/iharm_loader.cpp:420-505).  No reference ABI counterpart
 * (the loader runs it inside load_iharm_model_from_hdf5). */
kp_status kp_grmhd_derive_scalars(const kp_grmhd_model* model, void* stream);

/* Defaults of the reference's structs (field initialisers of CameraParams,
 * PassAParams and RIAFAnalyticRadiationModel) and of kp_run_options. */
void kp_pass_a_params_default(kp_pass_a_params* params);
void kp_riaf_model_default(kp_riaf_model* model);
void kp_run_options_default(kp_run_options* options);
/* GRMHDGridRadiationModel field initialisers (src/model/grmhd_grid.hpp:53-80). */
void kp_grmhd_model_default(kp_grmhd_model* model);

/* ---- Host option layer (the reference CLI's rules) ------------------------
 * Mirrors the image tool's Options struct subset on the hot path
 * (tools/kpolaris_model_image_options.hpp:18-99), key=value parsing with the
 * same key names (tools/kpolaris_model_image_options.cpp), camera-extent
 * resolution (tools/kpolaris_model_image_options.cpp:136-180), PassAParams
 * assembly (tools/kpolaris_model_image.cpp:3940-3990), RIAF model fill and
 * radius defaults (tools/kpolaris_model_image.cpp:4925-4962), and the flux
 * scale (tools/kpolaris_model_image.cpp:558-595). */
typedef struct kp_options kp_options;

kp_options* kp_options_create(void);
void kp_options_destroy(kp_options* options);
/* Set one key (e.g. "spin", "0.94").  Unknown keys return KP_UNSUPPORTED. */
kp_status kp_options_set(kp_options* options, const char* key, const char* value);
/* Parse a parameter file of key=value lines ('#' comments). */
kp_status kp_options_load_file(kp_options* options, const char* path);
/* Resolve camera extents and radii, then fill PassAParams and the RIAF model.
 * intensity_to_flux_jy_per_pixel receives the flux scale (NaN if invalid). */
kp_status kp_options_resolve_riaf(kp_options* options,
                                  kp_pass_a_params* params,
                                  kp_riaf_model* model,
                                  double* intensity_to_flux_jy_per_pixel,
                                  int32_t* evpa_qu_sign);

/* Packing rules of pack_transport_frequency_image
 * (tools/kpolaris_model_image.cpp:2668-2692, tools/kpolaris_evpa.hpp:15-17):
 * pixels with reason != 1 are zeroed, Q and U take evpa_qu_sign, and the
 * nu^3 factor turns invariant into observed intensity.  Host arrays.
 * Any of i_nu..v_nu may be NULL; flux_jy[4] receives the I,Q,U,V fluxes. */
kp_status kp_pack_image(const double* image_i, const double* image_q,
                        const double* image_u, const double* image_v,
                        const int32_t* reason, int64_t npix,
                        double frequency, int32_t evpa_qu_sign,
                        double intensity_to_flux_jy_per_pixel,
                        double* i_nu, double* q_nu, double* u_nu, double* v_nu,
                        double* flux_jy);

#ifdef __cplusplus
}
#endif

#endif /* KP_IMAGE_H */
