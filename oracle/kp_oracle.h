/*
 * kp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C11, IEEE double, libm) of the KPolaris imaging
 * path.  It is the checker for the HIP kernels: only tests, the smoke check
 * and the benchmark's CPU-baseline leg may load it.  The product library
 * never links, loads or falls back to it.
 *
 * Parity status of every function is recorded in the header of kp_oracle.c.
 */
#ifndef KP_ORACLE_H
#define KP_ORACLE_H

#include <stdint.h>
#include "../include/kp_image.h"

#ifdef __cplusplus
extern "C" {
#endif

/* State layout: x[0..3], k[0..3], e1[0..3], e2[0..3] (TransportState). */
/* Largest frequency list of one shared-geodesic image. */
#define KPO_MAX_FREQUENCIES 16

typedef struct kpo_state { double v[16]; } kpo_state;

typedef struct kpo_coeffs {
    double jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV;
} kpo_coeffs;

typedef struct kpo_stokes { double I, Q, U, V; } kpo_stokes;

/* Plasma perturbation (src/radiation/plasma_perturbation.hpp:9-33). */
typedef struct kpo_perturbation {
    int32_t parameter;   /* 0 none, 1 density, 2 temperature, 3 magnetic */
    int32_t reserved0;
    double log_scale;
} kpo_perturbation;

typedef struct kpo_stats {
    int64_t rhs_evals;
    int64_t coefficient_evals;
    int64_t transfer_steps;
    int64_t pass_a_steps;
    int64_t pass_b_steps;
} kpo_stats;

/* ---- geometry -------------------------------------------------------- */
void kpo_bl_gcov(double mass, double spin, const double x[4], double g[16]);
void kpo_bl_gcon(double mass, double spin, const double x[4], double g[16]);
void kpo_bl_dgcov(double mass, double spin, const double x[4], double dg[64]);
/* Generic connection from gcon and dgcov (src/geometry/connection.hpp:7-24). */
void kpo_bl_connection(double mass, double spin, const double x[4], double gamma[64]);
void kpo_bl_transport_rhs(double mass, double spin, const kpo_state* s, kpo_state* d);
void kpo_bl_transport_rhs_generic(double mass, double spin, const kpo_state* s, kpo_state* d);
double kpo_bl_dot(double mass, double spin, const double x[4],
                  const double u[4], const double v[4]);

/* FMKS coordinate map parameters (src/geometry/kerr_fmks.hpp:11-30). */
typedef struct kpo_fmks {
    double startx1, hslope, mks_smooth, poly_alpha, poly_xt, poly_norm;
} kpo_fmks;

double kpo_fmks_theta(const kpo_fmks* f, double x1, double x2);
double kpo_fmks_x2_from_theta(const kpo_fmks* f, double x1, double theta);
void kpo_fmks_gcov(double mass, double spin, const kpo_fmks* f, const double x[4], double g[16]);
void kpo_fmks_gcon(double mass, double spin, const kpo_fmks* f, const double x[4], double g[16]);
void kpo_fmks_transport_rhs(double mass, double spin, const kpo_fmks* f, const kpo_state* s, kpo_state* d);
/* Christoffel contraction from central differences of gcov with step h. */
void kpo_fmks_transport_rhs_fd(double mass, double spin, const kpo_fmks* f, const kpo_state* s, double h,
                               kpo_state* d);

/* ---- integrators ----------------------------------------------------- */
void kpo_bl_rk4_step(double mass, double spin, const kpo_state* s, double h, kpo_state* out);
/* Returns accepted flag; fills state, mid_state, used_h, next_h, error. */
int32_t kpo_bl_adaptive_rk4_step(double mass, double spin, const kpo_state* s,
                                 double requested_h, double tolerance,
                                 double min_step, double max_step, int32_t max_attempts,
                                 kpo_state* state, kpo_state* mid_state,
                                 double* used_h, double* next_h, double* error);
double kpo_state_error_norm(const kpo_state* a, const kpo_state* b);

/* ---- camera ---------------------------------------------------------- */
void kpo_pixel_offsets(int32_t pixel, const kp_camera_params* cam, double* sx, double* sy);
void kpo_bl_camera_ray(const kp_pass_a_params* params, int32_t pixel, kpo_state* out);
void kpo_bl_camera_tetrad(const kp_pass_a_params* params, double tetrad[16]);

/* ---- radiation ------------------------------------------------------- */
void kpo_thermal_coefficients(int32_t fit, double nu, double ne, double thetae,
                              double b_cgs, double theta, double sin_theta,
                              double cos_theta, double max_pol_frac,
                              double min_sin_theta, double emission_scale,
                              double absorption_scale, double faraday_scale,
                              kpo_coeffs* out);
/* NonthermalSynchrotronParams (src/radiation/nonthermal_synchrotron.hpp:17-36);
 * distribution 2 = kappa, 3 = power law. */
typedef struct kpo_nonthermal {
    int32_t distribution;
    int32_t reserved0;
    double max_pol_frac_emission, max_pol_frac_absorption, min_sin_theta;
    double emission_scale, absorption_scale, faraday_scale;
    double power_law_p, power_law_eta, power_law_gamma_min, power_law_gamma_max, power_law_gamma_cutoff;
    double kappa, kappa_width, kappa_interp_begin, kappa_interp_end;
} kpo_nonthermal;
void kpo_nonthermal_coefficients(const kpo_nonthermal* params, double nu, double ne, double thetae,
                                 double b_cgs, double sin_theta, double cos_theta, kpo_coeffs* out);
double kpo_hyp2f1(double a, double b, double c, double z);
double kpo_variable_kappa(double sigma, double beta, double kappa_default);
double kpo_symphony_rho_v(double nu, double ne, double thetae, double b_cgs,
                          double theta, double min_sin_theta);
void kpo_bessel_k012(double x, double* k0, double* k1, double* k2);
void kpo_riaf_coefficients(const kp_pass_a_params* params, const kp_riaf_model* model,
                           const kpo_state* s, const kpo_perturbation* pert,
                           kpo_coeffs* out);
void kpo_riaf_fluid_velocity(const kp_pass_a_params* params, const kp_riaf_model* model,
                             const double x[4], double u[4]);
double kpo_riaf_isco_radius(double spin);
void kpo_semi_analytic_step(kpo_stokes* s, const kpo_coeffs* c, double dlambda);
void kpo_transform_to_observer_basis(const kpo_stokes* in, double r11, double r12,
                                     double r21, double r22, kpo_stokes* out);
int32_t kpo_radiation_substep_count(const kpo_coeffs* c, double geometric_step,
                                    double dlambda_scale, int32_t requested_substeps,
                                    double max_radiation_step, double max_radiation_depth,
                                    double max_absorption_depth, double max_faraday_depth);

/* ---- passes and images ----------------------------------------------- */
/* Pass A segment endpoint for one pixel (BL). Returns valid flag. */
/* direct_only pieces (pass_a_core.hpp:202-275): the vertical position and
 * velocity of a state (fm NULL: Boyer-Lindquist, else FMKS with this map),
 * the vertical-turn bracket, and the truncation of one RK step (h and last
 * updated; returns 1 when the step brackets a turn). */
void kpo_vertical_pv(const kpo_fmks* fm, const kpo_state* s, double* z, double* dzds);
int32_t kpo_vertical_turn_bracket(double z0, double v0, double z1, double v1, double h, double radius_scale);
int32_t kpo_truncate_vertical_turn(double mass, double spin, const kpo_fmks* fm, const kpo_state* first,
                                   kpo_state* last, double* h, int32_t composed);
/* Equatorial emission wedge pieces (emission_selection.hpp:21-122) on metric
 * kind 0 (BL), 1 (FMKS with fm) or 2 (a flat test path: zero connection,
 * radius 1, height ratio x[3]). */
int32_t kpo_wedge_substeps(const kp_emission_selection* sel, int32_t kind, const kpo_fmks* fm,
                           const kpo_state* first, const kpo_state* middle, const kpo_state* last, double outer);
int32_t kpo_refine_wedge(const kp_emission_selection* sel, int32_t kind, const kpo_fmks* fm,
                         const kpo_state* first, kpo_state* middle, kpo_state* last, double* h, int32_t composed,
                         double outer);
void kpo_select_emission(const kp_emission_selection* sel, int32_t kind, const kpo_fmks* fm, const kpo_state* s,
                         kpo_coeffs* c);
double kpo_height_ratio(int32_t kind, const kpo_fmks* fm, const kpo_state* s);
void kpo_path_state(int32_t kind, double mass, double spin, const kpo_fmks* fm, const kpo_state* first, double h,
                    int32_t composed, kpo_state* out);
/* One RK4 step of the transport ODE (fm NULL: Boyer-Lindquist). */
void kpo_rk4_step(double mass, double spin, const kpo_fmks* fm, const kpo_state* s, double h, kpo_state* out);

int32_t kpo_bl_pass_a_endpoint(const kp_pass_a_params* params, int32_t pixel,
                               kpo_state* endpoint, int32_t* steps, int32_t* reason);

/* Whole image: pass A then pass B per pixel, OpenMP static schedule over
 * pixels.  pixel_begin/pixel_end select a sub-range (end<0: all).  Output
 * arrays are indexed by absolute pixel and may be NULL. */
kp_status kpo_run_riaf_image(const kp_pass_a_params* params, const kp_riaf_model* model,
                             const kpo_perturbation* pert,
                             int32_t pixel_begin, int32_t pixel_end,
                             int32_t nthreads, kp_image_data* out, kpo_stats* stats);

/* Multi-frequency image on shared geodesics; image_i/q/u/v point to
 * [nfreq][npix] arrays.  ctrl lists the step-control frequencies. */
kp_status kpo_run_riaf_multifrequency_image(const kp_pass_a_params* params, const kp_riaf_model* model,
                                            const kpo_perturbation* pert, const double* freqs,
                                            int32_t nfreq, const double* ctrl, int32_t nctrl,
                                            int32_t pixel_begin, int32_t pixel_end, int32_t nthreads,
                                            kp_image_data* out, kpo_stats* stats);

/* Parameter responses (src/diagnostics/response.hpp:25-191): one step of
 * the channels (bins source tags, four mechanism derivatives per bin, four
 * reruns; data[4 ch + k]), the coefficient scaling of the "coefficients"
 * family, and the RIAF image with one family's responses (parameter 1
 * density, 2 temperature, 3 magnetic; one radial bin; resp [36][npix]). */
void kpo_response_transfer_step(int32_t bins, int32_t responses, double step, double* data, int32_t bin,
                                const kpo_stokes* incoming, const kpo_coeffs* base, const kpo_coeffs* variants,
                                double dl);
void kpo_response_scale_coefficients(const kpo_coeffs* c, double q, kpo_coeffs* out);
kp_status kpo_run_riaf_response_image(const kp_pass_a_params* p, const kp_riaf_model* model, int32_t parameter,
                                      double step, int32_t nthreads, kp_image_data* out, double* resp,
                                      kpo_stats* stats);

kp_status kpo_run_riaf_image_strided(const kp_pass_a_params* params, const kp_riaf_model* model,
                                     const kpo_perturbation* pert, int32_t offset, int32_t stride,
                                     int32_t count, int32_t nthreads, kp_image_data* out,
                                     kpo_stats* stats);

/* ---- FMKS camera and the GRMHD grid model ------------------------------ */
void kpo_fmks_camera_tetrad(const kp_pass_a_params* params, const kpo_fmks* f, double xcam[4],
                            double tetrad[16]);
void kpo_fmks_camera_ray(const kp_pass_a_params* params, const kpo_fmks* f, int32_t pixel, kpo_state* out);
/* u^mu and b^mu at the centre of one cell from its primitives (host cells). */
void kpo_grmhd_cell_four_vectors(const kp_grmhd_model* model, const kp_grmhd_cell* cells, int32_t i, int32_t j,
                                 int32_t k, double ucon[4], double bcon[4]);
/* Derived scalars of every cell from its primitives (host cells). */
void kpo_grmhd_derive_scalars(const kp_grmhd_model* model, kp_grmhd_cell* cells);
double kpo_grmhd_sample_prim(const kp_grmhd_model* model, const kp_grmhd_cell* cells, int v, double x1,
                             double x2, double phi);
double kpo_grmhd_dlambda_scale(const kp_grmhd_model* model);
void kpo_grmhd_coefficients(const kp_grmhd_model* model, const kpo_state* s, const kpo_perturbation* pert,
                            int faraday_rotation, kpo_coeffs* out);
/* GRMHD image (FMKS); freqs/ctrl NULL: the model's freq_cgs.  model->cells: host. */
kp_status kpo_run_grmhd_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                              const kpo_perturbation* pert, const double* freqs, int32_t nfreq,
                              const double* ctrl, int32_t nctrl, int32_t pixel_begin, int32_t pixel_end,
                              int32_t nthreads, kp_image_data* out, kpo_stats* stats);

/* ---- Slow light -------------------------------------------------------- */
double kpo_interpolate_time_value(double a, double b, double fraction);
double kpo_slow_light_fraction(double t_lo, double t_hi, double t_obs, double x0);
int32_t kpo_slow_light_pair(const double* times, int32_t nsnap, double t);
void kpo_slow_light_coefficients(const kp_grmhd_model* model, const kp_slow_light* sequence, int32_t pair,
                                 const kpo_state* s, const kpo_perturbation* pert, int faraday_rotation,
                                 kpo_coeffs* out);
/* Slow-light GRMHD image; sequence->cells are HOST arrays.  waiting (may be
 * NULL) receives {count, min, max} of the times decoupled stepping needed
 * outside the sequence; KP_TIME_SUPPORT when the count is nonzero. */
kp_status kpo_run_grmhd_slow_light_image(const kp_pass_a_params* params, const kp_grmhd_model* model,
                                         const kp_slow_light* sequence, const kpo_perturbation* pert,
                                         const double* freqs, int32_t nfreq,
                                         int32_t pixel_begin, int32_t pixel_end, int32_t nthreads,
                                         kp_image_data* out, kpo_stats* stats, double waiting[3]);

#ifdef __cplusplus
}
#endif

#endif
