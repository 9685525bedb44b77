/*
 * kp_io.h -- HDF5 input and output around the imaging ABI (libkpolaris_io.so).
 *
 * Two on-disk contracts of the reference, in plain C:
 *  - iHARM GRMHD dumps as the reference's loader reads them
 *    (src/grmhd/iharm_loader.cpp:262-396): /header scalars, the FMKS (or
 *    MMKS) geometry group, and /prims [n1][n2][n3][nprim] read as float;
 *  - the `kpolaris_image` v1 image product (docs/hdf5_schema.md:81-200), as
 *    the image tool writes it (tools/kpolaris_model_image.cpp:4283-4880), and
 *    the publication norms of scripts/compare_stokes_hdf5.py:66-143.
 * Errors return a kp_status with the message in kp_last_error().
 */
#ifndef KP_IO_H
#define KP_IO_H

#include "kp_image.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- iHARM dumps ---------------------------------------------------------- */
typedef struct kp_iharm_header {
    int32_t n1, n2, n3;
    int32_t nprim;      /* primitives per cell in /prims (>= 8; the first 8 are read) */
    double a, gam;      /* /header/a, /header/gam (or /header/gamma) */
    double startx1, startx2, startx3, dx1, dx2, dx3;  /* /header/geom */
    /* /header/geom/fmks (or /header/geom/mmks) */
    double hslope, mks_smooth, poly_alpha, poly_xt, r_in, r_out;
    double poly_norm;   /* derived: 0.5 pi / (1 + 1 / (poly_alpha + 1) / poly_xt^poly_alpha) */
} kp_iharm_header;

/* Reads and validates the header (the loader's rules, iharm_loader.cpp:270-330). */
kp_status kp_iharm_read_header(const char* path, kp_iharm_header* header);
/* Reads /prims as float into n1*n2*n3 host cells (primitives only; derived
 * scalars zeroed), rejecting a shape that disagrees with the header or a
 * non-finite value among the first eight primitives (iharm_loader.cpp:363-396). */
kp_status kp_iharm_read_cells(const char* path, const kp_iharm_header* header, kp_grmhd_cell* cells);
/* Writes a dump in the same layout: header scalars, geom and geom/fmks
 * groups, /prims as float32 [n1][n2][n3][nprim] from `prims`. */
kp_status kp_iharm_write(const char* path, const kp_iharm_header* header, const float* prims);
/* Copies the header's grid and geometry into `model` (other fields kept). */
void kp_iharm_apply_header(const kp_iharm_header* header, kp_grmhd_model* model);

/* A time sequence of iHARM dumps streamed by kp_run_grmhd_slow_light_stream:
 * kp_iharm_load_snapshot is its kp_snapshot_loader, with a kp_iharm_sequence
 * as `user` and derive_scalars = 1.  Every dump must have the header's grid
 * and coordinate mapping (fluid time interpolation needs "a common coordinate
 * mapping/domain", src/image/slow_light_driver.hpp:31-42); the loader reads
 * each dump as kp_iharm_read_cells does (src/grmhd/iharm_loader.cpp:262-396).
 * HDF5 calls happen on the scheduler's read-ahead thread, one at a time. */
typedef struct kp_iharm_sequence {
    int32_t nsnap;
    int32_t reserved0;
    const char* const* paths;   /* [nsnap] dump files in time order */
    kp_iharm_header header;     /* the grid of every dump (kp_iharm_read_header of the first) */
} kp_iharm_sequence;
kp_status kp_iharm_load_snapshot(void* user, int32_t index, kp_grmhd_cell* cells, uint64_t ncells);

/* ---- kpolaris_image v1 ------------------------------------------------------ */
typedef struct kp_image_file_meta {
    const char* model;           /* "riaf", "iharm", ... */
    const char* camera;          /* "pinhole" */
    const char* coordinate;      /* "bl", "fmks", ... */
    const char* metric;          /* "kerr" */
    const char* evpa_0;          /* "N" (passive Q,U flip) or "W" */
    const char* parameter_file;  /* may be NULL */
    double mbh_solar;
    double dsource_pc;
    double spin;
    double intensity_to_flux_jy_per_pixel; /* NaN when the scale is invalid */
    int32_t nfreq;
    int32_t reserved0;
    const double* frequency_hz;  /* [nfreq] */
} kp_image_file_meta;

/* Writes the image: `image` holds HOST arrays, Stokes invariant and
 * frequency-major [nfreq][npix] as the kernels return them (pixel
 * p = iy * nx + ix), diagnostics [npix].  The Stokes datasets are packed as
 * the image tool packs them (kp_pack_image): zero unless reason == 1, Q and U
 * flipped for evpa_0 = N, times nu^3 for I..V and kept invariant for
 * I_inv..V_inv.  Fluxes use intensity_to_flux_jy_per_pixel. */
kp_status kp_image_write_hdf5(const char* path, const kp_image_file_meta* meta, const kp_pass_a_params* params,
                              const kp_image_data* image);

/* compare_stokes_hdf5.py norms of one Stokes array (candidate vs reference). */
typedef struct kp_stokes_norms {
    double reference_sum, candidate_sum, sum_relative, l1_relative, rms_relative, linf_over_peak;
} kp_stokes_norms;
typedef struct kp_image_comparison {
    int32_t nx, ny;
    double frequency_hz;
    kp_stokes_norms stokes[4];          /* I, Q, U, V, both read in the N basis */
    double worst_relative_l1, worst_linf_over_peak;
    double grid_max_abs_difference[2];  /* x, y (tolerance 64 eps x scale) */
    /* reason, pass_a_steps, steps, total_steps: differing pixels, max |diff| */
    int64_t discrete_mismatched[4];
    int64_t discrete_max_abs_difference[4];
    /* closure_x, closure_k, final_null, frame_error, basis_identity_error */
    double continuous_l1_relative[5];
    double continuous_linf[5];
} kp_image_comparison;

/* Loads frame `frame`, frequency `freq_index` of both files (schema must be
 * kpolaris_image; Stokes finite), checks shapes, grids and frequencies as
 * the script does (KP_INVALID_ARGUMENT on a mismatch) and fills the norms
 * and the diagnostics comparison. */
kp_status kp_image_compare_hdf5(const char* reference, const char* candidate, int32_t frame, int32_t freq_index,
                                kp_image_comparison* out);

#ifdef __cplusplus
}
#endif

#endif /* KP_IO_H */
