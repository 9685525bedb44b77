// kp_internal.h -- host-side helpers shared by the kernels' translation unit
// and the host option layer.  Not part of the public ABI.
#pragma once

#include "../../include/kp_image.h"
#include "kp_errors.h"
#include "kp_riaf.h"

#include <string>

namespace kp_internal {

// ipole camera tetrad for the BL pinhole camera at (r, theta, phi=0):
// tetrad[0..15] = e0,e1,e2,e3 and gcov5 = (g00, g03, g11, g22, g33) there.
void bl_camera_tetrad(double mass, double spin, double radius, double inclination, double tetrad[16],
                      double gcov5[5]);

// Ray-independent RIAF constants in double.  perturbation: 0 none,
// 1 density, 2 temperature, 3 magnetic (log-scale families).
void riaf_constants(const kp_pass_a_params& p, const kp_riaf_model& m, int perturbation, double log_scale,
                    kpd::RiafDev<double>& out);

double isco_radius(double spin);

// ipole camera position, tetrad and metric there for the FMKS pinhole camera;
// fm = (startx1, hslope, mks_smooth, poly_alpha, poly_xt, poly_norm).
void fmks_camera_tetrad(double mass, double spin, const double fm[6], double radius, double inclination,
                        double xcam[4], double tetrad[16], double gcov16[16]);

}  // namespace kp_internal
