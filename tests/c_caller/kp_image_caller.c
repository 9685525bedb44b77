/*
 * A plain-C caller of include/kp_image.h: the flow of the reference image
 * tool's RIAF path (tools/kpolaris_model_image.cpp:3396-3421, 3940-3990) over
 * the C ABI.  It loads a parameter file through the option layer, overrides
 * keys given as key=value arguments, sizes a caller-owned workspace with
 * kp_image_workspace_size, images with host outputs, packs the image with the
 * reference's rules and prints one line of results.
 *
 *   kp_image_caller <parameter file> [key=value ...]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kp_image.h"

static int check(kp_status s, const char* what) {
    if (s != KP_OK) {
        fprintf(stderr, "%s failed: status %d: %s\n", what, (int)s, kp_last_error());
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <parameter file> [key=value ...]\n", argv[0]);
        return 2;
    }
    if (kp_abi_version() != KP_ABI_VERSION) {
        fprintf(stderr, "library ABI %d, header ABI %d\n", kp_abi_version(), KP_ABI_VERSION);
        return 1;
    }
    kp_options* o = kp_options_create();
    if (check(kp_options_load_file(o, argv[1]), "kp_options_load_file")) return 1;
    for (int a = 2; a < argc; ++a) {
        char key[128];
        const char* eq = strchr(argv[a], '=');
        if (!eq || (size_t)(eq - argv[a]) >= sizeof(key)) {
            fprintf(stderr, "expected key=value, got %s\n", argv[a]);
            return 2;
        }
        memcpy(key, argv[a], (size_t)(eq - argv[a]));
        key[eq - argv[a]] = '\0';
        if (check(kp_options_set(o, key, eq + 1), "kp_options_set")) return 1;
    }
    kp_pass_a_params p;
    kp_riaf_model m;
    double flux_scale = 0.0;
    int32_t qu_sign = 1;
    if (check(kp_options_resolve_riaf(o, &p, &m, &flux_scale, &qu_sign), "kp_options_resolve_riaf")) return 1;
    kp_options_destroy(o);

    const int64_t npix = (int64_t)p.camera.nx * p.camera.ny;
    kp_run_options opt;
    kp_run_options_default(&opt);
    uint64_t ws_bytes = 0;
    if (check(kp_image_workspace_size(npix, 1, 0, &opt, &ws_bytes), "kp_image_workspace_size")) return 1;

    double* real = calloc((size_t)npix * 4, sizeof(double));
    int32_t* ints = calloc((size_t)npix * 3, sizeof(int32_t));
    double* packed = calloc((size_t)npix * 4, sizeof(double));
    if (!real || !ints || !packed) return 1;
    kp_image_data out;
    memset(&out, 0, sizeof(out));
    out.image_i = real;
    out.image_q = real + npix;
    out.image_u = real + 2 * npix;
    out.image_v = real + 3 * npix;
    out.pass_a_steps = ints;
    out.steps = ints + npix;
    out.reason = ints + 2 * npix;
    kp_image_stats st;
    if (check(kp_run_riaf_image_host(&p, &m, 1, &opt, &out, &st), "kp_run_riaf_image_host")) return 1;
    double flux[4] = {0, 0, 0, 0};
    if (check(kp_pack_image(out.image_i, out.image_q, out.image_u, out.image_v, out.reason, npix, m.freq_cgs, qu_sign,
                            flux_scale, packed, packed + npix, packed + 2 * npix, packed + 3 * npix, flux),
              "kp_pack_image"))
        return 1;
    long long camera = 0, steps = 0;
    for (int64_t i = 0; i < npix; ++i) {
        camera += out.reason[i] == KP_REASON_REACHED_CAMERA;
        steps += out.steps[i];
    }
    double sum_i = 0.0;
    for (int64_t i = 0; i < npix; ++i) sum_i += out.image_i[i];
    printf("npix %lld camera %lld steps %lld workspace %llu sum_i %.17g flux_i %.17g flux_q %.17g flux_u %.17g "
           "flux_v %.17g\n",
           (long long)npix, camera, steps, (unsigned long long)ws_bytes, sum_i, flux[0], flux[1], flux[2], flux[3]);
    free(real);
    free(ints);
    free(packed);
    return 0;
}
