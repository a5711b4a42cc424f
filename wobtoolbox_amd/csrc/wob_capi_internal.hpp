// Helpers shared by the C-ABI translation units (wob_capi.cpp, wob_cv.cpp, wob_hcv.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/wob_c.h"
#include "scene_build.hpp"
#include "wob_device.hpp"

// record a message for wob_last_error_message() and return the status
wob_status wob_fail(wob_status s, const std::string &msg);
wob_status wob_hip_fail(hipError_t e, const char *what);

#define WOB_HIP(call, what)                                   \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return wob_hip_fail(e_, what);  \
    } while (0)

const wobhost::HostScene &wob_scene_host(const wob_scene *scene);
// takes ownership of d_table (device, [R][n][dim] floats); frees the previous table
wob_status wob_scene_bind_cv_table(wob_scene *scene, float *d_table, uint32_t R);
const wobdev::DevScene &wob_scene_dev(const wob_scene *scene);
// side of the domain a bound harmonic control variate serves: fitted for the interior, fitted
// for the exterior, or given explicitly (wob_scene_set_harmonic_cv: either); an exterior fit
// with sum_j w_j = 0 serves exterior double-layer walks too
enum { wob_hcv_interior = 0, wob_hcv_exterior = 1, wob_hcv_any = 2, wob_hcv_exterior_zero_charge = 3 };
// takes ownership of d_src (device, m float4 records); frees the previous sources; m == 0
// (d_src == nullptr) unbinds
wob_status wob_scene_bind_hcv(wob_scene *scene, float4 *d_src, uint32_t m, float c0, int side);
// the scene's device workspace of harmonic control-variate fits, grown to at least `bytes`
// (kept between fits; a fit on one stream at a time)
wob_status wob_scene_hcv_workspace(wob_scene *scene, size_t bytes, char **out);
