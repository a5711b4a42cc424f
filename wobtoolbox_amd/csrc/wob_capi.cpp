// C ABI of the MI355X walk-on-boundary engine (declared in include/wob_c.h).
//
// Owns the device copy of each scene (wide BVH, leaf-ordered primitives, element
// attributes, CDFs and guide tables) and a per-scene scratch buffer for per-lane partial
// sums; validates every request, maps (model, problem, estimator) to the sampling flavour
// the reference's Estimator specialisation uses (estimator.cuh:158-168, 255-347, 423-448)
// and launches the kernels on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "../../include/wob_c.h"
#include "scene_build.hpp"
#include "wob_launch.hpp"

namespace {

thread_local std::string g_last_error;

wob_status fail(wob_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

wob_status hip_fail(hipError_t e, const char *what) {
    return fail(WOB_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define WOB_HIP(call, what)                               \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(e_, what);  \
    } while (0)

}  // namespace

struct wob_scene {
    wobhost::HostScene host;
    wobdev::DevScene dev{};
    void *bufs[8] = {nullptr};
    float *scratch = nullptr;
    size_t scratch_bytes = 0;
    uint64_t device_bytes = 0;
    int device = 0;
};

namespace {

template <typename T> wob_status upload(wob_scene *s, int slot, const std::vector<T> &v, const T **out) {
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc scene buffer: ") + hipGetErrorString(e));
    s->bufs[slot] = p;
    s->device_bytes += bytes;
    if (!v.empty()) WOB_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy scene");
    *out = (const T *)p;
    return WOB_OK;
}

void free_scene(wob_scene *s) {
    for (auto &b : s->bufs)
        if (b) { (void)hipFree(b); b = nullptr; }
    if (s->scratch) (void)hipFree(s->scratch);
    s->scratch = nullptr;
}

wob_status ensure_scratch(wob_scene *s, size_t bytes) {
    if (bytes <= s->scratch_bytes) return WOB_OK;
    if (s->scratch) WOB_HIP(hipFree(s->scratch), "hipFree scratch");
    s->scratch = nullptr;
    s->scratch_bytes = 0;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc scratch: ") + hipGetErrorString(e));
    s->scratch = (float *)p;
    s->scratch_bytes = bytes;
    return WOB_OK;
}

wob_status check_est(const wob_scene *scene, const wob_estimator_desc *est, int &flavour) {
    if (!scene || !est) return fail(WOB_INVALID_ARG, "null scene or estimator descriptor");
    if (est->equation_kernel != WOB_POISSON_KERNEL) return fail(WOB_INVALID_ENUM, "equation_kernel must be PoissonKernel");
    if (est->bie_model < 0 || est->bie_model > 2) return fail(WOB_INVALID_ENUM, "bie_model out of range");
    if (est->domain < 0 || est->domain > 1) return fail(WOB_INVALID_ENUM, "domain out of range");
    if (est->estimator < 0 || est->estimator > 1) return fail(WOB_INVALID_ENUM, "estimator out of range");
    if (est->problem < 0 || est->problem > 3) return fail(WOB_INVALID_ENUM, "problem out of range");
    if (est->path_length == 0) return fail(WOB_INVALID_ARG, "path_length must be >= 1");
    if (est->num_resampling_candidates == 0) return fail(WOB_INVALID_ARG, "num_resampling_candidates must be >= 1");
    if (est->num_volume_samples != 0)
        return fail(WOB_UNSUPPORTED_COMBO, "volume sources are not enabled in this build (num_volume_samples must be 0)");
    if (est->bie_model == WOB_DOUBLE_LAYER && est->problem != WOB_DIRICHLET_PROBLEM)
        return fail(WOB_UNSUPPORTED_COMBO,
                    "the double-layer model supports Dirichlet problems only (bie_models.cuh:383, 394)");
    if (!(est->eps > 0.f)) return fail(WOB_INVALID_ARG, "eps must be > 0");
    const bool needs_c = est->bie_model != WOB_DOUBLE_LAYER;
    if (needs_c && !(est->first_kind_equation_scaling_constant > 0.f))
        return fail(WOB_INVALID_ARG, "first_kind_equation_scaling_constant must be > 0");
    if (est->estimator == WOB_BACKWARD) {
        if (est->bie_model == WOB_DOUBLE_LAYER && est->problem == WOB_DIRICHLET_PROBLEM) flavour = wobdev::FL_LINE_LINE;
        else if (est->bie_model == WOB_DIRECT && est->problem == WOB_NEUMANN_PROBLEM) flavour = wobdev::FL_LINE_BV;
        else flavour = wobdev::FL_UNIFORM;
    } else {
        flavour = (est->bie_model == WOB_SINGLE_LAYER && est->problem == WOB_NEUMANN_PROBLEM) ? wobdev::FL_FWD_LINE
                                                                                           : wobdev::FL_FWD_UNIFORM;
    }
    const bool uses_bv = flavour == wobdev::FL_LINE_BV || flavour == wobdev::FL_FWD_LINE;
    if (uses_bv && !(scene->host.total_bv > 0.f))
        return fail(WOB_INVALID_ARG, "boundary-value sampling needs nonzero boundary data");
    if (!(scene->host.total_area > 0.f)) return fail(WOB_INVALID_ARG, "scene has zero total area");
    return WOB_OK;
}

wob_status run_walk(const wob_scene *cscene, const wob_estimator_desc *est, const float *d_points, uint32_t n_points,
                    uint32_t point_base, uint32_t paths, uint64_t seed, uint64_t path_base, uint32_t stream_id,
                    double *d_sum, double *d_sumsq, wob_launch_config *launch, void *stream, bool boundary) {
    int flavour = 0;
    wob_status st = check_est(cscene, est, flavour);
    if (st != WOB_OK) return st;
    if (n_points == 0 || paths == 0) return WOB_OK;
    if (!d_points || !d_sum || !d_sumsq) return fail(WOB_INVALID_ARG, "null device buffer");
    wob_scene *scene = const_cast<wob_scene *>(cscene);
    const int D = (int)scene->host.dim;
    const uint32_t V = boundary ? 1u : (uint32_t)D + 1u;

    uint32_t block = launch && launch->block_size ? launch->block_size : 128u;
    if (block < 64 || block > 256 || block % 64 != 0) return fail(WOB_INVALID_ARG, "block_size must be 64, 128, 192 or 256");
    uint32_t ppl = launch && launch->paths_per_lane ? launch->paths_per_lane : 0u;
    if (ppl == 0) {
        // about a million work items keeps every CU busy with room for load balancing
        const double total = (double)n_points * paths;
        ppl = (uint32_t)std::max(1.0, std::ceil(total / (1u << 20)));
    }
    ppl = std::min(ppl, paths);
    uint32_t chunks = (paths + ppl - 1) / ppl;
    ppl = (paths + chunks - 1) / chunks;     // balance the chunks
    chunks = (paths + ppl - 1) / ppl;
    const uint64_t n_items = (uint64_t)n_points * chunks;
    if (n_items >= (1ull << 31)) return fail(WOB_INVALID_ARG, "too many work items; split the point set");
    st = ensure_scratch(scene, (size_t)2 * V * n_items * sizeof(float));
    if (st != WOB_OK) return st;

    wobl::WalkArgs a;
    a.scene = scene->dev;
    a.points = d_points;
    a.partials = scene->scratch;
    a.n_points = n_points;
    a.chunks = chunks;
    a.paths_per_point = paths;
    a.paths_per_chunk = ppl;
    a.point_base = point_base;
    a.stream = stream_id;
    a.seed = seed;
    a.path_base = path_base;
    a.dsign = est->domain == WOB_INTERIOR ? 1.0f : -1.0f;
    a.eps = est->eps;
    a.fkesc = est->first_kind_equation_scaling_constant;
    a.path_length = est->path_length;
    a.nrc = est->num_resampling_candidates;
    hipStream_t hs = (hipStream_t)stream;
    const bool timed = launch && (launch->flags & WOB_LAUNCH_TIME_WALK);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (timed) {
        WOB_HIP(hipEventCreate(&ev[0]), "hipEventCreate");
        WOB_HIP(hipEventCreate(&ev[1]), "hipEventCreate");
        WOB_HIP(hipEventRecord(ev[0], hs), "hipEventRecord");
    }
    WOB_HIP(wobl::launch_walk(D, boundary, a, est->bie_model, est->estimator, flavour, block, hs), "walk launch");
    if (timed) WOB_HIP(hipEventRecord(ev[1], hs), "hipEventRecord");
    WOB_HIP(wobl::launch_reduce(scene->scratch, n_points, chunks, V, d_sum, d_sumsq, hs), "reduce launch");
    if (launch) {
        launch->used_paths_per_lane = ppl;
        launch->used_work_items = (uint32_t)n_items;
    }
    if (timed) {
        WOB_HIP(hipEventSynchronize(ev[1]), "hipEventSynchronize");
        float ms = 0.f;
        WOB_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
        launch->walk_ms = ms;
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    return WOB_OK;
}

}  // namespace

extern "C" {

wob_status wob_scene_create(const wob_scene_desc *desc, wob_scene **out_scene) {
    if (!desc || !out_scene) return fail(WOB_INVALID_ARG, "null argument");
    *out_scene = nullptr;
    wob_scene *s = new (std::nothrow) wob_scene();
    if (!s) return fail(WOB_OOM, "host allocation failed");
    std::string err;
    try {
        err = wobhost::build_scene(*desc, s->host);
    } catch (const std::bad_alloc &) {
        delete s;
        return fail(WOB_OOM, "host allocation failed during scene build");
    }
    if (!err.empty()) {
        delete s;
        return fail(WOB_INVALID_ARG, err);
    }
    if (s->host.stack_need > (uint32_t)wobl::kStackEntries) {
        const std::string m = "BVH traversal stack bound " + std::to_string(s->host.stack_need) + " exceeds " +
                              std::to_string(wobl::kStackEntries);
        delete s;
        return fail(WOB_INVALID_ARG, m);
    }
    (void)hipGetDevice(&s->device);
    const auto &h = s->host;
    wob_status st;
    const float *nodes = nullptr, *prims = nullptr, *attr = nullptr, *acdf = nullptr, *bcdf = nullptr;
    const uint32_t *pid = nullptr, *ag = nullptr, *bg = nullptr;
    if ((st = upload(s, 0, h.nodes, &nodes)) != WOB_OK || (st = upload(s, 1, h.prim_geom, &prims)) != WOB_OK ||
        (st = upload(s, 2, h.prim_id, &pid)) != WOB_OK || (st = upload(s, 3, h.elem_attr, &attr)) != WOB_OK ||
        (st = upload(s, 4, h.area_cdf, &acdf)) != WOB_OK || (st = upload(s, 5, h.area_guide, &ag)) != WOB_OK ||
        (st = upload(s, 6, h.bv_cdf, &bcdf)) != WOB_OK || (st = upload(s, 7, h.bv_guide, &bg)) != WOB_OK) {
        free_scene(s);
        delete s;
        return st;
    }
    s->dev.nodes = (const float4 *)nodes;
    s->dev.prims = (const float4 *)prims;
    s->dev.prim_id = pid;
    s->dev.attr = (const float4 *)attr;
    s->dev.area_cdf = acdf;
    s->dev.area_guide = ag;
    s->dev.bv_cdf = bcdf;
    s->dev.bv_guide = bg;
    s->dev.n = h.n;
    s->dev.total_area = h.total_area;
    s->dev.total_bv = h.total_bv;
    *out_scene = s;
    return WOB_OK;
}

wob_status wob_scene_destroy(wob_scene *scene) {
    if (!scene) return WOB_OK;
    free_scene(scene);
    delete scene;
    return WOB_OK;
}

wob_status wob_scene_get_info(const wob_scene *scene, wob_scene_info *info) {
    if (!scene || !info) return fail(WOB_INVALID_ARG, "null argument");
    info->dim = scene->host.dim;
    info->n_elements = scene->host.n;
    info->n_nodes = scene->host.n_nodes;
    info->depth = scene->host.depth;
    info->total_area = scene->host.total_area;
    info->total_boundary_value = scene->host.total_bv;
    info->device_bytes = scene->device_bytes;
    return WOB_OK;
}

wob_status wob_scene_set_volume_grid(wob_scene *scene, uint32_t, float, const float *) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    return fail(WOB_UNSUPPORTED_COMBO, "grid volume sources are not enabled in this build");
}

wob_status wob_estimate_domain(const wob_scene *scene, const wob_estimator_desc *est, const float *d_points,
                               uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                               uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                               wob_launch_config *launch, void *stream) {
    return run_walk(scene, est, d_points, n_points, point_base, paths_per_point, seed, path_base, stream_id, d_sum,
                    d_sumsq, launch, stream, false);
}

wob_status wob_estimate_boundary(const wob_scene *scene, const wob_estimator_desc *est, const float *d_records,
                                 uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                                 uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                                 wob_launch_config *launch, void *stream) {
    return run_walk(scene, est, d_records, n_points, point_base, paths_per_point, seed, path_base, stream_id, d_sum,
                    d_sumsq, launch, stream, true);
}

wob_status wob_intersect_lines(const wob_scene *scene, const float *d_origins, const float *d_dirs, uint32_t n,
                               uint32_t max_hits, uint32_t *d_counts, uint32_t *d_ids, float *d_t, void *stream) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n == 0) return WOB_OK;
    if (!d_origins || !d_dirs || !d_counts || (max_hits && (!d_ids || !d_t)))
        return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_intersect((int)scene->host.dim, scene->dev, n, d_origins, d_dirs, max_hits, d_counts, d_ids,
                                   d_t, (hipStream_t)stream),
            "intersect launch");
    return WOB_OK;
}

wob_status wob_eval_kernels(uint32_t dim, uint32_t n, const float *d_y, const float *d_x, const float *d_ny,
                            const float *d_nx, float eps, float *d_out, void *stream) {
    if (dim != 2 && dim != 3) return fail(WOB_INVALID_ARG, "dim must be 2 or 3");
    if (n == 0) return WOB_OK;
    if (!d_y || !d_x || !d_ny || !d_nx || !d_out) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_eval_kernels((int)dim, n, d_y, d_x, d_ny, d_nx, eps, d_out, (hipStream_t)stream),
            "eval_kernels launch");
    return WOB_OK;
}

wob_status wob_eval_weights(uint32_t dim, int32_t bie_model, int32_t domain, float eps, float fkesc, uint32_t n,
                            const float *d_x, const float *d_y, const float *d_inv_pdf, const float *d_xdomain,
                            const float *d_u_split, float *d_out, void *stream) {
    if (dim != 2 && dim != 3) return fail(WOB_INVALID_ARG, "dim must be 2 or 3");
    if (bie_model < 0 || bie_model > 2) return fail(WOB_INVALID_ENUM, "bie_model out of range");
    if (domain < 0 || domain > 1) return fail(WOB_INVALID_ENUM, "domain out of range");
    if (n == 0) return WOB_OK;
    if (!d_x || !d_y || !d_inv_pdf || !d_xdomain || !d_u_split || !d_out) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_eval_weights((int)dim, bie_model, domain == WOB_INTERIOR ? 1.f : -1.f, eps, fkesc, n, d_x, d_y,
                                      d_inv_pdf, d_xdomain, d_u_split, d_out, (hipStream_t)stream),
            "eval_weights launch");
    return WOB_OK;
}

wob_status wob_philox4x32_10(uint32_t n, const uint32_t *d_ctr, const uint32_t *d_key, uint32_t *d_out, void *stream) {
    if (n == 0) return WOB_OK;
    if (!d_ctr || !d_key || !d_out) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_philox(n, d_ctr, d_key, d_out, (hipStream_t)stream), "philox launch");
    return WOB_OK;
}

int32_t wob_string_to_enum(const char *s) {
    if (!s) return -1;
    static const struct { const char *name; int32_t v; } table[] = {
        {"PoissonKernel", 0},        {"IndirectSingleLayer", 0}, {"IndirectDoubleLayer", 1}, {"Direct", 2},
        {"InteriorDomain", 0},       {"ExteriorDomain", 1},      {"BackwardEstimator", 0},   {"ForwardEstimator", 1},
        {"DirichletProblem", 0},     {"NeumannProblem", 1},      {"RobinProblem", 2},        {"MixedProblem", 3},
        {"Dirichlet", 0},            {"Neumann", 1},             {"Robin", 2}};
    for (const auto &t : table)
        if (std::strcmp(s, t.name) == 0) return t.v;
    return -1;
}

const char *wob_enum_to_string(const char *kind, int32_t v) {
    if (!kind) return "Invalid";
    const std::string k(kind);
    if (k == "equation_kernel") return v == 0 ? "PoissonKernel" : "InvalidEqiuationKernelType";
    if (k == "bie_model") {
        static const char *n[] = {"IndirectSingleLayer", "IndirectDoubleLayer", "Direct"};
        return v >= 0 && v < 3 ? n[v] : "InvalidBIEModelType";
    }
    if (k == "domain") return v == 0 ? "InteriorDomain" : v == 1 ? "ExteriorDomain" : "InvalidDomainType";
    if (k == "estimator") return v == 0 ? "BackwardEstimator" : v == 1 ? "ForwardEstimator" : "InvalidEstimatorType";
    if (k == "problem") {
        static const char *n[] = {"DirichletProblem", "NeumannProblem", "RobinProblem", "MixedProblem"};
        return v >= 0 && v < 4 ? n[v] : "InvalidProblemType";
    }
    return "Invalid";
}

const char *wob_last_error_message(void) { return g_last_error.c_str(); }

const char *wob_version(void) { return "wob-mi355x 0.1"; }

}  // extern "C"
