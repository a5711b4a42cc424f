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
#include <mutex>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/wob_c.h"
#include "scene_build.hpp"
#include "wob_capi_internal.hpp"
#include "wob_launch.hpp"

namespace {
thread_local std::string g_last_error;
}  // namespace

wob_status wob_fail(wob_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

wob_status wob_hip_fail(hipError_t e, const char *what) {
    return wob_fail(WOB_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
inline wob_status fail(wob_status s, const std::string &msg) { return wob_fail(s, msg); }
inline wob_status hip_fail(hipError_t e, const char *what) { return wob_hip_fail(e, what); }
}  // namespace

struct wob_scene {
    wobhost::HostScene host;
    wobdev::DevScene dev{};
    void *bufs[11] = {nullptr};
    float *scratch = nullptr;
    size_t scratch_bytes = 0;
    uint32_t *counter = nullptr;   // work counter of wavefront launches
    uint64_t device_bytes = 0;
    int device = 0;
    float *cv_table = nullptr;
    float4 *hcv_src = nullptr;
    int hcv_side = 0;       // the bound harmonic control variate: kHcvInterior / kHcvExterior / kHcvAny
    void *vol_bufs[3] = {nullptr, nullptr, nullptr};
    uint64_t version = 0;   // bumped whenever device data referenced by launches changes
    double *aux = nullptr;  // FP64 accumulators of VPL path phases
    size_t aux_bytes = 0;
    char *vpl = nullptr;    // VPL aggregation records and partial sums (apart from the walks' buffers)
    size_t vpl_bytes = 0;
    char *hcv_work = nullptr;   // harmonic control-variate fits: targets, Gram matrix, partial sums
    size_t hcv_work_bytes = 0;
};

const wobhost::HostScene &wob_scene_host(const wob_scene *scene) { return scene->host; }
const wobdev::DevScene &wob_scene_dev(const wob_scene *scene) { return scene->dev; }

wob_status wob_scene_hcv_workspace(wob_scene *s, size_t bytes, char **out) {
    if (bytes > s->hcv_work_bytes) {
        if (s->hcv_work) WOB_HIP(hipFree(s->hcv_work), "hipFree fit workspace");
        s->hcv_work = nullptr;
        s->hcv_work_bytes = 0;
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc fit workspace: ") + hipGetErrorString(e));
        s->hcv_work = (char *)p;
        s->hcv_work_bytes = bytes;
    }
    *out = s->hcv_work;
    return WOB_OK;
}

wob_status wob_scene_bind_hcv(wob_scene *scene, float4 *d_src, uint32_t m, float c0, int side) {
    if (scene->hcv_src) (void)hipFree(scene->hcv_src);
    scene->hcv_src = d_src;
    scene->hcv_side = side;
    scene->dev.hcv_src = d_src;
    scene->dev.hcv_m = d_src ? m : 0u;
    scene->dev.hcv_c0 = d_src ? c0 : 0.f;
    scene->version += 1;
    return WOB_OK;
}

wob_status wob_scene_bind_cv_table(wob_scene *scene, float *d_table, uint32_t R) {
    if (scene->cv_table) (void)hipFree(scene->cv_table);
    scene->cv_table = d_table;
    scene->dev.cv_table = d_table;
    scene->dev.cv_R = d_table ? R : 0u;
    scene->version += 1;
    return WOB_OK;
}

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

void free_volume(wob_scene *s) {
    for (auto &b : s->vol_bufs)
        if (b) { (void)hipFree(b); b = nullptr; }
    s->dev.vol = wobdev::VolSrc{};
    s->version += 1;
}

void free_scene(wob_scene *s) {
    free_volume(s);
    for (auto &b : s->bufs)
        if (b) { (void)hipFree(b); b = nullptr; }
    if (s->scratch) (void)hipFree(s->scratch);
    s->scratch = nullptr;
    if (s->counter) (void)hipFree(s->counter);
    s->counter = nullptr;
    if (s->cv_table) (void)hipFree(s->cv_table);
    s->cv_table = nullptr;
    if (s->hcv_src) (void)hipFree(s->hcv_src);
    s->hcv_src = nullptr;
    if (s->aux) (void)hipFree(s->aux);
    s->aux = nullptr;
    if (s->vpl) (void)hipFree(s->vpl);
    s->vpl = nullptr;
    if (s->hcv_work) (void)hipFree(s->hcv_work);
    s->hcv_work = nullptr;
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

wob_status ensure_aux(wob_scene *s, size_t bytes) {
    if (bytes <= s->aux_bytes) return WOB_OK;
    if (s->aux) WOB_HIP(hipFree(s->aux), "hipFree aux");
    s->aux = nullptr;
    s->aux_bytes = 0;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc aux: ") + hipGetErrorString(e));
    s->aux = (double *)p;
    s->aux_bytes = bytes;
    return WOB_OK;
}

wob_status ensure_vpl(wob_scene *s, size_t bytes) {
    if (bytes <= s->vpl_bytes) return WOB_OK;
    if (s->vpl) WOB_HIP(hipFree(s->vpl), "hipFree VPL buffer");
    s->vpl = nullptr;
    s->vpl_bytes = 0;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc VPL buffer: ") + hipGetErrorString(e));
    s->vpl = (char *)p;
    s->vpl_bytes = bytes;
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
    if (est->flags & ~(uint32_t)(WOB_EST_EXCLUDE_ORIGIN_ELEMENT | WOB_EST_CONTROL_VARIATE | WOB_EST_BOUNDARY_DENSITY |
                                 WOB_EST_INDEPENDENT_VOLUME_SAMPLES | WOB_EST_HARMONIC_CV))
        return fail(WOB_INVALID_ARG, "unknown estimator flags");
    if (est->flags & WOB_EST_CONTROL_VARIATE) {
        if (est->bie_model != WOB_SINGLE_LAYER || est->estimator != WOB_BACKWARD)
            return fail(WOB_UNSUPPORTED_COMBO, "the control variate needs the single-layer model and the backward estimator");
        if (scene->dev.cv_R < est->path_length)
            return fail(WOB_INVALID_ARG, "no control-variate table with path_length value functions is bound to the scene");
    }
    if (est->flags & WOB_EST_HARMONIC_CV) {
        if (est->bie_model == WOB_DIRECT && est->domain != WOB_INTERIOR)
            return fail(WOB_UNSUPPORTED_COMBO, "the harmonic control variate takes the direct model in interior domains "
                                               "only");
        if (scene->dev.hcv_m == 0)
            return fail(WOB_INVALID_ARG, "no harmonic control variate is bound to the scene "
                                         "(wob_scene_fit_harmonic_cv / wob_scene_set_harmonic_cv)");
        const int side = scene->hcv_side;
        if (side != wob_hcv_any) {
            if (est->domain == WOB_INTERIOR && side != wob_hcv_interior)
                return fail(WOB_INVALID_ARG, "the scene's harmonic control variate was fitted for the exterior domain "
                                             "(WOB_HCV_EXTERIOR); interior walks need an interior fit");
            if (est->domain != WOB_INTERIOR && side == wob_hcv_interior)
                return fail(WOB_INVALID_ARG, "the scene's harmonic control variate was fitted for the interior domain; "
                                             "exterior walks need a WOB_HCV_EXTERIOR fit");
            if (est->domain != WOB_INTERIOR && est->bie_model == WOB_DOUBLE_LAYER && side != wob_hcv_exterior_zero_charge)
                return fail(WOB_INVALID_ARG, "exterior double-layer walks need a WOB_HCV_EXTERIOR | WOB_HCV_ZERO_CHARGE "
                                             "fit (a double-layer potential decays at infinity)");
        }
    }
    if ((est->flags & WOB_EST_BOUNDARY_DENSITY) && est->estimator != WOB_BACKWARD)
        return fail(WOB_UNSUPPORTED_COMBO, "boundary density estimates need the backward estimator");
    if (est->num_resampling_candidates == 0) return fail(WOB_INVALID_ARG, "num_resampling_candidates must be >= 1");
    if (est->num_volume_samples != 0 && scene->dev.vol.kind == wobdev::kVolNone)
        return fail(WOB_INVALID_ARG, "num_volume_samples > 0 needs a volume source on the scene "
                                     "(wob_scene_set_volume_grid / wob_scene_set_volume_gaussians)");
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

uint32_t auto_paths_per_lane(double total_paths) {
    // about a million work items keeps every CU busy with room for load balancing
    return (uint32_t)std::max(1.0, std::ceil(total_paths / (1u << 20)));
}

// chunking of one entry: paths per chunk (balanced) and chunks per point
void chunking(uint32_t paths, uint32_t ppl, uint32_t &paths_per_chunk, uint32_t &chunks) {
    ppl = std::max(1u, std::min(ppl, paths));
    chunks = (paths + ppl - 1) / ppl;
    paths_per_chunk = (paths + chunks - 1) / chunks;
    chunks = (paths + paths_per_chunk - 1) / paths_per_chunk;
}

wobl::WalkEntry make_entry(const wob_scene *scene, const wob_estimator_desc *est, const float *points,
                           uint32_t n_points, uint32_t paths, uint32_t ppl, uint32_t point_base, uint32_t stream_id,
                           uint64_t path_base, double *sum, double *sumsq) {
    wobl::WalkEntry e{};
    e.scene = scene->dev;
    e.points = points;
    e.sum = sum;
    e.sumsq = sumsq;
    e.n_points = n_points;
    chunking(paths, ppl, e.paths_per_chunk, e.chunks);
    e.paths_per_point = paths;
    e.point_base = point_base;
    e.stream = stream_id;
    e.path_length = est->path_length;
    e.nrc = est->num_resampling_candidates;
    e.dsign = est->domain == WOB_INTERIOR ? 1.0f : -1.0f;
    e.eps = est->eps;
    e.fkesc = est->first_kind_equation_scaling_constant;
    e.path_base = path_base;
    e.flags = est->flags;
    e.cv_samples = est->cv_samples ? est->cv_samples : 64u;
    e.nvs = est->num_volume_samples;
    return e;
}

constexpr uint32_t kDefaultRegen = 32;

// Resolve a requested schedule: true = wavefront walk.  Automatic: the per-lane walk, whose
// line queries share traversal work across the wave; measured on MI355X it is the faster one
// for every line walk (2D circle, 1,024 segments, double layer: 15.0 vs 19.2 ms for 51,468
// points x 256 paths; level-5 icosphere, double layer: 24.8 vs 27.1 ms for 100k x 64).
wob_status resolve_schedule(uint32_t schedule, bool supported, int flavour, bool &wavefront) {
    (void)flavour;
    if (schedule > WOB_SCHEDULE_WAVEFRONT) return fail(WOB_INVALID_ARG, "unknown schedule");
    if (schedule == WOB_SCHEDULE_WAVEFRONT && !supported)
        return fail(WOB_UNSUPPORTED_COMBO, "the wavefront walk covers line-sampled backward domain walks without a "
                                           "volume term only");
    wavefront = supported && schedule == WOB_SCHEDULE_WAVEFRONT;
    return WOB_OK;
}

wob_status check_block(uint32_t block) {
    if (block < 64 || block > 256 || block % 64 != 0) return fail(WOB_INVALID_ARG, "block_size must be 64, 128, 192 or 256");
    return WOB_OK;
}

// Paths are numbered per point by a 32-bit counter word of the Philox stream: a call may use
// path indices up to 2^32 - 1 (path_base + paths <= 2^32).  Later paths of the same point
// need another seed or stream id, whose streams are independent of these.
constexpr uint64_t kPathLimit = 1ull << 32;
wob_status check_paths(uint64_t path_base, uint64_t paths) {
    if (path_base > kPathLimit || paths > kPathLimit - path_base)
        return fail(WOB_INVALID_ARG, "path_base + paths exceeds 2^32 paths per point: the Philox stream's path word is "
                                     "32 bits; continue with another seed or stream_id");
    return WOB_OK;
}

// Device memory a walk launch works in: the work counter of resident launches and the
// per-work-item partial sums (2 V floats per item).  run_walk takes them from the scene
// (allocated on first need) unless the caller passes its own.
struct WalkScratch {
    uint32_t *counter;
    float *partials;
    size_t partial_floats;
};

wob_status run_walk(const wob_scene *cscene, const wob_estimator_desc *est, const float *d_points, uint32_t n_points,
                    uint32_t point_base, uint32_t paths, uint64_t seed, uint64_t path_base, uint32_t stream_id,
                    double *d_sum, double *d_sumsq, wob_launch_config *launch, void *stream, bool boundary,
                    const WalkScratch *ws = nullptr) {
    // the path range first: it needs no scene, so the rule holds (and is testable) on its own
    wob_status st = check_paths(path_base, paths);
    if (st != WOB_OK) return st;
    int flavour = 0;
    if ((st = check_est(cscene, est, flavour)) != WOB_OK) return st;
    if (boundary && (est->flags & WOB_EST_HARMONIC_CV) && !(est->flags & WOB_EST_BOUNDARY_DENSITY))
        return fail(WOB_UNSUPPORTED_COMBO, "the harmonic control variate applies to domain estimates (and to boundary "
                                           "density walks, which then walk the residual)");
    if (n_points == 0 || paths == 0) return WOB_OK;
    if (!d_points || !d_sum || !d_sumsq) return fail(WOB_INVALID_ARG, "null device buffer");
    wob_scene *scene = const_cast<wob_scene *>(cscene);
    const int D = (int)scene->host.dim;
    const uint32_t V = boundary ? 1u : (uint32_t)D + 1u;

    const uint32_t block = launch && launch->block_size ? launch->block_size : 128u;
    if ((st = check_block(block)) != WOB_OK) return st;
    uint32_t ppl = launch && launch->paths_per_lane ? launch->paths_per_lane : 0u;
    if (ppl == 0) ppl = auto_paths_per_lane((double)n_points * paths);

    wobl::WalkLaunch L{};
    L.e0 = make_entry(scene, est, d_points, n_points, paths, ppl, point_base, stream_id, path_base, d_sum, d_sumsq);
    const uint64_t n_items = (uint64_t)n_points * L.e0.chunks;
    if (n_items >= (1ull << 31)) return fail(WOB_INVALID_ARG, "too many work items; split the point set");
    const size_t partial_floats = (size_t)2 * V * n_items;
    if (ws) {
        if (partial_floats > ws->partial_floats)
            return fail(WOB_INVALID_ARG, "the launch needs " + std::to_string(partial_floats * sizeof(float)) +
                                             " bytes of partial sums; the allocation-free scratch holds " +
                                             std::to_string(ws->partial_floats * sizeof(float)) +
                                             " (split the point set or raise paths_per_lane)");
        if (launch && (launch->flags & WOB_LAUNCH_TIME_WALK))
            return fail(WOB_INVALID_ARG, "WOB_LAUNCH_TIME_WALK creates events; the allocation-free entry point does not time");
    } else if ((st = ensure_scratch(scene, partial_floats * sizeof(float))) != WOB_OK) {
        return st;
    }
    L.table = nullptr;
    L.n_entries = 1;
    L.total_items = (uint32_t)n_items;
    L.total_points = n_points;
    L.persistent = 0;
    L.work_counter = nullptr;
    L.pass_counter = nullptr;
    L.seed = seed;
    L.partials = ws ? ws->partials : scene->scratch;
    const bool volume = est->num_volume_samples > 0;
    bool wavefront = false;
    st = resolve_schedule(launch ? launch->schedule : 0u,
                          wobl::supports_wavefront(D, boundary, volume, est->bie_model, est->estimator, flavour),
                          flavour, wavefront);
    if (st != WOB_OK) return st;
    uint32_t grid = 0;
    hipStream_t hs = (hipStream_t)stream;
    // resident waves pulling 64-item blocks from a counter (the per-lane walk) or single items
    // (the wavefront walk): no partly filled last round of waves
    const bool resident = wavefront || !(launch && (launch->flags & WOB_LAUNCH_ONE_LANE_PER_ITEM));
    if (resident) {
        uint32_t *counter = ws ? ws->counter : scene->counter;
        if (!counter) {
            void *p = nullptr;
            if (hipMalloc(&p, 256) != hipSuccess) return fail(WOB_OOM, "hipMalloc work counter");
            scene->counter = (uint32_t *)p;
            counter = scene->counter;
        }
        WOB_HIP(wobl::launch_fill_u32(counter, 1, 0u, hs), "counter reset");
        L.work_counter = counter;
        L.persistent = 1;
        L.wavefront = wavefront ? 1u : 0u;
        L.regen_threshold = launch && launch->regen_threshold ? launch->regen_threshold : kDefaultRegen;
        grid = wobl::persistent_grid(D, boundary, volume, est->bie_model, est->estimator, flavour, block, wavefront,
                                     false);
    }
    const bool timed = launch && (launch->flags & WOB_LAUNCH_TIME_WALK);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (timed) {
        WOB_HIP(hipEventCreate(&ev[0]), "hipEventCreate");
        WOB_HIP(hipEventCreate(&ev[1]), "hipEventCreate");
        WOB_HIP(hipEventRecord(ev[0], hs), "hipEventRecord");
    }
    WOB_HIP(wobl::launch_walk(D, boundary, volume, L, est->bie_model, est->estimator, flavour, block, grid, hs),
            "walk launch");
    if (timed) WOB_HIP(hipEventRecord(ev[1], hs), "hipEventRecord");
    WOB_HIP(wobl::launch_reduce(L, V, hs), "reduce launch");
    if (launch) {
        launch->used_paths_per_lane = L.e0.paths_per_chunk;
        launch->used_work_items = (uint32_t)n_items;
        launch->used_schedule = wavefront ? WOB_SCHEDULE_WAVEFRONT : WOB_SCHEDULE_LANE;
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

struct wob_batch {
    std::vector<wob_batch_entry> entries;
    wob_batch_config cfg{};
    int dim = 0, bie = 0, est = 0, flavour = 0;
    bool volume = false;
    uint32_t V = 1;
    std::vector<uint64_t> versions;         // scene versions the entry table was built from
    wobl::WalkLaunch L{};
    wobl::WalkEntry *d_table = nullptr;     // device entry table
    float *partials = nullptr;
    size_t partial_bytes = 0;
    uint32_t *counters = nullptr;           // [0] work counter, [1] pass counter
    uint64_t passes_run = 0;                // passes run so far (the device pass counter's value)
    uint32_t grid = 0;                      // 0: one lane per item
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    double *tune_out = nullptr;             // scratch outputs for the autotuner
};

namespace {

void batch_release_graph(wob_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void batch_free(wob_batch *b) {
    batch_release_graph(b);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    if (b->d_table) (void)hipFree(b->d_table);
    if (b->partials) (void)hipFree(b->partials);
    if (b->counters) (void)hipFree(b->counters);
    if (b->tune_out) (void)hipFree(b->tune_out);
}

// (Re)build the device entry table and partial buffer for the batch's launch configuration.
// `redirect`: write every entry's outputs into the tuning scratch instead.
wob_status batch_configure(wob_batch *b, bool redirect) {
    const uint32_t n = (uint32_t)b->entries.size();
    std::vector<wobl::WalkEntry> tab(n);
    uint64_t items = 0, points = 0, partial = 0, out_off = 0;
    uint32_t ppl = b->cfg.paths_per_lane;
    if (ppl == 0) {
        double total = 0;
        for (const auto &e : b->entries) total += (double)e.n_points * e.paths_per_point;
        ppl = auto_paths_per_lane(total);
    }
    for (uint32_t i = 0; i < n; ++i) {
        const auto &be = b->entries[i];
        double *s = be.d_sum, *q = be.d_sumsq;
        if (redirect) {
            s = b->tune_out + out_off;
            q = b->tune_out + out_off;
            out_off += (uint64_t)be.n_points * b->V;
        }
        tab[i] = make_entry(be.scene, &be.est, be.d_points, be.n_points, be.paths_per_point, ppl, be.point_base,
                            be.stream_id, be.path_base, s, q);
        tab[i].item_begin = (uint32_t)items;
        tab[i].point_begin = (uint32_t)points;
        tab[i].partial_base = (uint32_t)partial;
        const uint64_t ni = (uint64_t)be.n_points * tab[i].chunks;
        items += ni;
        points += be.n_points;
        partial += 2ull * b->V * ni;
        if (items >= (1ull << 31) || partial >= (1ull << 32)) return fail(WOB_INVALID_ARG, "batch too large for one launch");
    }
    const size_t pbytes = (size_t)partial * sizeof(float);
    if (pbytes > b->partial_bytes) {
        if (b->partials) WOB_HIP(hipFree(b->partials), "hipFree");
        b->partials = nullptr;
        b->partial_bytes = 0;
        if (hipMalloc(&b->partials, pbytes) != hipSuccess) return fail(WOB_OOM, "hipMalloc batch partials");
        b->partial_bytes = pbytes;
    }
    WOB_HIP(hipMemcpy(b->d_table, tab.data(), n * sizeof(wobl::WalkEntry), hipMemcpyHostToDevice), "hipMemcpy table");
    wobl::WalkLaunch &L = b->L;
    L = wobl::WalkLaunch{};
    L.e0 = tab[0];
    L.table = b->d_table;
    L.n_entries = n;
    L.total_items = (uint32_t)items;
    L.total_points = (uint32_t)points;
    L.persistent = b->cfg.persistent ? 1u : 0u;
    L.work_counter = b->counters;
    L.pass_counter = b->counters + 1;
    L.partials = b->partials;
    bool wavefront = false;
    wob_status st = resolve_schedule(
        b->cfg.schedule, wobl::supports_wavefront(b->dim, b->cfg.boundary != 0, b->volume, b->bie, b->est, b->flavour),
        b->flavour, wavefront);
    if (st != WOB_OK) return st;
    L.wavefront = wavefront ? 1u : 0u;
    L.regen_threshold = b->cfg.regen_threshold ? b->cfg.regen_threshold : kDefaultRegen;
    const uint32_t block = b->cfg.block_size ? b->cfg.block_size : 128u;
    b->grid = wavefront ? wobl::persistent_grid(b->dim, b->cfg.boundary != 0, b->volume, b->bie, b->est, b->flavour,
                                                block, true, true)
              : L.persistent ? wobl::persistent_grid(b->dim, b->cfg.boundary != 0, b->volume, b->bie, b->est, b->flavour,
                                                     block)
                             : 0u;
    b->versions.resize(n);
    for (uint32_t i = 0; i < n; ++i) b->versions[i] = b->entries[i].scene->version;
    batch_release_graph(b);
    return WOB_OK;
}

wob_status batch_pass(wob_batch *b, uint64_t seed, hipStream_t hs) {
    b->L.seed = seed;
    const uint32_t block = b->cfg.block_size ? b->cfg.block_size : 128u;
    WOB_HIP(wobl::launch_walk(b->dim, b->cfg.boundary != 0, b->volume, b->L, b->bie, b->est, b->flavour, block, b->grid,
                              hs),
            "batch walk launch");
    WOB_HIP(wobl::launch_reduce(b->L, b->V, hs), "batch reduce launch");
    WOB_HIP(wobl::launch_finalize(b->counters, b->counters + 1, hs), "batch finalize launch");
    return WOB_OK;
}

}  // namespace

extern "C" {

wob_status wob_scene_create(const wob_scene_desc *desc, wob_scene **out_scene) {
    return wob_scene_create_ex(desc, 0u, out_scene);
}

wob_status wob_scene_create_ex(const wob_scene_desc *desc, uint32_t flags, wob_scene **out_scene) {
    if (!desc || !out_scene) return fail(WOB_INVALID_ARG, "null argument");
    *out_scene = nullptr;
    if (flags & ~(uint32_t)(WOB_SCENE_DEVICE_BVH | WOB_SCENE_BVH_WIDTH_MASK))
        return fail(WOB_INVALID_ARG, "unknown scene flags");
    bool device_bvh = (flags & WOB_SCENE_DEVICE_BVH) != 0u;
    const uint32_t wfield = (flags & WOB_SCENE_BVH_WIDTH_MASK) >> WOB_SCENE_BVH_WIDTH_SHIFT;
    const int width = wfield == 0 ? 4 : (int)wfield;
    if (width < 2 || width > 4) return fail(WOB_INVALID_ARG, "BVH width must be 2, 3 or 4");
    wob_scene *s = new (std::nothrow) wob_scene();
    if (!s) return fail(WOB_OOM, "host allocation failed");
    (void)hipGetDevice(&s->device);
    // The device BVH build reads only the vertices: it runs on a second host thread (on this
    // thread's device) while this one builds the attributes, CDFs and guides.
    wobl::DeviceBvh dv;
    std::string dev_err;
    std::thread dev_build;
    const bool shapes_ok = (desc->dim == 2u || desc->dim == 3u) && desc->n_elements > 0u &&
                           desc->n_elements < (1u << 27) && desc->vertices != nullptr;
    if (device_bvh && shapes_ok) {
        try {
            dev_build = std::thread([&dv, &dev_err, desc, width, device = s->device] {
                if (hipSetDevice(device) != hipSuccess) {
                    dev_err = "hipSetDevice on the build thread failed";
                    return;
                }
                dev_err = wobl::build_bvh_device((int)desc->dim, desc->n_elements, desc->vertices, dv, width);
            });
        } catch (const std::system_error &) {   // no thread: build in sequence below
        }
    }
    std::string err;
    bool oom = false;
    try {
        err = wobhost::build_scene(*desc, s->host, !device_bvh, width);
    } catch (const std::bad_alloc &) {
        oom = true;
    }
    if (dev_build.joinable()) dev_build.join();
    else if (device_bvh && !oom && err.empty())
        dev_err = wobl::build_bvh_device((int)desc->dim, desc->n_elements, desc->vertices, dv, width);
    if (oom || !err.empty() || !dev_err.empty()) {
        (void)hipFree(dv.fnodes);
        (void)hipFree(dv.qnodes);
        (void)hipFree(dv.prims);
        (void)hipFree(dv.prim_id);
        delete s;
        if (oom) return fail(WOB_OOM, "host allocation failed during scene build");
        if (!err.empty()) return fail(WOB_INVALID_ARG, err);
        return fail(WOB_DEVICE, dev_err);
    }
    s->host.bvh_width = (uint32_t)width;
    if (device_bvh && dv.stack_need > (uint32_t)wobl::kStackEntries) {
        // a tree too deep for the traversal stack: the host SAH build (which falls back to
        // median splits) replaces it
        (void)hipFree(dv.fnodes);
        (void)hipFree(dv.qnodes);
        (void)hipFree(dv.prims);
        (void)hipFree(dv.prim_id);
        dv = wobl::DeviceBvh{};
        try {
            err = wobhost::build_scene(*desc, s->host, true, width);
        } catch (const std::bad_alloc &) {
            delete s;
            return fail(WOB_OOM, "host allocation failed during scene build");
        }
        if (!err.empty()) {
            delete s;
            return fail(WOB_INVALID_ARG, err);
        }
        s->host.bvh_width = (uint32_t)width;
        device_bvh = false;
    }
    if (device_bvh) {
        s->bufs[0] = dv.fnodes;
        s->bufs[1] = dv.prims;
        s->bufs[2] = dv.prim_id;
        s->bufs[8] = dv.qnodes;
        s->device_bytes += dv.bytes;
        s->host.n_nodes = dv.n_nodes;
        s->host.depth = dv.depth;
        s->host.stack_need = dv.stack_need;
    }
    if (s->host.stack_need > (uint32_t)wobl::kStackEntries) {
        const std::string m = "BVH traversal stack bound " + std::to_string(s->host.stack_need) + " exceeds " +
                              std::to_string(wobl::kStackEntries);
        free_scene(s);
        delete s;
        return fail(WOB_INVALID_ARG, m);
    }
    const auto &h = s->host;
    wob_status st = WOB_OK;
    const float *nodes = dv.fnodes, *prims = dv.prims, *attr = nullptr, *geo = nullptr, *acdf = nullptr, *bcdf = nullptr;
    const float *ag = nullptr, *bg = nullptr;
    const uint32_t *pid = dv.prim_id, *qn = dv.qnodes;
    if (!device_bvh &&
        ((st = upload(s, 0, h.nodes, &nodes)) != WOB_OK || (st = upload(s, 1, h.prim_geom, &prims)) != WOB_OK ||
         (st = upload(s, 2, h.prim_id, &pid)) != WOB_OK || (st = upload(s, 8, h.qnodes, &qn)) != WOB_OK)) {
        free_scene(s);
        delete s;
        return st;
    }
    const float *fused = nullptr;
    const uint32_t *fused_guide = nullptr;
    const uint32_t n_fused = (uint32_t)h.area_fused_guide.size();
    if ((st = upload(s, 3, h.elem_attr, &attr)) != WOB_OK || (st = upload(s, 9, h.elem_geo, &geo)) != WOB_OK ||
        (st = upload(s, 10, h.area_fused_guide, &fused_guide)) != WOB_OK ||
        (st = upload(s, 4, h.area_cdf, &acdf)) != WOB_OK ||
        (st = upload(s, 5, wobhost::guide_records(h.area_cdf, h.area_guide), &ag)) != WOB_OK ||
        (st = upload(s, 6, h.bv_cdf, &bcdf)) != WOB_OK ||
        (st = upload(s, 7, wobhost::guide_records(h.bv_cdf, h.bv_guide), &bg)) != WOB_OK) {
        free_scene(s);
        delete s;
        return st;
    }
    s->dev.nodes = (const float4 *)qn;
    s->dev.fnodes = (const float4 *)nodes;
    s->dev.prims = (const float4 *)prims;
    s->dev.prim_id = pid;
    s->dev.attr = (const float4 *)attr;
    s->dev.geo = (const float4 *)geo;
    if (n_fused > 0) {   // slot 10 holds the guide; the records replace it
        float4 *rec = nullptr;
        const size_t bytes = (size_t)n_fused * wobhost::kGeoWords3 * sizeof(float);
        if (hipMalloc(&rec, bytes) != hipSuccess) {
            free_scene(s);
            delete s;
            return fail(WOB_OOM, "hipMalloc fused area records");
        }
        hipError_t e = wobl::launch_fused_records(fused_guide, n_fused, acdf, h.n, (const float4 *)geo, rec, 0);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(s->bufs[10]);
        s->device_bytes -= std::max<size_t>(n_fused * sizeof(uint32_t), 16);
        s->bufs[10] = rec;
        s->device_bytes += bytes;
        fused = (const float *)rec;
        if (e != hipSuccess) {
            free_scene(s);
            delete s;
            return fail(WOB_DEVICE, std::string("fused area records: ") + hipGetErrorString(e));
        }
    }
    s->dev.area_fused = (const float4 *)fused;
    s->dev.n_fused = n_fused;
    s->dev.area_cdf = acdf;
    s->dev.area_guide = (const float4 *)ag;
    s->dev.bv_cdf = bcdf;
    s->dev.bv_guide = (const float4 *)bg;
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
    info->stack_need = scene->host.stack_need;
    info->bvh_width = scene->host.bvh_width;
    return WOB_OK;
}

wob_status wob_tune_bvh_width(const wob_scene_desc *desc, uint32_t flags, const wob_estimator_desc *est,
                              const float *d_points, uint32_t n_points, uint32_t paths_per_point, uint64_t seed,
                              uint32_t *best_width, float *ms_by_width, void *stream) {
    if (!desc || !est || !best_width) return fail(WOB_INVALID_ARG, "null argument");
    if (flags & WOB_SCENE_BVH_WIDTH_MASK) return fail(WOB_INVALID_ARG, "the tuner chooses the BVH width");
    if (est->num_volume_samples != 0)
        return fail(WOB_UNSUPPORTED_COMBO, "the BVH width tuner times walks without a volume source");
    if (est->flags & (WOB_EST_CONTROL_VARIATE | WOB_EST_HARMONIC_CV))
        return fail(WOB_UNSUPPORTED_COMBO, "the BVH width tuner times plain walks");
    if (n_points == 0 || paths_per_point == 0 || !d_points) return fail(WOB_INVALID_ARG, "empty point set");
    const size_t V = (size_t)desc->dim + 1;
    double *acc = nullptr;
    WOB_HIP(hipMalloc((void **)&acc, 2 * (size_t)n_points * V * sizeof(double)), "hipMalloc tuner outputs");
    float ms[3];
    std::string last;
    for (int w = 2; w <= 4; ++w) {
        ms[w - 2] = NAN;
        wob_scene *sc = nullptr;
        if (wob_scene_create_ex(desc, flags | WOB_SCENE_BVH_WIDTH(w), &sc) != WOB_OK) {
            last = g_last_error;
            continue;
        }
        wob_launch_config lc{};
        wob_status st = wob_estimate_domain(sc, est, d_points, n_points, 0, paths_per_point, seed, 0, 0, acc,
                                            acc + (size_t)n_points * V, &lc, stream);
        lc = wob_launch_config{};
        lc.flags = WOB_LAUNCH_TIME_WALK;
        if (st == WOB_OK)
            st = wob_estimate_domain(sc, est, d_points, n_points, 0, paths_per_point, seed, paths_per_point, 0, acc,
                                     acc + (size_t)n_points * V, &lc, stream);
        wob_scene_destroy(sc);
        if (st != WOB_OK) {
            (void)hipFree(acc);
            return st;
        }
        ms[w - 2] = lc.walk_ms;
    }
    (void)hipFree(acc);
    int best = 0;
    for (int w = 2; w <= 4; ++w)
        if (ms[w - 2] == ms[w - 2] && (best == 0 || ms[w - 2] < ms[best - 2])) best = w;
    if (ms_by_width)
        for (int i = 0; i < 3; ++i) ms_by_width[i] = ms[i];
    if (best == 0) return fail(WOB_INVALID_ARG, "no BVH width gives a usable scene: " + last);
    *best_width = (uint32_t)best;
    return WOB_OK;
}

wob_status wob_scene_set_volume_grid(wob_scene *scene, uint32_t grid_res, float domain_size, const float *cache) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    WOB_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    free_volume(scene);
    if (grid_res == 0 || !cache) return WOB_OK;
    if (!(domain_size > 0.f)) return fail(WOB_INVALID_ARG, "domain_size must be > 0");
    const uint32_t D = scene->host.dim;
    const uint64_t cells = (uint64_t)grid_res * grid_res * (D == 3 ? grid_res : 1u);
    if (cells >= (1ull << 31)) return fail(WOB_INVALID_ARG, "grid too large");
    // scene.cuh:215-228: |cache| normalised by its sum, inclusive scan (here sequential FP32)
    std::vector<float> cdf(cells);
    float total = 0.f;
    for (uint64_t i = 0; i < cells; ++i) { cdf[i] = std::fabs(cache[i]); total += cdf[i]; }
    if (total == 0.f) for (auto &c : cdf) c = 1.f / (float)cells;
    else for (auto &c : cdf) c /= total;
    for (uint64_t i = 1; i < cells; ++i) cdf[i] += cdf[i - 1];
    std::vector<uint32_t> guide;
    wobhost::build_guide(cdf, guide);
    const std::vector<float> vals(cache, cache + cells);
    void *p[3] = {nullptr, nullptr, nullptr};
    const size_t bytes[3] = {cells * 4, cells * 4, guide.size() * 4};
    for (int i = 0; i < 3; ++i) {
        if (hipMalloc(&p[i], bytes[i]) != hipSuccess) {
            for (int j = 0; j < i; ++j) {
                (void)hipFree(p[j]);
                scene->vol_bufs[j] = nullptr;
            }
            return fail(WOB_OOM, "hipMalloc volume grid");
        }
        scene->vol_bufs[i] = p[i];
    }
    WOB_HIP(hipMemcpy(p[0], vals.data(), bytes[0], hipMemcpyHostToDevice), "hipMemcpy volume");
    WOB_HIP(hipMemcpy(p[1], cdf.data(), bytes[1], hipMemcpyHostToDevice), "hipMemcpy volume");
    WOB_HIP(hipMemcpy(p[2], guide.data(), bytes[2], hipMemcpyHostToDevice), "hipMemcpy volume");
    wobdev::VolSrc &v = scene->dev.vol;
    v.kind = wobdev::kVolGrid;
    v.n = (uint32_t)cells;
    v.grid_res = grid_res;
    v.domain_size = domain_size;
    v.cache = (const float *)p[0];
    v.cdf = (const float *)p[1];
    v.guide = (const uint32_t *)p[2];
    scene->version += 1;
    return WOB_OK;
}

wob_status wob_scene_set_volume_gaussians(wob_scene *scene, uint32_t n, const float *centers, const float *widths,
                                          const float *amplitudes, uint32_t flags) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    WOB_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    free_volume(scene);
    if (n == 0) return WOB_OK;
    if (!centers || !widths || !amplitudes) return fail(WOB_INVALID_ARG, "null source arrays");
    if (flags & ~(uint32_t)WOB_VOLUME_NO_MASK) return fail(WOB_INVALID_ARG, "unknown volume flags");
    const int D = (int)scene->host.dim;
    // mixture weights |A_j| w_j^D; constants in double, rounded once
    std::vector<double> wt(n);
    double tot = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
        if (!(widths[j] > 0.f)) return fail(WOB_INVALID_ARG, "Gaussian widths must be > 0");
        wt[j] = std::fabs((double)amplitudes[j]) * std::pow((double)widths[j], D);
        tot += wt[j];
    }
    if (!(tot > 0.0)) return fail(WOB_INVALID_ARG, "all Gaussian amplitudes are zero");
    std::vector<float> comp((size_t)8 * n), cdf(n);
    double acc = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
        const double w = widths[j], pj = wt[j] / tot;
        acc += pj;
        float *c = &comp[(size_t)8 * j];
        for (int k = 0; k < 3; ++k) c[k] = k < D ? centers[(size_t)j * D + k] : 0.f;
        c[3] = widths[j];
        c[4] = amplitudes[j];
        c[5] = (float)(pj * std::pow(2.0 * M_PI * w * w, -0.5 * D));
        c[6] = (float)(1.0 / (2.0 * w * w));
        c[7] = 0.f;
        cdf[j] = j + 1 == n ? 1.0f : (float)acc;
    }
    void *p[2] = {nullptr, nullptr};
    if (hipMalloc(&p[0], comp.size() * 4) != hipSuccess) return fail(WOB_OOM, "hipMalloc volume source");
    scene->vol_bufs[0] = p[0];
    if (hipMalloc(&p[1], cdf.size() * 4) != hipSuccess) return fail(WOB_OOM, "hipMalloc volume source");
    scene->vol_bufs[1] = p[1];
    WOB_HIP(hipMemcpy(p[0], comp.data(), comp.size() * 4, hipMemcpyHostToDevice), "hipMemcpy volume");
    WOB_HIP(hipMemcpy(p[1], cdf.data(), cdf.size() * 4, hipMemcpyHostToDevice), "hipMemcpy volume");
    wobdev::VolSrc &v = scene->dev.vol;
    v.kind = wobdev::kVolGauss;
    v.n = n;
    v.mask = (flags & WOB_VOLUME_NO_MASK) ? 0u : 1u;
    v.comp = (const float4 *)p[0];
    v.cdf = (const float *)p[1];
    v.gflag = nullptr;
    v.gres = 0u;
    scene->version += 1;
    // the inside test's occupancy grid: near-boundary cells marked on the host, the others
    // classified on the device by ray parity
    std::vector<uint8_t> gflag;
    const uint32_t res = D == 2 ? wobhost::kInsideGridRes2 : wobhost::kInsideGridRes3;
    float glo[3], gh[3], ginv[3];
    if (v.mask && wobhost::inside_grid_marks(scene->host, res, glo, gh, ginv, gflag)) {
        void *g = nullptr;
        if (hipMalloc(&g, gflag.size()) != hipSuccess) return fail(WOB_OOM, "hipMalloc inside-test grid");
        scene->vol_bufs[2] = g;
        WOB_HIP(hipMemcpy(g, gflag.data(), gflag.size(), hipMemcpyHostToDevice), "hipMemcpy inside-test grid");
        v.gres = res;
        for (int k = 0; k < 3; ++k) { v.glo[k] = glo[k]; v.gh[k] = gh[k]; v.ginv[k] = ginv[k]; }
        const size_t lds = wobl::walk_lds_bytes(wobdev::FL_UNIFORM, true, 256);
        hipError_t e = D == 2 ? wobl::launch_inside_grid_2(scene->dev, (uint8_t *)g, (uint32_t)gflag.size(), lds, nullptr)
                              : wobl::launch_inside_grid_3(scene->dev, (uint8_t *)g, (uint32_t)gflag.size(), lds, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            v.gres = 0u;
            return fail(WOB_DEVICE, std::string("inside-test grid: ") + hipGetErrorString(e));
        }
        v.gflag = (const uint8_t *)g;
    }
    return WOB_OK;
}

wob_status wob_estimate_domain(const wob_scene *scene, const wob_estimator_desc *est, const float *d_points,
                               uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                               uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                               wob_launch_config *launch, void *stream) {
    return run_walk(scene, est, d_points, n_points, point_base, paths_per_point, seed, path_base, stream_id, d_sum,
                    d_sumsq, launch, stream, false);
}

// The allocation-free entry point's device memory: one file-scope array per GPU, held until
// the process ends.  The first 256 bytes are the work counter, the rest the partial sums.
// In host code `scratch` names a host-side stand-in, not device memory: its device address
// comes from hipGetSymbolAddress, once per device (with that device current), and only that
// address is ever passed on.
constexpr size_t kNoAllocScratchBytes = (size_t)WOB_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

wob_status noalloc_scratch(WalkScratch &ws) {
    int dev = 0;
    WOB_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(WOB_DEVICE, "device index out of range for the static scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(WOB_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = p;
    }
    unsigned char *base = (unsigned char *)g_scratch_ptr[dev];
    ws.counter = (uint32_t *)base;
    ws.partials = (float *)(base + 256);
    ws.partial_floats = (kNoAllocScratchBytes - 256) / sizeof(float);
    return WOB_OK;
}
}  // namespace

wob_status wob_estimate_domain_noalloc(const wob_scene *scene, const wob_estimator_desc *est, const float *d_points,
                                       uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                                       uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                                       wob_launch_config *launch, void *stream) {
    wob_status st = check_paths(path_base, paths_per_point);
    if (st != WOB_OK) return st;
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    WalkScratch ws{};
    if ((st = noalloc_scratch(ws)) != WOB_OK) return st;
    return run_walk(scene, est, d_points, n_points, point_base, paths_per_point, seed, path_base, stream_id, d_sum,
                    d_sumsq, launch, stream, false, &ws);
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

wob_status wob_sample_lines(const wob_scene *scene, const float *d_points, uint32_t n, uint64_t seed, uint32_t path,
                            uint32_t stream_id, float *d_records, float *d_inv_pdf, void *stream) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n == 0) return WOB_OK;
    if (!d_points || !d_records || !d_inv_pdf) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_sample_lines((int)scene->host.dim, scene->dev, n, d_points, seed, path, stream_id, d_records,
                                      d_inv_pdf, (hipStream_t)stream),
            "sample_lines launch");
    return WOB_OK;
}

wob_status wob_sample_boundary(const wob_scene *scene, uint32_t n, uint64_t seed, uint32_t stream_id, float *d_records,
                               void *stream) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n == 0) return WOB_OK;
    if (!d_records) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_sample_boundary((int)scene->host.dim, scene->dev, n, seed, stream_id, d_records,
                                         (hipStream_t)stream),
            "sample_boundary launch");
    return WOB_OK;
}

// Stream word of the VPL start-point samples: distinct from every walk stream word (walks
// use the caller's stream_id; the control variate's independent draws flip bit 31).
constexpr uint32_t kVplSampleStream = 0x40000000u;

wob_status wob_vpl_paths(const wob_scene *cscene, const wob_estimator_desc *est, uint32_t n_paths, uint64_t seed,
                         uint64_t path_base, uint32_t stream_id, float *d_records, float *d_contrib,
                         wob_launch_config *launch, void *stream) {
    if (!cscene || !est) return fail(WOB_INVALID_ARG, "null scene or estimator descriptor");
    if (est->estimator != WOB_BACKWARD) return fail(WOB_UNSUPPORTED_COMBO, "VPL paths need the backward estimator");
    if (est->flags & (WOB_EST_CONTROL_VARIATE | WOB_EST_HARMONIC_CV))
        return fail(WOB_UNSUPPORTED_COMBO, "VPL paths do not take a control variate");
    if (n_paths == 0) return WOB_OK;
    if (!d_records || !d_contrib) return fail(WOB_INVALID_ARG, "null device buffer");
    if (stream_id & (kVplSampleStream | 0x80000000u)) return fail(WOB_INVALID_ARG, "stream_id must be below 2^30");
    if (path_base > 0xFFFFFFFFull) return fail(WOB_INVALID_ARG, "path_base must fit in 32 bits");
    wob_scene *scene = const_cast<wob_scene *>(cscene);
    wob_estimator_desc e = *est;
    e.flags |= WOB_EST_BOUNDARY_DENSITY;
    int flavour = 0;
    wob_status st = check_est(scene, &e, flavour);
    if (st != WOB_OK) return st;
    hipStream_t hs = (hipStream_t)stream;
    if ((st = ensure_aux(scene, (size_t)2 * n_paths * sizeof(double))) != WOB_OK) return st;
    // start points: VPL i of the pass uses the stream (seed, path_base, i, stream_id | 2^30)
    WOB_HIP(wobl::launch_sample_boundary((int)scene->host.dim, scene->dev, n_paths, seed, stream_id | kVplSampleStream,
                                         d_records, hs, (int64_t)path_base),
            "VPL start-point launch");
    WOB_HIP(wobl::launch_fill_u32((uint32_t *)scene->aux, 4u * n_paths, 0u, hs), "aux clear");
    // one boundary-density walk per start point, stream (seed, path_base, i, stream_id)
    st = run_walk(scene, &e, d_records, n_paths, 0u, 1u, seed, path_base, stream_id, scene->aux, scene->aux + n_paths,
                  launch, stream, true);
    if (st != WOB_OK) return st;
    WOB_HIP(wobl::launch_f64_to_f32(scene->aux, n_paths, d_contrib, hs), "VPL contribution launch");
    return WOB_OK;
}

wob_status wob_vpl_aggregate(const wob_scene *cscene, const wob_estimator_desc *est, const float *d_records,
                             const float *d_contrib, uint32_t n_vpl, const float *d_points, uint32_t n_points,
                             double *d_sum, void *stream) {
    if (!cscene || !est) return fail(WOB_INVALID_ARG, "null scene or estimator descriptor");
    if (est->equation_kernel != WOB_POISSON_KERNEL) return fail(WOB_INVALID_ENUM, "equation_kernel must be PoissonKernel");
    if (est->bie_model < 0 || est->bie_model > 2) return fail(WOB_INVALID_ENUM, "bie_model out of range");
    if (est->domain < 0 || est->domain > 1) return fail(WOB_INVALID_ENUM, "domain out of range");
    if (!(est->eps > 0.f)) return fail(WOB_INVALID_ARG, "eps must be > 0");
    if (n_vpl == 0 || n_points == 0) return WOB_OK;
    if (!d_records || !d_contrib || !d_points || !d_sum) return fail(WOB_INVALID_ARG, "null device buffer");
    wob_scene *scene = const_cast<wob_scene *>(cscene);
    const int D = (int)scene->host.dim;
    const uint32_t V = (uint32_t)D + 1u;
    uint32_t splits = 0, chunk = 0;
    wobl::vpl_plan(n_points, n_vpl, splits, chunk);
    const size_t rec_bytes = ((size_t)n_vpl * wobl::vpl_record_floats(D, est->bie_model) * sizeof(float) + 255) & ~(size_t)255;
    const size_t part_bytes = (size_t)splits * n_points * V * sizeof(double);
    // its own buffer: an aggregation may run beside a path phase of the same scene (another stream)
    wob_status st = ensure_vpl(scene, rec_bytes + part_bytes);
    if (st != WOB_OK) return st;
    float *rec = (float *)scene->vpl;
    double *partial = (double *)(scene->vpl + rec_bytes);
    hipStream_t hs = (hipStream_t)stream;
    // inv_pdf = total boundary area, as apps/potential_flow_with_vpl.cu:61-63 passes it
    WOB_HIP(wobl::launch_vpl_stage(D, est->bie_model, d_records, d_contrib, n_vpl, scene->host.total_area,
                                   est->domain == WOB_INTERIOR ? 1.f : -1.f, rec, hs),
            "VPL stage launch");
    WOB_HIP(wobl::launch_vpl_aggregate(D, est->bie_model, d_points, n_points, rec, n_vpl, splits, chunk, est->eps,
                                       partial, hs),
            "VPL aggregate launch");
    WOB_HIP(wobl::launch_vpl_reduce(partial, n_points, V, splits, d_sum, hs), "VPL reduce launch");
    return WOB_OK;
}

wob_status wob_closest_points(const wob_scene *scene, const float *d_points, uint32_t n, uint32_t *d_elem, float *d_dist,
                              float *d_value, void *stream) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n == 0) return WOB_OK;
    if (!d_points || !d_elem || !d_dist || !d_value) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_closest((int)scene->host.dim, scene->dev, n, d_points, d_elem, d_dist, d_value,
                                 (hipStream_t)stream),
            "closest-point launch");
    return WOB_OK;
}

wob_status wob_wos_estimate(const wob_scene *cscene, const float *d_points, uint32_t n_points, uint32_t point_base,
                            uint32_t paths_per_point, uint64_t seed, uint64_t path_base, uint32_t stream_id, float eps,
                            uint32_t max_steps, double *d_sum, double *d_sumsq, void *stream) {
    if (const wob_status ps = check_paths(path_base, paths_per_point); ps != WOB_OK) return ps;
    if (!cscene) return fail(WOB_INVALID_ARG, "null scene");
    if (!(eps > 0.f)) return fail(WOB_INVALID_ARG, "eps must be > 0");
    if (n_points == 0 || paths_per_point == 0) return WOB_OK;
    if (!d_points || !d_sum || !d_sumsq) return fail(WOB_INVALID_ARG, "null device buffer");
    wob_scene *scene = const_cast<wob_scene *>(cscene);
    const uint32_t ppl = auto_paths_per_lane((double)n_points * paths_per_point);
    wobl::WalkLaunch L{};
    wobl::WalkEntry &e = L.e0;
    e.scene = scene->dev;
    e.points = d_points;
    e.sum = d_sum;
    e.sumsq = d_sumsq;
    e.n_points = n_points;
    chunking(paths_per_point, ppl, e.paths_per_chunk, e.chunks);
    e.paths_per_point = paths_per_point;
    e.point_base = point_base;
    e.stream = stream_id;
    e.path_base = path_base;
    const uint64_t n_items = (uint64_t)n_points * e.chunks;
    if (n_items >= (1ull << 31)) return fail(WOB_INVALID_ARG, "too many work items; split the point set");
    wob_status st = ensure_scratch(scene, (size_t)2 * n_items * sizeof(float));
    if (st != WOB_OK) return st;
    L.n_entries = 1;
    L.total_items = (uint32_t)n_items;
    L.total_points = n_points;
    L.seed = seed;
    L.partials = scene->scratch;
    hipStream_t hs = (hipStream_t)stream;
    WOB_HIP(wobl::launch_wos((int)scene->host.dim, L, eps, max_steps ? max_steps : 1u << 16, hs), "WoS launch");
    WOB_HIP(wobl::launch_reduce(L, 1u, hs), "reduce launch");
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

wob_status wob_batch_create(const wob_batch_entry *entries, uint32_t n, const wob_batch_config *cfg,
                            wob_batch **out) {
    if (!entries || !out || n == 0) return fail(WOB_INVALID_ARG, "null or empty batch");
    *out = nullptr;
    wob_batch_config c{};
    if (cfg) c = *cfg;
    if (c.block_size && check_block(c.block_size) != WOB_OK) return WOB_INVALID_ARG;
    auto *b = new (std::nothrow) wob_batch();
    if (!b) return fail(WOB_OOM, "host allocation failed");
    b->entries.assign(entries, entries + n);
    b->cfg = c;
    for (uint32_t i = 0; i < n; ++i) {
        const auto &e = entries[i];
        int fl = 0;
        wob_status st = check_est(e.scene, &e.est, fl);
        if (st != WOB_OK) { delete b; return fail(st, "entry " + std::to_string(i) + ": " + g_last_error); }
        if (c.boundary && (e.est.flags & WOB_EST_HARMONIC_CV) && !(e.est.flags & WOB_EST_BOUNDARY_DENSITY)) {
            delete b;
            return fail(WOB_UNSUPPORTED_COMBO, "entry " + std::to_string(i) +
                                                   ": the harmonic control variate applies to domain estimates and "
                                                   "boundary density walks");
        }
        if (e.n_points == 0 || e.paths_per_point == 0 || !e.d_points || !e.d_sum || !e.d_sumsq) {
            delete b;
            return fail(WOB_INVALID_ARG, "entry " + std::to_string(i) + ": empty point set or null buffer");
        }
        const int dim = (int)e.scene->host.dim;
        const bool vol = e.est.num_volume_samples > 0;
        if (i == 0) { b->dim = dim; b->bie = e.est.bie_model; b->est = e.est.estimator; b->flavour = fl; b->volume = vol; }
        else if (dim != b->dim || e.est.bie_model != b->bie || e.est.estimator != b->est || fl != b->flavour ||
                 vol != b->volume) {
            delete b;
            return fail(WOB_UNSUPPORTED_COMBO, "entry " + std::to_string(i) +
                                                   ": all batch entries must share dimension, model, estimator, "
                                                   "sampling strategy and the use of a volume term (split the batch)");
        }
    }
    b->V = c.boundary ? 1u : (uint32_t)b->dim + 1u;
    uint64_t out_elems = 0;
    for (const auto &e : b->entries) out_elems += (uint64_t)e.n_points * b->V;
    if (hipMalloc(&b->d_table, n * sizeof(wobl::WalkEntry)) != hipSuccess || hipMalloc(&b->counters, 64) != hipSuccess ||
        hipMalloc(&b->tune_out, out_elems * sizeof(double)) != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(WOB_OOM, "hipMalloc batch buffers");
    }
    if (hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(WOB_DEVICE, "hipStreamCreate");
    }
    wob_status st = batch_configure(b, false);
    if (st == WOB_OK) {
        hipError_t e = wobl::launch_fill_u32(b->counters, 16, 0u, b->capture_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->capture_stream);
        if (e != hipSuccess) st = hip_fail(e, "counter init");
    }
    if (st != WOB_OK) {
        batch_free(b);
        delete b;
        return st;
    }
    *out = b;
    return WOB_OK;
}

wob_status wob_batch_run(wob_batch *b, uint64_t seed, uint32_t n_passes, void *stream) {
    if (!b) return fail(WOB_INVALID_ARG, "null batch");
    for (size_t i = 0; i < b->entries.size(); ++i) {
        const auto &e = b->entries[i];
        const wob_status st = check_paths(e.path_base, (b->passes_run + n_passes) * (uint64_t)e.paths_per_point);
        if (st != WOB_OK) return fail(st, "entry " + std::to_string(i) + ": " + g_last_error);
    }
    b->passes_run += n_passes;
    hipStream_t hs = (hipStream_t)stream;
    // a scene whose device data changed since the entry table was built (volume source, control
    // variate) invalidates the table and any captured graph
    for (size_t i = 0; i < b->entries.size(); ++i) {
        if (b->entries[i].scene->version != b->versions[i]) {
            WOB_HIP(hipStreamSynchronize(hs), "hipStreamSynchronize");
            wob_status st = batch_configure(b, false);
            if (st != WOB_OK) return st;
            break;
        }
    }
    if (!b->cfg.use_graph) {
        for (uint32_t p = 0; p < n_passes; ++p) {
            wob_status st = batch_pass(b, seed, hs);
            if (st != WOB_OK) return st;
        }
        return WOB_OK;
    }
    if (b->exec && b->L.seed != seed) batch_release_graph(b);
    if (!b->exec) {
        // capture one pass (walk, reduction, counter update) on the batch's own stream
        WOB_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        wob_status st = batch_pass(b, seed, b->capture_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
        if (st != WOB_OK) { if (g) (void)hipGraphDestroy(g); return st; }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        b->graph = g;
        WOB_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    for (uint32_t p = 0; p < n_passes; ++p) WOB_HIP(hipGraphLaunch(b->exec, hs), "hipGraphLaunch");
    return WOB_OK;
}

wob_status wob_batch_tune(wob_batch *b, uint64_t seed, wob_batch_config *best, float *best_ms, void *stream) {
    if (!b) return fail(WOB_INVALID_ARG, "null batch");
    hipStream_t hs = (hipStream_t)stream;
    const wob_batch_config orig = b->cfg;
    // the pass counter is not advanced by tuning passes
    uint32_t saved_pass = 0;
    WOB_HIP(hipStreamSynchronize(hs), "hipStreamSynchronize");
    WOB_HIP(hipMemcpy(&saved_pass, b->counters + 1, 4, hipMemcpyDeviceToHost), "hipMemcpy");
    double total = 0;
    uint32_t min_paths = 0xFFFFFFFFu;
    for (const auto &e : b->entries) {
        total += (double)e.n_points * e.paths_per_point;
        min_paths = std::min(min_paths, e.paths_per_point);
    }
    const uint32_t base = auto_paths_per_lane(total);
    std::vector<uint32_t> ppls;
    for (uint32_t f : {1u, 2u, 4u, 8u, 16u, 32u, 64u}) ppls.push_back(f);
    ppls.push_back(base);
    std::sort(ppls.begin(), ppls.end());
    ppls.erase(std::unique(ppls.begin(), ppls.end()), ppls.end());
    // launch modes: per-lane walk (one lane per item, or resident waves pulling items) and,
    // where it exists, the wavefront walk with a few regeneration thresholds
    struct Mode { uint32_t schedule, persistent, regen; };
    std::vector<Mode> modes;
    const bool wf_ok = wobl::supports_wavefront(b->dim, orig.boundary != 0, b->volume, b->bie, b->est, b->flavour);
    if (orig.schedule != WOB_SCHEDULE_WAVEFRONT) {
        modes.push_back({WOB_SCHEDULE_LANE, 0u, 0u});
        modes.push_back({WOB_SCHEDULE_LANE, 1u, 0u});
    }
    if (wf_ok && orig.schedule != WOB_SCHEDULE_LANE)
        for (uint32_t r : {16u, 32u, 48u}) modes.push_back({WOB_SCHEDULE_WAVEFRONT, 1u, r});
    hipEvent_t e0, e1;
    WOB_HIP(hipEventCreate(&e0), "hipEventCreate");
    WOB_HIP(hipEventCreate(&e1), "hipEventCreate");
    float best_t = 1e30f;
    wob_batch_config best_c = orig;
    for (uint32_t ppl : ppls) {
        if (ppl > std::max(1u, min_paths) && ppl != base) continue;
        if ((double)total / ppl < 65536.0 && ppl != base) continue;   // too few items to fill the GPU
        for (uint32_t block : {64u, 128u, 256u}) {
            for (const Mode &md : modes) {
                b->cfg = orig;
                b->cfg.paths_per_lane = ppl;
                b->cfg.block_size = block;
                b->cfg.persistent = md.persistent;
                b->cfg.schedule = md.schedule;
                b->cfg.regen_threshold = md.regen;
                b->cfg.use_graph = 0;
                wob_status st = batch_configure(b, true);
                if (st != WOB_OK) { b->cfg = orig; (void)batch_configure(b, false); return st; }
                st = batch_pass(b, seed, hs);                       // warm-up
                if (st != WOB_OK) { b->cfg = orig; (void)batch_configure(b, false); return st; }
                WOB_HIP(hipEventRecord(e0, hs), "hipEventRecord");
                st = batch_pass(b, seed, hs);
                if (st != WOB_OK) { b->cfg = orig; (void)batch_configure(b, false); return st; }
                WOB_HIP(hipEventRecord(e1, hs), "hipEventRecord");
                WOB_HIP(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.f;
                WOB_HIP(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                if (ms < best_t) { best_t = ms; best_c = b->cfg; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->cfg = best_c;
    b->cfg.use_graph = orig.use_graph;
    wob_status st = batch_configure(b, false);
    if (st != WOB_OK) return st;
    WOB_HIP(wobl::launch_fill_u32(b->counters, 1, 0u, hs), "counter reset");
    WOB_HIP(wobl::launch_fill_u32(b->counters + 1, 1, saved_pass, hs), "counter restore");
    WOB_HIP(hipStreamSynchronize(hs), "hipStreamSynchronize");
    if (best) *best = b->cfg;
    if (best_ms) *best_ms = best_t;
    return WOB_OK;
}

wob_status wob_batch_destroy(wob_batch *b) {
    if (!b) return WOB_OK;
    batch_free(b);
    delete b;
    return WOB_OK;
}

const char *wob_last_error_message(void) { return g_last_error.c_str(); }

const char *wob_version(void) { return "wob-mi355x 0.1"; }

}  // extern "C"
