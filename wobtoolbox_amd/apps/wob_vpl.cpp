// wob_vpl -- the reference's potential-flow virtual-point-light app on the MI355X engine.
//
// Same command line, configuration keys, scene and output files as
// apps/potential_flow_with_vpl.cu (rsugimoto/WoBToolbox):
//   wob_vpl config.json [--max-iterations N] [--dry-run] [--seed S]
// reads output_dir, path_length, num_sample_paths, num_resampling_candidates (default 1),
// grid_res, eps (default FLT_EPSILON) and max_time (minutes) (:104-113); builds the square
// with flux -1 / +1 on its left / right sides and the four zero-flux circular obstacles
// (:141-175); masks the grid points inside the obstacles (:177-184); and repeats
// iterations until max_time minutes have passed (:227-265).  One iteration is
//   * the path phase: num_sample_paths uniform boundary points y_i and the single-layer
//     Neumann walk value c_i at each (PathComputeFunc, :73-88) -> wob_vpl_paths;
//   * the aggregation over every unmasked grid point (AggregateFunc, :44-71) ->
//     wob_vpl_aggregate.
// After iteration i it writes potential<i>.scalar and gradient<i>.vector2 in output_dir
// (uint32 grid_res, then grid_res^2 float records; runtime_utils.cuh:35-57), with NaN at masked
// points and the sums divided by the number of paths so far, and a copy of the configuration
// as config.json.
//
// Differences, on purpose: the walks use the engine's Philox streams (iteration i draws
// path index i - 1 of every VPL slot), and the per-point sums are accumulated in FP64.
#include <hip/hip_runtime_api.h>

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/wob/wob.hpp"
#include "app_common.hpp"

namespace {

using namespace wobapp;
using Elem = wob::ElementXs<float, 2>;

struct Circle {
    double cx, cy, r;
    unsigned segments;
};

Elem segment(double ax, double ay, double bx, double by, double nx, double ny, double g) {
    Elem e{};
    e.a[0] = (float)ax; e.a[1] = (float)ay;
    e.b[0] = (float)bx; e.b[1] = (float)by;
    e.n[0] = (float)nx; e.n[1] = (float)ny;
    e.boundary_value_a = e.boundary_value_b = (float)g;
    e.boundary_type = wob::Neumann;
    return e;
}

int run_app(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s config.json [--max-iterations N] [--dry-run] [--seed S]\n", argv[0]);
        return 1;
    }
    const std::string config_path = argv[1];
    long max_iter = -1;
    bool dry = false;
    uint64_t seed = 0;
    for (int i = 2; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--max-iterations") && i + 1 < argc) max_iter = std::atol(argv[++i]);
        else if (!std::strcmp(argv[i], "--dry-run")) dry = true;
        else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else throw std::runtime_error(std::string("unknown option ") + argv[i]);
    }
    std::ifstream cf(config_path);
    if (!cf) {
        std::cout << "Failed to load config file: " << config_path << std::endl;
        return 0;
    }
    std::stringstream ss;
    ss << cf.rdbuf();
    const Config cfg = parse_config(ss.str());
    const std::string out_dir = cfg.str("output_dir");
    const unsigned L = (unsigned)cfg.num("path_length");
    const uint32_t n_paths = (uint32_t)cfg.num("num_sample_paths");
    const unsigned nrc = (unsigned)cfg.num_or("num_resampling_candidates", 1);
    const unsigned res = (unsigned)cfg.num("grid_res");
    const float eps = (float)cfg.num_or("eps", FLT_EPSILON);
    const double max_time = cfg.num("max_time");

    std::cout << "------------------------------------------------------------" << std::endl
              << "input json file\t: " << config_path << std::endl
              << "output directory\t: " << out_dir << std::endl
              << "path length\t: " << L << std::endl
              << "#sample paths per iteration\t: " << n_paths << std::endl
              << "#resampling candidates\t: " << nrc << std::endl
              << "grid resolution\t: " << res << std::endl
              << "eps\t: " << eps << std::endl
              << "precision\t: float" << std::endl
              << "max time\t: " << max_time << " minutes" << std::endl
              << "------------------------------------------------------------" << std::endl;
    std::filesystem::create_directories(out_dir);
    std::filesystem::copy(config_path, out_dir + "/config.json", std::filesystem::copy_options::overwrite_existing);

    // scene (potential_flow_with_vpl.cu:141-175): the square, then the obstacles' segments
    std::vector<Elem> elems;
    elems.push_back(segment(-0.5, -0.5, -0.5, 0.5, -1.0, 0.0, -1.0));
    elems.push_back(segment(-0.5, 0.5, 0.5, 0.5, 0.0, 1.0, 0.0));
    elems.push_back(segment(0.5, 0.5, 0.5, -0.5, 1.0, 0.0, 1.0));
    elems.push_back(segment(0.5, -0.5, -0.5, -0.5, 0.0, -1.0, 0.0));
    const std::vector<Circle> circles = {
        {0.2, 0.1, 0.25, 360}, {-0.25, 0.3, 0.15, 180}, {-0.20, -0.25, 0.2, 180}, {0.25, -0.3, 0.05, 90}};
    for (const Circle &c : circles) {
        // vertices and normal in FP32 as the app forms them (Eigen float vectors; the angle's
        // cosine and sine in double)
        for (unsigned i = 0; i < c.segments; ++i) {
            const float cx = (float)c.cx, cy = (float)c.cy, r = (float)c.r;
            const double t0 = (2 * M_PI) * i / c.segments, t1 = (2 * M_PI) * (i + 1) / c.segments;
            const float v1x = cx + r * (float)std::cos(t0), v1y = cy + r * (float)std::sin(t0);
            const float v2x = cx + r * (float)std::cos(t1), v2y = cy + r * (float)std::sin(t1);
            const float dx = v1x - v2x, dy = v1y - v2y, len = std::sqrt(dx * dx + dy * dy);
            const float tx = dx / len, ty = dy / len;
            elems.push_back(segment(v1x, v1y, v2x, v2y, ty, -tx, 0.0));
        }
    }

    // evaluation mask (:177-184)
    const unsigned npts = res * res;
    std::vector<float> pts;
    std::vector<unsigned> idx_of;
    for (unsigned i = 0; i < npts; ++i) {
        const float dx = 1.0f / (float)res;
        const float x = ((float)(i % res) + 0.5f) * dx - 0.5f, y = ((float)(i / res) + 0.5f) * dx - 0.5f;
        bool keep = true;
        for (const Circle &c : circles) {
            const float ex = x - (float)c.cx, ey = y - (float)c.cy, r = (float)c.r;
            if (ex * ex + ey * ey < r * r) keep = false;
        }
        if (!keep) continue;
        pts.push_back(x);
        pts.push_back(y);
        idx_of.push_back(i);
    }
    const uint32_t m = (uint32_t)idx_of.size();
    std::cout << "elements\t: " << elems.size() << "\nevaluation points\t: " << m << " of " << npts << std::endl;
    if (dry) return 0;

    wob::SceneHost<float, 2> scene(elems);
    wob::Estimator<float, 2, wob::PoissonKernel, wob::IndirectSingleLayer, wob::InteriorDomain,
                   wob::BackwardEstimator, wob::NeumannProblem>
        est(scene, L, nrc, eps);
    float *d_pts = nullptr, *d_contrib = nullptr;
    wob::BoundaryPoint<float, 2> *d_rec = nullptr;
    double *d_acc = nullptr;
    HIP_OK(hipMalloc(&d_pts, std::max<size_t>(pts.size(), 1) * sizeof(float)));
    HIP_OK(hipMalloc(&d_rec, std::max<size_t>(n_paths, 1) * sizeof(wob::BoundaryPoint<float, 2>)));
    HIP_OK(hipMalloc(&d_contrib, std::max<size_t>(n_paths, 1) * sizeof(float)));
    HIP_OK(hipMalloc(&d_acc, std::max<size_t>((size_t)m * 3, 1) * sizeof(double)));
    HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_acc, 0, (size_t)m * 3 * sizeof(double)));

    std::vector<double> acc((size_t)m * 3);
    const float nan = std::numeric_limits<float>::quiet_NaN();
    double elapsed = 0.0;
    uint64_t total = 0;
    for (long it = 1; elapsed < max_time * 60.0 && (max_iter < 0 || it <= max_iter); ++it) {
        HIP_OK(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        est.compute_vpl_paths(n_paths, seed, (uint64_t)(it - 1), d_rec, d_contrib);
        est.aggregate_vpl(d_rec, d_contrib, n_paths, reinterpret_cast<const wob::DomainPoint<float, 2> *>(d_pts), m,
                          d_acc);
        HIP_OK(hipDeviceSynchronize());
        elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += n_paths;
        std::cout << "\33[2K\rtime " << elapsed << " s\t" << it << " iterations (" << total << " samples)\t"
                  << std::flush;
        HIP_OK(hipMemcpy(acc.data(), d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> sol(npts, nan), grad((size_t)npts * 2, 0.f);
        for (uint32_t j = 0; j < m; ++j) {
            sol[idx_of[j]] = (float)(acc[(size_t)j * 3] / (double)total);
            grad[(size_t)idx_of[j] * 2] = (float)(acc[(size_t)j * 3 + 1] / (double)total);
            grad[(size_t)idx_of[j] * 2 + 1] = (float)(acc[(size_t)j * 3 + 2] / (double)total);
        }
        save_field(out_dir + "/potential" + std::to_string(it) + ".scalar", res, sol);
        save_field(out_dir + "/gradient" + std::to_string(it) + ".vector2", res, grad);
    }
    std::cout << std::endl;
    HIP_OK(hipFree(d_pts));
    HIP_OK(hipFree(d_rec));
    HIP_OK(hipFree(d_contrib));
    HIP_OK(hipFree(d_acc));
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        return run_app(argc, argv);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
}
