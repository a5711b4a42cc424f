// wob_wos -- the reference's simple walk-on-boundary / walk-on-spheres app on the MI355X
// engine.
//
// Same command line, configuration keys and output files as apps/simple_wob_and_wos.cu
// (rsugimoto/WoBToolbox):
//   wob_wos config.json [--max-iterations N] [--dry-run] [--seed S]
// reads input_obj_file (a 2D polyline, "v x y z" and 2-vertex "f"/"l" records), output_dir,
// is_wos, evaluate_at_all_points (default false), path_length (walk on boundary only),
// num_sample_paths, grid_res, eps (default FLT_EPSILON) and max_time (minutes) (:153-168).
// Dirichlet data alternate between -1 and +1 over 40 equal arc-length pieces of the boundary
// (:200-221); the grid points with winding number < -0.5 are evaluated unless
// evaluate_at_all_points (:224-234).  Each iteration adds num_sample_paths paths per point:
//   * is_wos: walk on spheres (WoSFunc, :116-141) -> wob_wos_estimate;
//   * otherwise the double-layer walk on boundary of WoBFunc (:83-113): a line sample from x,
//     then path_length - 1 more bounces, contributions m sign(...) g(y) with the last one
//     halved.  Path by path that is the estimator of the general apps' interior Dirichlet
//     double layer with the same path length (estimator.cuh:154-191, 232-263; for path_length
//     1 twice it, as WoBFunc's halving sits in its bounce loop), so it runs as
//     Estimator<..., IndirectDoubleLayer, InteriorDomain, BackwardEstimator, DirichletProblem>.
// After iteration i it writes potential<i>.scalar in output_dir (uint32 grid_res, then
// grid_res^2 floats, NaN at points not evaluated; runtime_utils.cuh:35-57) and a copy of the
// configuration as config.json.
//
// Differences, on purpose: the walks use the engine's Philox streams (iteration i draws paths
// [(i - 1) N, i N) of every point); a walk on spheres stops after 65536 steps (the reference
// loops until it reaches the boundary).
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
    const std::string obj = cfg.str("input_obj_file");
    const std::string out_dir = cfg.str("output_dir");
    const bool is_wos = cfg.num("is_wos") != 0.0;
    const bool all_points = cfg.num_or("evaluate_at_all_points", 0.0) != 0.0;
    const unsigned L = is_wos ? 1u : (unsigned)cfg.num("path_length");
    const uint32_t paths = (uint32_t)cfg.num("num_sample_paths");
    const unsigned res = (unsigned)cfg.num("grid_res");
    const float eps = (float)cfg.num_or("eps", FLT_EPSILON);
    const double max_time = cfg.num("max_time");

    std::vector<double> V;
    std::vector<int> F;
    int vpf = 0;
    read_obj(obj, V, F, vpf);
    if (vpf != 2) throw std::runtime_error("the mesh must be a 2D polyline (2 vertices per face)");
    const size_t nseg = F.size() / 2;

    std::cout << "------------------------------------------------------------" << std::endl
              << "input json file\t: " << config_path << std::endl
              << "input obj file\t: " << obj << std::endl
              << " ( #line segments: " << nseg << " )" << std::endl
              << "output directory\t: " << out_dir << std::endl;
    if (!is_wos) std::cout << "path length\t: " << L << std::endl;
    std::cout << "#sample paths per iteration\t: " << paths << std::endl
              << "grid resolution\t: " << res << std::endl
              << "eps\t: " << eps << std::endl
              << "precision\t: float" << std::endl
              << "max time\t: " << max_time << " minutes" << std::endl
              << "is WoS\t: " << is_wos << std::endl
              << "evaluate at all points\t: " << all_points << std::endl
              << "------------------------------------------------------------" << std::endl;
    std::filesystem::create_directories(out_dir);
    std::filesystem::copy(config_path, out_dir + "/config.json", std::filesystem::copy_options::overwrite_existing);

    // boundary data (:200-221), in FP32 as the app computes it
    auto vert = [&](int k, float *p) { p[0] = (float)V[3 * k]; p[1] = (float)V[3 * k + 1]; };
    float total_length = 0.f;
    for (size_t f = 0; f < nseg; ++f) {
        float a[2], b[2];
        vert(F[2 * f], a);
        vert(F[2 * f + 1], b);
        const float dx = a[0] - b[0], dy = a[1] - b[1];
        total_length += std::sqrt(dx * dx + dy * dy);
    }
    const unsigned num_segments = 40;
    float length_sum = 0.f;
    std::vector<Elem> elems;
    for (size_t f = 0; f < nseg; ++f) {
        float a[2], b[2];
        vert(F[2 * f], a);
        vert(F[2 * f + 1], b);
        const float dx = a[0] - b[0], dy = a[1] - b[1], len = std::sqrt(dx * dx + dy * dy);
        const float tx = dx / len, ty = dy / len;
        const int segment_id = (int)(length_sum * num_segments / total_length);
        const float g = segment_id % 2 == 1 ? 1.0f : -1.0f;
        Elem e{};
        e.a[0] = a[0]; e.a[1] = a[1];
        e.b[0] = b[0]; e.b[1] = b[1];
        e.n[0] = ty; e.n[1] = -tx;
        e.boundary_value_a = e.boundary_value_b = g;
        e.boundary_type = wob::Dirichlet;
        elems.push_back(e);
        length_sum += len;
    }

    // evaluation points (:224-234)
    const unsigned npts = res * res;
    std::vector<float> pts;
    std::vector<unsigned> idx_of;
    for (unsigned i = 0; i < npts; ++i) {
        const float dx = 1.0f / (float)res;
        const float x = ((float)(i % res) + 0.5f) * dx - 0.5f, y = ((float)(i / res) + 0.5f) * dx - 0.5f;
        const double xd[2] = {x, y};
        if (!all_points && !(winding_2d(V, F, xd) < -0.5)) continue;
        pts.push_back(x);
        pts.push_back(y);
        idx_of.push_back(i);
    }
    const uint32_t m = (uint32_t)idx_of.size();
    std::cout << "evaluation points\t: " << m << " of " << npts << std::endl;
    if (dry) return 0;

    wob::SceneHost<float, 2> scene(elems);
    wob::Estimator<float, 2, wob::PoissonKernel, wob::IndirectDoubleLayer, wob::InteriorDomain,
                   wob::BackwardEstimator, wob::DirichletProblem>
        wob_est(scene, L, 1, eps);
    float *d_pts = nullptr;
    double *d_acc = nullptr;
    const size_t V1 = is_wos ? 1 : 3;   // the walk on boundary also returns grad u
    HIP_OK(hipMalloc(&d_pts, std::max<size_t>(pts.size(), 1) * sizeof(float)));
    HIP_OK(hipMalloc(&d_acc, std::max<size_t>(2 * (size_t)m * V1, 1) * sizeof(double)));
    HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_acc, 0, 2 * (size_t)m * V1 * sizeof(double)));

    std::vector<double> acc((size_t)m * V1);
    const float nan = std::numeric_limits<float>::quiet_NaN();
    double elapsed = 0.0;
    uint64_t total = 0;
    for (long it = 1; elapsed < max_time * 60.0 && (max_iter < 0 || it <= max_iter); ++it) {
        HIP_OK(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        if (m > 0) {
            if (is_wos)
                wob::check(wob_wos_estimate(scene.get_device_repr(), d_pts, m, 0, paths, seed, total, 0, eps, 0,
                                            d_acc, d_acc + m, nullptr));
            else
                wob_est.compute_domain(reinterpret_cast<const wob::DomainPoint<float, 2> *>(d_pts), m, paths, seed,
                                       total, d_acc, d_acc + (size_t)m * V1);
        }
        HIP_OK(hipDeviceSynchronize());
        elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += paths;
        std::cout << "\33[2K\rtime " << elapsed << " s\t" << it << " iterations (" << total << " samples)\t"
                  << std::flush;
        HIP_OK(hipMemcpy(acc.data(), d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
        // with path_length 1 WoBFunc's single term is not halved (its halving sits in the bounce
        // loop), while the estimator halves its last term: scale by 2
        const double scale = (!is_wos && L == 1) ? 2.0 : 1.0;
        std::vector<float> sol(npts, nan);
        for (uint32_t j = 0; j < m; ++j) sol[idx_of[j]] = (float)(scale * acc[(size_t)j * V1] / (double)total);
        save_field(out_dir + "/potential" + std::to_string(it) + ".scalar", res, sol);
    }
    std::cout << std::endl;
    HIP_OK(hipFree(d_pts));
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
