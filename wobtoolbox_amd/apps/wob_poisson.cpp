// wob_poisson -- the reference's Poisson-equation app on the MI355X engine.
//
// Same command line, configuration keys, scene and output files as apps/poisson_equation.cu
// (rsugimoto/WoBToolbox):
//   wob_poisson config.json [--max-iterations N] [--dry-run] [--seed S]
// reads output_dir, path_length, num_sample_paths, num_volume_samples, grid_res, eps (default
// FLT_EPSILON) and max_time (minutes) (:77-85); builds the square (-1, 1)^2 of four Dirichlet
// segments with data 1 (:112-131) and the constant source 5 on a grid_res^2 volume cache over
// it (:133-135); and runs progressive passes of num_sample_paths double-layer walks (interior,
// backward, Dirichlet, one resampling candidate, first-kind constant 1, num_volume_samples
// volume samples; :42-46) from every grid cell centre (utils.cuh:92-99 with domain size 2;
// :49) until max_time minutes of walk time have passed (:165-197).  After pass i it writes
// potential<i>.scalar and gradient<i>.vector2 (uint32 grid_res, then grid_res^2 float records;
// runtime_utils.cuh:35-57) with the sums divided by the paths so far, and a copy of the
// configuration as config.json.
//
// Differences, on purpose: the walks use the engine's Philox streams (pass i walks paths
// [(i-1) N, i N) of every point) and the per-point sums are accumulated in FP64.
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

Elem side(float ax, float ay, float bx, float by, float nx, float ny) {
    Elem e{};
    e.a[0] = ax; e.a[1] = ay;
    e.b[0] = bx; e.b[1] = by;
    e.n[0] = nx; e.n[1] = ny;
    e.boundary_value_a = e.boundary_value_b = 1.0f;
    e.boundary_type = wob::Dirichlet;
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
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--max-iterations") && a + 1 < argc) max_iter = std::atol(argv[++a]);
        else if (!std::strcmp(argv[a], "--dry-run")) dry = true;
        else if (!std::strcmp(argv[a], "--seed") && a + 1 < argc) seed = (uint64_t)std::atoll(argv[++a]);
    }
    std::ifstream cf(config_path);
    if (!cf) {
        std::cout << "Failed to load config file: " << config_path << std::endl;   // poisson_equation.cu:68-72
        return 0;
    }
    std::stringstream buf;
    buf << cf.rdbuf();
    const Config cfg = parse_config(buf.str());
    const std::string out_dir = cfg.str("output_dir");
    const unsigned L = (unsigned)cfg.num("path_length"), paths = (unsigned)cfg.num("num_sample_paths");
    const unsigned nvs = (unsigned)cfg.num("num_volume_samples");
    const unsigned res = (unsigned)cfg.num("grid_res");
    const float eps = (float)cfg.num_or("eps", FLT_EPSILON);
    const double max_time = cfg.num("max_time");

    std::cout << "------------------------------------------------------------\n"
              << "input json file\t: " << config_path << "\noutput directory\t: " << out_dir
              << "\npath length\t: " << L << "\n#sample paths per iteration\t: " << paths
              << "\n#num volume samples\t: " << nvs << "\ngrid resolution\t: " << res << "\neps\t: " << eps
              << "\nprecision\t: float\nmax time\t: " << max_time << " minutes"
              << "\n------------------------------------------------------------" << std::endl;
    std::filesystem::create_directories(out_dir);
    std::filesystem::copy(config_path, out_dir + "/config.json", std::filesystem::copy_options::overwrite_existing);

    // poisson_equation.cu:112-131: the square's sides, outward normals, data 1
    const float h = 1.0f;
    const std::vector<Elem> elems = {side(-h, -h, -h, h, -1.f, 0.f), side(-h, h, h, h, 0.f, 1.f),
                                     side(h, h, h, -h, 1.f, 0.f), side(h, -h, -h, -h, 0.f, -1.f)};
    // every cell centre of the 2 x 2 grid (utils.cuh:92-99 with domain size 2)
    const unsigned npts = res * res;
    std::vector<float> pts(2 * (size_t)npts);
    const float dx = 2.0f / (float)res;   // in FP32, as the reference evaluates it
    for (unsigned i = 0; i < npts; ++i) {
        pts[2 * (size_t)i] = ((float)(i % res) + 0.5f) * dx - 1.0f;
        pts[2 * (size_t)i + 1] = ((float)(i / res) + 0.5f) * dx - 1.0f;
    }
    std::cout << "evaluation points\t: " << npts << std::endl;
    if (dry) return 0;

    wob::SceneHost<float, 2> scene(elems);
    scene.set_volume_cache(res, 2.0f, std::vector<float>(npts, 5.0f));   // poisson_equation.cu:133-135
    wob::Estimator<float, 2, wob::PoissonKernel, wob::IndirectDoubleLayer, wob::InteriorDomain,
                   wob::BackwardEstimator, wob::DirichletProblem>
        est(scene, L, 1, eps, 1.0f, nvs);

    float *d_pts = nullptr;
    double *d_acc = nullptr;
    HIP_OK(hipMalloc(&d_pts, pts.size() * sizeof(float)));
    HIP_OK(hipMalloc(&d_acc, 2 * (size_t)npts * 3 * sizeof(double)));
    HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_acc, 0, 2 * (size_t)npts * 3 * sizeof(double)));
    std::vector<double> acc((size_t)npts * 3);
    double elapsed = 0.0;
    uint64_t total = 0;
    for (long it = 1; elapsed < max_time * 60.0 && (max_iter < 0 || it <= max_iter); ++it) {
        HIP_OK(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        est.compute_domain(reinterpret_cast<const wob::DomainPoint<float, 2> *>(d_pts), npts, paths, seed, total,
                           d_acc, d_acc + (size_t)npts * 3);
        HIP_OK(hipDeviceSynchronize());
        elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += paths;
        std::cout << "\33[2K\rtime " << elapsed << " s\t" << it << " iterations (" << total << " samples)\t"
                  << std::flush;
        HIP_OK(hipMemcpy(acc.data(), d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> sol(npts), grad(2 * (size_t)npts);
        for (unsigned i = 0; i < npts; ++i) {
            sol[i] = (float)(acc[3 * (size_t)i] / (double)total);
            grad[2 * (size_t)i] = (float)(acc[3 * (size_t)i + 1] / (double)total);
            grad[2 * (size_t)i + 1] = (float)(acc[3 * (size_t)i + 2] / (double)total);
        }
        save_field(out_dir + "/potential" + std::to_string(it) + ".scalar", res, sol);
        save_field(out_dir + "/gradient" + std::to_string(it) + ".vector2", res, grad);
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
