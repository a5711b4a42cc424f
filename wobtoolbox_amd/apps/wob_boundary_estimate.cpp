// wob_boundary_estimate -- the reference's boundary-estimate app on the MI355X engine.
//
// Same command line, configuration keys, scene and output files as apps/boundary_estimate.cu
// (rsugimoto/WoBToolbox):
//   wob_boundary_estimate config.json [--max-iterations N] [--dry-run] [--seed S]
// reads output_dir, path_length, num_sample_paths_domain, num_sample_paths_boundary,
// num_resampling_candidates (read, unused: the app's estimators take one candidate), grid_res,
// eps (default FLT_EPSILON) and max_time (minutes) (:115-126).  The scene is the circle of
// radius 0.5 cut into 720 segments, element normals -n for n = (t_y, -t_x), t = normalised
// v1 - v2, with Neumann data (1, 200) on segments i mod 240 in {0, 239}, (-1, -200) on
// {119, 120} and 0 elsewhere (:153-172).  The solver is the single layer, interior, forward
// estimator, Neumann problem, one resampling candidate, first-kind constant 1 (:44-48, :67-71).
// Two progressive loops of max_time minutes each (:215-247, :258-293):
//   * domain estimates at the grid_res^2 cell centres of (-1/2, 1/2)^2 (utils.cuh:92-99) with
//     |x| <= 1/2 (:175-179), num_sample_paths_domain paths per pass; after pass i
//     potential<i>.scalar (NaN outside the disc) and gradient<i>.vector2 (0 outside);
//   * boundary estimates at the 720 segment midpoints, BoundaryPoint{x, n, 1/2, 0, Neumann}
//     with the segment's n (not the element's -n) (:74-97), num_sample_paths_boundary paths
//     per pass; after pass i boundary_solution<i>.scalar (uint32 720, then 720 floats).
// Fields are written as runtime_utils.cuh:35-57 writes them, with the sums divided by the
// paths so far; config.json is a copy of the configuration.  --max-iterations N caps each loop.
//
// Differences, on purpose: the walks use the engine's Philox streams (pass i walks paths
// [(i-1) N, i N) of every point; the boundary loop uses stream 1) and the per-point sums are
// accumulated in FP64.
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
using Est = wob::Estimator<float, 2, wob::PoissonKernel, wob::IndirectSingleLayer, wob::InteriorDomain,
                           wob::ForwardEstimator, wob::NeumannProblem>;

constexpr unsigned kSegments = 720;

// boundary_estimate.cu:155-163 (and :75-83): vertex i of the circle, the product in double
// rounded to FP32 as Eigen's float vector stores it; t = normalised v1 - v2 in FP32
void circle_segment(unsigned i, float *v1, float *v2, float *n) {
    const double r = 0.5f;
    v1[0] = (float)(r * std::sin((2 * M_PI) * i / kSegments));
    v1[1] = (float)(r * std::cos((2 * M_PI) * i / kSegments));
    v2[0] = (float)(r * std::sin((2 * M_PI) * (i + 1) / kSegments));
    v2[1] = (float)(r * std::cos((2 * M_PI) * (i + 1) / kSegments));
    const float tx = v1[0] - v2[0], ty = v1[1] - v2[1];
    const float l = std::sqrt(tx * tx + ty * ty);
    n[0] = ty / l;
    n[1] = -(tx / l);
}

template <typename F> double timed(F &&f) {
    HIP_OK(hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    f();
    HIP_OK(hipDeviceSynchronize());
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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
        std::cout << "Failed to load config file: " << config_path << std::endl;   // boundary_estimate.cu:105-109
        return 0;
    }
    std::stringstream buf;
    buf << cf.rdbuf();
    const Config cfg = parse_config(buf.str());
    const std::string out_dir = cfg.str("output_dir");
    const unsigned L = (unsigned)cfg.num("path_length");
    const unsigned paths_d = (unsigned)cfg.num("num_sample_paths_domain");
    const unsigned paths_b = (unsigned)cfg.num("num_sample_paths_boundary");
    const unsigned nrc = (unsigned)cfg.num_or("num_resampling_candidates", 1);
    const unsigned res = (unsigned)cfg.num("grid_res");
    const float eps = (float)cfg.num_or("eps", FLT_EPSILON);
    const double max_time = cfg.num("max_time");

    std::cout << "------------------------------------------------------------\n"
              << "input json file\t: " << config_path << "\noutput directory\t: " << out_dir
              << "\npath length\t: " << L << "\n#sample paths domain per iteration\t: " << paths_d
              << "\n#sample paths boundary per iteration\t: " << paths_b << "\n#resampling candidates\t: " << nrc
              << "\ngrid resolution\t: " << res << "\neps\t: " << eps << "\nprecision\t: float\nmax time\t: "
              << max_time << " minutes\n------------------------------------------------------------" << std::endl;
    std::filesystem::create_directories(out_dir);
    std::filesystem::copy(config_path, out_dir + "/config.json", std::filesystem::copy_options::overwrite_existing);

    // the scene (boundary_estimate.cu:153-172) and the boundary points (:74-92)
    std::vector<Elem> elems(kSegments);
    std::vector<wob::BoundaryPoint<float, 2>> bpts(kSegments);
    for (unsigned i = 0; i < kSegments; ++i) {
        float v1[2], v2[2], n[2];
        circle_segment(i, v1, v2, n);
        Elem &e = elems[i];
        e.a[0] = v1[0]; e.a[1] = v1[1];
        e.b[0] = v2[0]; e.b[1] = v2[1];
        e.n[0] = -n[0]; e.n[1] = -n[1];
        const float g = 200.0f;
        const unsigned k = i % 240;
        e.boundary_value_a = k == 0 || k == 239 ? 1.0f : k == 120 || k == 119 ? -1.0f : 0.0f;
        e.boundary_value_b = k == 0 || k == 239 ? g : k == 120 || k == 119 ? -g : 0.0f;
        e.boundary_type = wob::Neumann;
        auto &b = bpts[i];
        b.p[0] = (v1[0] + v2[0]) / 2.0f;
        b.p[1] = (v1[1] + v2[1]) / 2.0f;
        b.n[0] = n[0]; b.n[1] = n[1];
        b.interior_angle = 0.5f;
        b.boundary_value = 0.0f;
        b.boundary_type = (float)wob::Neumann;
        b.robin_alpha = 0.0f;
    }
    // the grid (utils.cuh:92-99 in FP32, domain size 1) and its mask |x| <= 1/2 (:175-179)
    const unsigned npts = res * res;
    const float dx = 1.0f / (float)res;
    std::vector<float> pts;
    std::vector<unsigned> idx_of;
    for (unsigned i = 0; i < npts; ++i) {
        const float x = ((float)(i % res) + 0.5f) * dx - 0.5f, y = ((float)(i / res) + 0.5f) * dx - 0.5f;
        if (std::sqrt(x * x + y * y) <= 0.5f) {
            pts.push_back(x);
            pts.push_back(y);
            idx_of.push_back(i);
        }
    }
    const uint32_t m = (uint32_t)idx_of.size();
    std::cout << "evaluation points\t: " << m << " of " << npts << std::endl;
    if (dry) return 0;

    wob::SceneHost<float, 2> scene(elems);
    Est est(scene, L, 1, eps, 1.0f);
    float *d_pts = nullptr, *d_bpts = nullptr;
    double *d_acc = nullptr, *d_bacc = nullptr;
    HIP_OK(hipMalloc(&d_pts, std::max<size_t>(pts.size(), 1) * sizeof(float)));
    HIP_OK(hipMalloc(&d_bpts, sizeof(bpts[0]) * kSegments));
    HIP_OK(hipMalloc(&d_acc, std::max<size_t>(6 * (size_t)m, 1) * sizeof(double)));
    HIP_OK(hipMalloc(&d_bacc, 2 * kSegments * sizeof(double)));
    HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_bpts, bpts.data(), sizeof(bpts[0]) * kSegments, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_acc, 0, std::max<size_t>(6 * (size_t)m, 1) * sizeof(double)));
    HIP_OK(hipMemset(d_bacc, 0, 2 * kSegments * sizeof(double)));

    const float nan = std::numeric_limits<float>::quiet_NaN();
    std::cout << "computing solution in the domain" << std::endl;
    double elapsed = 0.0;
    uint64_t total = 0;
    std::vector<double> acc(3 * (size_t)m);
    for (long it = 1; elapsed < max_time * 60.0 && (max_iter < 0 || it <= max_iter); ++it) {
        if (m > 0)
            elapsed += timed([&] {
                est.compute_domain(reinterpret_cast<const wob::DomainPoint<float, 2> *>(d_pts), m, paths_d, seed,
                                   total, d_acc, d_acc + 3 * (size_t)m);
            });
        total += paths_d;
        std::cout << "\33[2K\rtime " << elapsed << " s\t" << it << " iterations (" << total << " samples)\t"
                  << std::flush;
        HIP_OK(hipMemcpy(acc.data(), d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> sol(npts, nan), grad(2 * (size_t)npts, 0.0f);
        for (uint32_t j = 0; j < m; ++j) {
            const unsigned i = idx_of[j];
            sol[i] = (float)(acc[3 * (size_t)j] / (double)total);
            grad[2 * (size_t)i] = (float)(acc[3 * (size_t)j + 1] / (double)total);
            grad[2 * (size_t)i + 1] = (float)(acc[3 * (size_t)j + 2] / (double)total);
        }
        save_field(out_dir + "/potential" + std::to_string(it) + ".scalar", res, sol);
        save_field(out_dir + "/gradient" + std::to_string(it) + ".vector2", res, grad);
    }
    std::cout << std::endl << "computing solution on boundary" << std::endl;
    elapsed = 0.0;
    total = 0;
    std::vector<double> bacc(kSegments);
    for (long it = 1; elapsed < max_time * 60.0 && (max_iter < 0 || it <= max_iter); ++it) {
        elapsed += timed([&] {
            wob::check(wob_estimate_boundary(scene.get_device_repr(), &est.desc(), d_bpts, kSegments, 0, paths_b,
                                             seed, total, 1u, d_bacc, d_bacc + kSegments, nullptr, nullptr));
        });
        total += paths_b;
        std::cout << "\33[2K\rtime " << elapsed << " s\t" << it << " iterations (" << total << " samples)\t"
                  << std::flush;
        HIP_OK(hipMemcpy(bacc.data(), d_bacc, kSegments * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> bs(kSegments);
        for (unsigned i = 0; i < kSegments; ++i) bs[i] = (float)(bacc[i] / (double)total);
        save_field(out_dir + "/boundary_solution" + std::to_string(it) + ".scalar", kSegments, bs);
    }
    std::cout << std::endl;
    HIP_OK(hipFree(d_pts));
    HIP_OK(hipFree(d_bpts));
    HIP_OK(hipFree(d_acc));
    HIP_OK(hipFree(d_bacc));
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
