// wob_general_2d / wob_general_3d -- the reference's general apps on the MI355X engine.
//
// Same command line, configuration keys, input meshes and output files as
// apps/general_2d.cu and apps/general_3d.cu (rsugimoto/WoBToolbox):
//   wob_general_<D>d config.json [--max-iterations N] [--dry-run] [--seed S]
//                    [--harmonic-cv N_SOURCES [RADIUS_SCALE]]
// reads input_obj_file, output_dir, path_length, num_sample_paths, num_resampling_candidates,
// grid_res, eps, first_kind_equation_scaling_constant, max_time (minutes), equation_kernel,
// bie_model, domain, estimator, problem (general_2d.cu:174-195); builds the boundary with
// data from the apps' analytic solutions (general_2d.cu:73-94, general_3d.cu:75-98); masks
// the grid by winding number; and runs progressive passes of num_sample_paths paths per
// point until max_time minutes of walk time have passed (general_2d.cu:318-367).  Outputs in
// output_dir: config.json (a copy), rmse_list.csv ("samples,seconds,rmse_u,rmse_du/dx,..."),
// potential_analytical.scalar, gradient_analytical.vector<D>, and potential<i>.scalar /
// gradient<i>.vector<D> after pass i (uint32 grid_res, then grid_res^2 float records;
// runtime_utils.cuh:35-57), which visualize.py reads.
//
// Differences, on purpose:
//   - mixed problems draw the per-element boundary type from std::mt19937(seed) (default 0);
//     the apps use std::random_device (general_2d.cu:242-253), so their runs cannot be repeated;
//   - the RMSE is computed on the normalised estimate; general_3d.cu:367-368 passes the raw
//     sums instead;
//   - paths use the engine's Philox streams (pass i walks paths [(i-1) N, i N) of every point).
// Addition: --harmonic-cv N fits the engine's harmonic control variate (N fundamental solutions
// fitted to 32768 boundary samples of the walk's data; wob_scene_fit_harmonic_cv) before the
// first pass and walks the residual (WOB_EST_HARMONIC_CV).  Interior domains: sources on a
// sphere or circle of RADIUS_SCALE x the scene radius (default 2); exterior domains: inside
// the body at RADIUS_SCALE x the smallest centre-to-element distance (default 0.75).  The fit's
// time counts in the "seconds" column.  It takes every model and estimator but the direct
// model in an exterior domain; exterior double-layer runs fit with sum_j w_j = 0
// (WOB_HCV_ZERO_CHARGE: a double-layer potential decays at infinity).
// The 3D Robin elements keep the apps' robin alpha (1, 1, 0): the aggregate initialisation at
// general_3d.cu:279-284 leaves the third vertex's alpha at 0.
#include <hip/hip_runtime_api.h>

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <array>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/wob/wob.hpp"
#include "app_common.hpp"

#ifndef WOB_APP_DIM
#error "WOB_APP_DIM must be 2 or 3"
#endif
constexpr unsigned D = WOB_APP_DIM;

namespace {

using namespace wobapp;


// ---------------------------------------------------------------------------------------
// the apps' analytic solutions: [u, grad u]
// ---------------------------------------------------------------------------------------
enum SolutionType { kInterior2, kExterior1 };

template <unsigned DD> std::array<double, DD + 1> analytic_d(SolutionType st, const double *y) {
    std::array<double, DD + 1> r{};
    const double pi = M_PI;
    if constexpr (DD == 2) {
        if (st == kExterior1) {                                           // general_2d.cu:78-82
            const double s = y[0] * y[0] + y[1] * y[1];
            r = {y[0] / s, 1.0 / s - 2.0 * y[0] * y[0] / (s * s), -2.0 * y[0] * y[1] / (s * s)};
        } else {                                                          // general_2d.cu:86-91
            r = {std::cos(pi * y[0]) * std::sinh(pi * y[1]), -pi * std::sin(pi * y[0]) * std::sinh(pi * y[1]),
                 pi * std::cos(pi * y[0]) * std::cosh(pi * y[1])};
        }
    } else {
        if (st == kExterior1) {                                           // general_3d.cu:80-85
            const double s = y[0] * y[0] + y[1] * y[1] + y[2] * y[2], n = std::sqrt(s);
            r = {y[2] / (s * n), -3.0 * y[2] * y[0] / (s * s * n), -3.0 * y[2] * y[1] / (s * s * n),
                 1.0 / (s * n) - 3.0 * y[2] * y[2] / (s * s * n)};
        } else {                                                          // general_3d.cu:89-94
            r = {-std::cos(pi * y[1]) * std::sinh(pi * y[2]), 0.0, pi * std::sin(pi * y[1]) * std::sinh(pi * y[2]),
                 -pi * std::cos(pi * y[1]) * std::cosh(pi * y[2])};
        }
    }
    return r;
}
std::array<double, D + 1> analytic(SolutionType st, const double *y) { return analytic_d<D>(st, y); }

// winding number of the boundary around x (2D: signed angle sum; 3D: solid-angle sum,
// Van Oosterom and Strackee)
double winding(const std::vector<double> &V, const std::vector<int> &F, const double *x) {
    double w = 0.0;
    if constexpr (D == 2) {
        return winding_2d(V, F, x);
    } else {
        for (size_t f = 0; f < F.size(); f += 3) {
            double a[3], b[3], c[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = V[3 * F[f] + k] - x[k];
                b[k] = V[3 * F[f + 1] + k] - x[k];
                c[k] = V[3 * F[f + 2] + k] - x[k];
            }
            const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            const double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            const double lc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                               a[2] * (b[0] * c[1] - b[1] * c[0]);
            const double den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc +
                               (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la +
                               (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]) * lb;
            w += 2.0 * std::atan2(det, den);
        }
        return w / (4.0 * M_PI);
    }
}

// third vertex of a triangle; the Robin alpha of that vertex stays 0 as in the apps
inline void set_third_vertex(wob::ElementXs<float, 2> &, const double *, double) {}
inline void set_third_vertex(wob::ElementXs<float, 3> &e, const double *v, double bv) {
    for (int k = 0; k < 3; ++k) e.c[k] = (float)v[k];
    e.boundary_value_c = (float)bv;
    e.robin_alpha_c = 0.0f;
}

struct Run {
    const wob::SceneHost<float, D> *scene;
    unsigned L, nrc;
    float eps, fkesc;
    const float *d_pts;
    uint32_t n, paths;
    uint64_t path_base;
    double *d_sum, *d_sq;
    uint32_t flags;
    template <wob::BIEModelType B, wob::DomainType Dm, wob::EstimatorType E, wob::ProblemType P> void operator()() {
        wob::Estimator<float, D, wob::PoissonKernel, B, Dm, E, P> est(*scene, L, nrc, eps, fkesc);
        est.desc().flags |= flags;
        est.compute_domain(reinterpret_cast<const wob::DomainPoint<float, D> *>(d_pts), n, paths, 0, path_base, d_sum,
                           d_sq);
    }
};

int run_app(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr,
                     "usage: %s config.json [--max-iterations N] [--dry-run] [--seed S] "
                     "[--harmonic-cv N_SOURCES [RADIUS_SCALE]]\n",
                     argv[0]);
        return 1;
    }
    const std::string config_path = argv[1];
    long max_iter = -1;
    bool dry = false;
    unsigned seed = 0;
    unsigned hcv_sources = 0;
    float hcv_radius = 0.0f;   // 0: the side's default
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--max-iterations") && a + 1 < argc) max_iter = std::atol(argv[++a]);
        else if (!std::strcmp(argv[a], "--dry-run")) dry = true;
        else if (!std::strcmp(argv[a], "--seed") && a + 1 < argc) seed = (unsigned)std::atol(argv[++a]);
        else if (!std::strcmp(argv[a], "--harmonic-cv") && a + 1 < argc) {
            hcv_sources = (unsigned)std::atol(argv[++a]);
            if (a + 1 < argc && argv[a + 1][0] != '-') hcv_radius = (float)std::atof(argv[++a]);
        }
    }
    std::ifstream cf(config_path);
    if (!cf) {
        std::cout << "Failed to load config file: " << config_path << std::endl;   // general_2d.cu:166-169
        return 0;
    }
    std::stringstream buf;
    buf << cf.rdbuf();
    const Config cfg = parse_config(buf.str());

    const std::string obj = cfg.str("input_obj_file"), out_dir = cfg.str("output_dir");
    const unsigned L = (unsigned)cfg.num("path_length"), paths = (unsigned)cfg.num("num_sample_paths");
    const unsigned nrc = (unsigned)cfg.num_or("num_resampling_candidates", 1);
    const unsigned res = (unsigned)cfg.num("grid_res");
    const float eps = (float)cfg.num_or("eps", std::numeric_limits<float>::epsilon());
    const float fkesc = (float)cfg.num_or("first_kind_equation_scaling_constant", 1.0);
    const double max_time = cfg.num("max_time");
    if (wob::string_to_equation_kernel_type(cfg.str("equation_kernel")) != wob::PoissonKernel)
        throw std::runtime_error("equation_kernel must be PoissonKernel");
    const auto bie = wob::string_to_bie_model_type(cfg.str("bie_model"));
    const auto domain = wob::string_to_domain_type(cfg.str("domain"));
    const auto est = wob::string_to_estimator_type(cfg.str("estimator"));
    const auto problem = wob::string_to_problem_type(cfg.str("problem"));
    if (hcv_sources > 0 && bie == wob::Direct && domain != wob::InteriorDomain)
        throw std::runtime_error("--harmonic-cv needs the InteriorDomain with the Direct model");
    const bool hcv_exterior = domain == wob::ExteriorDomain;
    if (hcv_radius == 0.0f) hcv_radius = hcv_exterior ? 0.75f : 2.0f;

    std::vector<double> V;
    std::vector<int> F;
    int vpf = 0;
    read_obj(obj, V, F, vpf);
    if (vpf < (int)D) throw std::runtime_error("mesh faces have fewer than " + std::to_string(D) + " vertices");
    if (vpf != (int)D) {   // keep the first D vertices of each face (the 2D app reads triangles' first two)
        std::vector<int> G;
        for (size_t f = 0; f < F.size(); f += vpf) G.insert(G.end(), F.begin() + f, F.begin() + f + D);
        F.swap(G);
    }
    const size_t nv = V.size() / 3, nf = F.size() / D;
    if constexpr (D == 3) {   // general_3d.cu:205-211: fit into [-1/2, 1/2]^3
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (size_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], V[3 * i + k]); hi[k] = std::max(hi[k], V[3 * i + k]); }
        const double size = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        for (size_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) V[3 * i + k] = (V[3 * i + k] - (lo[k] + 0.5 * (hi[k] - lo[k]))) / size;
    }
    if (domain == wob::ExteriorDomain) {   // general_2d.cu:204-207, general_3d.cu:213-216
        for (size_t i = 0; i < nv; ++i) {
            for (int k = 0; k < 3; ++k) V[3 * i + k] *= 0.5;
            V[3 * i + 1] += D == 2 ? 0.0625 : 0.05;
        }
    }
    // mesh vertices are FP32 in the apps (ScalarType = float)
    for (auto &v : V) v = (double)(float)v;

    std::cout << "------------------------------------------------------------\n"
              << "input json file\t: " << config_path << "\ninput obj file\t: " << obj << "\n ( #faces: " << nf
              << " )\noutput directory\t: " << out_dir << "\npath length\t: " << L << "\n#sample paths\t: " << paths
              << "\n#resampling candidates\t: " << nrc << "\ngrid resolution\t: " << res << "\neps\t: " << eps
              << "\nfirst kind equation scaling constant\t: " << fkesc << "\nprecision\t: float\nmax time\t: "
              << max_time << " minutes\nBIE model type\t: " << wob::bie_model_type_to_string(bie)
              << "\ndomain type\t: " << wob::domain_type_to_string(domain)
              << "\nestimator type\t: " << wob::estimator_type_to_string(est)
              << "\nproblem type\t: " << wob::problem_type_to_string(problem)
              << "\n------------------------------------------------------------" << std::endl;

    std::filesystem::create_directories(out_dir);
    std::filesystem::copy(config_path, out_dir + "/config.json", std::filesystem::copy_options::overwrite_existing);
    std::ofstream rmse_out(out_dir + "/rmse_list.csv");

    const SolutionType st = domain == wob::InteriorDomain ? kInterior2 : kExterior1;
    std::mt19937 mt(seed);
    std::uniform_int_distribution<int> dist(0, 2);
    std::vector<wob::ElementXs<float, D>> elems(nf);
    for (size_t f = 0; f < nf; ++f) {
        double v[3][3] = {};
        for (unsigned j = 0; j < D; ++j)
            for (int k = 0; k < 3; ++k) v[j][k] = V[3 * F[D * f + j] + k];
        double n[3] = {0, 0, 0};
        if constexpr (D == 2) {   // general_2d.cu:248-249: t = normalize(v1 - v2), n = (t_y, -t_x)
            const float tx = (float)(v[0][0] - v[1][0]), ty = (float)(v[0][1] - v[1][1]);
            const float tl = std::sqrt(tx * tx + ty * ty);
            n[0] = ty / tl; n[1] = -(tx / tl);
        } else {                  // igl::per_face_normals: normalize((v1 - v0) x (v2 - v0))
            const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
            const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
            double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const double l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            for (int k = 0; k < 3; ++k) n[k] = c[k] / l;
        }
        const int type = problem == wob::MixedProblem ? dist(mt) : (int)problem;
        double bv[3] = {0, 0, 0};
        for (unsigned j = 0; j < D; ++j) {
            const auto a = analytic(st, v[j]);
            double dn = 0;
            for (unsigned k = 0; k < D; ++k) dn += a[1 + k] * (double)(float)n[k];
            bv[j] = type == wob::Dirichlet ? a[0] : type == wob::Neumann ? dn : a[0] + dn;
        }
        auto &e = elems[f];
        for (unsigned k = 0; k < D; ++k) { e.a[k] = (float)v[0][k]; e.b[k] = (float)v[1][k]; e.n[k] = (float)n[k]; }
        e.boundary_value_a = (float)bv[0];
        e.boundary_value_b = (float)bv[1];
        e.boundary_type = (wob::BoundaryType)type;
        const float alpha = type == wob::Robin ? 1.0f : 0.0f;
        e.robin_alpha_a = alpha;
        e.robin_alpha_b = alpha;
        set_third_vertex(e, v[2], bv[2]);
    }

    // evaluation grid: the plane (2D) or the x = 0 slice with (y, z) = (row, column) (3D)
    const unsigned npts = res * res;
    const float nan = std::numeric_limits<float>::quiet_NaN();
    std::vector<float> ana_u(npts, nan), ana_g((size_t)npts * D, 0.f);
    std::vector<double> ana_full((size_t)npts * (D + 1), 0.0);
    std::vector<char> mask(npts, 0);
    std::vector<float> pts;
    std::vector<unsigned> idx_of;
    for (unsigned i = 0; i < npts; ++i) {
        double p2[2], x[3] = {0, 0, 0};
        grid_point(i, res, p2);
        if constexpr (D == 2) { x[0] = (float)p2[0]; x[1] = (float)p2[1]; }
        else { x[0] = 0.0; x[1] = (float)p2[1]; x[2] = (float)p2[0]; }
        const double w = winding(V, F, x);
        const bool in = D == 2 ? (domain == wob::InteriorDomain ? w < -0.5 : w >= -0.5)
                               : (domain == wob::InteriorDomain ? w > 0.5 : w <= 0.5);
        if (!in) continue;
        mask[i] = 1;
        const auto a = analytic(st, x);
        ana_u[i] = (float)a[0];
        for (unsigned k = 0; k < D; ++k) ana_g[(size_t)i * D + k] = (float)a[1 + k];
        for (unsigned k = 0; k <= D; ++k) ana_full[(size_t)i * (D + 1) + k] = a[k];
        for (unsigned k = 0; k < D; ++k) pts.push_back((float)x[k]);
        idx_of.push_back(i);
    }
    const std::string gext = ".vector" + std::to_string(D);
    save_field(out_dir + "/potential_analytical.scalar", res, ana_u);
    save_field(out_dir + "/gradient_analytical" + gext, res, ana_g);
    const uint32_t m = (uint32_t)idx_of.size();
    std::cout << "evaluation points\t: " << m << " of " << npts << std::endl;
    if (dry) return 0;

    wob::SceneHost<float, D> scene(elems);
    float *d_pts = nullptr;
    double *d_acc = nullptr;
    const size_t V1 = D + 1;
    HIP_OK(hipMalloc(&d_pts, std::max<size_t>(pts.size(), 1) * sizeof(float)));
    HIP_OK(hipMalloc(&d_acc, std::max<size_t>(2 * m * V1, 1) * sizeof(double)));
    HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(d_acc, 0, 2 * m * V1 * sizeof(double)));

    double elapsed = 0.0;
    if (hcv_sources > 0) {
        HIP_OK(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        const float rms = scene.fit_harmonic_cv(hcv_sources, hcv_radius, 32768, 0, eps, 1e-10f, seed, nullptr,
                                                hcv_exterior, hcv_exterior && bie == wob::IndirectDoubleLayer);
        HIP_OK(hipDeviceSynchronize());
        elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "harmonic control variate\t: " << hcv_sources << " sources at " << hcv_radius
                  << (hcv_exterior ? " x the inner radius (exterior)" : " x the scene radius") << ", fit residual "
                  << rms << ", " << elapsed << " s" << std::endl;
    }

    std::vector<double> acc(2 * m * V1);
    const unsigned ref_idx = domain == wob::InteriorDomain ? res * (res / 2) + res / 2 : 0u;   // shift_solution
    uint64_t total = 0;
    for (long it = 1; elapsed < max_time * 60.0 && (max_iter < 0 || it <= max_iter); ++it) {
        HIP_OK(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        Run r{&scene, L, nrc, eps, fkesc, d_pts, m, paths, total, d_acc, d_acc + m * V1,
              hcv_sources > 0 ? (uint32_t)WOB_EST_HARMONIC_CV : 0u};
        if (m > 0 && !wob::run(bie, domain, est, problem, r)) throw std::runtime_error("invalid solver types");
        HIP_OK(hipDeviceSynchronize());
        elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += paths;
        HIP_OK(hipMemcpy(acc.data(), d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));

        std::vector<float> sol(npts, nan), grad((size_t)npts * D, 0.f);
        for (uint32_t j = 0; j < m; ++j) {
            sol[idx_of[j]] = (float)(acc[(size_t)j * V1] / (double)total);
            for (unsigned k = 0; k < D; ++k) grad[(size_t)idx_of[j] * D + k] = (float)(acc[(size_t)j * V1 + 1 + k] / (double)total);
        }
        if (problem == wob::NeumannProblem && mask[ref_idx]) {   // general_2d.cu:116-131, general_3d.cu:118-133
            // the reference value is taken at idx_to_domain_point<Dim>(ref_idx): for 3D the index
            // is unflattened over a 3D grid, as in the apps
            double x[3] = {0, 0, 0};
            const double dx = 1.0 / res;
            const unsigned u3[3] = {ref_idx % res, (ref_idx / res) % res, ref_idx / (res * res)};
            for (unsigned k = 0; k < D; ++k) x[k] = ((double)(D == 2 && k == 1 ? ref_idx / res : u3[k]) + 0.5) * dx - 0.5;
            const float off = sol[ref_idx] - (float)analytic(st, x)[0];
            for (unsigned i = 0; i < npts; ++i)
                if (mask[i]) sol[i] -= off;
        }
        double se[1 + D] = {0};
        for (uint32_t j = 0; j < m; ++j) {
            const unsigned i = idx_of[j];
            se[0] += std::pow((double)sol[i] - ana_full[(size_t)i * V1], 2);
            for (unsigned k = 0; k < D; ++k) se[1 + k] += std::pow((double)grad[(size_t)i * D + k] - ana_full[(size_t)i * V1 + 1 + k], 2);
        }
        std::ostringstream line;
        line << total << "," << elapsed;
        for (unsigned k = 0; k <= D; ++k) line << "," << std::sqrt(se[k] / std::max<uint32_t>(m, 1));
        rmse_out << line.str() << std::endl;
        std::cout << "\33[2K\rtime " << elapsed << " s\t" << it << " iterations (" << total << " samples)\tRMSE : "
                  << line.str() << std::flush;
        save_field(out_dir + "/potential" + std::to_string(it) + ".scalar", res, sol);
        save_field(out_dir + "/gradient" + std::to_string(it) + gext, res, grad);
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
