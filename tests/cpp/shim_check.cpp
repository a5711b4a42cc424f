// Drives the C++ host interface (include/wob/wob.hpp) the way a reference app drives its
// templates: builds a SceneHost from ElementXs records, picks the Estimator instantiation at
// run time with wob::run (utils::run, runtime_utils.cuh:106-153) and accumulates
// compute_sample_path_contribution_domain over device points.
//
// usage: shim_check <in.bin> <out.bin> [device]   (device: BVH built on the device)
//   in:  uint32 dim, n_elements, n_points, paths, path_length, bie, domain, estimator, problem,
//        nrc; float eps, fkesc; uint64 seed; then float vertices[n*dim*dim], normals[n*dim],
//        boundary_values[n*dim]; int32 types[n]; float alphas[n*dim], points[n_points*dim]
//   out: double sum[n_points*(dim+1)], sumsq[n_points*(dim+1)] from the templated Estimator,
//        then the same from wob::run_domain, then the scene's node count as uint32.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/wob/wob.hpp"

#define HIP_OK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            return 2;                                                              \
        }                                                                          \
    } while (0)

struct Header {
    uint32_t dim, n, n_pts, paths, path_length, bie, domain, est, problem, nrc;
    float eps, fkesc;
    uint64_t seed;
};

template <unsigned Dim> struct RunEstimator {
    const wob::SceneHost<float, Dim> *scene;
    const Header *h;
    const wob::DomainPoint<float, Dim> *d_pts;
    double *d_sum, *d_sq;
    template <wob::BIEModelType B, wob::DomainType Dm, wob::EstimatorType E, wob::ProblemType P> void operator()() {
        wob::Estimator<float, Dim, wob::PoissonKernel, B, Dm, E, P> est(*scene, h->path_length, h->nrc, h->eps,
                                                                          h->fkesc);
        est.compute_domain(d_pts, h->n_pts, h->paths, h->seed, 0, d_sum, d_sq);
    }
};

template <unsigned Dim> int run(const Header &h, FILE *f, FILE *out, bool device_bvh) {
    const size_t n = h.n;
    std::vector<float> v(n * Dim * Dim), nr(n * Dim), bv(n * Dim), al(n * Dim), pts((size_t)h.n_pts * Dim);
    std::vector<int32_t> ty(n);
    bool ok = std::fread(v.data(), 4, v.size(), f) == v.size() && std::fread(nr.data(), 4, nr.size(), f) == nr.size() &&
              std::fread(bv.data(), 4, bv.size(), f) == bv.size() && std::fread(ty.data(), 4, n, f) == n &&
              std::fread(al.data(), 4, al.size(), f) == al.size() &&
              std::fread(pts.data(), 4, pts.size(), f) == pts.size();
    if (!ok) {
        std::fprintf(stderr, "short input\n");
        return 1;
    }
    using El = wob::ElementXs<float, Dim>;
    std::vector<El> els(n);
    for (size_t e = 0; e < n; ++e) {
        El &x = els[e];
        for (unsigned k = 0; k < Dim; ++k) {
            x.a[k] = v[(e * Dim + 0) * Dim + k];
            x.b[k] = v[(e * Dim + 1) * Dim + k];
            x.n[k] = nr[e * Dim + k];
        }
        x.boundary_value_a = bv[e * Dim];
        x.boundary_value_b = bv[e * Dim + 1];
        x.robin_alpha_a = al[e * Dim];
        x.robin_alpha_b = al[e * Dim + 1];
        if constexpr (Dim == 3) {
            for (unsigned k = 0; k < 3; ++k) x.c[k] = v[(e * 3 + 2) * 3 + k];
            x.boundary_value_c = bv[e * 3 + 2];
            x.robin_alpha_c = al[e * 3 + 2];
        }
        x.boundary_type = (wob::BoundaryType)ty[e];
    }
    try {
        wob::SceneHost<float, Dim> scene(els, device_bvh);
        const size_t m = (size_t)h.n_pts * (Dim + 1);
        float *d_pts = nullptr;
        double *d_acc = nullptr;
        HIP_OK(hipMalloc(&d_pts, pts.size() * 4));
        HIP_OK(hipMalloc(&d_acc, 4 * m * 8));
        HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size() * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemset(d_acc, 0, 4 * m * 8));
        RunEstimator<Dim> r{&scene, &h, reinterpret_cast<const wob::DomainPoint<float, Dim> *>(d_pts), d_acc,
                            d_acc + m};
        if (!wob::run((wob::BIEModelType)h.bie, (wob::DomainType)h.domain, (wob::EstimatorType)h.est,
                      (wob::ProblemType)h.problem, r)) {
            std::fprintf(stderr, "invalid enum\n");
            return 1;
        }
        wob::run_domain(scene.get_device_repr(), (wob::BIEModelType)h.bie, (wob::DomainType)h.domain,
                        (wob::EstimatorType)h.est, (wob::ProblemType)h.problem, h.path_length, h.nrc, h.eps, h.fkesc,
                        d_pts, h.n_pts, h.paths, h.seed, 0, d_acc + 2 * m, d_acc + 3 * m);
        std::vector<double> acc(4 * m);
        HIP_OK(hipMemcpy(acc.data(), d_acc, acc.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipFree(d_pts));
        HIP_OK(hipFree(d_acc));
        const uint32_t nodes = scene.info().n_nodes;
        std::fwrite(acc.data(), 8, acc.size(), out);
        std::fwrite(&nodes, 4, 1, out);
        std::printf("shim_check: %s %s %s %s, %u points x %u paths, %u BVH nodes\n",
                    wob::bie_model_type_to_string((wob::BIEModelType)h.bie).c_str(),
                    wob::domain_type_to_string((wob::DomainType)h.domain).c_str(),
                    wob::estimator_type_to_string((wob::EstimatorType)h.est).c_str(),
                    wob::problem_type_to_string((wob::ProblemType)h.problem).c_str(), h.n_pts, h.paths, nodes);
    } catch (const wob::Error &e) {
        std::fprintf(stderr, "wob error %d: %s\n", (int)e.status, e.what());
        return 3;
    }
    return 0;
}

int main(int argc, char **argv) {
    const bool device_bvh = argc == 4 && std::strcmp(argv[3], "device") == 0;
    if (argc != 3 && !device_bvh) {
        std::fprintf(stderr, "usage: %s in.bin out.bin [device]\n", argv[0]);
        return 1;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 1;
    Header h;
    if (std::fread(&h, sizeof h, 1, f) != 1) return 1;
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) return 1;
    const int rc = h.dim == 2 ? run<2>(h, f, out, device_bvh) : h.dim == 3 ? run<3>(h, f, out, device_bvh) : 1;
    std::fclose(f);
    std::fclose(out);
    return rc;
}
