// wob/wob.hpp -- C++ host interface over the C ABI (wob_c.h), with the reference's names.
//
// The reference's host surface is a set of templates (rsugimoto/WoBToolbox):
//   enum BoundaryType / BIEModelType / DomainType / EstimatorType / ProblemType
//                                                   points.hpp:22, bie_models.cuh:21-22,
//                                                   estimator.cuh:24, 38
//   string_to_*_type / *_type_to_string             bie_models.cuh:24-51, estimator.cuh:25-54
//   ElementXs<S, 2 | 3, IsVectorProblem>            scene.cuh:378-382, 646-650
//   BoundaryPoint, DomainPoint                      points.hpp:32-43
//   SceneHost<S, Dim, IsVectorProblem>(elements)    scene.cuh:264-299
//   Estimator<S, Dim, Kernel, BIE, Domain, Est, Problem>(rng, scene, path_length
//             [, nrc [, eps [, fkesc [, num_volume_samples]]]])   estimator.cuh:118-152
//   compute_sample_path_contribution_domain / _boundary           estimator.cuh:154-222
// Its estimators are called per point from a user __device__ functor launched by
// utils::run (runtime_utils.cuh:106-153).  A functor cannot cross a C ABI, so here the
// Estimator runs the library's walk over a device array of points and accumulates per-point
// FP64 sums (the app functors' loop, apps/general_2d.cu:40-71) in one launch.
// Only scalar problems in FP32 are supported (the reference's apps instantiate float).
#pragma once

#include <cfloat>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../wob_c.h"

namespace wob {

enum BoundaryType { Dirichlet, Neumann, Robin };
enum EquationKernelType { PoissonKernel, EquationKernelTypeCount };
enum BIEModelType { IndirectSingleLayer, IndirectDoubleLayer, Direct, BIEModelTypeCount };
enum DomainType { InteriorDomain, ExteriorDomain, DomainTypeCount };
enum EstimatorType { BackwardEstimator, ForwardEstimator, EstimatorTypeCount };
enum ProblemType { DirichletProblem, NeumannProblem, RobinProblem, MixedProblem, ProblemTypeCount };

// The library's own string tables (wob_string_to_enum) back these; unknown strings give the
// *TypeCount value, as in the reference.
inline int enum_or(const std::string &s, int count) {
    const int32_t v = wob_string_to_enum(s.c_str());
    return v < 0 ? count : v;
}
inline BIEModelType string_to_bie_model_type(const std::string &s) {
    return s == "IndirectSingleLayer" || s == "IndirectDoubleLayer" || s == "Direct"
               ? (BIEModelType)enum_or(s, BIEModelTypeCount) : BIEModelTypeCount;
}
inline DomainType string_to_domain_type(const std::string &s) {
    return s == "InteriorDomain" || s == "ExteriorDomain" ? (DomainType)enum_or(s, DomainTypeCount) : DomainTypeCount;
}
inline EstimatorType string_to_estimator_type(const std::string &s) {
    return s == "BackwardEstimator" || s == "ForwardEstimator" ? (EstimatorType)enum_or(s, EstimatorTypeCount)
                                                                : EstimatorTypeCount;
}
inline ProblemType string_to_problem_type(const std::string &s) {
    return s == "DirichletProblem" || s == "NeumannProblem" || s == "RobinProblem" || s == "MixedProblem"
               ? (ProblemType)enum_or(s, ProblemTypeCount) : ProblemTypeCount;
}
inline EquationKernelType string_to_equation_kernel_type(const std::string &s) {
    return s == "PoissonKernel" ? PoissonKernel : EquationKernelTypeCount;
}
inline std::string bie_model_type_to_string(BIEModelType v) { return wob_enum_to_string("bie_model", v); }
inline std::string domain_type_to_string(DomainType v) { return wob_enum_to_string("domain", v); }
inline std::string estimator_type_to_string(EstimatorType v) { return wob_enum_to_string("estimator", v); }
inline std::string problem_type_to_string(ProblemType v) { return wob_enum_to_string("problem", v); }

// Errors from the C ABI surface as exceptions on the C++ side.
struct Error : std::runtime_error {
    wob_status status;
    Error(wob_status s, const char *msg) : std::runtime_error(msg ? msg : "wob error"), status(s) {}
};
inline void check(wob_status s) {
    if (s != WOB_OK) throw Error(s, wob_last_error_message());
}

template <typename S, unsigned Dim, bool IsVectorProblem = false> struct ElementXs;

template <> struct ElementXs<float, 2, false> {
    float a[2], b[2], n[2];
    float boundary_value_a, boundary_value_b;
    BoundaryType boundary_type;
    float robin_alpha_a, robin_alpha_b;
};

template <> struct ElementXs<float, 3, false> {
    float a[3], b[3], c[3], n[3];
    float boundary_value_a, boundary_value_b, boundary_value_c;
    BoundaryType boundary_type;
    float robin_alpha_a, robin_alpha_b, robin_alpha_c;
};

// Device record layouts.  A BoundaryPoint is 2*Dim+4 floats; its boundary type is stored as
// a float (0, 1, 2) so the record is one row of floats.
template <typename S, unsigned Dim> struct DomainPoint { float p[Dim]; };
template <typename S, unsigned Dim, bool IsVectorProblem = false> struct BoundaryPoint {
    float p[Dim];
    float n[Dim];
    float interior_angle;
    float boundary_value;
    float boundary_type;
    float robin_alpha;
};
static_assert(sizeof(DomainPoint<float, 3>) == 12, "DomainPoint layout");
static_assert(sizeof(BoundaryPoint<float, 3>) == 40, "BoundaryPoint layout");

template <typename S, unsigned Dim, bool IsVectorProblem = false> class SceneHost {
    static_assert(!IsVectorProblem && (Dim == 2 || Dim == 3), "scalar 2D / 3D scenes only");

  public:
    using Element = ElementXs<float, Dim, false>;

    // scene.cuh:265: the reference builds its BVH on the device (lbvh); device_bvh = true does
    // the same (wob_scene_create_ex with WOB_SCENE_DEVICE_BVH), false uses the host SAH build.
    // Results do not depend on which tree is used.
    explicit SceneHost(const std::vector<Element> &elems, bool device_bvh = false) {
        const size_t n = elems.size();
        std::vector<float> v(n * Dim * Dim), nr(n * Dim), bv(n * Dim), al(n * Dim);
        std::vector<int32_t> ty(n);
        for (size_t e = 0; e < n; ++e) {
            const Element &x = elems[e];
            for (unsigned k = 0; k < Dim; ++k) {
                v[(e * Dim + 0) * Dim + k] = x.a[k];
                v[(e * Dim + 1) * Dim + k] = x.b[k];
                nr[e * Dim + k] = x.n[k];
            }
            bv[e * Dim + 0] = x.boundary_value_a;
            bv[e * Dim + 1] = x.boundary_value_b;
            al[e * Dim + 0] = x.robin_alpha_a;
            al[e * Dim + 1] = x.robin_alpha_b;
            if constexpr (Dim == 3) {
                for (unsigned k = 0; k < 3; ++k) v[(e * 3 + 2) * 3 + k] = x.c[k];
                bv[e * 3 + 2] = x.boundary_value_c;
                al[e * 3 + 2] = x.robin_alpha_c;
            }
            ty[e] = (int32_t)x.boundary_type;
        }
        const wob_scene_desc d{Dim, (uint32_t)n, v.data(), nr.data(), bv.data(), ty.data(), al.data()};
        check(wob_scene_create_ex(&d, device_bvh ? WOB_SCENE_DEVICE_BVH : 0u, &scene_));
    }
    SceneHost(const SceneHost &) = delete;
    SceneHost &operator=(const SceneHost &) = delete;
    SceneHost(SceneHost &&o) noexcept : scene_(o.scene_) { o.scene_ = nullptr; }
    ~SceneHost() {
        if (scene_) wob_scene_destroy(scene_);
    }

    // scene.cuh:301-306: grid source over (-domain_size/2, domain_size/2)^Dim, grid_res^Dim
    // host values (x fastest); estimators with num_volume_samples > 0 sample it.
    void set_volume_cache(unsigned grid_res, float domain_size, const std::vector<float> &cache) {
        check(wob_scene_set_volume_grid(scene_, grid_res, domain_size, cache.data()));
    }

    // Learned harmonic control variate (no counterpart in the reference; wob_c.h
    // wob_scene_fit_harmonic_cv): fitted to the walk's known-term data at n_train boundary
    // samples and bound to the scene; estimators whose desc().flags carry WOB_EST_HARMONIC_CV
    // then walk the residual.  exterior = true fits the exterior domain's (radius_scale in
    // (0, 1), sources inside the body); zero_charge = true with it keeps sum_j w_j = 0, as
    // exterior double-layer walks need.  Returns the fit's RMS residual.
    float fit_harmonic_cv(uint32_t n_sources = 64, float radius_scale = 1.5f, uint32_t n_train = 16384,
                          uint32_t volume_samples = 0, float eps = 1e-4f, float ridge = 1e-10f, uint64_t seed = 0,
                          void *stream = nullptr, bool exterior = false, bool zero_charge = false) {
        const wob_hcv_desc d{n_sources, radius_scale, n_train, volume_samples, eps, ridge, seed, 0u,
                             (exterior ? WOB_HCV_EXTERIOR : 0u) | (zero_charge ? WOB_HCV_ZERO_CHARGE : 0u)};
        float rms = 0.f;
        check(wob_scene_fit_harmonic_cv(scene_, &d, &rms, stream));
        return rms;
    }

    // The device representation is the opaque handle (scene.cuh:308-318 returns a POD view).
    const wob_scene *get_device_repr() const { return scene_; }
    wob_scene_info info() const {
        wob_scene_info i{};
        check(wob_scene_get_info(scene_, &i));
        return i;
    }

  private:
    wob_scene *scene_ = nullptr;
};

// Estimator<S, Dim, Kernel, BIE, Domain, Est, Problem>.  The constructor arguments follow
// estimator.cuh:118-152 (the RNG state pointer is replaced by a seed given per call).
template <typename S, unsigned Dim, EquationKernelType equationKernel, BIEModelType bieModel, DomainType domainType,
          EstimatorType estimatorType, ProblemType problem>
class Estimator {
    static_assert(equationKernel == PoissonKernel, "PoissonKernel only");

  public:
    Estimator(const SceneHost<S, Dim, false> &scene, unsigned path_length, unsigned num_resampling_candidates = 1,
              float eps = FLT_EPSILON, float first_kind_equation_scaling_constant = 1.0f,
              unsigned num_volume_samples = 0)
        : scene_(scene.get_device_repr()) {
        desc_ = wob_estimator_desc{WOB_POISSON_KERNEL, (int32_t)bieModel, (int32_t)domainType, (int32_t)estimatorType,
                                   (int32_t)problem, path_length, num_resampling_candidates, num_volume_samples, eps,
                                   first_kind_equation_scaling_constant, 0u, 0u};
    }

    wob_estimator_desc &desc() { return desc_; }

    // compute_sample_path_contribution_domain, num_sample_paths times per point:
    // d_sum / d_sumsq (n * (Dim + 1) doubles, device) += sums of [u, grad u] and of squares.
    void compute_domain(const DomainPoint<S, Dim> *d_points, uint32_t n, uint32_t num_sample_paths, uint64_t seed,
                        uint64_t path_base, double *d_sum, double *d_sumsq, void *stream = nullptr,
                        wob_launch_config *launch = nullptr) const {
        check(wob_estimate_domain(scene_, &desc_, reinterpret_cast<const float *>(d_points), n, 0, num_sample_paths,
                                  seed, path_base, 0, d_sum, d_sumsq, launch, stream));
    }

    // compute_sample_path_contribution_boundary: d_sum / d_sumsq n doubles each (device).
    void compute_boundary(const BoundaryPoint<S, Dim> *d_records, uint32_t n, uint32_t num_sample_paths, uint64_t seed,
                          uint64_t path_base, double *d_sum, double *d_sumsq, void *stream = nullptr,
                          wob_launch_config *launch = nullptr) const {
        check(wob_estimate_boundary(scene_, &desc_, reinterpret_cast<const float *>(d_records), n, 0,
                                    num_sample_paths, seed, path_base, 0, d_sum, d_sumsq, launch, stream));
    }

    // Virtual point lights (apps/potential_flow_with_vpl.cu:73-88): n uniform boundary start
    // points y_i and c_i = compute_sample_unknown_boundary(y_i) (estimator.cuh:232-357);
    // iteration k of a progressive loop passes path_base = k.
    void compute_vpl_paths(uint32_t n, uint64_t seed, uint64_t path_base, BoundaryPoint<S, Dim> *d_records,
                           S *d_contrib, void *stream = nullptr, wob_launch_config *launch = nullptr) const {
        check(wob_vpl_paths(scene_, &desc_, n, seed, path_base, 0, reinterpret_cast<float *>(d_records), d_contrib,
                            launch, stream));
    }

    // AggregateFunc (apps/potential_flow_with_vpl.cu:44-71) over all points: d_sum (n_points *
    // (Dim + 1) doubles, device) += sum_i c_i [solution_domain_unknown, gradient_domain_unknown].
    void aggregate_vpl(const BoundaryPoint<S, Dim> *d_records, const S *d_contrib, uint32_t n_vpl,
                       const DomainPoint<S, Dim> *d_points, uint32_t n_points, double *d_sum,
                       void *stream = nullptr) const {
        check(wob_vpl_aggregate(scene_, &desc_, reinterpret_cast<const float *>(d_records), d_contrib, n_vpl,
                                reinterpret_cast<const float *>(d_points), n_points, d_sum, stream));
    }

  private:
    const wob_scene *scene_;
    wob_estimator_desc desc_;
};

// Runtime -> compile-time dispatch (utils::run, runtime_utils.cuh:106-153): calls
// f.template operator()<BIE, Domain, Estimator, Problem>() for the values picked at run time.
// The reference silently launches nothing for an invalid value; here run returns false.
namespace detail {
template <BIEModelType B, DomainType D, EstimatorType E, typename F> bool by_problem(ProblemType p, F &f) {
    switch (p) {
    case DirichletProblem: f.template operator()<B, D, E, DirichletProblem>(); return true;
    case NeumannProblem: f.template operator()<B, D, E, NeumannProblem>(); return true;
    case RobinProblem: f.template operator()<B, D, E, RobinProblem>(); return true;
    case MixedProblem: f.template operator()<B, D, E, MixedProblem>(); return true;
    default: return false;
    }
}
template <BIEModelType B, DomainType D, typename F> bool by_estimator(EstimatorType e, ProblemType p, F &f) {
    switch (e) {
    case BackwardEstimator: return by_problem<B, D, BackwardEstimator>(p, f);
    case ForwardEstimator: return by_problem<B, D, ForwardEstimator>(p, f);
    default: return false;
    }
}
template <BIEModelType B, typename F> bool by_domain(DomainType d, EstimatorType e, ProblemType p, F &f) {
    switch (d) {
    case InteriorDomain: return by_estimator<B, InteriorDomain>(e, p, f);
    case ExteriorDomain: return by_estimator<B, ExteriorDomain>(e, p, f);
    default: return false;
    }
}
}  // namespace detail

template <typename F> bool run(BIEModelType bie, DomainType domain, EstimatorType est, ProblemType problem, F &&f) {
    switch (bie) {
    case IndirectSingleLayer: return detail::by_domain<IndirectSingleLayer>(domain, est, problem, f);
    case IndirectDoubleLayer: return detail::by_domain<IndirectDoubleLayer>(domain, est, problem, f);
    case Direct: return detail::by_domain<Direct>(domain, est, problem, f);
    default: return false;
    }
}

// The same with the choice kept as data: run_domain validates the values and reports
// invalid ones as WOB_INVALID_ENUM.
inline void run_domain(const wob_scene *scene, BIEModelType bie, DomainType domain, EstimatorType est,
                       ProblemType problem, unsigned path_length, unsigned nrc, float eps, float fkesc,
                       const float *d_points, uint32_t n, uint32_t num_sample_paths, uint64_t seed,
                       uint64_t path_base, double *d_sum, double *d_sumsq, void *stream = nullptr) {
    const wob_estimator_desc d{WOB_POISSON_KERNEL, (int32_t)bie, (int32_t)domain, (int32_t)est, (int32_t)problem,
                               path_length, nrc, 0u, eps, fkesc, 0u, 0u};
    check(wob_estimate_domain(scene, &d, d_points, n, 0, num_sample_paths, seed, path_base, 0, d_sum, d_sumsq,
                              nullptr, stream));
}

}  // namespace wob
