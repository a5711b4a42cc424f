// Walk kernels of the MI355X walk-on-boundary engine, one translation unit per (dimension,
// volume term): the build compiles this file four times with WOB_WALK_DIM = 2 / 3 and
// WOB_WALK_VOL = 0 / 1, so the instantiations compile in parallel.
//
// walk_kernel   one lane = one (point, chunk of paths) work item; paths of the chunk run
//               back to back (a path that stops early is followed at once by the next), the
//               lane Kahan-sums [u, grad u] and their squares in FP32 and writes per-item
//               partials (structure of arrays, coalesced).  Items are chunk-major so that
//               neighbouring lanes walk from neighbouring points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "wob_device.hpp"
#include "wob_launch.hpp"

#ifndef WOB_WALK_DIM
#error "WOB_WALK_DIM must be 2 or 3"
#endif
#ifndef WOB_WALK_VOL
#error "WOB_WALK_VOL must be 0 or 1"
#endif

using namespace wobdev;

namespace wobk {

using wobl::WalkEntry;
using wobl::WalkLaunch;

struct Kahan {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float v) {   // utils.cuh:158-164
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// entry of a global work item / point (binary search over the entry table)
__device__ __forceinline__ const WalkEntry &entry_of_item(const WalkLaunch &L, uint32_t item) {
    if (L.table == nullptr) return L.e0;
    uint32_t lo = 0, hi = L.n_entries - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (L.table[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    return L.table[lo];
}

__device__ __forceinline__ const WalkEntry &entry_of_point(const WalkLaunch &L, uint32_t gp) {
    if (L.table == nullptr) return L.e0;
    uint32_t lo = 0, hi = L.n_entries - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (L.table[mid].point_begin <= gp) lo = mid; else hi = mid - 1;
    }
    return L.table[lo];
}

// One work item: `paths_per_chunk` consecutive paths of one point, back to back.
template <int D, int BIE, int EST, int FL, int STACK, bool BOUNDARY, bool VOL>
__device__ __forceinline__ void walk_item(const WalkLaunch &L, uint32_t item, uint32_t pass, uint32_t *stack,
                                          uint32_t stride) {
    constexpr int V = BOUNDARY ? 1 : D + 1;
    const WalkEntry &e = entry_of_item(L, item);
    const uint32_t local = item - e.item_begin;
    const uint32_t chunk = local / e.n_points;
    const uint32_t pt = local - chunk * e.n_points;
    const uint32_t j0 = chunk * e.paths_per_chunk;
    const uint32_t j1 = min(e.paths_per_point, j0 + e.paths_per_chunk);
    const uint64_t path_base = e.path_base + (uint64_t)pass * e.paths_per_point;
    const DevScene S = e.scene;

    Model<D, BIE> m;
    m.k.eps = e.eps;
    m.dsign = e.dsign;
    m.c = e.fkesc;
    const WalkParams wp{e.path_length, e.nrc, e.flags & 1u, (e.flags >> 1) & 1u, e.cv_samples, (e.flags >> 2) & 1u};
    using VT = typename std::conditional<VOL, VolCtx<D, STACK>, NoVol>::type;
    VT vt;
    if constexpr (VOL) vt = VolCtx<D, STACK>{&S, e.nvs, (e.flags >> 3) & 1u, e.dsign > 0.f, stack, stride};

    Kahan ks[V], kq[V];
#pragma unroll
    for (int k = 0; k < V; ++k) { ks[k].init(); kq[k].init(); }

    if constexpr (BOUNDARY) {
        const float *r = e.points + (size_t)pt * (2 * D + 4);
        BP<D> x;
#pragma unroll
        for (int k = 0; k < D; ++k) { x.p[k] = r[k]; x.n[k] = r[D + k]; }
        x.bv = r[2 * D + 1];
        x.type = (int)r[2 * D + 2];
        x.alpha = r[2 * D + 3];
        x.elem = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < D; ++k) x.bw[k] = 0.f;
        for (uint32_t j = j0; j < j1; ++j) {
            Rng rng;
            rng.init(L.seed, (uint32_t)(path_base + j), e.point_base + pt, e.stream);
            const float v = boundary_path<D, BIE, EST, FL, STACK>(S, m, wp, vt, x, rng, stack, stride);
            ks[0].add(v);
            kq[0].add(v * v);
        }
    } else {
        float x[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = e.points[(size_t)pt * D + k];
        for (uint32_t j = j0; j < j1; ++j) {
            Rng rng;
            rng.init(L.seed, (uint32_t)(path_base + j), e.point_base + pt, e.stream);
            float out[D + 1];
            domain_path<D, BIE, EST, FL, STACK>(S, m, wp, vt, x, rng, stack, stride, out);
#pragma unroll
            for (int k = 0; k < V; ++k) { ks[k].add(out[k]); kq[k].add(out[k] * out[k]); }
        }
    }
    const size_t n_items = (size_t)e.n_points * e.chunks;
    float *part = L.partials + e.partial_base;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        part[(size_t)k * n_items + local] = ks[k].s;
        part[(size_t)(V + k) * n_items + local] = kq[k].s;
    }
}

// Work items are numbered entry by entry, chunk-major inside an entry so that neighbouring
// lanes walk from neighbouring points.  Persistent mode: each wave takes 64 items at a time
// from a device counter until the items run out, so lanes of a finished wave move on to new
// (point, chunk) work at once.
template <int D, int BIE, int EST, int FL, int STACK, bool BOUNDARY, bool VOL>
__global__ __launch_bounds__(256) void walk_kernel(WalkLaunch L) {
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t stride = blockDim.x;
    const uint32_t pass = L.pass_counter ? *L.pass_counter : 0u;
    if (!L.persistent) {
        const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
        if (item < L.total_items) walk_item<D, BIE, EST, FL, STACK, BOUNDARY, VOL>(L, item, pass, stack, stride);
        return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    while (true) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(L.work_counter, 64u);
        base = __shfl(base, 0, 64);
        if (base >= L.total_items) break;
        const uint32_t item = base + lane;
        if (item < L.total_items) walk_item<D, BIE, EST, FL, STACK, BOUNDARY, VOL>(L, item, pass, stack, stride);
    }
}

}  // namespace wobk

namespace wobl {

using namespace wobk;

template <int D, int BIE, int EST, int FL, bool BOUNDARY, bool VOL> struct WalkKernel {
    // line flavours traverse the BVH; volume kernels also run the inside test of mixture samples
    static constexpr bool lines = FL == FL_LINE_LINE || FL == FL_LINE_BV || FL == FL_FWD_LINE || VOL;
    static constexpr int STACK = lines ? kStackEntries : 0;
    static size_t lds(uint32_t block) { return lines ? (size_t)STACK * block * sizeof(uint32_t) : 0; }
    static const void *fn() { return (const void *)&walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY, VOL>; }
    static hipError_t launch(const WalkLaunch &a, uint32_t block, uint32_t grid, hipStream_t st) {
        if (grid == 0) grid = (a.total_items + block - 1) / block;
        hipLaunchKernelGGL((walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY, VOL>), dim3(grid), dim3(block), lds(block), st,
                           a);
        return hipGetLastError();
    }
    static uint32_t resident_grid(uint32_t block) {
        int per_cu = 0, dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn(), (int)block, lds(block)) != hipSuccess)
            per_cu = 1;
        return (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    }
};

// dispatch (model, estimator, flavour) -> instantiation; F(kernel-type) is called with the chosen one
template <int D, bool BOUNDARY, bool VOL, typename F> static auto dispatch(int bie, int est, int fl, F &&f) {
    if (est == kBackward) {
        if (bie == kDL && fl == FL_LINE_LINE) return f(WalkKernel<D, kDL, kBackward, FL_LINE_LINE, BOUNDARY, VOL>{});
        if (bie == kDL) return f(WalkKernel<D, kDL, kBackward, FL_UNIFORM, BOUNDARY, VOL>{});
        if (bie == kDirect && fl == FL_LINE_BV) return f(WalkKernel<D, kDirect, kBackward, FL_LINE_BV, BOUNDARY, VOL>{});
        if (bie == kDirect) return f(WalkKernel<D, kDirect, kBackward, FL_UNIFORM, BOUNDARY, VOL>{});
        return f(WalkKernel<D, kSL, kBackward, FL_UNIFORM, BOUNDARY, VOL>{});
    }
    if (bie == kSL && fl == FL_FWD_LINE) return f(WalkKernel<D, kSL, kForward, FL_FWD_LINE, BOUNDARY, VOL>{});
    if (bie == kSL) return f(WalkKernel<D, kSL, kForward, FL_FWD_UNIFORM, BOUNDARY, VOL>{});
    if (bie == kDL) return f(WalkKernel<D, kDL, kForward, FL_FWD_UNIFORM, BOUNDARY, VOL>{});
    return f(WalkKernel<D, kDirect, kForward, FL_FWD_UNIFORM, BOUNDARY, VOL>{});
}

#define WOB_CAT3(a, b, c) a##b##c
#define WOB_NAME(base, d, v) WOB_CAT3(base, d, v)

// this translation unit's entry points: launch_walk_<dim>_<vol>, persistent_grid_<dim>_<vol>
hipError_t WOB_NAME(launch_walk_, WOB_WALK_DIM, WOB_WALK_VOL)(bool boundary, const WalkLaunch &a, int bie, int est,
                                                                int fl, uint32_t block, uint32_t grid, hipStream_t st) {
    auto go = [&](auto k) { return decltype(k)::launch(a, block, grid, st); };
    return boundary ? dispatch<WOB_WALK_DIM, true, WOB_WALK_VOL != 0>(bie, est, fl, go)
                    : dispatch<WOB_WALK_DIM, false, WOB_WALK_VOL != 0>(bie, est, fl, go);
}

uint32_t WOB_NAME(persistent_grid_, WOB_WALK_DIM, WOB_WALK_VOL)(bool boundary, int bie, int est, int fl,
                                                                  uint32_t block) {
    auto go = [&](auto k) { return decltype(k)::resident_grid(block); };
    return boundary ? dispatch<WOB_WALK_DIM, true, WOB_WALK_VOL != 0>(bie, est, fl, go)
                    : dispatch<WOB_WALK_DIM, false, WOB_WALK_VOL != 0>(bie, est, fl, go);
}

}  // namespace wobl
