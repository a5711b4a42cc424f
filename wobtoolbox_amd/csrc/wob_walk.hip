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
    const WalkParams wp{e.path_length, e.nrc,           e.flags & 1u,        (e.flags >> 1) & 1u,
                        e.cv_samples,  (e.flags >> 2) & 1u, (e.flags >> 4) & 1u};
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
// Waves per SIMD the compiler must fit (the register budget): 4 (128 VGPRs) for the 2D
// single-layer RIS domain walk (the batched exterior scenes), which fits them without spills;
// 3 (168 VGPRs, as many as a 48-entry LDS stack allows) for the walks with line-sampled known
// terms (direct Neumann, forward single layer) and every walk with a volume term, whose line
// queries and volume samples wait on memory.  Measured against the compiler's choice of 2
// waves (some spill at 3): 3D double layer with the Gaussian source a third faster, 2D double
// layer with a grid source (the Poisson app) an eighth, 2D direct Neumann 29%, 3D forward
// single layer 21%.  The double-layer line walks without a volume term fit 3 waves by
// themselves and run no faster with the bound; the others take the compiler's choice.
template <int D, int BIE, int EST, int FL, bool BOUNDARY, bool VOL> constexpr int walk_min_waves() {
    if (D == 2 && BIE == kSL && EST == kBackward && FL == FL_UNIFORM && !BOUNDARY && !VOL) return 4;
    if (FL == FL_LINE_BV || FL == FL_FWD_LINE || VOL) return 3;
    return 1;
}

template <int D, int BIE, int EST, int FL, int STACK, bool BOUNDARY, bool VOL>
__global__ __launch_bounds__(256, (walk_min_waves<D, BIE, EST, FL, BOUNDARY, VOL>())) void walk_kernel(WalkLaunch L) {
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

// ---------------------------------------------------------------------------------------
// Wavefront walk for the line-sampled backward domain walks (double layer / Dirichlet and
// direct / Neumann).  The walk above runs each lane's query to its end before the lane's walk
// continues, so every wave waits, query after query, for its slowest traversal.  Here a lane
// is a small state machine: each iteration the wave either runs one traversal step (a node
// or a primitive step, chosen by ballot as in traverse) for the lanes with a query in flight,
// or -- once at least `regen` lanes have finished their query, or none is traversing --
// advances those lanes: consume the hit, run the walk up to its next query (or finish the
// path, start the next path of the item, or take a new item from the work counter).  Every
// path keeps its own Philox stream and the operation order of walk_item, so the per-item
// sums are bit-identical to it.
// ---------------------------------------------------------------------------------------
enum { kLaneIdle = 0, kLaneTrav = 1, kLaneAdv = 2 };

template <int D, int BIE, int FL, bool MULTI>
__global__ __launch_bounds__(256) void walk_kernel_wf(WalkLaunch L) {
    constexpr int STACK = wobl::kStackEntries;
    constexpr int V = D + 1;
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t stride = blockDim.x;
    const uint32_t pass = L.pass_counter ? *L.pass_counter : 0u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t regen = L.regen_threshold;

    // item
    uint32_t ent = 0, local = 0, pt = 0, j = 0, j1 = 0, pbase = 0;
    float x[D];
    Kahan ks[V], kq[V];
    // path
    uint32_t phase = 0, it = 0;
    float W = 0.f, spc = 0.f, known = 0.f, unk = 0.f, gk[D], gu[D];
    BP<D> yp = bp_zero<D>();
    Rng rng;
    rng.init(0, 0, 0, 0);
    // query
    float o[D], d[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { x[k] = gk[k] = gu[k] = 0.f; o[k] = 0.f; d[k] = 1.f; }
    uint32_t pick = 0, skipv = 0xFFFFFFFFu;
    Trav<D> tr;
    tr.begin(o, d);
    tr.cur = kTravEmpty;
    LineHit<D> h;
    h.count = 0; h.id = 0xFFFFFFFFu; h.t = 0.f; h.a = 0.f; h.b = 0.f;
    uint32_t hbest = 0xFFFFFFFFu;
    const float4 *nodes = L.e0.scene.nodes, *prims = L.e0.scene.prims;
    const uint32_t *pid = L.e0.scene.prim_id;
    uint32_t state = kLaneAdv;
    bool need_item = true;

    auto entry = [&]() -> const WalkEntry & { return MULTI ? L.table[ent] : L.e0; };
    auto model = [&](const WalkEntry &e) {
        Model<D, BIE> m;
        m.k.eps = e.eps;
        m.dsign = e.dsign;
        m.c = e.fkesc;
        return m;
    };
    auto issue_query = [&](const float *origin, uint32_t skip) {
#pragma unroll
        for (int k = 0; k < D; ++k) o[k] = origin[k];
        line_dir<D>(rng, d);
        pick = rng.next_u32();
        skipv = skip;
        tr.begin(o, d);
        h.count = 0; h.id = 0xFFFFFFFFu; h.t = 0.f; h.a = 0.f; h.b = 0.f;
        hbest = 0xFFFFFFFFu;
        state = kLaneTrav;
    };
    auto start_bounce = [&](const WalkEntry &e) {
        const Model<D, BIE> m = model(e);
        if (it == e.path_length - 1) W = 0.5f * W;
        if constexpr (FL == FL_LINE_BV) {   // as in walk_backward, with or without the control variate
            if (e.flags & 16u) {
                const Sample<D> qb = sample_uniform<D>(e.scene, rng);
                spc += W * m.fredholm_known(hcv_residual<D>(e.scene, yp), hcv_residual<D>(e.scene, qb.y), qb.inv_pdf,
                                            NoVol{}, rng);
            } else {
                const Sample<D> qb = sample_bv<D>(e.scene, rng);
                spc += W * m.fredholm_known(yp, qb.y, qb.inv_pdf, NoVol{}, rng);
            }
        }
        issue_query(yp.p, (e.flags & 1u) ? yp.elem : 0xFFFFFFFFu);
    };
    auto start_path = [&](const WalkEntry &e) {
        rng.init(L.seed, pbase + j, e.point_base + pt, e.stream);
        phase = 0;
        issue_query(x, 0xFFFFFFFFu);
    };
    auto advance = [&]() {
        if (!need_item) {
            const WalkEntry &e = entry();
            const Model<D, BIE> m = model(e);
            const Sample<D> q = line_finish<D>(e.scene, o, d, h);
            bool done = false;
            if (phase == 0) {                                              // domain_path, first samples
                Sample<D> y1 = q;
                if constexpr (FL == FL_LINE_BV) {
                    if (e.flags & 16u) {
                        y1 = sample_uniform<D>(e.scene, rng);
                        y1.y = hcv_residual<D>(e.scene, y1.y);
                    } else {
                        y1 = sample_bv<D>(e.scene, rng);
                    }
                }
                known = m.solution_domain_known(x, y1.y, y1.inv_pdf, NoVol{}, rng);
                m.gradient_domain_known(x, y1.y, y1.inv_pdf, NoVol{}, rng, gk);
                unk = m.solution_domain_unknown(x, q.y, q.inv_pdf);
                m.gradient_domain_unknown(x, q.y, q.inv_pdf, gu);
                W = 1.f; spc = 0.f; yp = q.y; it = 0; phase = 1;
            } else {                                                       // walk_backward, one bounce
                Sample<D> qq = q;
                if constexpr (FL == FL_LINE_LINE) {
                    if (e.flags & 16u)   // harmonic control variate, as in walk_backward
                        spc += W * m.fredholm_known(hcv_residual<D>(e.scene, yp), qq.y, qq.inv_pdf, NoVol{}, rng);
                    else
                        spc += W * m.fredholm_known(yp, qq.y, qq.inv_pdf, NoVol{}, rng);
                    W *= m.fredholm_unknown_backward(yp, qq.y, qq.inv_pdf, rng);
                } else {
                    W = W * m.fredholm_unknown_backward(yp, qq.y, qq.inv_pdf, rng);
                }
                yp = qq.y;
                it += 1;
                done = W == 0.0f || it == e.path_length;
            }
            if (!done) { start_bounce(e); return; }
            float out[V];
            out[0] = known + spc * unk;
#pragma unroll
            for (int k = 0; k < D; ++k) out[1 + k] = gk[k] + spc * gu[k];
            {
                if (e.flags & 16u) {
                    float gh[D];
                    hcv_gradient<D>(e.scene, x, gh);
                    out[0] += hcv_value<D>(e.scene, x);
#pragma unroll
                    for (int k = 0; k < D; ++k) out[1 + k] += gh[k];
                }
            }
#pragma unroll
            for (int k = 0; k < V; ++k) { ks[k].add(out[k]); kq[k].add(out[k] * out[k]); }
            j += 1;
            if (j < j1) { start_path(e); return; }
            const size_t n_items = (size_t)e.n_points * e.chunks;
            float *part = L.partials + e.partial_base;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                part[(size_t)k * n_items + local] = ks[k].s;
                part[(size_t)(V + k) * n_items + local] = kq[k].s;
            }
            need_item = true;
        }
        // take the next item: one counter update for all lanes of the wave that need one
        const uint64_t want = __ballot(1);
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)want) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(L.work_counter, (uint32_t)__popcll(want));
        base = __shfl(base, (int)leader, 64);
        const uint32_t item = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
        if (item >= L.total_items) { state = kLaneIdle; return; }
        if constexpr (MULTI) {
            uint32_t lo = 0, hi = L.n_entries - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (L.table[mid].item_begin <= item) lo = mid; else hi = mid - 1;
            }
            ent = lo;
        }
        const WalkEntry &e = entry();
        local = item - e.item_begin;
        const uint32_t chunk = local / e.n_points;
        pt = local - chunk * e.n_points;
        j = chunk * e.paths_per_chunk;
        j1 = min(e.paths_per_point, j + e.paths_per_chunk);
        pbase = (uint32_t)(e.path_base + (uint64_t)pass * e.paths_per_point);
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = e.points[(size_t)pt * D + k];
#pragma unroll
        for (int k = 0; k < V; ++k) { ks[k].init(); kq[k].init(); }
        nodes = e.scene.nodes;
        prims = e.scene.prims;
        pid = e.scene.prim_id;
        need_item = false;
        start_path(e);
    };

    while (true) {
        if (state == kLaneTrav) {
            tr.settle(stack, stride);
            if (tr.done()) state = kLaneAdv;                    // query finished
        }
        const uint64_t badv = __ballot(state == kLaneAdv);
        const uint64_t btrav = __ballot(state == kLaneTrav);
        if ((badv | btrav) == 0) break;
        if (btrav == 0 || (uint32_t)__popcll(badv) >= regen) {
            if (state == kLaneAdv) advance();
            continue;
        }
        const bool want_leaf = state == kLaneTrav && tr.want_prim();
        const bool want_node = state == kLaneTrav && tr.want_node();
        const uint64_t bl = __ballot(want_leaf);
        const uint64_t bn = __ballot(want_node);
        if (__popcll(bl) >= __popcll(bn)) {
            if (want_leaf) {
                DevScene Sq{};
                Sq.prims = prims;
#pragma unroll
                for (int u = 0; u < kPrimsPerRound; ++u) {
                    if (u > 0 && tr.lp >= tr.le) break;
                    float t, a, b;
                    const bool hit = prim_test<D>(Sq, tr.lp, o, d, t, a, b);
                    hit_update<D>(hit, gldu(pid + tr.lp), skipv, pick, t, a, b, h, hbest);
                    ++tr.lp;
                }
            }
        } else if (want_node) {
            tr.template node_step<STACK>(nodes, stack, stride);
        }
    }
}

#if WOB_WALK_VOL
// Training targets of the harmonic control variate.  Sample i is a uniform boundary point y
// (Philox stream (seed, path i, point 0, stream)); its target is the quantity the walk's known
// term carries at y (bie_models.cuh:301-316, 389-399): the datum g(y) plus the volume term at
// y, here the mean of nvs independent volume samples (nvs = 0: no volume source).  Output: a
// BoundaryPoint record [p, n, 0.5, target, type, alpha] per sample.
template <int D>
__global__ __launch_bounds__(256) void hcv_targets_kernel(DevScene S, uint32_t n, uint64_t seed, uint32_t stream,
                                                          uint32_t nvs, float eps, uint32_t interior, float *rec) {
    extern __shared__ uint32_t lds_stack[];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng rng;
    rng.init(seed, i, 0u, stream);
    const Sample<D> y = sample_uniform<D>(S, rng);
    Kern<D> k;
    k.eps = eps;
    const VolCtx<D, wobl::kStackEntries> vt{&S, nvs, 1u, interior != 0u, lds_stack + threadIdx.x, blockDim.x};
    const float t = y.y.bv + vt.term(y.y, k, rng);
    float *r = rec + (size_t)i * (2 * D + 4);
#pragma unroll
    for (int c = 0; c < D; ++c) { r[c] = y.y.p[c]; r[D + c] = y.y.n[c]; }
    r[2 * D] = 0.5f;
    r[2 * D + 1] = t;
    r[2 * D + 2] = (float)y.y.type;
    r[2 * D + 3] = y.y.alpha;
}

// The inside test's occupancy grid: every cell the host left unmarked (255) is classified by
// the rays from its centre and from the point a quarter cell towards its low corner: their
// common parity, or 2 (near: samples there cast their own rays) when they disagree.
template <int D>
__global__ __launch_bounds__(256) void inside_grid_kernel(DevScene S, uint8_t *flag, uint32_t n_cells) {
#pragma clang fp contract(off)
    extern __shared__ uint32_t lds_stack[];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells || flag[c] != 255u) return;
    const VolSrc &v = S.vol;
    uint32_t rem = c;
    float pc[3] = {0.f, 0.f, 0.f}, pq[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float ik = (float)(rem % v.gres);
        rem /= v.gres;
        pc[k] = v.glo[k] + (ik + 0.5f) * v.gh[k];
        pq[k] = v.glo[k] + (ik + 0.25f) * v.gh[k];
    }
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t a = ray_crossings<D, wobl::kStackEntries>(S.nodes, S.prims, pc[0], pc[1], pc[2], stack, blockDim.x) & 1u;
    const uint32_t b = ray_crossings<D, wobl::kStackEntries>(S.nodes, S.prims, pq[0], pq[1], pq[2], stack, blockDim.x) & 1u;
    flag[c] = (uint8_t)(a == b ? a : 2u);
}
#endif

}  // namespace wobk

namespace wobl {

using namespace wobk;

template <int D, int BIE, int EST, int FL, bool BOUNDARY, bool VOL> struct WalkKernel {
    // line flavours traverse the BVH; volume kernels also run the inside test of mixture samples
    static constexpr bool lines = FL == FL_LINE_LINE || FL == FL_LINE_BV || FL == FL_FWD_LINE || VOL;
    static constexpr int STACK = lines ? kStackEntries : 0;
    // per lane: the stack of STACK entries plus the spare slot the node step writes above a full
    // stack, and the line query's result slots
    static size_t lds(uint32_t block) {
        return lines ? (size_t)line_query_lds_words<STACK>() * block * sizeof(uint32_t) : 0;
    }
    static const void *fn() { return (const void *)&walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY, VOL>; }
    // the wavefront walk covers the line-sampled backward domain walks without a volume term
    static constexpr bool wf = EST == kBackward && !BOUNDARY && !VOL && (FL == FL_LINE_LINE || FL == FL_LINE_BV);
    static const void *wf_fn(bool multi) {
        if constexpr (wf) {
            return multi ? (const void *)&walk_kernel_wf<D, BIE, FL, true> : (const void *)&walk_kernel_wf<D, BIE, FL, false>;
        } else {
            return nullptr;
        }
    }
    static hipError_t launch(const WalkLaunch &a, uint32_t block, uint32_t grid, hipStream_t st) {
        if constexpr (wf) {
            if (a.wavefront) {
                if (a.table)
                    hipLaunchKernelGGL((walk_kernel_wf<D, BIE, FL, true>), dim3(grid), dim3(block), lds(block), st, a);
                else
                    hipLaunchKernelGGL((walk_kernel_wf<D, BIE, FL, false>), dim3(grid), dim3(block), lds(block), st, a);
                return hipGetLastError();
            }
        }
        if (grid == 0) grid = (a.total_items + block - 1) / block;
        hipLaunchKernelGGL((walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY, VOL>), dim3(grid), dim3(block), lds(block), st,
                           a);
        return hipGetLastError();
    }
    static bool supports_wavefront() { return wf; }
    static uint32_t resident_grid(uint32_t block, bool wavefront = false, bool multi = false) {
        int per_cu = 0, dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const void *f = wavefront && wf ? wf_fn(multi) : fn();
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, (int)block, lds(block)) != hipSuccess)
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
                                                                  uint32_t block, bool wavefront, bool multi) {
    auto go = [&](auto k) { return decltype(k)::resident_grid(block, wavefront, multi); };
    return boundary ? dispatch<WOB_WALK_DIM, true, WOB_WALK_VOL != 0>(bie, est, fl, go)
                    : dispatch<WOB_WALK_DIM, false, WOB_WALK_VOL != 0>(bie, est, fl, go);
}

#if WOB_WALK_VOL
hipError_t WOB_NAME(launch_hcv_targets_, WOB_WALK_DIM, )(const wobdev::DevScene &S, uint32_t n, uint64_t seed,
                                                          uint32_t stream, uint32_t nvs, float eps, bool interior,
                                                          float *rec, size_t lds_bytes, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 256;
    hipLaunchKernelGGL((hcv_targets_kernel<WOB_WALK_DIM>), dim3((n + block - 1) / block), dim3(block), lds_bytes, st, S,
                       n, seed, stream, nvs, eps, interior ? 1u : 0u, rec);
    return hipGetLastError();
}

hipError_t WOB_NAME(launch_inside_grid_, WOB_WALK_DIM, )(const wobdev::DevScene &S, uint8_t *flag, uint32_t n_cells,
                                                         size_t lds_bytes, hipStream_t st) {
    if (n_cells == 0) return hipSuccess;
    const uint32_t block = 256;
    hipLaunchKernelGGL((inside_grid_kernel<WOB_WALK_DIM>), dim3((n_cells + block - 1) / block), dim3(block), lds_bytes,
                       st, S, flag, n_cells);
    return hipGetLastError();
}
#endif

bool WOB_NAME(supports_wavefront_, WOB_WALK_DIM, WOB_WALK_VOL)(bool boundary, int bie, int est, int fl) {
    auto go = [&](auto k) { return decltype(k)::supports_wavefront(); };
    return boundary ? dispatch<WOB_WALK_DIM, true, WOB_WALK_VOL != 0>(bie, est, fl, go)
                    : dispatch<WOB_WALK_DIM, false, WOB_WALK_VOL != 0>(bie, est, fl, go);
}

}  // namespace wobl
