// Device BVH build for the MI355X walk-on-boundary engine.
//
// The reference builds its lbvh binary BVH on the device when a SceneHost is constructed
// (scene.cuh:264-299, deps/lbvh: Morton codes, a radix tree, bottom-up boxes).  This build
// follows the same published pipeline (Karras 2012, "Maximizing parallelism in the
// construction of BVHs, octrees, and k-d trees") and then writes the 4-wide layouts the
// traversal reads (scene_build.hpp), so a scene built here walks exactly like a host-built
// one: hit counts and picks do not depend on the tree.
//   1. per element: the inflated box (4e-6 |coordinate| + 1e-7, as the host build) and the
//      Morton code of its centre in the vertex bounds (3 x 10 bits in 3D, 2 x 16 in 2D);
//   2. a stable LSD radix sort of (code, element) pairs (8-bit digits, this file's own);
//   3. the binary radix tree, one internal node per thread; equal codes are ordered by
//      index (the delta function falls back to the index bits);
//   4. boxes bottom-up: each leaf walks to the root, the second child to arrive at a node
//      writes its box (device-scope fences and coherent loads), and restructures the
//      treelet below it by the SAH (two passes; the first one also writes the boxes); then
//      every node's first primitive in depth-first order (primitive records follow that
//      order);
//   5. top-down, one level per launch: each binary node of the level becomes a wide node
//      whose children are its binary children, expanded largest-measure-first until there
//      are `width` (four by default; the host's collapse rule); child indices of the next level come from an
//      exclusive scan of the internal-child counts, so the layout is deterministic; the
//      float node and its quantized copy are written in the same kernel;
//   6. the traversal stack bound, bottom-up over the levels;
//   7. primitive records in depth-first leaf order.
// Subtrees of at most kLeafPrims primitives become leaves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "scene_build.hpp"
#include "wob_launch.hpp"

namespace wobb {

using wobhost::kNodeWords2;
using wobhost::kNodeWords3;
using wobhost::kPrimWords2;
using wobhost::kPrimWords3;
using wobhost::kQNodeWords2;
using wobhost::kQNodeWords3;

struct Box6 {
    float lo[3], hi[3];
};

// spread the low bits of v so that there are (gap) zero bits between them
__device__ __forceinline__ uint32_t spread3(uint32_t v) {   // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {   // 16 bits -> every second bit
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// 1: element boxes (inflated as scene_build.cpp), Morton codes of their centres
template <int D>
__global__ __launch_bounds__(256) void leaf_kernel(uint32_t n, const float *__restrict__ V, float3 vmin, float3 vscale,
                                                   uint32_t *__restrict__ code, uint32_t *__restrict__ idx,
                                                   Box6 *__restrict__ pbox) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    Box6 b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = INFINITY; b.hi[k] = -INFINITY; }
    for (int j = 0; j < D; ++j)
        for (int k = 0; k < 3; ++k) {
            const float p = k < D ? V[((size_t)e * D + j) * D + k] : 0.f;
            b.lo[k] = fminf(b.lo[k], p);
            b.hi[k] = fmaxf(b.hi[k], p);
        }
    for (int k = 0; k < 3; ++k) {
        const float m = fmaxf(fabsf(b.lo[k]), fabsf(b.hi[k]));
        const float pad = m * 4.0e-6f + 1.0e-7f;
        b.lo[k] -= pad;
        b.hi[k] += pad;
    }
    pbox[e] = b;
    const float c[3] = {0.5f * (b.lo[0] + b.hi[0]), 0.5f * (b.lo[1] + b.hi[1]), 0.5f * (b.lo[2] + b.hi[2])};
    const float mn[3] = {vmin.x, vmin.y, vmin.z}, sc[3] = {vscale.x, vscale.y, vscale.z};
    uint32_t q[3];
    const float cells = D == 3 ? 1024.f : 65536.f;
    for (int k = 0; k < 3; ++k) {
        const float u = fminf(fmaxf((c[k] - mn[k]) * sc[k], 0.f), 1.f);
        q[k] = (uint32_t)fminf(u * cells, cells - 1.f);
    }
    code[e] = D == 3 ? (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]) : (spread2(q[0]) << 1) | spread2(q[1]);
    idx[e] = e;
}

// length of the common prefix of sorted codes i and j (index bits break ties); -1 outside
__device__ __forceinline__ int delta(const uint32_t *code, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = code[i], b = code[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// 3: binary radix tree.  Node ids: internal i in [0, n-1), leaf j as (n - 1) + j.
__global__ __launch_bounds__(256) void tree_kernel(int n, const uint32_t *__restrict__ code, uint32_t *__restrict__ left,
                                                   uint32_t *__restrict__ right, uint32_t *__restrict__ parent,
                                                   uint32_t *__restrict__ first, uint32_t *__restrict__ count) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    const int d = delta(code, n, i, i + 1) - delta(code, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(code, n, i, i - d);
    int lmax = 2;
    while (delta(code, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(code, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (delta(code, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int g = i + s * d + min(d, 0);
    const uint32_t lc = min(i, j) == g ? (uint32_t)(n - 1 + g) : (uint32_t)g;
    const uint32_t rc = max(i, j) == g + 1 ? (uint32_t)(n - 1 + g + 1) : (uint32_t)(g + 1);
    left[i] = lc;
    right[i] = rc;
    first[i] = (uint32_t)min(i, j);
    count[i] = (uint32_t)(max(i, j) - min(i, j) + 1);
    parent[lc] = (uint32_t)i;
    parent[rc] = (uint32_t)i;
}

__device__ __forceinline__ float coherent_load(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 4: boxes bottom-up; box[id] for every node id (leaves in sorted order)
__global__ __launch_bounds__(256) void bounds_kernel(int n, const uint32_t *__restrict__ idx, const Box6 *__restrict__ pbox,
                                                     const uint32_t *__restrict__ left, const uint32_t *__restrict__ right,
                                                     const uint32_t *__restrict__ parent, uint32_t *__restrict__ flag,
                                                     Box6 *box) {
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= n) return;
    const uint32_t leaf = (uint32_t)(n - 1 + j);
    box[leaf] = pbox[idx[j]];
    if (n == 1) return;
    uint32_t p = parent[leaf];
    for (int up = 0; up < 128; ++up) {   // the radix tree over 64-bit (code, index) keys is at most 64 deep
        __threadfence();
        if (atomicAdd(&flag[p], 1u) == 0u) return;   // the sibling subtree is not done yet
        Box6 b;
        const float *bl = box[left[p]].lo, *br = box[right[p]].lo;
        for (int k = 0; k < 6; ++k) {
            const float u = coherent_load(bl + k), v = coherent_load(br + k);
            (k < 3 ? b.lo[k] : b.hi[k - 3]) = k < 3 ? fminf(u, v) : fmaxf(u, v);
        }
        box[p] = b;
        if (p == 0u) return;
        p = parent[p];
    }
}

template <int D> __device__ __forceinline__ float measure(const Box6 &b) {
    float e[3];
    for (int k = 0; k < 3; ++k) e[k] = fmaxf(0.f, b.hi[k] - b.lo[k]);
    return D == 2 ? e[0] + e[1] : e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
}

__device__ __forceinline__ uint32_t coherent_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ Box6 coherent_box(const Box6 *b) {
    Box6 r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = coherent_load(b->lo + k);
        r.hi[k] = coherent_load(b->hi + k);
    }
    return r;
}
__device__ __forceinline__ void grow(Box6 &a, const Box6 &b) {
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = fminf(a.lo[k], b.lo[k]);
        a.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
}

// 4b: treelet restructuring (Karras and Aila 2013, "Fast parallel construction of
// high-quality bounding volume hierarchies"), bottom-up with the arrival counters of step 4.
// At each binary node of at least three primitives the treelet of up to kTreeletLeaves
// leaves is grown by opening the leaf of largest measure, then rebuilt with the topology of
// least SAH cost over all binary trees on those leaves (dynamic programming over the leaf
// subsets), reusing the treelet's own internal node ids.  Cost model: kTreeCi * measure per
// internal node, kTreeCt * measure per primitive.  Boxes of rebuilt nodes are unions of
// their leaves' boxes, so every box stays the exact union of its primitives' boxes.  A node
// only changes when the optimum is strictly cheaper, and each treelet depends only on its
// own subtree, so the result does not depend on the order in which threads arrive.
//
// Each lane climbs from one leaf and grows the treelets it reaches in its own registers
// (their loads are in flight across the lanes); the dynamic program then runs over the
// whole wave, one ready treelet at a time: a lane per subset for the unions, and for each
// subset size the (subset, partition) pairs spread over the lanes with an LDS minimum.
constexpr int kTreeletLeaves = 7;
constexpr int kTreeletPasses = 2;
constexpr float kTreeCi = 1.2f, kTreeCt = 1.0f;
constexpr int kTreeletSets = 1 << kTreeletLeaves;   // leaf subsets as bit masks
constexpr int kTreeletLanes = 64;                   // one wave per block

constexpr int binom7(int m) {
    int c = 1;
    for (int i = 0; i < m; ++i) c = c * (kTreeletLeaves - i) / (i + 1);
    return c;
}
// partitions {q, s ^ q} of a subset of m leaves, each once
constexpr int partitions(int m) { return m < 2 ? 0 : (1 << (m - 1)) - 1; }
constexpr int pair_begin(int m) {
    int b = 0;
    for (int i = 2; i < m; ++i) b += binom7(i) * partitions(i);
    return b;
}
constexpr int set_begin(int m) {
    int b = 0;
    for (int i = 1; i < m; ++i) b += binom7(i);
    return b;
}
constexpr int kTreeletPairs = pair_begin(kTreeletLeaves + 1);
constexpr int pair_rounds(int m) { return (binom7(m) * partitions(m) + kTreeletLanes - 1) / kTreeletLanes; }
constexpr int pair_slot(int m) {
    int b = 0;
    for (int i = 2; i < m; ++i) b += pair_rounds(i);
    return b;
}
constexpr int kPairSlots = pair_slot(kTreeletLeaves + 1);
static_assert(kTreeletPairs == 966 && kPairSlots == 17, "seven-leaf treelet plan");

// The dynamic program's work list for seven leaves: the subsets grouped by size, and for
// each subset of at least two leaves its partitions {q, s ^ q}, q holding the lowest leaf
// of s and q = low | r for the submasks r of the other leaves below all of them, largest
// first (the tie order of the sequential search).
struct TreeletPlan {
    uint8_t pair_s[kTreeletPairs], pair_q[kTreeletPairs];
    uint8_t set[kTreeletSets - 1];
};
constexpr int popcount7(int s) {
    int c = 0;
    for (int t = 0; t < kTreeletLeaves; ++t) c += (s >> t) & 1;
    return c;
}
constexpr TreeletPlan make_treelet_plan() {
    TreeletPlan t{};
    int np = 0, ns = 0;
    for (int m = 1; m <= kTreeletLeaves; ++m)
        for (int s = 1; s < kTreeletSets; ++s) {
            if (popcount7(s) != m) continue;
            t.set[ns++] = (uint8_t)s;
            if (m < 2) continue;
            const int low = s & -s, rest = s ^ low;
            for (int r = (rest - 1) & rest;; r = (r - 1) & rest) {
                t.pair_s[np] = (uint8_t)s;
                t.pair_q[np] = (uint8_t)(low | r);
                ++np;
                if (r == 0) break;
            }
        }
    return t;
}
__constant__ TreeletPlan d_treelet_plan = make_treelet_plan();

// a treelet as its climbing lane grows it: leaf ids, boxes, measures, primitive counts and
// subtree costs; the internal node ids it reuses (the root first)
template <int D> struct TreeletLeaves {
    uint32_t L[kTreeletLeaves], I[kTreeletLeaves - 1], LC[kTreeletLeaves];
    Box6 LB[kTreeletLeaves];
    float LA[kTreeletLeaves], LCost[kTreeletLeaves];
    int k;       // leaves
    float cur;   // the present topology's cost
};

struct TreeletShared {
    unsigned long long best[kTreeletSets];   // (cost bits, 255 - q) of the best partition so far
    Box6 u[kTreeletSets];                    // union box of subset s
    float c[kTreeletSets];                   // least cost over subset s
    float area[kTreeletSets];
    uint32_t cnt[kTreeletSets];
    uint8_t split[kTreeletSets];             // q of the best partition of s
    Box6 lb[kTreeletLeaves];
    uint32_t lc[kTreeletLeaves];
    int k;
};

// Leaf arrays are only indexed by compile-time constants (unrolled loops with a predicate),
// so they stay in registers.
template <int D>
__device__ __forceinline__ void grow_treelet(uint32_t n_int, uint32_t p, const uint32_t *left, const uint32_t *right,
                                             const uint32_t *count, Box6 *box, const float *cost, bool write_box,
                                             TreeletLeaves<D> &T) {
    constexpr int K = kTreeletLeaves;
    auto load_leaf = [&](int i, uint32_t id) {
        const Box6 b = coherent_box(box + id);
        const float a = measure<D>(b);
        const uint32_t cn = id < n_int ? coherent_load(count + id) : 1u;
        const float cc = id < n_int ? coherent_load(cost + id) : kTreeCt * a;
#pragma unroll
        for (int t = 0; t < K; ++t)
            if (t == i) { T.L[t] = id; T.LB[t] = b; T.LA[t] = a; T.LC[t] = cn; T.LCost[t] = cc; }
    };
    load_leaf(0, coherent_load(left + p));
    load_leaf(1, coherent_load(right + p));
    Box6 bp = T.LB[0];
    grow(bp, T.LB[1]);
    if (write_box) box[p] = bp;   // the first pass forms the boxes bottom-up as it climbs
    T.cur = kTreeCi * measure<D>(bp) + T.LCost[0] + T.LCost[1];
    int k = 2;
    T.I[0] = p;
    if (T.LC[0] + T.LC[1] >= 3u) {
        while (k < K) {
            int best = -1;
            float bestA = -1.f;
            uint32_t x = 0;
#pragma unroll
            for (int t = 0; t < K; ++t)
                if (t < k && T.L[t] < n_int && T.LA[t] > bestA) { bestA = T.LA[t]; best = t; x = T.L[t]; }
            if (best < 0) break;
#pragma unroll
            for (int t = 1; t < K - 1; ++t)
                if (t == k - 1) T.I[t] = x;
            load_leaf(best, coherent_load(left + x));
            load_leaf(k, coherent_load(right + x));
            ++k;
        }
    }
    T.k = k;
}

// The least-cost topology over the leaves of lane `owner`'s treelet, by the whole wave.
// On return sh.c / sh.split / sh.u / sh.cnt hold every subset's optimum.
template <int D>
__device__ __forceinline__ void treelet_dp(TreeletShared &sh, const TreeletLeaves<D> &T, bool owner,
                                           const uint32_t (&pw)[kPairSlots], const uint32_t (&sets)[kTreeletLeaves - 1]) {
    constexpr int K = kTreeletLeaves;
    const int lane = (int)threadIdx.x;
    if (owner) {
#pragma unroll
        for (int t = 0; t < K; ++t)
            if (t < T.k) { sh.lb[t] = T.LB[t]; sh.lc[t] = T.LC[t]; sh.c[1u << t] = T.LCost[t]; }
        sh.k = T.k;
    }
    __syncthreads();
    const int k = sh.k;
    const uint32_t full = (1u << k) - 1u;
    // unions (in leaf order, as a sequential build would form them), one lane per subset
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t s = (uint32_t)(lane + 1 + kTreeletLanes * h);
        if (s < (uint32_t)kTreeletSets && (s & ~full) == 0u) {
            Box6 u;
            for (int q = 0; q < 3; ++q) { u.lo[q] = INFINITY; u.hi[q] = -INFINITY; }
            uint32_t cn = 0;
#pragma unroll
            for (int t = 0; t < K; ++t)
                if ((s >> t) & 1u) { grow(u, sh.lb[t]); cn += sh.lc[t]; }
            sh.u[s] = u;
            sh.area[s] = measure<D>(u);
            sh.cnt[s] = cn;
            sh.best[s] = ~0ull;
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 2; m <= K; ++m) {
        if (m > k) break;
#pragma unroll
        for (int r = 0; r < pair_rounds(m); ++r) {
            const uint32_t w = pw[pair_slot(m) + r];
            const uint32_t s = w & 0xFFu, q = (w >> 8) & 0xFFu;
            if (w != 0u && (s & ~full) == 0u) {
                const float v = sh.c[q] + sh.c[s ^ q];
                atomicMin(&sh.best[s], ((unsigned long long)__float_as_uint(v) << 32) | (255u - q));
            }
        }
        __syncthreads();
        const uint32_t s = sets[m - 2];
        if (s != 0u && (s & ~full) == 0u) {
            const unsigned long long b = sh.best[s];
            sh.c[s] = kTreeCi * sh.area[s] + __uint_as_float((uint32_t)(b >> 32));
            sh.split[s] = (uint8_t)(255u - (uint32_t)(b & 0xFFu));
        }
        __syncthreads();
    }
}

// rebuild top-down from the subset tables; internal node ids come from T.I in order (k - 1
// of them, p first)
template <int D>
__device__ __forceinline__ void treelet_rebuild(const TreeletShared &sh, const TreeletLeaves<D> &T, uint32_t p,
                                                uint32_t *left, uint32_t *right, uint32_t *parent, uint32_t *count,
                                                Box6 *box, float *cost) {
    constexpr int K = kTreeletLeaves;
    const uint32_t full = (1u << T.k) - 1u;
    if (!(sh.c[full] < T.cur)) {
        cost[p] = T.cur;
        return;
    }
    auto leaf_of = [&](uint32_t bit_set) {   // the leaf of a one-leaf subset
        uint32_t id = 0;
#pragma unroll
        for (int t = 0; t < K; ++t)
            if (bit_set == (1u << t)) id = T.L[t];
        return id;
    };
    uint32_t stk_node[K - 1], stk_set[K - 1];
    int sp = 0, used = 1;
    stk_node[sp] = p;
    stk_set[sp++] = full;
    while (sp > 0) {
        --sp;
        const uint32_t node = stk_node[sp], s = stk_set[sp];
        const uint32_t sub[2] = {sh.split[s], s ^ sh.split[s]};
        uint32_t ch[2];
        for (int t = 0; t < 2; ++t) {
            if ((sub[t] & (sub[t] - 1u)) == 0u) {
                ch[t] = leaf_of(sub[t]);
            } else {
                uint32_t id = 0;
#pragma unroll
                for (int q = 1; q < K - 1; ++q)
                    if (q == used) id = T.I[q];
                ++used;
                ch[t] = id;
                stk_node[sp] = id;
                stk_set[sp++] = sub[t];
            }
        }
        left[node] = ch[0];
        right[node] = ch[1];
        parent[ch[0]] = node;
        parent[ch[1]] = node;
        box[node] = sh.u[s];
        count[node] = sh.cnt[s];
        cost[node] = sh.c[s];
    }
}

// With pbox (the first pass) each lane first writes its leaf's box and every node's box is
// written on arrival, as the union of its children's (step 4 folded into the climb).
template <int D>
__global__ __launch_bounds__(kTreeletLanes) void treelet_kernel(int n, uint32_t *left, uint32_t *right, uint32_t *parent,
                                                                uint32_t *count, Box6 *box, float *cost, uint32_t *flag,
                                                                const uint32_t *__restrict__ idx,
                                                                const Box6 *__restrict__ pbox) {
    __shared__ TreeletShared sh;
    const int lane = (int)threadIdx.x;
    const int j = (int)(blockIdx.x * blockDim.x) + lane;
    if (n < 3) return;
    const uint32_t n_int = (uint32_t)(n - 1);
    const bool write_box = pbox != nullptr;
    if (write_box && j < n) box[n_int + (uint32_t)j] = pbox[idx[j]];
    // this lane's share of the work list, fixed for every treelet
    uint32_t pw[kPairSlots], sets[kTreeletLeaves - 1];
#pragma unroll
    for (int m = 2; m <= kTreeletLeaves; ++m) {
#pragma unroll
        for (int r = 0; r < pair_rounds(m); ++r) {
            const int g = lane + kTreeletLanes * r;
            pw[pair_slot(m) + r] = g < binom7(m) * partitions(m)
                                       ? (uint32_t)d_treelet_plan.pair_s[pair_begin(m) + g] |
                                             ((uint32_t)d_treelet_plan.pair_q[pair_begin(m) + g] << 8)
                                       : 0u;
        }
        sets[m - 2] = lane < binom7(m) ? (uint32_t)d_treelet_plan.set[set_begin(m) + lane] : 0u;
    }
    bool live = j < n;
    uint32_t p = live ? parent[n_int + (uint32_t)j] : 0u;
    // every step goes to an ancestor in the tree as it was at launch (a node's parent only
    // changes when a treelet above it is rebuilt, after this lane has passed), so a climb
    // is at most the launch-time depth
    for (uint32_t up = 0; up < n_int; ++up) {
        bool ready = false;
        if (live) {
            __threadfence();
            ready = atomicAdd(&flag[p], 1u) != 0u;   // the sibling subtree is done
            live = ready;
        }
        TreeletLeaves<D> T;
        T.k = 0;
        if (ready) grow_treelet<D>(n_int, p, left, right, count, box, cost, write_box, T);
        const bool dp = ready && T.k >= 3;
        if (ready && !dp) cost[p] = T.cur;
        for (unsigned long long todo = __ballot(dp); todo != 0ull; todo &= todo - 1ull) {
            const int owner = __ffsll((long long)todo) - 1;
            treelet_dp<D>(sh, T, lane == owner, pw, sets);
            if (lane == owner) treelet_rebuild<D>(sh, T, p, left, right, parent, count, box, cost);
            __syncthreads();   // the tables are the next treelet's
        }
        if (ready) {
            if (p == 0u) live = false;
            else p = coherent_load(parent + p);
        }
        if (__ballot(live) == 0ull) break;
    }
}

// 4c: position of every node's first primitive in depth-first (left before right) order, so
// each subtree is a contiguous range of primitive records after the restructuring
__global__ __launch_bounds__(256) void order_kernel(int n, const uint32_t *__restrict__ left,
                                                    const uint32_t *__restrict__ right, const uint32_t *__restrict__ parent,
                                                    const uint32_t *__restrict__ count, uint32_t *__restrict__ first,
                                                    uint32_t *__restrict__ bad) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_int = (uint32_t)(n - 1);
    if (id >= n_int + (uint32_t)n) return;
    uint32_t x = id, pos = 0;
    for (uint32_t up = 0; up < n_int && x != 0u; ++up) {
        const uint32_t pp = parent[x];
        if (pp >= n_int) break;   // parents are internal nodes
        if (right[pp] == x) {
            const uint32_t lc = left[pp];
            pos += lc < n_int ? count[lc] : 1u;
        }
        x = pp;
    }
    // a tree over the n primitives: parent links lead to the root, child and parent links
    // agree, and the ranges fit
    const uint32_t own = id < n_int ? count[id] : 1u;
    bool ok = x == 0u && pos + own <= (uint32_t)n;
    if (id < n_int) ok = ok && left[id] < n_int + (uint32_t)n && right[id] < n_int + (uint32_t)n &&
                         parent[left[id]] == id && parent[right[id]] == id;
    if (id != 0u) ok = ok && parent[id] < n_int && (left[parent[id]] == id || right[parent[id]] == id);
    if (!ok) atomicOr(bad, 1u);
    first[id] = pos;
}

// children of binary node b as a wide node (host rule: expand the internal child of largest
// measure, first one on ties, until there are `width` (2..4) or only leaves)
// A subtree of at most kLeafPrims primitives (a contiguous range in depth-first order)
// becomes one leaf.
constexpr uint32_t kLeafPrims = 2;
__device__ __forceinline__ bool is_inner(uint32_t id, uint32_t n_int, const uint32_t *count) {
    return id < n_int && count[id] > kLeafPrims;
}

template <int D>
__device__ __forceinline__ int expand(int n, uint32_t b, const uint32_t *left, const uint32_t *right, const Box6 *box,
                                      const uint32_t *count, int width, uint32_t kids[4]) {
    const uint32_t n_int = (uint32_t)(n - 1);
    int nk;
    if (!is_inner(b, n_int, count)) { kids[0] = b; nk = 1; }
    else { kids[0] = left[b]; kids[1] = right[b]; nk = 2; }
    while (nk < width) {
        int best = -1;
        float bestA = -1.f;
        for (int i = 0; i < nk; ++i)
            if (is_inner(kids[i], n_int, count)) {
                const float A = measure<D>(box[kids[i]]);
                if (A > bestA) { bestA = A; best = i; }
            }
        if (best < 0) break;
        const uint32_t k = kids[best];
        kids[best] = left[k];
        kids[nk++] = right[k];
    }
    return nk;
}

template <int D>
__global__ __launch_bounds__(256) void count_kernel(int n, uint32_t m, const uint32_t *__restrict__ level,
                                                    const uint32_t *__restrict__ left, const uint32_t *__restrict__ right,
                                                    const Box6 *__restrict__ box, const uint32_t *__restrict__ count,
                                                    int width, uint32_t *__restrict__ counts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    uint32_t kids[4];
    const int nk = expand<D>(n, level[t], left, right, box, count, width, kids);
    uint32_t c = 0;
    for (int i = 0; i < nk; ++i) c += is_inner(kids[i], (uint32_t)(n - 1), count) ? 1u : 0u;
    counts[t] = c;
}

// 8-bit child bounds of one axis (the rule of wobhost::quantize_nodes, checked in double)
__device__ void quantize_axis(const float *clo, const float *chi, int nk, float &lo, float &s, uint32_t &qlo,
                              uint32_t &qhi) {
    float a = clo[0], b = chi[0];
    for (int c = 1; c < nk; ++c) { a = fminf(a, clo[c]); b = fmaxf(b, chi[c]); }
    float st = (float)(((double)b - (double)a) / 255.0);
    while ((double)a + 255.0 * (double)st < (double)b) st = nextafterf(st, INFINITY);
    lo = a;
    s = st;
    qlo = 0u;
    qhi = 0u;
    for (int c = 0; c < nk; ++c) {
        uint32_t ql = 0u, qh = 0u;
        if (st > 0.f) {
            const double l = clo[c], h = chi[c];
            double fl = floor((l - (double)a) / (double)st), ch = ceil((h - (double)a) / (double)st);
            fl = fmin(fmax(fl, 0.0), 255.0);
            ch = fmin(fmax(ch, 0.0), 255.0);
            while (fl > 0.0 && (double)a + fl * (double)st > l) fl -= 1.0;
            while (ch < 255.0 && (double)a + ch * (double)st < h) ch += 1.0;
            ql = (uint32_t)fl;
            qh = (uint32_t)ch;
        }
        qlo |= ql << (8 * c);
        qhi |= qh << (8 * c);
    }
}

// 5: wide node w = base + t for binary node level[t]; next-level entries at next[offs[t] + r]
template <int D>
__global__ __launch_bounds__(256) void write_kernel(int n, uint32_t m, uint32_t base, uint32_t next_base,
                                                    const uint32_t *__restrict__ level, const uint32_t *__restrict__ offs,
                                                    const uint32_t *__restrict__ left, const uint32_t *__restrict__ right,
                                                    const Box6 *__restrict__ box, const uint32_t *__restrict__ first,
                                                    const uint32_t *__restrict__ count, int width,
                                                    uint32_t *__restrict__ next, float *__restrict__ fnodes,
                                                    uint32_t *__restrict__ qnodes) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    constexpr int NW = D == 2 ? kNodeWords2 : kNodeWords3;
    constexpr int QW = D == 2 ? kQNodeWords2 : kQNodeWords3;
    const uint32_t n_int = (uint32_t)(n - 1);
    uint32_t kids[4];
    const int nk = expand<D>(n, level[t], left, right, box, count, width, kids);
    const uint32_t w = base + t;
    float *nd = fnodes + (size_t)w * NW;
    uint32_t *q = qnodes + (size_t)w * QW;
    uint32_t r = 0, child[4], mask = 0;
    float clo[3][4], chi[3][4];
    for (int c = 0; c < 4; ++c) {
        Box6 b;
        for (int k = 0; k < 3; ++k) { b.lo[k] = 0.f; b.hi[k] = 0.f; }
        child[c] = 0x80000000u;
        if (c < nk) {
            b = box[kids[c]];
            mask |= 1u << c;
            if (is_inner(kids[c], n_int, count)) {
                const uint32_t slot = offs[t] + r++;
                next[slot] = kids[c];
                child[c] = next_base + slot;
            } else if (kids[c] < n_int) {   // a small subtree: its sorted range is one leaf
                child[c] = ~((first[kids[c]] << 4) | (count[kids[c]] - 1u));
            } else {
                child[c] = ~(first[kids[c]] << 4);   // one primitive: ~((first << 4) | 0)
            }
        }
        for (int k = 0; k < D; ++k) {
            clo[k][c] = b.lo[k];
            chi[k][c] = b.hi[k];
            nd[4 * k + c] = b.lo[k];
            nd[4 * (D + k) + c] = b.hi[k];
        }
        nd[4 * 2 * D + c] = __uint_as_float(c < nk ? child[c] : 0u);
    }
    nd[NW - 4] = __uint_as_float(mask);
    nd[NW - 3] = 0.f;
    nd[NW - 2] = 0.f;
    nd[NW - 1] = 0.f;
    float lo[3] = {0.f, 0.f, 0.f}, s[3] = {0.f, 0.f, 0.f};
    uint32_t qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
    for (int k = 0; k < D; ++k) quantize_axis(clo[k], chi[k], nk, lo[k], s[k], qlo[k], qhi[k]);
    if constexpr (D == 2) {
        q[0] = __float_as_uint(lo[0]); q[1] = __float_as_uint(lo[1]); q[2] = __float_as_uint(s[0]); q[3] = __float_as_uint(s[1]);
        q[4] = qlo[0]; q[5] = qlo[1]; q[6] = qhi[0]; q[7] = qhi[1];
    } else {
        q[0] = __float_as_uint(lo[0]); q[1] = __float_as_uint(lo[1]); q[2] = __float_as_uint(lo[2]); q[3] = __float_as_uint(s[0]);
        q[4] = __float_as_uint(s[1]); q[5] = __float_as_uint(s[2]); q[6] = qlo[0]; q[7] = qlo[1];
        q[8] = qlo[2]; q[9] = qhi[0]; q[10] = qhi[1]; q[11] = qhi[2];
    }
    for (int c = 0; c < 4; ++c) q[QW - 4 + c] = child[c];
}

// 6: stack bound of wide node w (children of the next level are done):
// need = (#children - 1) + max need(internal child)
template <int D>
__global__ __launch_bounds__(256) void need_kernel(uint32_t base, uint32_t m, const float *__restrict__ fnodes,
                                                   uint32_t *need) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    constexpr int NW = D == 2 ? kNodeWords2 : kNodeWords3;
    const float *nd = fnodes + (size_t)(base + t) * NW;
    const uint32_t mask = __float_as_uint(nd[NW - 4]);
    uint32_t nk = 0, mx = 0;
    for (int c = 0; c < 4; ++c) {
        if (!((mask >> c) & 1u)) continue;
        ++nk;
        const int32_t ch = (int32_t)__float_as_uint(nd[4 * 2 * D + c]);
        if (ch >= 0) mx = max(mx, need[ch]);
    }
    need[base + t] = (nk > 0 ? nk - 1 : 0) + mx;
}

// 7: primitive records in depth-first leaf order, in the layout scene_build.cpp writes
template <int D>
__global__ __launch_bounds__(256) void prims_kernel(uint32_t n, const float *__restrict__ V, const uint32_t *__restrict__ idx,
                                                    const uint32_t *__restrict__ first, float *__restrict__ prims,
                                                    uint32_t *__restrict__ prim_id, uint32_t *__restrict__ seen,
                                                    uint32_t *__restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = idx[i], j = first[n - 1 + i];   // sorted leaf i sits at depth-first position j
    // the positions must be a permutation of [0, n): in range and each taken once
    if (j >= n || atomicAdd(seen + j, 1u) != 0u) {
        atomicOr(bad, 2u);
        return;
    }
    prim_id[j] = e;
    auto v = [&](int a, int k) { return V[((size_t)e * D + a) * D + k]; };
    if constexpr (D == 2) {
        float *g = prims + (size_t)j * kPrimWords2;
        g[0] = v(0, 0); g[1] = v(0, 1);
        g[2] = v(0, 0) - v(1, 0); g[3] = v(0, 1) - v(1, 1);
    } else {
        float *g = prims + (size_t)j * kPrimWords3;
        for (int k = 0; k < 3; ++k) {
            g[k] = v(0, k);
            g[3 + k] = v(0, k) - v(1, k);
            g[6 + k] = v(0, k) - v(2, k);
        }
        g[9] = __uint_as_float(e);
        g[10] = 0.f;
        g[11] = 0.f;
    }
}

}  // namespace wobb

namespace wobl {

using namespace wobb;

// ---------------------------------------------------------------------------------------
// Exclusive scan of uint32 counts and a stable LSD radix sort of (key, value) pairs.
// ---------------------------------------------------------------------------------------
// One block scans 1024 entries (4 per thread): the threads' sums are scanned in LDS, each
// thread writes its entries' exclusive prefixes and the block's total goes to block_sums.
// In place is allowed (every thread reads its entries before writing them).
constexpr uint32_t kScanItems = 1024;
__global__ __launch_bounds__(256) void scan_block_kernel(uint32_t n, const uint32_t *in, uint32_t *out,
                                                         uint32_t *block_sums) {
    __shared__ uint32_t s[256];
    const uint32_t tid = threadIdx.x;
    const uint32_t base = blockIdx.x * kScanItems + tid * 4;
    uint32_t a[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = base + k < n ? in[base + k] : 0u;
        sum += a[k];
    }
    s[tid] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t t = tid >= off ? s[tid - off] : 0u;
        __syncthreads();
        s[tid] += t;
        __syncthreads();
    }
    uint32_t run = s[tid] - sum;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += a[k];
    }
    if (tid == 255 && block_sums) block_sums[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t n, uint32_t *out, const uint32_t *block_offsets) {
    const uint32_t add = block_offsets[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanItems + threadIdx.x * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

// Radix sort tiles: 256 threads x 8 rounds; in round r thread t holds element
// base + 256 r + t, so a tile is visited in element order.
constexpr uint32_t kSortRounds = 8, kSortTile = 256 * kSortRounds;

// per tile and digit: the count of the tile's keys with that digit, stored digit-major
// (counts[d * tiles + tile]) so that one exclusive scan gives every (digit, tile) its start
__global__ __launch_bounds__(256) void radix_count_kernel(uint32_t n, uint32_t shift, const uint32_t *keys,
                                                          uint32_t *counts) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (uint32_t r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: an element goes to its (digit, tile) start plus the number of earlier
// elements of the tile with the same digit.  Within a round, a wave finds the lanes that share
// a lane's digit with eight ballots; the waves' per-digit counts in LDS order the waves.
__global__ __launch_bounds__(256) void radix_scatter_kernel(uint32_t n, uint32_t shift, const uint32_t *kin,
                                                            const uint32_t *vin, const uint32_t *starts,
                                                            uint32_t *kout, uint32_t *vout) {
    __shared__ uint32_t next[256];        // next free position of each digit for this tile
    __shared__ uint32_t wcnt[4][256];     // per wave and digit: elements of the current round
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    next[tid] = starts[tid * gridDim.x + blockIdx.x];
    const uint32_t base = blockIdx.x * kSortTile;
    const uint64_t below = ((uint64_t)1 << lane) - 1u;
    for (uint32_t r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * 256 + tid;
        const bool valid = i < n;
        const uint32_t k = valid ? kin[i] : 0u, v = valid ? vin[i] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = ((d >> b) & 1u) != 0u;
            const uint64_t bb = __ballot(bit);
            same &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(same & below);
#pragma unroll
        for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0u;
        __syncthreads();
        if (valid && rank == 0u) wcnt[wave][d] = (uint32_t)__popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t pos = next[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        next[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        __syncthreads();
    }
}

namespace {

struct DevBuf {
    std::vector<void *> ptrs;
    hipError_t err = hipSuccess;
    template <typename T> T *get(size_t count) {
        void *p = nullptr;
        if (err == hipSuccess) err = hipMalloc(&p, std::max<size_t>(count * sizeof(T), 16));
        if (err == hipSuccess) ptrs.push_back(p);
        return (T *)p;
    }
    ~DevBuf() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }

// Scratch entries exclusive_scan needs for n entries: the block sums of every level.
size_t scan_scratch(uint64_t n) {
    size_t total = 0;
    while (n > kScanItems) {
        n = (n + kScanItems - 1) / kScanItems;
        total += n;
    }
    return total;
}

// out[i] = in[0] + ... + in[i - 1]; `scratch` holds scan_scratch(n) entries.
hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *scratch, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t nb = (n + kScanItems - 1) / kScanItems;
    hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(256), 0, st, n, in, out, nb > 1 ? scratch : nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 1) return e;
    if ((e = exclusive_scan(scratch, scratch, nb, scratch + nb, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(256), 0, st, n, out, scratch);
    return hipGetLastError();
}

// Stable sort of n (key, value) pairs by the key bits below `bits`, 8 bits per pass, ping-
// ponging between (k0, v0) and (k1, v1); returns in `which` the pair that holds the result
// (0: k0/v0, 1: k1/v1).  counts: 256 * tiles entries; scan scratch: scan_scratch(256 * tiles).
hipError_t radix_sort(uint32_t n, uint32_t bits, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1,
                      uint32_t *counts, uint32_t *scratch, int &which, hipStream_t st) {
    which = 0;
    if (n <= 1) return hipSuccess;
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(radix_count_kernel, dim3(tiles), dim3(256), 0, st, n, shift, ki, counts);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = exclusive_scan(counts, counts, 256 * tiles, scratch, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(tiles), dim3(256), 0, st, n, shift, ki, vi, counts, ko, vo);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(ki, ko);
        std::swap(vi, vo);
        which ^= 1;
    }
    return hipSuccess;
}

template <int D> std::string build(uint32_t n, const float *h_vertices, int width, DeviceBvh &out) {
    constexpr int NW = D == 2 ? kNodeWords2 : kNodeWords3;
    constexpr int QW = D == 2 ? kQNodeWords2 : kQNodeWords3;
    constexpr int PW = D == 2 ? kPrimWords2 : kPrimWords3;
    const hipStream_t st = nullptr;
    // vertex bounds (the Morton grid)
    float mn[3] = {0.f, 0.f, 0.f}, mx[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < D; ++k) { mn[k] = INFINITY; mx[k] = -INFINITY; }
    for (size_t i = 0; i < (size_t)n * D; ++i)
        for (int k = 0; k < D; ++k) {
            mn[k] = std::min(mn[k], h_vertices[i * D + k]);
            mx[k] = std::max(mx[k], h_vertices[i * D + k]);
        }
    float sc[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < D; ++k) sc[k] = mx[k] > mn[k] ? 1.0f / (mx[k] - mn[k]) : 0.f;
    if (!std::isfinite(mn[0]) || !std::isfinite(mx[0])) return "non-finite vertex coordinates";

    DevBuf tmp;
    float *V = tmp.get<float>((size_t)n * D * D);
    uint32_t *code = tmp.get<uint32_t>(n), *code2 = tmp.get<uint32_t>(n);
    uint32_t *idx = tmp.get<uint32_t>(n), *idx2 = tmp.get<uint32_t>(n);
    Box6 *pbox = tmp.get<Box6>(n);
    const uint32_t n_int = n - 1, n_all = 2 * n - 1;
    uint32_t *left = tmp.get<uint32_t>(std::max<uint32_t>(n_int, 1)), *right = tmp.get<uint32_t>(std::max<uint32_t>(n_int, 1));
    uint32_t *parent = tmp.get<uint32_t>(n_all), *flag = tmp.get<uint32_t>(std::max<uint32_t>(n_int, 1));
    uint32_t *first = tmp.get<uint32_t>(n_all), *count = tmp.get<uint32_t>(std::max<uint32_t>(n_int, 1));
    float *cost = tmp.get<float>(std::max<uint32_t>(n_int, 1));
    uint32_t *bad = tmp.get<uint32_t>(1);
    Box6 *box = tmp.get<Box6>(n_all);
    // wide nodes: at most one per binary internal node (or the single leaf root)
    const uint32_t max_wide = std::max<uint32_t>(n_int, 1);
    uint32_t *lv_a = tmp.get<uint32_t>(max_wide), *lv_b = tmp.get<uint32_t>(max_wide);
    uint32_t *counts = tmp.get<uint32_t>(max_wide), *offs = tmp.get<uint32_t>(max_wide);
    uint32_t *need = tmp.get<uint32_t>(max_wide);
    if (tmp.err != hipSuccess) return std::string("hipMalloc during the device BVH build: ") + hipGetErrorString(tmp.err);

#define WOB_TRY(x)                                                                              \
    do {                                                                                        \
        const hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return std::string("device BVH build: ") + hipGetErrorString(e_); \
    } while (0)
    WOB_TRY(hipMemsetAsync(bad, 0, sizeof(uint32_t), st));
    WOB_TRY(hipMemcpyAsync(V, h_vertices, (size_t)n * D * D * sizeof(float), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((leaf_kernel<D>), dim3(blocks(n)), dim3(256), 0, st, n, V, make_float3(mn[0], mn[1], mn[2]),
                       make_float3(sc[0], sc[1], sc[2]), code, idx, pbox);
    WOB_TRY(hipGetLastError());
    const uint32_t sort_tiles = (n + kSortTile - 1) / kSortTile;
    uint32_t *sort_counts = tmp.get<uint32_t>((size_t)256 * sort_tiles);
    uint32_t *scan_tmp = tmp.get<uint32_t>(std::max(scan_scratch((size_t)256 * sort_tiles), scan_scratch(max_wide)));
    if (tmp.err != hipSuccess) return std::string("hipMalloc during the device BVH build: ") + hipGetErrorString(tmp.err);
    // Morton codes have 3 x 10 bits (3D) or 2 x 16 bits (2D)
    int sorted_in = 0;
    WOB_TRY(radix_sort(n, D == 3 ? 30u : 32u, code, idx, code2, idx2, sort_counts, scan_tmp, sorted_in, st));
    if (sorted_in == 0) {   // the result is in (code, idx): the later steps read (code2, idx2)
        std::swap(code, code2);
        std::swap(idx, idx2);
    }
    if (n > 1) {
        hipLaunchKernelGGL(tree_kernel, dim3(blocks(n_int)), dim3(256), 0, st, (int)n, code2, left, right, parent, first,
                           count);
        WOB_TRY(hipGetLastError());
        WOB_TRY(hipMemsetAsync(flag, 0, (size_t)n_int * sizeof(uint32_t), st));
    }
    if (n > 2 && kTreeletPasses > 0) {
        // boxes come from the first treelet pass
        for (int pass = 0; pass < kTreeletPasses; ++pass) {
            if (pass > 0) WOB_TRY(hipMemsetAsync(flag, 0, (size_t)n_int * sizeof(uint32_t), st));
            hipLaunchKernelGGL((treelet_kernel<D>), dim3((n + kTreeletLanes - 1) / kTreeletLanes), dim3(kTreeletLanes), 0, st,
                               (int)n, left, right, parent, count, box, cost, flag, idx2, pass == 0 ? pbox : nullptr);
            WOB_TRY(hipGetLastError());
        }
    } else {
        hipLaunchKernelGGL(bounds_kernel, dim3(blocks(n)), dim3(256), 0, st, (int)n, idx2, pbox, left, right, parent, flag,
                           box);
        WOB_TRY(hipGetLastError());
    }
    if (n > 1) {
        hipLaunchKernelGGL(order_kernel, dim3(blocks(n_all)), dim3(256), 0, st, (int)n, left, right, parent, count, first,
                           bad);
        WOB_TRY(hipGetLastError());
    } else {
        WOB_TRY(hipMemsetAsync(first, 0, sizeof(uint32_t), st));
    }
    uint32_t bad0 = 0;
    WOB_TRY(hipMemcpyAsync(&bad0, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    WOB_TRY(hipStreamSynchronize(st));
    if (bad0 != 0u) return "device BVH build: internal error (tree check " + std::to_string(bad0) + ")";

    // outputs
    float *fnodes = nullptr, *prims = nullptr;
    uint32_t *qnodes = nullptr, *prim_id = nullptr;
    WOB_TRY(hipMalloc(&fnodes, (size_t)max_wide * NW * sizeof(float)));
    out.fnodes = fnodes;
    WOB_TRY(hipMalloc(&qnodes, (size_t)max_wide * QW * sizeof(uint32_t)));
    out.qnodes = qnodes;
    WOB_TRY(hipMalloc(&prims, (size_t)n * PW * sizeof(float)));
    out.prims = prims;
    WOB_TRY(hipMalloc(&prim_id, (size_t)n * sizeof(uint32_t)));
    out.prim_id = prim_id;
    out.bytes = (size_t)max_wide * (NW + QW) * 4 + (size_t)n * (PW + 1) * 4;

    // top-down levels
    const uint32_t root = 0u;   // internal node 0, or the single leaf (id n - 1 = 0) when n == 1
    WOB_TRY(hipMemcpyAsync(lv_a, &root, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    std::vector<uint32_t> level_base{0u}, level_size{1u};
    uint32_t *cur = lv_a, *nxt = lv_b;
    uint32_t base = 0, m = 1;
    while (m > 0) {
        if (level_base.size() > 256) return "device BVH build: tree deeper than 256 levels";
        hipLaunchKernelGGL((count_kernel<D>), dim3(blocks(m)), dim3(256), 0, st, (int)n, m, cur, left, right, box, count,
                           width, counts);
        WOB_TRY(hipGetLastError());
        WOB_TRY(exclusive_scan(counts, offs, m, scan_tmp, st));
        uint32_t last[2] = {0u, 0u};
        WOB_TRY(hipMemcpyAsync(&last[0], counts + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        WOB_TRY(hipMemcpyAsync(&last[1], offs + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        const uint32_t next_base = base + m;
        hipLaunchKernelGGL((write_kernel<D>), dim3(blocks(m)), dim3(256), 0, st, (int)n, m, base, next_base, cur, offs, left,
                           right, box, first, count, width, nxt, fnodes, qnodes);
        WOB_TRY(hipGetLastError());
        WOB_TRY(hipStreamSynchronize(st));
        const uint32_t m_next = last[0] + last[1];
        if (next_base + m_next > max_wide) return "device BVH build: internal error (wide node count)";
        std::swap(cur, nxt);
        base = next_base;
        m = m_next;
        if (m > 0) {
            level_base.push_back(base);
            level_size.push_back(m);
        }
    }
    const uint32_t n_wide = base;
    for (size_t l = level_base.size(); l-- > 0;) {
        hipLaunchKernelGGL((need_kernel<D>), dim3(blocks(level_size[l])), dim3(256), 0, st, level_base[l], level_size[l],
                           fnodes, need);
        WOB_TRY(hipGetLastError());
    }
    WOB_TRY(hipMemsetAsync(code, 0, (size_t)n * sizeof(uint32_t), st));   // the sort's other buffer, reused as position marks
    hipLaunchKernelGGL((prims_kernel<D>), dim3(blocks(n)), dim3(256), 0, st, n, V, idx2, first, prims, prim_id, code, bad);
    WOB_TRY(hipGetLastError());
    uint32_t need0 = 0;
    WOB_TRY(hipMemcpyAsync(&need0, need, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    WOB_TRY(hipMemcpyAsync(&bad0, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    WOB_TRY(hipStreamSynchronize(st));
#undef WOB_TRY
    if (bad0 != 0u) return "device BVH build: internal error (tree check " + std::to_string(bad0) + ")";
    out.n_nodes = n_wide;
    out.depth = (uint32_t)level_base.size();
    out.stack_need = need0 + 1;
    return "";
}

}  // namespace

std::string build_bvh_device(int dim, uint32_t n, const float *h_vertices, DeviceBvh &out, int width) {
    out = DeviceBvh{};
    if (width < 2 || width > 4) return "BVH width must be 2, 3 or 4";
    std::string err = dim == 2 ? build<2>(n, h_vertices, width, out) : build<3>(n, h_vertices, width, out);
    if (!err.empty()) {
        (void)hipFree(out.fnodes);
        (void)hipFree(out.qnodes);
        (void)hipFree(out.prims);
        (void)hipFree(out.prim_id);
        out = DeviceBvh{};
    }
    return err;
}

}  // namespace wobl
