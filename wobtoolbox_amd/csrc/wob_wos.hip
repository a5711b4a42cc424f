// Closest-point queries and walk on spheres (apps/simple_wob_and_wos.cu).
//
// closest_point: nearest element to a point.  The reference calls lbvh's nearest query with a
// point-segment squared distance (simple_wob_and_wos.cu:48-79).  Here it is a depth-first
// search of the 4-wide BVH: the children of a node are ordered by the squared distance from
// the point to their boxes, the nearest is visited next and the others are pushed with their
// distances on the lane's LDS stack; a popped entry whose box is farther than the best
// element so far is dropped.  Element distances use the element's vertices from the attribute
// rows (caller order) with the reference's formula, so the result does not depend on the
// tree; ties go to the smaller element id.
//
// wos_kernel: WoSFunc (simple_wob_and_wos.cu:116-141) -- from x, jump to a uniform point on
// the sphere of radius R = distance to the boundary until R < eps, then score the boundary
// value at the closest point.  Per lane a chunk of one point's paths, FP32 Kahan sums (and
// sums of squares) into partials that reduce_kernel adds in FP64 in a fixed order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wob_device.hpp"
#include "wob_launch.hpp"

#pragma clang fp contract(off)

using namespace wobdev;

namespace wobw {

using wobl::WalkEntry;
using wobl::WalkLaunch;

template <int D> struct Closest {
    float d2;
    uint32_t elem;
    float w[3];   // 2D: t along a -> b in w[0]; 3D: barycentric weights of a, b, c
};

// squared distance from p to segment (a, b) and the clamped parameter t (the LineSqrD functor,
// simple_wob_and_wos.cu:53-62, with Eigen's left-to-right dot products)
__device__ __forceinline__ float seg_dist2(const float *p, const float *v, float &t) {
    const float l0 = v[2] - v[0], l1 = v[3] - v[1];
    const float q0 = p[0] - v[0], q1 = p[1] - v[1];
    t = (q0 * l0 + q1 * l1) / (l0 * l0 + l1 * l1);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float e0 = (v[0] + l0 * t) - p[0], e1 = (v[1] + l1 * t) - p[1];
    return e0 * e0 + e1 * e1;
}

// squared distance from p to triangle (a, b, c) and the barycentric weights of the closest
// point: the Voronoi-region test of Ericson, Real-Time Collision Detection, 5.1.5.
__device__ __forceinline__ float tri_dist2(const float *p, const float *v, float *w) {
    const float *a = v, *b = v + 3, *c = v + 6;
    float ab[3], ac[3], ap[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
    auto dot3 = [](const float *x, const float *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    float wa, wb, wc;
    if (d1 <= 0.f && d2 <= 0.f) {
        wa = 1.f; wb = 0.f; wc = 0.f;
    } else {
        float bp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) bp[k] = p[k] - b[k];
        const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
        float cp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) cp[k] = p[k] - c[k];
        const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
        const float vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
        if (d3 >= 0.f && d4 <= d3) {
            wa = 0.f; wb = 1.f; wc = 0.f;
        } else if (d6 >= 0.f && d5 <= d6) {
            wa = 0.f; wb = 0.f; wc = 1.f;
        } else if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
            const float s = d1 / (d1 - d3);
            wa = 1.f - s; wb = s; wc = 0.f;
        } else if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
            const float s = d2 / (d2 - d6);
            wa = 1.f - s; wb = 0.f; wc = s;
        } else if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
            const float s = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            wa = 0.f; wb = 1.f - s; wc = s;
        } else {
            const float den = 1.f / (va + vb + vc);
            wb = vb * den; wc = vc * den; wa = (1.f - wb) - wc;
        }
    }
    w[0] = wa; w[1] = wb; w[2] = wc;
    float e2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float q = ((wa * a[k] + wb * b[k]) + wc * c[k]) - p[k];
        e2 = e2 + q * q;
    }
    return e2;
}

template <int D> __device__ __forceinline__ void elem_vertices(const DevScene &S, uint32_t e, float *v) {
    if constexpr (D == 2) {
        const float4 f0 = S.attr[3 * e];
        v[0] = f0.x; v[1] = f0.y; v[2] = f0.z; v[3] = f0.w;
    } else {
        const float4 f0 = S.geo[4 * e], f1 = S.geo[4 * e + 1], f2 = S.geo[4 * e + 2];
        v[0] = f0.x; v[1] = f0.y; v[2] = f0.z; v[3] = f0.w; v[4] = f1.x; v[5] = f1.y; v[6] = f1.z; v[7] = f1.w;
        v[8] = f2.x;
    }
}

// boundary value at the closest point: linear along the segment (simple_wob_and_wos.cu:74),
// barycentric on a triangle
template <int D> __device__ __forceinline__ float closest_value(const DevScene &S, const Closest<D> &c) {
    if constexpr (D == 2) {
        const float4 f1 = S.attr[3 * c.elem + 1];
        return (f1.w - f1.z) * c.w[0] + f1.z;
    } else {
        const float4 f3 = S.attr[5 * c.elem + 3];
        return (c.w[0] * f3.x + c.w[1] * f3.y) + c.w[2] * f3.z;
    }
}

// Stack: STACK (child, distance) pairs per lane, entries `stride` words apart.  The builder
// bounds the pushes of a depth-first descent (at most three per level) by STACK.
//
// Element distances carry rounding errors of order eps * |coordinates|, which can exceed the
// distance itself near the boundary, so a box is pruned only when it is farther than the best
// element by more than delta = 16 eps (|x|_inf + the root boxes' largest |coordinate|): every
// element whose computed distance could tie or beat the best is tested, and the result equals
// a brute-force scan with the same formulas.
template <int D, int STACK>
__device__ Closest<D> closest_point(const DevScene &S, const float *x, uint32_t *stack, uint32_t stride) {
    constexpr int NQ = D == 2 ? 6 : 8;
    Closest<D> best;
    best.d2 = __builtin_inff();
    best.elem = 0xFFFFFFFFu;
    best.w[0] = best.w[1] = best.w[2] = 0.f;
    float delta;
    {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float4 l = S.fnodes[k], u = S.fnodes[D + k];
            m = fmaxf(m, fmaxf(fmaxf(fmaxf(fabsf(l.x), fabsf(l.y)), fmaxf(fabsf(l.z), fabsf(l.w))),
                               fmaxf(fmaxf(fabsf(u.x), fabsf(u.y)), fmaxf(fabsf(u.z), fabsf(u.w)))));
            m = fmaxf(m, fabsf(x[k]));
        }
        delta = 16.0f * kFltEps * m;
    }
    float lim = __builtin_inff();   // (sqrt(best) + delta)^2
    int cur = 0;
    uint32_t sp = 0;
    while (true) {
        if (cur >= 0) {
            const float4 *np = S.fnodes + (size_t)cur * NQ;
            float4 lo[D], hi[D];
#pragma unroll
            for (int k = 0; k < D; ++k) { lo[k] = np[k]; hi[k] = np[D + k]; }
            const float4 chf = np[2 * D];
            const uint32_t valid = __float_as_uint(np[2 * D + 1].x);
            int ch[4] = {(int)__float_as_uint(chf.x), (int)__float_as_uint(chf.y), (int)__float_as_uint(chf.z),
                         (int)__float_as_uint(chf.w)};
            float bd[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float d2 = 0.f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float l = c == 0 ? lo[k].x : c == 1 ? lo[k].y : c == 2 ? lo[k].z : lo[k].w;
                    const float u = c == 0 ? hi[k].x : c == 1 ? hi[k].y : c == 2 ? hi[k].z : hi[k].w;
                    const float e = fmaxf(fmaxf(l - x[k], x[k] - u), 0.f);
                    d2 = d2 + e * e;
                }
                bd[c] = ((valid >> c) & 1u) && d2 <= lim ? d2 : __builtin_inff();
            }
            // order the four children by box distance (nearest first)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3 - i; ++j) {
                    const bool sw = bd[j + 1] < bd[j];
                    const float tb = bd[j]; bd[j] = sw ? bd[j + 1] : bd[j]; bd[j + 1] = sw ? tb : bd[j + 1];
                    const int tc = ch[j]; ch[j] = sw ? ch[j + 1] : ch[j]; ch[j + 1] = sw ? tc : ch[j + 1];
                }
            // push the farther ones (farthest first), continue with the nearest
#pragma unroll
            for (int c = 3; c >= 1; --c)
                if (bd[c] != __builtin_inff() && sp < (uint32_t)STACK) {
                    stack[(2 * sp) * stride] = (uint32_t)ch[c];
                    stack[(2 * sp + 1) * stride] = __float_as_uint(bd[c]);
                    ++sp;
                }
            if (bd[0] != __builtin_inff()) {
                cur = ch[0];
                continue;
            }
        } else {
            // leaf: test its elements
            const uint32_t v = ~(uint32_t)cur;
            const uint32_t p0 = v >> 4, p1 = p0 + (v & 15u) + 1u;
            for (uint32_t p = p0; p < p1; ++p) {
                const uint32_t e = S.prim_id[p];
                float vv[D * D], w[3] = {0.f, 0.f, 0.f}, d2;
                elem_vertices<D>(S, e, vv);
                if constexpr (D == 2) d2 = seg_dist2(x, vv, w[0]);
                else d2 = tri_dist2(x, vv, w);
                if (d2 < best.d2 || (d2 == best.d2 && e < best.elem)) {
                    best.d2 = d2; best.elem = e; best.w[0] = w[0]; best.w[1] = w[1]; best.w[2] = w[2];
                    const float r = sqrtf(d2) + delta;
                    lim = r * r;
                }
            }
        }
        // pop the next entry whose box can still hold the closest element
        bool found = false;
        while (sp > 0) {
            --sp;
            const float d2 = __uint_as_float(stack[(2 * sp + 1) * stride]);
            if (d2 <= lim) {
                cur = (int)stack[(2 * sp) * stride];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return best;
}

template <int D, int STACK>
__global__ __launch_bounds__(256) void closest_kernel(DevScene S, uint32_t n, const float *pts, uint32_t *elem,
                                                      float *dist, float *value) {
    extern __shared__ uint32_t lds_stack[];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = pts[(size_t)i * D + k];
    const Closest<D> c = closest_point<D, STACK>(S, x, lds_stack + threadIdx.x, blockDim.x);
    elem[i] = c.elem;
    dist[i] = sqrtf(c.d2);
    value[i] = closest_value<D>(S, c);
}

// One lane per (point, chunk of paths).  Partials: [0][item] sums, [1][item] sums of squares,
// item = chunk * n_points + point (reduce_kernel's layout for V = 1).
template <int D, int STACK>
__global__ __launch_bounds__(256) void wos_kernel(WalkLaunch L, float eps, uint32_t max_steps) {
    extern __shared__ uint32_t lds_stack[];
    const WalkEntry &e = L.e0;
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= L.total_items) return;
    const uint32_t chunk = item / e.n_points, pt = item - chunk * e.n_points;
    const uint32_t j0 = chunk * e.paths_per_chunk;
    const uint32_t j1 = min(j0 + e.paths_per_chunk, e.paths_per_point);
    float x0[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x0[k] = e.points[(size_t)pt * D + k];
    float s = 0.f, sc = 0.f, q = 0.f, qc = 0.f;
    for (uint32_t j = j0; j < j1; ++j) {
        Rng rng;
        rng.init(L.seed, (uint32_t)(e.path_base + j), e.point_base + pt, e.stream);
        float x[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = x0[k];
        float val = 0.f;
        for (uint32_t step = 0;; ++step) {
            const Closest<D> c = closest_point<D, STACK>(e.scene, x, lds_stack + threadIdx.x, blockDim.x);
            const float R = sqrtf(c.d2);
            if (R < eps || step + 1 >= max_steps) {
                val = closest_value<D>(e.scene, c);
                break;
            }
            if constexpr (D == 2) {
                float sn, cs;
                sincospif(2.0f * rng.uniform(), &sn, &cs);
                x[0] = x[0] + R * cs;
                x[1] = x[1] + R * sn;
            } else {
                const float z = 1.0f - 2.0f * rng.uniform();
                float sn, cs;
                sincospif(2.0f * rng.uniform(), &sn, &cs);
                const float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
                x[0] = x[0] + R * (r * cs);
                x[1] = x[1] + R * (r * sn);
                x[2] = x[2] + R * z;
            }
        }
        // Kahan sums of the values and their squares
        const float y = val - sc, t = s + y;
        sc = (t - s) - y; s = t;
        const float y2 = val * val - qc, t2 = q + y2;
        qc = (t2 - q) - y2; q = t2;
    }
    L.partials[item] = s;
    L.partials[(size_t)L.total_items + item] = q;
}

}  // namespace wobw

namespace wobl {

size_t closest_lds_bytes(uint32_t block) { return (size_t)(2 * kStackEntries + 2) * block * sizeof(uint32_t); }

hipError_t launch_closest(int dim, const wobdev::DevScene &S, uint32_t n, const float *pts, uint32_t *elem, float *dist,
                          float *value, hipStream_t st) {
    const uint32_t block = 128, grid = (n + block - 1) / block;
    const size_t lds = closest_lds_bytes(block);
    if (dim == 2)
        hipLaunchKernelGGL((wobw::closest_kernel<2, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, pts, elem, dist, value);
    else
        hipLaunchKernelGGL((wobw::closest_kernel<3, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, pts, elem, dist, value);
    return hipGetLastError();
}

hipError_t launch_wos(int dim, const WalkLaunch &L, float eps, uint32_t max_steps, hipStream_t st) {
    const uint32_t block = 128, grid = (L.total_items + block - 1) / block;
    const size_t lds = closest_lds_bytes(block);
    if (dim == 2)
        hipLaunchKernelGGL((wobw::wos_kernel<2, kStackEntries>), dim3(grid), dim3(block), lds, st, L, eps, max_steps);
    else
        hipLaunchKernelGGL((wobw::wos_kernel<3, kStackEntries>), dim3(grid), dim3(block), lds, st, L, eps, max_steps);
    return hipGetLastError();
}

}  // namespace wobl
