// Host-side scene construction: wide BVH, leaf-ordered primitives, element attributes,
// CDFs and guide tables.  See scene_build.hpp.
//
// Built with -ffp-contract=off so the FP32 element areas and Kahan running sums are the
// operation sequence written here (and in scene.cuh:271-298, 384, 652-654), not a fused one.

#include "scene_build.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace wobhost {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) { lo[k] = std::numeric_limits<float>::infinity(); hi[k] = -lo[k]; }
    }
    void grow(const Box &b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow(const float *p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    // Probability that a random line meets the box is proportional to its surface area (3D)
    // or perimeter (2D) (Cauchy-Crofton); this is the SAH weight for all-hit line queries.
    float measure(int dim) const {
        float e[3];
        for (int k = 0; k < 3; ++k) e[k] = std::max(0.f, hi[k] - lo[k]);
        if (dim == 2) return e[0] + e[1];
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
    }
};

struct BNode {
    Box box;
    int left = -1, right = -1;   // children (binary build)
    int start = 0, count = 0;    // leaf range
    bool leaf() const { return left < 0; }
};

struct Builder {
    int dim;
    std::vector<Box> pbox;
    std::vector<float> cent;      // 3 per prim
    std::vector<uint32_t> order;  // permutation
    std::vector<BNode> nodes;
    bool median_only = false;

    int build(int start, int count, int depth) {
        const int id = (int)nodes.size();
        nodes.emplace_back();
        Box b; b.reset();
        Box cb; cb.reset();
        for (int i = start; i < start + count; ++i) {
            b.grow(pbox[order[i]]);
            cb.grow(&cent[3 * order[i]]);
        }
        nodes[id].box = b;
        nodes[id].start = start;
        nodes[id].count = count;
        if (count <= 2) return id;
        int axis = 0;
        float ext = cb.hi[0] - cb.lo[0];
        for (int k = 1; k < dim; ++k)
            if (cb.hi[k] - cb.lo[k] > ext) { ext = cb.hi[k] - cb.lo[k]; axis = k; }
        int split = -1;
        if (ext > 0 && !median_only && depth < 48) {
            constexpr int B = 32;
            // binned SAH over every axis with extent (the best split of all three)
            float best = std::numeric_limits<float>::infinity();
            int bestj = -1, bestax = axis;
            for (int ax = 0; ax < dim; ++ax) {
                const float ea = cb.hi[ax] - cb.lo[ax];
                if (!(ea > 0)) continue;
                Box bb[B];
                int bc[B] = {0};
                for (int jj = 0; jj < B; ++jj) bb[jj].reset();
                const float scale = B / ea;
                for (int i = start; i < start + count; ++i) {
                    int jb = (int)((cent[3 * order[i] + ax] - cb.lo[ax]) * scale);
                    jb = std::min(std::max(jb, 0), B - 1);
                    bb[jb].grow(pbox[order[i]]);
                    bc[jb]++;
                }
                float rightA[B];
                int rightN[B];
                Box acc; acc.reset();
                int an = 0;
                for (int jj = B - 1; jj > 0; --jj) { acc.grow(bb[jj]); an += bc[jj]; rightA[jj] = acc.measure(dim); rightN[jj] = an; }
                acc.reset(); an = 0;
                for (int jj = 1; jj < B; ++jj) {
                    acc.grow(bb[jj - 1]); an += bc[jj - 1];
                    if (an == 0 || rightN[jj] == 0) continue;
                    const float c = acc.measure(dim) * an + rightA[jj] * rightN[jj];
                    if (c < best) { best = c; bestj = jj; bestax = ax; }
                }
            }
            const float leaf_cost = b.measure(dim) * count;
            // traversal cost ~ one node visit per box test; stop at small leaves when splitting does not pay
            if (bestj > 0 && !(count <= kMaxLeaf && best + 0.5f * b.measure(dim) >= leaf_cost)) {
                const float ea = cb.hi[bestax] - cb.lo[bestax];
                const float scale = B / ea;
                auto bin_of = [&](uint32_t p) {
                    int jb = (int)((cent[3 * p + bestax] - cb.lo[bestax]) * scale);
                    return std::min(std::max(jb, 0), B - 1);
                };
                auto mid = std::partition(order.begin() + start, order.begin() + start + count,
                                          [&](uint32_t p) { return bin_of(p) < bestj; });
                split = (int)(mid - order.begin()) - start;
            } else if (count <= kMaxLeaf) {
                return id;
            }
        } else if (count <= kMaxLeaf && !median_only) {
            return id;
        }
        if (split <= 0 || split >= count) {
            if (count <= kMaxLeaf && !median_only) return id;
            // median split by centroid along the widest axis (balanced)
            split = count / 2;
            std::nth_element(order.begin() + start, order.begin() + start + split, order.begin() + start + count,
                             [&](uint32_t a, uint32_t c) { return cent[3 * a + axis] < cent[3 * c + axis]; });
        }
        const int l = build(start, split, depth + 1);
        const int r = build(start + split, count - split, depth + 1);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

struct WNode {
    int child_bnode[4];
    int nchild = 0;
};

}  // namespace

void build_cdf(const std::vector<float> &w, std::vector<float> &cdf, float &total, std::vector<uint32_t> &guide) {
    const size_t n = w.size();
    cdf.resize(n);
    float sum = 0.f, c = 0.f;   // KahanSum (utils.cuh:148-166)
    for (size_t i = 0; i < n; ++i) {
        const float y = w[i] - c;
        const float t = sum + y;
        c = (t - sum) - y;
        sum = t;
        cdf[i] = sum;
    }
    total = sum;
    for (size_t i = 0; i < n; ++i) cdf[i] /= total;
    build_guide(cdf, guide);
}

void build_guide(const std::vector<float> &cdf, std::vector<uint32_t> &guide, size_t entries) {
    // guide[k] = first i with cdf[i] >= k / entries; the device search walks forward /
    // backward from there, so it returns the first i with u <= cdf[i] (clamped to n-1) for
    // any u.
    const size_t n = cdf.size();
    const size_t m = entries ? entries : n;
    guide.resize(m);
    size_t i = 0;
    for (size_t k = 0; k < m; ++k) {
        const float lo = (float)k / (float)m;
        while (i + 1 < n && cdf[i] < lo) ++i;
        guide[k] = (uint32_t)i;
    }
}

std::vector<float> guide_records(const std::vector<float> &cdf, const std::vector<uint32_t> &guide) {
    const size_t n = cdf.size();
    std::vector<float> r(4 * guide.size(), 0.f);
    for (size_t k = 0; k < guide.size(); ++k) {
        const uint32_t i = guide[k];
        std::memcpy(&r[4 * k], &i, 4);
        r[4 * k + 1] = i > 0 ? cdf[i - 1] : 0.f;
        r[4 * k + 2] = cdf[i];
        r[4 * k + 3] = i + 1 < n ? cdf[i + 1] : 0.f;
    }
    return r;
}

std::string build_scene(const wob_scene_desc &d, HostScene &S, bool with_bvh, int width) {
    if (d.dim != 2 && d.dim != 3) return "dim must be 2 or 3";
    if (d.n_elements == 0) return "scene has no elements";
    if (!d.vertices || !d.normals || !d.boundary_values || !d.boundary_types) return "null element array";
    if (d.n_elements >= (1u << 27)) return "too many elements (max 2^27 - 1)";
    const int D = (int)d.dim;
    const uint32_t n = d.n_elements;
    S.dim = d.dim;
    S.n = n;
    for (uint32_t i = 0; i < n; ++i) {
        const int t = d.boundary_types[i];
        if (t < 0 || t > 2) return "boundary type out of range at element " + std::to_string(i);
    }

    auto V = [&](uint32_t e, int j, int k) { return d.vertices[((size_t)e * D + j) * D + k]; };

    // ---- element attributes, areas, boundary-value weights (element order)
    const int AW = D == 2 ? kAttrWords2 : kAttrWords3;
    S.elem_attr.assign((size_t)n * AW, 0.f);
    S.elem_geo.assign(D == 3 ? (size_t)n * kGeoWords3 : 0u, 0.f);
    std::vector<float> area(n), bvw(n);
    for (uint32_t e = 0; e < n; ++e) {
        float *a = &S.elem_attr[(size_t)e * AW];
        const float *bv = &d.boundary_values[(size_t)e * D];
        const float al0 = d.robin_alphas ? d.robin_alphas[(size_t)e * D + 0] : 0.f;
        const float al1 = d.robin_alphas ? d.robin_alphas[(size_t)e * D + 1] : 0.f;
        if (D == 2) {
            const float dx = V(e, 1, 0) - V(e, 0, 0), dy = V(e, 1, 1) - V(e, 0, 1);
            const float ar = std::sqrt(dx * dx + dy * dy);                                    // scene.cuh:384
            area[e] = ar;
            bvw[e] = ar * (std::abs(bv[0]) + std::abs(bv[1])) / 2.0f;                        // scene.cuh:386-392
            a[0] = V(e, 0, 0); a[1] = V(e, 0, 1); a[2] = V(e, 1, 0); a[3] = V(e, 1, 1);
            a[4] = d.normals[2 * e]; a[5] = d.normals[2 * e + 1]; a[6] = bv[0]; a[7] = bv[1];
            a[8] = al0; a[9] = al1; a[10] = (float)d.boundary_types[e]; a[11] = ar;
        } else {
            const float al2 = d.robin_alphas ? d.robin_alphas[(size_t)e * D + 2] : 0.f;
            const float ca[3] = {V(e, 2, 0) - V(e, 0, 0), V(e, 2, 1) - V(e, 0, 1), V(e, 2, 2) - V(e, 0, 2)};
            const float ba[3] = {V(e, 1, 0) - V(e, 0, 0), V(e, 1, 1) - V(e, 0, 1), V(e, 1, 2) - V(e, 0, 2)};
            const float cr[3] = {ca[1] * ba[2] - ca[2] * ba[1], ca[2] * ba[0] - ca[0] * ba[2], ca[0] * ba[1] - ca[1] * ba[0]};
            const float ar = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 2.0f;  // scene.cuh:652-654
            area[e] = ar;
            bvw[e] = ar * (std::abs(bv[0]) + std::abs(bv[1]) + std::abs(bv[2])) / 3.0f;      // scene.cuh:656-663
            a[0] = V(e, 0, 0); a[1] = V(e, 0, 1); a[2] = V(e, 0, 2); a[3] = V(e, 1, 0);
            a[4] = V(e, 1, 1); a[5] = V(e, 1, 2); a[6] = V(e, 2, 0); a[7] = V(e, 2, 1);
            a[8] = V(e, 2, 2); a[9] = d.normals[3 * e]; a[10] = d.normals[3 * e + 1]; a[11] = d.normals[3 * e + 2];
            a[12] = bv[0]; a[13] = bv[1]; a[14] = bv[2]; a[15] = (float)d.boundary_types[e];
            a[16] = al0; a[17] = al1; a[18] = al2; a[19] = ar;
            std::copy(a, a + 12, &S.elem_geo[(size_t)e * kGeoWords3]);
        }
    }
    build_cdf(area, S.area_cdf, S.total_area, S.area_guide);
    if (D == 3) build_guide(S.area_cdf, S.area_fused_guide, (size_t)kFusedPerElement * n);
    build_cdf(bvw, S.bv_cdf, S.total_bv, S.bv_guide);
    if (!with_bvh) return "";

    // ---- binary SAH build
    Builder B;
    B.dim = D;
    B.pbox.resize(n);
    B.cent.assign((size_t)3 * n, 0.f);
    B.order.resize(n);
    for (uint32_t e = 0; e < n; ++e) {
        Box b; b.reset();
        for (int j = 0; j < D; ++j) {
            float p[3] = {V(e, j, 0), V(e, j, 1), D == 3 ? V(e, j, 2) : 0.f};
            b.grow(p);
        }
        if (D == 2) { b.lo[2] = b.hi[2] = 0.f; }
        // conservative inflation: the primitive test, not the box, decides every hit
        for (int k = 0; k < 3; ++k) {
            const float m = std::max(std::abs(b.lo[k]), std::abs(b.hi[k]));
            const float pad = m * 4.0e-6f + 1.0e-7f;
            b.lo[k] -= pad; b.hi[k] += pad;
        }
        B.pbox[e] = b;
        for (int k = 0; k < 3; ++k) B.cent[3 * e + k] = 0.5f * (b.lo[k] + b.hi[k]);
        B.order[e] = e;
    }
    int root = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        B.nodes.clear();
        B.nodes.reserve(2 * (size_t)n);
        std::iota(B.order.begin(), B.order.end(), 0u);
        B.median_only = attempt == 1;
        root = B.build(0, (int)n, 0);

        // ---- collapse to 4-wide nodes with at most `width` children
        std::vector<WNode> wn;
        std::vector<int> wn_of_bnode;
        // Each wide node is created for a binary internal node (or the root even if it is a leaf).
        std::vector<std::pair<int, int>> todo;  // (binary node, wide depth)
        wn.reserve(n);
        auto make_wide = [&](int bn) {
            WNode w;
            std::vector<int> kids;
            if (B.nodes[bn].leaf()) kids.push_back(bn);
            else { kids.push_back(B.nodes[bn].left); kids.push_back(B.nodes[bn].right); }
            while ((int)kids.size() < width) {
                int best = -1;
                float bestA = -1.f;
                for (int i = 0; i < (int)kids.size(); ++i)
                    if (!B.nodes[kids[i]].leaf()) {
                        const float A = B.nodes[kids[i]].box.measure(D);
                        if (A > bestA) { bestA = A; best = i; }
                    }
                if (best < 0) break;
                const int k = kids[best];
                kids[best] = B.nodes[k].left;
                kids.push_back(B.nodes[k].right);
            }
            w.nchild = (int)kids.size();
            for (int i = 0; i < w.nchild; ++i) w.child_bnode[i] = kids[i];
            return w;
        };
        // BFS to assign wide node indices (children of a node are contiguous-ish, root = 0)
        std::vector<int> queue{root};
        std::vector<int> depth_of{0};
        std::vector<WNode> wides;
        std::vector<int> wdepth;
        size_t head = 0;
        while (head < queue.size()) {
            const int bn = queue[head];
            const int dep = depth_of[head];
            ++head;
            WNode w = make_wide(bn);
            wides.push_back(w);
            wdepth.push_back(dep);
            for (int i = 0; i < w.nchild; ++i)
                if (!B.nodes[w.child_bnode[i]].leaf()) { queue.push_back(w.child_bnode[i]); depth_of.push_back(dep + 1); }
        }
        // map binary internal node -> wide index (BFS order)
        std::vector<int> widx(B.nodes.size(), -1);
        for (size_t i = 0; i < queue.size(); ++i) widx[queue[i]] = (int)i;

        // leaf prims: emit in DFS order of wide nodes so leaves are contiguous in memory
        const int NW = D == 2 ? kNodeWords2 : kNodeWords3;
        const int PW = D == 2 ? kPrimWords2 : kPrimWords3;
        S.nodes.assign(wides.size() * NW, 0.f);
        S.prim_geom.assign((size_t)n * PW, 0.f);
        S.prim_id.assign(n, 0);
        uint32_t next_prim = 0;
        std::vector<int> need(wides.size(), 0);
        uint32_t maxdepth = 0;
        for (size_t w = 0; w < wides.size(); ++w) maxdepth = std::max<uint32_t>(maxdepth, wdepth[w]);
        for (size_t w = 0; w < wides.size(); ++w) {
            float *nd = &S.nodes[w * NW];
            uint32_t mask = 0;
            for (int c = 0; c < 4; ++c) {
                int32_t child = 0;
                Box b;
                if (c < wides[w].nchild) {
                    const int bn = wides[w].child_bnode[c];
                    b = B.nodes[bn].box;
                    mask |= 1u << c;
                    if (B.nodes[bn].leaf()) {
                        const int cnt = B.nodes[bn].count;
                        // split oversized leaves is impossible here: count <= kMaxLeaf unless all
                        // centroids coincide; such leaves are emitted in chunks of 16 below.
                        const uint32_t first = next_prim;
                        for (int i = 0; i < cnt; ++i) {
                            const uint32_t e = B.order[B.nodes[bn].start + i];
                            S.prim_id[next_prim] = e;
                            float *g = &S.prim_geom[(size_t)next_prim * PW];
                            if (D == 2) {
                                g[0] = V(e, 0, 0); g[1] = V(e, 0, 1);
                                g[2] = V(e, 0, 0) - V(e, 1, 0); g[3] = V(e, 0, 1) - V(e, 1, 1);
                            } else {
                                float e1[3], e2[3];
                                for (int k = 0; k < 3; ++k) {
                                    g[k] = V(e, 0, k);
                                    e1[k] = V(e, 0, k) - V(e, 1, k);
                                    e2[k] = V(e, 0, k) - V(e, 2, k);
                                    g[3 + k] = e1[k];
                                    g[6 + k] = e2[k];
                                }
                                std::memcpy(&g[9], &e, 4);   // element id; N = e1 x e2 is formed in the test
                                g[10] = 0.f;
                                g[11] = 0.f;
                            }
                            ++next_prim;
                        }
                        if (cnt > 16 || cnt < 1) return "leaf with more than 16 primitives (degenerate geometry)";
                        child = ~(int32_t)((first << 4) | (uint32_t)(cnt - 1));
                    } else {
                        child = widx[bn];
                    }
                } else {
                    b.reset();
                    for (int k = 0; k < 3; ++k) { b.lo[k] = 0.f; b.hi[k] = 0.f; }
                }
                if (D == 2) {
                    nd[0 + c] = b.lo[0]; nd[4 + c] = b.lo[1]; nd[8 + c] = b.hi[0]; nd[12 + c] = b.hi[1];
                    std::memcpy(&nd[16 + c], &child, 4);
                } else {
                    nd[0 + c] = b.lo[0]; nd[4 + c] = b.lo[1]; nd[8 + c] = b.lo[2];
                    nd[12 + c] = b.hi[0]; nd[16 + c] = b.hi[1]; nd[20 + c] = b.hi[2];
                    std::memcpy(&nd[24 + c], &child, 4);
                }
            }
            std::memcpy(&nd[NW - 4], &mask, 4);
        }
        if (next_prim != n) return "internal error: primitive count mismatch";
        quantize_nodes(D, S.nodes, S.qnodes);
        // stack bound for DFS "continue with one hit child, push the others" (leaves are
        // pushed too): need(w) = (#children - 1) + max need(internal child); bottom-up.  The
        // traversal pushes all hit children before popping one, one entry more: stack_need
        // below is need + 1.
        for (size_t w = wides.size(); w-- > 0;) {
            int m = 0;
            for (int c = 0; c < wides[w].nchild; ++c) {
                const int bn = wides[w].child_bnode[c];
                if (!B.nodes[bn].leaf()) m = std::max(m, need[widx[bn]]);
            }
            need[w] = std::max(0, wides[w].nchild - 1) + m;
        }
        S.n_nodes = (uint32_t)wides.size();
        S.depth = maxdepth + 1;
        S.stack_need = (uint32_t)need[0] + 1;
        if (S.stack_need <= 48) break;
        if (attempt == 1) return "BVH stack bound exceeds 48 entries";
    }
    return "";
}

// Child bounds in 8-bit steps of the node box: lower bounds rounded down, upper bounds up,
// checked in extended precision so that the dequantized box lo + q s (exact reals) contains
// the float box.  The line traversal evaluates the slab parameters as
// q (s inv) + (lo inv - o inv) with about three roundings, well inside the 4e-6 |coordinate|
// inflation of the primitive boxes.
void quantize_nodes(int dim, const std::vector<float> &nodes, std::vector<uint32_t> &qnodes) {
    const int NW = dim == 2 ? kNodeWords2 : kNodeWords3;
    const int QW = dim == 2 ? kQNodeWords2 : kQNodeWords3;
    const size_t nn = nodes.size() / NW;
    qnodes.assign(nn * QW, 0u);
    auto bits = [](float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; };
    for (size_t w = 0; w < nn; ++w) {
        const float *nd = &nodes[w * NW];
        uint32_t mask;
        std::memcpy(&mask, &nd[NW - 4], 4);
        const int CH = dim == 2 ? 16 : 24;   // word offset of the child entries
        float lo[3] = {0.f, 0.f, 0.f}, s[3] = {0.f, 0.f, 0.f};
        uint32_t qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
        for (int k = 0; k < dim; ++k) {
            float a = 0.f, b = 0.f;
            bool any = false;
            for (int c = 0; c < 4; ++c) {
                if (!((mask >> c) & 1u)) continue;
                const float l = nd[4 * k + c], h = nd[4 * (dim + k) + c];
                a = any ? std::min(a, l) : l;
                b = any ? std::max(b, h) : h;
                any = true;
            }
            // step: smallest float with a + 255 s >= b
            float st = (float)(((double)b - (double)a) / 255.0);
            while ((long double)a + 255.0L * (long double)st < (long double)b) st = std::nextafter(st, INFINITY);
            lo[k] = a;
            s[k] = st;
            for (int c = 0; c < 4; ++c) {
                uint32_t ql = 0u, qh = 0u;
                if ((mask >> c) & 1u) {
                    const long double l = nd[4 * k + c], h = nd[4 * (dim + k) + c];
                    if (st > 0.f) {
                        long double fl = std::floor((l - (long double)a) / (long double)st);
                        long double ch = std::ceil((h - (long double)a) / (long double)st);
                        fl = std::min(std::max(fl, 0.0L), 255.0L);
                        ch = std::min(std::max(ch, 0.0L), 255.0L);
                        while (fl > 0.0L && (long double)a + fl * (long double)st > l) fl -= 1.0L;
                        while (ch < 255.0L && (long double)a + ch * (long double)st < h) ch += 1.0L;
                        ql = (uint32_t)fl;
                        qh = (uint32_t)ch;
                    }
                }
                qlo[k] |= ql << (8 * c);
                qhi[k] |= qh << (8 * c);
            }
        }
        uint32_t *q = &qnodes[w * QW];
        if (dim == 2) {
            q[0] = bits(lo[0]); q[1] = bits(lo[1]); q[2] = bits(s[0]); q[3] = bits(s[1]);
            q[4] = qlo[0]; q[5] = qlo[1]; q[6] = qhi[0]; q[7] = qhi[1];
        } else {
            q[0] = bits(lo[0]); q[1] = bits(lo[1]); q[2] = bits(lo[2]); q[3] = bits(s[0]);
            q[4] = bits(s[1]); q[5] = bits(s[2]); q[6] = qlo[0]; q[7] = qlo[1];
            q[8] = qlo[2]; q[9] = qhi[0]; q[10] = qhi[1]; q[11] = qhi[2];
        }
        for (int c = 0; c < 4; ++c) {
            uint32_t ch;
            std::memcpy(&ch, &nd[CH + c], 4);
            q[QW - 4 + c] = ((mask >> c) & 1u) ? ch : 0x80000000u;
        }
    }
}

bool inside_grid_marks(const HostScene &S, uint32_t res, float lo[3], float h[3], float inv[3],
                       std::vector<uint8_t> &flag) {
    const int D = (int)S.dim;
    const int AW = D == 2 ? kAttrWords2 : kAttrWords3;
    float hi[3];
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; h[k] = 0.f; inv[k] = 0.f; }
    for (uint32_t e = 0; e < S.n; ++e)
        for (int j = 0; j < D; ++j)
            for (int k = 0; k < D; ++k) {
                const float x = S.elem_attr[(size_t)e * AW + j * D + k];
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
    for (int k = 0; k < D; ++k)
        if (!(hi[k] > lo[k])) return false;
    for (int k = 0; k < D; ++k) {
        h[k] = (hi[k] - lo[k]) / (float)res;
        inv[k] = (float)res / (hi[k] - lo[k]);
    }
    for (int k = D; k < 3; ++k) lo[k] = 0.f;
    size_t n = 1;
    for (int k = 0; k < D; ++k) n *= res;
    flag.assign(n, 255u);
    for (uint32_t e = 0; e < S.n; ++e) {
        int i0[3] = {0, 0, 0}, i1[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            float bl = S.elem_attr[(size_t)e * AW + k], bh = bl;
            for (int j = 1; j < D; ++j) {
                const float x = S.elem_attr[(size_t)e * AW + j * D + k];
                bl = std::min(bl, x);
                bh = std::max(bh, x);
            }
            const float pad = std::max(std::abs(bl), std::abs(bh)) * 4.0e-6f + 1.0e-7f;
            bl -= pad;
            bh += pad;
            i0[k] = std::max(0, (int)std::floor((bl - lo[k]) * inv[k]) - 1);
            i1[k] = std::min((int)res - 1, (int)std::floor((bh - lo[k]) * inv[k]) + 1);
        }
        for (int z = i0[2]; z <= i1[2]; ++z)
            for (int y = i0[1]; y <= i1[1]; ++y)
                for (int x = i0[0]; x <= i1[0]; ++x) flag[(size_t)x + ((size_t)y + (size_t)z * res) * res] = 2u;
    }
    return true;
}

}  // namespace wobhost
