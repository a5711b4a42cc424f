// Host-side scene construction for the MI355X walk-on-boundary engine.
//
// Replaces SceneHost's constructor (wob/wob/scene.cuh:264-299), which builds an lbvh binary
// BVH over the elements and the area / boundary-value CDFs in BVH leaf order.  Here:
//   * a binned-SAH binary BVH is built top-down and collapsed into a 4-wide BVH whose nodes
//     hold their four child boxes as structure-of-arrays (one 128-byte line per 3D node),
//     plus a copy with the child boxes quantized to 8 bits per bound (64 bytes per 3D node)
//     for line traversal;
//   * primitive geometry is re-laid out in leaf order in the form the intersection test
//     consumes (2D: a, a-b; 3D: a, a-b, a-c and the element id);
//   * element attributes stay in the caller's order, as do the CDFs (FP32 Kahan running
//     sums, scene.cuh:271-298), each with a guide table for O(1) inverse-CDF sampling that
//     returns exactly what binary_search (scene.cuh:30-43) returns.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wob_c.h"

namespace wobhost {

// Node layouts (float words).  Child boxes are SoA; `child` words are int32: >= 0 internal
// node index, < 0 leaf ~((first_prim << 4) | (count - 1)); `meta[0]` is the valid-child mask.
constexpr int kNodeWords2 = 24;   // lox[4] loy[4] hix[4] hiy[4] child[4] meta[4]       (96 B)
constexpr int kNodeWords3 = 32;   // lox[4] loy[4] loz[4] hix[4] hiy[4] hiz[4] child[4] meta[4] (128 B)
// Quantized nodes for line traversal (32-bit words): the node box's lower corner lo and a
// step s per axis (floats), then per axis one word of four 8-bit child lower bounds and one of
// four upper bounds (byte c = child c), then the four child words (an invalid child is
// 0x80000000).  Child c's box along axis k is [lo_k + qlo s_k, lo_k + qhi s_k], built to
// contain the float box exactly.
constexpr int kQNodeWords2 = 12;  // lo.xy s.xy | qlo.x qlo.y qhi.x qhi.y | child[4]           (48 B)
constexpr int kQNodeWords3 = 16;  // lo.xyz s.x | s.yz qlo.x qlo.y | qlo.z qhi.xyz | child[4] (64 B)
constexpr int kPrimWords2 = 4;    // a.x a.y e.x e.y                 (e = a - b)
constexpr int kPrimWords3 = 12;   // a.xyz e1.xyz e2.xyz id 0 0      (e1 = a - b, e2 = a - c, id = element id)
constexpr int kAttrWords2 = 12;   // a.xy b.xy | n.xy bv0 bv1 | al0 al1 type area
constexpr int kAttrWords3 = 20;   // a.xyz b.x | b.yz c.xy | c.z n.xyz | bv0 bv1 bv2 type | al0 al1 al2 area
// 3D element geometry again, 64 bytes per element: the uniform boundary samples (RIS
// candidates) read only vertices and normal, which then lie in one 64-byte memory block
// (in the 80-byte attribute records they straddle two half the time).
constexpr int kGeoWords3 = 16;    // a.xyz b.x | b.yz c.xy | c.z n.xyz | 4 pad
constexpr int kMaxLeaf = 8;

struct HostScene {
    uint32_t dim = 0, n = 0;
    std::vector<float> nodes;          // n_nodes * kNodeWords
    std::vector<uint32_t> qnodes;      // n_nodes * kQNodeWords
    std::vector<float> prim_geom;      // n * kPrimWords (leaf order)
    std::vector<uint32_t> prim_id;     // leaf order -> element id
    std::vector<float> elem_attr;      // n * kAttrWords (element order)
    std::vector<float> elem_geo;       // 3D: n * kGeoWords3 (element order)
    std::vector<uint32_t> area_fused_guide;   // 3D: the element of each fused area-guide record
    std::vector<float> area_cdf, bv_cdf;
    std::vector<uint32_t> area_guide, bv_guide;
    float total_area = 0, total_bv = 0;
    uint32_t n_nodes = 0, depth = 0, stack_need = 0;
    uint32_t bvh_width = 4;     // the most children a node has (2..4)
};

// Returns an empty string on success, otherwise an error message.  Without with_bvh only the
// element attributes and the CDFs are built (the BVH arrays stay empty: the device build
// supplies them).  width: the most children a node takes (2..4; the layout stays 4-wide).
std::string build_scene(const wob_scene_desc &desc, HostScene &out, bool with_bvh = true, int width = 4);

// Quantized copy of float nodes (kNodeWords{2,3} each) in the kQNodeWords{2,3} layout.
void quantize_nodes(int dim, const std::vector<float> &nodes, std::vector<uint32_t> &qnodes);

// FP32 Kahan CDF (scene.cuh:271-298) and its guide table.
void build_cdf(const std::vector<float> &weights, std::vector<float> &cdf, float &total,
               std::vector<uint32_t> &guide);

// Guide table of a normalised CDF: guide[k] = first i with cdf[i] >= k / entries (entries
// defaults to n).
void build_guide(const std::vector<float> &cdf, std::vector<uint32_t> &guide, size_t entries = 0);

// Fused area-guide records of a 3D scene: entries = kFusedPerElement * n; record k holds
// (i = guide[k], cdf[i - 1], cdf[i], cdf[i + 1]) and then the geometry (the first three rows
// of its kGeoWords3 record) of i or i + 1, whichever covers more of the entry's interval
// (flagged in bit 31 of the index word), 64 bytes in all.  A uniform boundary sample whose
// search ends at that element (about seven times in eight with two entries per element)
// reads this one record.  Two entries per element (25.6 MB at 200k triangles) walked faster
// than four (51 MB, one sample in sixteen reading a second record) and than one.  The host builds the guide (area_fused_guide); the device assembles the records
// from it, the area CDF and the geometry table (launch_fused_records).
constexpr int kFusedPerElement = 2;

// Guide records for the device search: per k four words (i = guide[k], cdf[i - 1], cdf[i],
// cdf[i + 1]) (0 where the index is outside the CDF), so that the usual search that ends
// within one step of guide[k] reads one 16-byte record and no CDF entry.
std::vector<float> guide_records(const std::vector<float> &cdf, const std::vector<uint32_t> &guide);

// The inside test's occupancy grid (volume samples of a masked Gaussian source): res^D cells
// over the vertex box (res = kInsideGridRes2 or kInsideGridRes3), cell (x, y, z) at
// x + res (y + res z).  inside_grid_marks fills the layout (corner lo, cell size h, inv =
// res / extent) and marks with 2 every cell within one cell of an element's inflated box (the
// BVH's inflation rule), 255 the rest (the device classifies those by ray parity).  It returns
// false, and no grid is used, when the vertex box is flat along an axis.
constexpr uint32_t kInsideGridRes2 = 512, kInsideGridRes3 = 64;
bool inside_grid_marks(const HostScene &S, uint32_t res, float lo[3], float h[3], float inv[3],
                       std::vector<uint8_t> &flag);

}  // namespace wobhost
