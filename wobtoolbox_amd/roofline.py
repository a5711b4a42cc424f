"""Algorithmic work per walk path (a fixed model, independent of the implementation).

Counts are FP32 operations (add, mul, compare, min/max = 1; FMA = 2; sqrt, rcp, log,
sin/cos = 1) and compulsory record bytes for an ideal binary BVH descent:
  per line query with k hits: 2 L k node tests (F_node ops, B_node bytes each) and 2 k
  primitive tests (F_prim, B_prim), plus F_q of per-query work (direction, pick, inv_pdf);
  per bounce F_w of weight / kernel evaluation; per CDF sample L + 15 compares and
  interpolation, 4 L + B_prim bytes (binary descent), or B_GUIDE + B_prim bytes for the RIS
  candidates, whose sampler is an O(1) guide-table lookup.
L = ceil(log2 N_elements).  The mean hit counts k were measured once with the CPU oracle on
20,000 random lines per origin type (seed 0) and are frozen here:
  circle_dirichlet_2d    k = 2.000 from interior points, 1.008 from boundary points
  icosphere_neumann_3d   k = 2.012 from interior points, 1.116 from boundary points
"""
import math

MEASURED_K = {
    "circle_dirichlet_2d": (2.000, 1.008),
    "icosphere_neumann_3d": (2.012, 1.116),
}

PARAMS = {2: dict(F_node=14, B_node=32, F_prim=24, B_prim=44, F_q=20, F_w=30),
          3: dict(F_node=20, B_node=32, F_prim=60, B_prim=76, F_q=50, F_w=35)}


def per_path(dim, n_elements, path_length, k_interior, k_boundary, cdf_samples_per_bounce=0):
    """(flop, bytes) per path for line-sampled walks: one domain query + path_length bounce queries."""
    p = PARAMS[dim]
    L = math.ceil(math.log2(n_elements))
    sum_k = k_interior + path_length * k_boundary
    queries = 1 + path_length
    samples = queries * cdf_samples_per_bounce
    flop = (2 * L * sum_k * p["F_node"] + 2 * sum_k * p["F_prim"] + queries * p["F_q"] + queries * p["F_w"]
            + samples * (L + 15))
    byts = 2 * L * sum_k * p["B_node"] + 2 * sum_k * p["B_prim"] + samples * (4 * L + p["B_prim"])
    return flop, byts


# Bytes of one guide-table record of an O(1) inverse-CDF sample: the element index and the CDF
# values around it (i, cdf[i-1], cdf[i], cdf[i+1]).
B_GUIDE = 16


def ris_per_path(dim, n_elements, path_length, candidates, descent=False):
    """(flop, bytes) per path of a single-layer walk with RIS over `candidates` uniform samples
    per bounce: each candidate is one CDF sample (L + 15 ops) and a kernel evaluation with its
    reservoir update (25 ops); path_length bounces plus the domain sample.  Bytes per
    candidate: the element's record (B_prim) and, for the CDF search, one guide-table record
    (B_GUIDE; the sampler is O(1)) or, with descent=True, the L 4-byte CDF entries of a binary
    descent (the fixed figure of an ideal binary search)."""
    p = PARAMS[dim]
    L = math.ceil(math.log2(n_elements))
    bounces = path_length + 1
    flop = bounces * candidates * (L + 15 + 25)
    byts = bounces * candidates * ((4 * L if descent else B_GUIDE) + p["B_prim"])
    return flop, byts


def workload_model(name, n_elements, path_length, candidates=64):
    if name == "geodesic_mixed_3d":
        return ris_per_path(3, n_elements, path_length, candidates)
    ki, kb = MEASURED_K[name]
    if name == "icosphere_neumann_3d":
        # direct Neumann: one boundary-value CDF sample per query (y1 and one per bounce)
        return per_path(3, n_elements, path_length, ki, kb, cdf_samples_per_bounce=1)
    return per_path(2, n_elements, path_length, ki, kb)


# HBM bytes per walk-kernel launch measured with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE,
# each counter scaled by what it reads for a 512 MiB device copy in the same run (FETCH_SIZE
# 262,161 KiB: x 2.0; WRITE_SIZE 524,288 KiB: x 1.0).  FETCH_SIZE counts L2 misses, served by
# the MALL or HBM.
#   geodesic_mixed_3d (1,000,000 points x 16 paths, single layer, RIS 64, 4 bounces):
#     FETCH_SIZE 257,465,393 KiB, WRITE_SIZE 31,834 KiB per launch (69.2 ms)
#   icosphere_neumann_3d (100,000 points x 256 paths, direct Neumann, 6 bounces):
#     FETCH_SIZE 122,634 KiB, WRITE_SIZE 70,919 KiB per launch
MEASURED_TRAFFIC = {
    "geodesic_mixed_3d": (2.0 * 257465393 + 31834) * 1024,
    "icosphere_neumann_3d": (2.0 * 122634 + 70919) * 1024,
}

# What bounds each workload's walk kernel (rocprofv3 counters): the RIS walk's scattered
# candidate loads keep the texture data unit 98 % busy and read L2 misses at 95 % of the HBM
# peak (0.8 lines of 128 bytes per candidate against 92 algorithmic bytes); the line-sampled
# walk keeps the vector ALU about 90 % busy.
BOUND = {"geodesic_mixed_3d": "hbm", "icosphere_neumann_3d": "mfma", "circle_dirichlet_2d": "mfma"}

MI355X_PEAKS = {"hbm_GBps": 8000.0, "fp32_TFLOPs": 157.3, "bf16_TFLOPs": 2500.0}
