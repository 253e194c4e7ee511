// The octree handle behind adeco_octree and the host/device split of its construction.
#pragma once

#include <cstdint>
#include <vector>

#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "runtime.hpp"

namespace adeco {

constexpr int kMaxOctreeDepth = 20;

struct RootBox {
    float b[6];
};

struct ONode {
    float ll[3];
    float ext;
    int level;
    int parent;
    int first_child;           // -1 when unrefined
    unsigned long long count;  // num_tets_covering_node_estimate (== num_tets_on_proc_estimate[0] on one rank)
    int ix[3];                 // integer position among the 2^level nodes per axis of its level
};

}  // namespace adeco

struct adeco_octree_s {
    adeco_lattice_s* L = nullptr;
    float root_bbox[6];  // global bbox the root is built from
    std::vector<adeco::ONode> nodes;  // reference (DFS) numbering
    std::vector<int32_t> leaves_dfs, leaves_order;
    std::vector<int64_t> leaf_mc_offset;
    std::vector<int32_t> leaf_mc_ids;
    int n_meta = 0;
    // binning: per-leaf lists of lattice tet ids (device), offsets (host copy)
    int binned_decomposition = -1;
    std::vector<int64_t> leaf_tet_offset;
    adeco::DevBuf d_mc_box, d_mc_cells, d_node_box, d_first, d_leaf_index, d_leaf_count, d_leaf_offset, d_pairs;
    adeco::DevBuf d_leaf_map, d_overflow, d_tleaf;  // binning: leaf of every finest-level cell; tets left
                                                    // to the descent; each tet's single leaf
    // leaf grids of the last deposit (device, FP64, leaf-major in DFS leaf order)
    int grid_dims = 0;
    adeco::DevBuf d_grids, d_leaf_geom, d_work, d_tiles, d_counters;
    int64_t last_candidates = 0;            // (tet, candidate) pairs of the last batched deposit
    unsigned long long* d_fallback = nullptr;  // FP64 recomputes of the last mixed-precision deposit

    // (ll, ll + ext) per axis in float, as AdOctreeGeneration::getBBox (AdOctreeGeneration.h:67-73)
    void node_box(int id, float out[6]) const {
        const adeco::ONode& n = nodes[id];
        for (int a = 0; a < 3; ++a) {
            out[2 * a] = n.ll[a];
            out[2 * a + 1] = n.ll[a] + n.ext;
        }
    }

    // The node's box as the exact dyadic subdivision of the root, in FP64: root_ll + ix * h,
    // h = root_ext / 2^level.  Every such bound is exactly representable (a float times a power
    // of two plus a float needs < 53 bits for level <= 20), so leaves share faces bit for bit and
    // tile the root exactly, unlike the float boxes above whose neighbours at different levels can
    // be an ulp apart.
    void node_box_exact(int id, double out[6]) const {
        const adeco::ONode& n = nodes[id];
        const double h = double(nodes[0].ext) / double(1u << n.level);
        for (int a = 0; a < 3; ++a) {
            out[2 * a] = double(nodes[0].ll[a]) + double(n.ix[a]) * h;
            out[2 * a + 1] = double(nodes[0].ll[a]) + double(n.ix[a] + 1) * h;
        }
    }

    // Grid geometry of a leaf: float reference boxes with the reference's scaling, or the exact
    // dyadic boxes with per-axis scaling.
    adeco::GridGeom leaf_geom(int dfs_leaf_index, int dims, int scaling) const {
        const int id = leaves_dfs[dfs_leaf_index];
        if (scaling == adeco::kScalingReference) {
            float b[6];
            node_box(id, b);
            return adeco::make_grid_geom(b, dims, scaling);
        }
        double b[6];
        node_box_exact(id, b);
        return adeco::make_grid_geom_d(b, dims, scaling);
    }
};

namespace adeco {

adeco_status octree_meta_bbox(adeco_lattice_s* L, const int md[3], float* d_box, unsigned* d_cells);
adeco_status octree_root_count(DeviceState& ds, const float* d_box, const unsigned* d_cells, int M, const float box[6],
                               unsigned long long* d_count, unsigned long long* h_count);
adeco_status octree_refine_level(DeviceState& ds, const float* d_box, const unsigned* d_cells,
                                 const std::vector<int32_t>& slot, const std::vector<int32_t>& mc,
                                 const std::vector<float>& child_boxes, std::vector<unsigned char>& flags,
                                 std::vector<unsigned long long>& child_count);
adeco_status octree_bin_tets(adeco_octree_s* T, int decomposition);

// The leaf of every cell of the finest leaf level, for the per-tet binning.
struct LeafMap {
    const int32_t* map;       // F^3 leaf indices, i fastest
    const float* leaf_box;    // per leaf (DFS order), widened as the node boxes
    int F;
    double ll[3];
    double inv_h;
};

// Everything the binning passes need, resolved once per octree and decomposition (node boxes,
// leaf map and work buffers uploaded by bin_plan); the passes then run with every size on the
// device, so they can be captured in a graph and replayed on moved lattice positions.
struct BinPlan {
    LatticeView lv{};
    PackedCorners ct{};
    int64_t n_tets = 0;
    int n_leaves = 0;
    bool use_map = false;
    LeafMap lm{};
    const float* node_box = nullptr;
    const int32_t* first = nullptr;
    const int32_t* leaf_index = nullptr;
    int* leaf_count = nullptr;  // n_leaves counters
    int32_t* over = nullptr;    // tets left to the descent, and their count
    int* n_over = nullptr;
    int32_t* tleaf = nullptr;   // per tet: its single leaf, or a recompute mark
};
adeco_status bin_plan(adeco_octree_s* T, int decomposition, BinPlan& p);
// Count pass: leaf_count[l] = the tets whose AABB strictly overlaps leaf l.
hipError_t bin_count_async(const BinPlan& p, hipStream_t stream);
// Fill pass: the tets of leaf l at d_pairs[d_leaf_offset[l] ...]; entries at or beyond pair_cap
// are not written.
hipError_t bin_fill_async(const BinPlan& p, const int64_t* d_leaf_offset, int32_t* d_pairs, int64_t pair_cap,
                          hipStream_t stream);

}  // namespace adeco
