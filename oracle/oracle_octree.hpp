// Test infrastructure only: the oracle (see oracle_core.hpp for the rules on its use).
//
// Float-faithful restatement of ADECO's one-rank octree construction:
//   MetaMesh meta-cells and their float AABBs      AdTypeDefinitions.h:715-1022
//   get_estimate_num_tets (float estimate -> uint64) AdTypeDefinitions.h:812-825
//   generate_tree DFS refinement                     AdOctreeGeneration.h:76-208
//   distribute_tets_to_children                      AdOctreeGeneration.h:213-258
//   Octree::setRootNode / refineNode / getLeafNodes  AdOctree.h:187-305
//   find_best_node_and_remove_it_from_list           AdDynamicLoadBalancing.h:706-730
// On one rank MPI_Allgather of the node count is the identity.  Compile with -ffp-contract=off.
#pragma once

#include <algorithm>
#include <cstdint>
#include <list>
#include <stack>
#include <vector>

namespace oracle {

struct FBox {
    float mn[3], mx[3];
    float volume() const { return ((mx[0] - mn[0]) * (mx[1] - mn[1])) * (mx[2] - mn[2]); }
};

inline bool fbox_strict(const FBox& a, const FBox& b) {  // AdTypeDefinitions.h:339-345
    return a.mn[0] < b.mx[0] && b.mn[0] < a.mx[0] && a.mn[1] < b.mx[1] && b.mn[1] < a.mx[1] &&
           a.mn[2] < b.mx[2] && b.mn[2] < a.mx[2];
}

struct MetaCells {
    int np[3];
    int md[3];  // meta-cells per axis
    std::vector<FBox> bbox;
    std::vector<uint32_t> cells;  // (dx-1)(dy-1)(dz-1) lattice cubes in the meta-cell

    void range(int64_t id, int lo[3], int d[3]) const {  // AdTypeDefinitions.h:917-953
        const int c[3] = {int(id % md[0]), int((id / md[0]) % md[1]), int(id / (int64_t(md[0]) * md[1]))};
        for (int a = 0; a < 3; ++a) {
            lo[a] = 7 * c[a];
            d[a] = (c[a] + 1 == md[a]) ? (np[a] - lo[a]) : 8;
        }
    }
};

inline MetaCells build_meta_cells(const float* x, const float* y, const float* z, const int np[3]) {
    MetaCells mc;
    for (int a = 0; a < 3; ++a) {
        mc.np[a] = np[a];
        mc.md[a] = (np[a] - 1) / 7 + ((np[a] - 1) % 7 > 0);  // ceil_div, AdUtils.h:182-186
    }
    const int64_t n = int64_t(mc.md[0]) * mc.md[1] * mc.md[2];
    mc.bbox.resize(n);
    mc.cells.resize(n);
    for (int64_t id = 0; id < n; ++id) {
        int lo[3], d[3];
        mc.range(id, lo, d);
        mc.cells[id] = uint32_t(d[0] - 1) * uint32_t(d[1] - 1) * uint32_t(d[2] - 1);
        FBox& b = mc.bbox[id];
        const int64_t v0 = lo[0] + int64_t(np[0]) * (lo[1] + int64_t(np[1]) * lo[2]);
        b.mn[0] = b.mx[0] = x[v0];
        b.mn[1] = b.mx[1] = y[v0];
        b.mn[2] = b.mx[2] = z[v0];
        for (int k = 0; k < d[2]; ++k)
            for (int j = 0; j < d[1]; ++j)
                for (int i = 0; i < d[0]; ++i) {
                    const int64_t v = (lo[0] + i) + int64_t(np[0]) * ((lo[1] + j) + int64_t(np[1]) * (lo[2] + k));
                    const float p[3] = {x[v], y[v], z[v]};
                    for (int a = 0; a < 3; ++a) {
                        b.mn[a] = std::min(b.mn[a], p[a]);
                        b.mx[a] = std::max(b.mx[a], p[a]);
                    }
                }
    }
    return mc;
}

// est = uint64( float(6*cells) * vol(box ∩ bbox_m) / vol(bbox_m) ); no strict overlap -> the
// default (all-zero) box, volume 0.
inline uint64_t estimate(const MetaCells& mc, const FBox& box, int64_t m) {
    const FBox& b = mc.bbox[m];
    FBox it;
    for (int a = 0; a < 3; ++a) it.mn[a] = it.mx[a] = 0.f;
    if (fbox_strict(box, b))
        for (int a = 0; a < 3; ++a) {
            it.mn[a] = (box.mn[a] < b.mn[a]) ? b.mn[a] : box.mn[a];
            it.mx[a] = (box.mx[a] < b.mx[a]) ? box.mx[a] : b.mx[a];
        }
    const float e = float(uint64_t(6) * uint64_t(mc.cells[m])) * it.volume() / b.volume();
    return uint64_t(e);
}

struct ONode {
    float ll[3];
    float ext;
    int level;
    int parent;
    int first_child = -1;
    uint64_t count = 0;        // num_tets_covering_node_estimate
    uint64_t local_count = 0;  // num_tets_on_proc_estimate[0]
    std::vector<int64_t> meta;

    FBox box() const {  // AdOctreeGeneration.h:67-73
        FBox b;
        for (int a = 0; a < 3; ++a) {
            b.mn[a] = ll[a];
            b.mx[a] = ll[a] + ext;
        }
        return b;
    }
};

struct OTree {
    FBox root_box;  // the global bbox the root is built from
    std::vector<ONode> nodes;
    std::vector<int> leaves_dfs, leaves_order;
};

inline FBox lattice_bbox(const float* x, const float* y, const float* z, int64_t n) {
    FBox b;  // AdPointDataReaders.h:90-151 (first point, then std::min / std::max)
    b.mn[0] = b.mx[0] = x[0];
    b.mn[1] = b.mx[1] = y[0];
    b.mn[2] = b.mx[2] = z[0];
    for (int64_t i = 0; i < n; ++i) {
        b.mn[0] = std::min(x[i], b.mn[0]); b.mn[1] = std::min(y[i], b.mn[1]); b.mn[2] = std::min(z[i], b.mn[2]);
        b.mx[0] = std::max(x[i], b.mx[0]); b.mx[1] = std::max(y[i], b.mx[1]); b.mx[2] = std::max(z[i], b.mx[2]);
    }
    return b;
}

inline void refine(OTree& t, int id) {  // AdOctree.h:201-237 + AdOctreeGeneration.h:213-258
    const int first = int(t.nodes.size());
    t.nodes[id].first_child = first;
    const float child_ext = float(double(t.nodes[id].ext) / 2.);
    const float ll[3] = {t.nodes[id].ll[0], t.nodes[id].ll[1], t.nodes[id].ll[2]};
    const int level = t.nodes[id].level;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                ONode c;
                c.ll[0] = ll[0] + float(i) * child_ext;
                c.ll[1] = ll[1] + float(j) * child_ext;
                c.ll[2] = ll[2] + float(k) * child_ext;
                c.ext = child_ext;
                c.level = level + 1;
                c.parent = id;
                t.nodes.push_back(c);
            }
}

inline OTree build_octree(const MetaCells& mc, const FBox& root_box, const FBox* roi,
                          uint64_t max_tets, int max_level) {
    OTree t;
    t.root_box = root_box;
    const FBox region = roi ? *roi : root_box;  // invalid ROI -> whole domain
    FBox eff = region;
    if (region.mn[0] > region.mx[0] || region.mn[1] > region.mx[1] || region.mn[2] > region.mx[2]) eff = root_box;
    ONode root;
    for (int a = 0; a < 3; ++a) root.ll[a] = root_box.mn[a];
    root.ext = std::max(std::max(root_box.mx[0] - root_box.mn[0], root_box.mx[1] - root_box.mn[1]),
                        root_box.mx[2] - root_box.mn[2]);
    root.level = 0;
    root.parent = -1;
    const int64_t nmeta = int64_t(mc.bbox.size());
    root.meta.resize(nmeta);
    for (int64_t m = 0; m < nmeta; ++m) root.meta[m] = m;
    t.nodes.push_back(root);
    std::stack<int> st;
    st.push(0);
    while (!st.empty()) {
        const int id = st.top();
        st.pop();
        const FBox box = t.nodes[id].box();
        uint64_t count = 0;
        for (int64_t m : t.nodes[id].meta) count += estimate(mc, box, m);
        t.nodes[id].count = count;
        t.nodes[id].local_count = count;
        if (fbox_strict(eff, box) && count > 0 && count > max_tets && t.nodes[id].level < max_level) {
            refine(t, id);
            const int first = t.nodes[id].first_child;
            FBox cbox[8];
            for (int c = 0; c < 8; ++c) cbox[c] = t.nodes[first + c].box();
            for (int64_t m : t.nodes[id].meta)
                for (int c = 0; c < 8; ++c)
                    if (fbox_strict(mc.bbox[m], cbox[c])) {
                        t.nodes[first + c].meta.push_back(m);
                        t.nodes[first + c].local_count += estimate(mc, cbox[c], m);
                    }
            std::vector<int64_t>().swap(t.nodes[id].meta);
            for (int c = 0; c < 8; ++c) st.push(first + c);
        }
    }
    std::stack<int> s2;  // AdOctree.h:265-305
    s2.push(0);
    while (!s2.empty()) {
        const int id = s2.top();
        s2.pop();
        if (t.nodes[id].first_child > 0)
            for (int c = 0; c < 8; ++c) s2.push(t.nodes[id].first_child + c);
        else
            t.leaves_dfs.push_back(id);
    }
    std::list<int> lst(t.leaves_dfs.begin(), t.leaves_dfs.end());  // AdDynamicLoadBalancing.h:706-730
    while (!lst.empty()) {
        auto best = lst.begin();
        uint64_t bc = t.nodes[*best].local_count;
        for (auto it = std::next(lst.begin()); it != lst.end(); ++it)
            if (t.nodes[*it].local_count > bc) {
                bc = t.nodes[*it].local_count;
                best = it;
            }
        t.leaves_order.push_back(*best);
        lst.erase(best);
    }
    return t;
}

}  // namespace oracle
