// Octree construction and the single-rank leaf deposit loop (C ABI part 4).
//
// adeco_octree_build follows generate_tree (AdOctreeGeneration.h:76-208) on one rank: root =
// (bbox.min, max extent), node count = sum of meta-cell estimates, refine iff the node strictly
// meets the ROI, count > 0, count > max_tets and level < max_level; children c = i + 2j + 4k at
// ll + (i,j,k) * ext/2 (AdOctree.h:201-237).  The float estimates and overlap tests run on the GPU
// (octree.hip); this file keeps the bookkeeping and reproduces the reference's node numbering,
// DFS leaf list (AdOctree.h:265-305) and leaf processing order
// (find_best_node_and_remove_it_from_list, AdDynamicLoadBalancing.h:706-730).
//
// adeco_deposit_leaves replaces deposit_leaf_nodes (AdDynamicLoadBalancing.h:983-1323) for one
// rank: every leaf gets a dims^3 FP64 grid in HBM holding the mass of all lattice tets binned to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <list>
#include <stack>
#include <vector>

#include "../../include/adeco.h"
#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "batched_impl.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

using namespace adeco;

namespace {

struct BNode {
    float ll[3];
    float ext;
    int level;
    int ix[3];
    int bfs_parent;
    int child0 = -1;  // BFS index of child 0 when refined
    unsigned long long count = 0;
    std::vector<int32_t> list;
};

bool strict6(const float* a, const float* b) {
    return a[0] < b[1] && b[0] < a[1] && a[2] < b[3] && b[2] < a[3] && a[4] < b[5] && b[4] < a[5];
}

void box_of(const BNode& n, float out[6]) {
    for (int a = 0; a < 3; ++a) {
        out[2 * a] = n.ll[a];
        out[2 * a + 1] = n.ll[a] + n.ext;
    }
}

adeco_status build_tree(adeco_octree_s* T, const float* roi6, uint64_t max_tets_param, int max_level) {
    adeco_lattice_s* L = T->L;
    DeviceState& ds = *L->ds;
    adeco_status st = ADECO_OK;
    if (L->has_domain)
        std::memcpy(T->root_bbox, L->domain, sizeof(T->root_bbox));
    else if ((st = lattice_bbox_device(L, T->root_bbox)))
        return st;
    int md[3];
    for (int a = 0; a < 3; ++a) md[a] = (L->np[a] - 1) / 7 + ((L->np[a] - 1) % 7 > 0);  // ceil_div(np-1, 7)
    const int M = md[0] * md[1] * md[2];
    T->n_meta = M;
    if ((st = T->d_mc_box.reserve(size_t(M) * 24)) || (st = T->d_mc_cells.reserve(size_t(M) * 4))) return st;
    if ((st = octree_meta_bbox(L, md, T->d_mc_box.as<float>(), T->d_mc_cells.as<unsigned>()))) return st;

    const uint64_t num_tets = 6ull * uint64_t(L->np[0] - 1) * uint64_t(L->np[1] - 1) * uint64_t(L->np[2] - 1);
    const uint64_t max_tets = std::min(num_tets, max_tets_param);  // AdDepositWorkFlow.h:169 (num_procs = 1)
    const float* gb = T->root_bbox;
    float eff[6];
    std::memcpy(eff, roi6 ? roi6 : gb, sizeof(eff));
    if (eff[0] > eff[1] || eff[2] > eff[3] || eff[4] > eff[5]) std::memcpy(eff, gb, sizeof(eff));  // invalid ROI

    std::vector<BNode> bn;
    bn.reserve(1024);
    BNode root;
    for (int a = 0; a < 3; ++a) root.ll[a] = gb[2 * a];
    root.ext = std::max(std::max(gb[1] - gb[0], gb[3] - gb[2]), gb[5] - gb[4]);
    root.level = 0;
    root.ix[0] = root.ix[1] = root.ix[2] = 0;
    root.bfs_parent = -1;
    root.list.resize(M);
    for (int m = 0; m < M; ++m) root.list[m] = m;
    float rbox[6];
    box_of(root, rbox);
    DevBuf d_cnt;
    if ((st = d_cnt.reserve(8))) return st;
    if ((st = octree_root_count(ds, T->d_mc_box.as<float>(), T->d_mc_cells.as<unsigned>(), M, rbox,
                                d_cnt.as<unsigned long long>(), &root.count)))
        return st;
    bn.push_back(std::move(root));

    size_t lvl_begin = 0, lvl_end = 1;
    while (lvl_begin < lvl_end) {
        std::vector<int> refined;
        for (size_t i = lvl_begin; i < lvl_end; ++i) {
            float box[6];
            box_of(bn[i], box);
            const BNode& n = bn[i];
            if (strict6(eff, box) && n.count > 0 && n.count > max_tets && n.level < max_level) refined.push_back(int(i));
        }
        if (refined.empty()) break;
        std::vector<int32_t> slot, mc;
        std::vector<float> cboxes(size_t(48) * refined.size());
        std::vector<BNode> kids(8 * refined.size());
        for (size_t s = 0; s < refined.size(); ++s) {
            const BNode& p = bn[refined[s]];
            const float child_ext = float(double(p.ext) / 2.);  // AdOctree.h:213
            for (int k = 0; k < 2; ++k)
                for (int j = 0; j < 2; ++j)
                    for (int i = 0; i < 2; ++i) {
                        BNode& c = kids[8 * s + i + 2 * j + 4 * k];
                        c.ll[0] = p.ll[0] + float(i) * child_ext;
                        c.ll[1] = p.ll[1] + float(j) * child_ext;
                        c.ll[2] = p.ll[2] + float(k) * child_ext;
                        c.ext = child_ext;
                        c.level = p.level + 1;
                        c.ix[0] = 2 * p.ix[0] + i;
                        c.ix[1] = 2 * p.ix[1] + j;
                        c.ix[2] = 2 * p.ix[2] + k;
                        c.bfs_parent = refined[s];
                        box_of(c, &cboxes[6 * (8 * s + i + 2 * j + 4 * k)]);
                    }
            for (int32_t m : p.list) {
                slot.push_back(int32_t(s));
                mc.push_back(m);
            }
        }
        std::vector<unsigned char> flags;
        std::vector<unsigned long long> ccount;
        if ((st = octree_refine_level(ds, T->d_mc_box.as<float>(), T->d_mc_cells.as<unsigned>(), slot, mc, cboxes,
                                      flags, ccount)))
            return st;
        size_t p = 0;
        for (size_t s = 0; s < refined.size(); ++s) {
            const size_t len = bn[refined[s]].list.size();
            for (int c = 0; c < 8; ++c) {
                BNode& kc = kids[8 * s + c];
                kc.count = ccount[8 * s + c];
                for (size_t q = 0; q < len; ++q)
                    if (flags[p + q] & (1u << c)) kc.list.push_back(mc[p + q]);
            }
            p += len;
        }
        const size_t base = bn.size();
        for (size_t s = 0; s < refined.size(); ++s) {
            bn[refined[s]].child0 = int(base + 8 * s);
            std::vector<int32_t>().swap(bn[refined[s]].list);
        }
        for (auto& k : kids) bn.push_back(std::move(k));
        lvl_begin = base;
        lvl_end = bn.size();
        if (bn.size() > size_t(1) << 26) return fail(ADECO_EOVERFLOW, "octree_build(): too many nodes");
    }

    // replay the reference's DFS: ids in creation order (AdOctreeGeneration.h:103-197)
    std::vector<int> dfs_of(bn.size(), -1);
    T->nodes.clear();
    T->nodes.reserve(bn.size());
    auto push_node = [&](int b, int parent_dfs) {
        ONode o;
        std::memcpy(o.ll, bn[b].ll, sizeof(o.ll));
        o.ext = bn[b].ext;
        o.level = bn[b].level;
        o.parent = parent_dfs;
        o.first_child = -1;
        o.count = bn[b].count;
        std::memcpy(o.ix, bn[b].ix, sizeof(o.ix));
        dfs_of[b] = int(T->nodes.size());
        T->nodes.push_back(o);
    };
    push_node(0, -1);
    std::stack<int> stk;
    stk.push(0);
    while (!stk.empty()) {
        const int b = stk.top();
        stk.pop();
        if (bn[b].child0 < 0) continue;
        const int first = int(T->nodes.size());
        T->nodes[dfs_of[b]].first_child = first;
        for (int c = 0; c < 8; ++c) push_node(bn[b].child0 + c, dfs_of[b]);
        for (int c = 0; c < 8; ++c) stk.push(bn[b].child0 + c);
    }
    std::vector<int> bfs_of(bn.size());
    for (size_t b = 0; b < bn.size(); ++b) bfs_of[dfs_of[b]] = int(b);
    // leaves (AdOctree.h:265-305)
    T->leaves_dfs.clear();
    std::stack<int> s2;
    s2.push(0);
    while (!s2.empty()) {
        const int id = s2.top();
        s2.pop();
        if (T->nodes[id].first_child > 0)
            for (int c = 0; c < 8; ++c) s2.push(T->nodes[id].first_child + c);
        else
            T->leaves_dfs.push_back(id);
    }
    std::list<int> lst(T->leaves_dfs.begin(), T->leaves_dfs.end());
    T->leaves_order.clear();
    while (!lst.empty()) {
        auto best = lst.begin();
        unsigned long long bc = T->nodes[*best].count;
        for (auto it = std::next(lst.begin()); it != lst.end(); ++it)
            if (T->nodes[*it].count > bc) {
                bc = T->nodes[*it].count;
                best = it;
            }
        T->leaves_order.push_back(*best);
        lst.erase(best);
    }
    T->leaf_mc_offset.assign(1, 0);
    T->leaf_mc_ids.clear();
    for (int id : T->leaves_dfs) {
        const auto& l = bn[bfs_of[id]].list;
        T->leaf_mc_ids.insert(T->leaf_mc_ids.end(), l.begin(), l.end());
        T->leaf_mc_offset.push_back(int64_t(T->leaf_mc_ids.size()));
    }
    return ADECO_OK;
}

adeco_status deposit_leaves_plain(adeco_octree_s* T, int dims, int decomposition, double tet_mass) {
    adeco_lattice_s* L = T->L;
    DeviceState& ds = *L->ds;
    const int64_t nl = int64_t(T->leaves_dfs.size());
    const int64_t cells = int64_t(dims) * dims * dims;
    const LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    const int scaling = cell_scaling();
    for (int64_t l = 0; l < nl; ++l) {
        const int64_t b = T->leaf_tet_offset[l], n = T->leaf_tet_offset[l + 1] - b;
        if (!n) continue;
        const GridGeom g = T->leaf_geom(int(l), dims, scaling);
        hipError_t e = launch_deposit_lattice_plain(lv, decomposition, T->d_pairs.as<int32_t>() + b, n, g, tet_mass,
                                                    T->d_grids.as<double>() + l * cells, nullptr, ds.stream);
        if (e != hipSuccess) return hip_fail(e, "leaf deposit launch");
    }
    return ADECO_OK;
}

}  // namespace

namespace adeco {
adeco_status deposit_leaves_batched(adeco_octree_s* T, int dims, int decomposition, double tet_mass, int precision);
}

extern "C" {

adeco_status adeco_octree_build(adeco_lattice L, const float* roi6, uint64_t max_tets_per_node, int max_level,
                                adeco_octree* out) {
    if (!L || !out) return fail(ADECO_EINVAL, "octree_build(): null argument");
    *out = nullptr;
    if (max_level < 0 || max_level > kMaxOctreeDepth)
        return fail(ADECO_EINVAL, "octree_build(): max_refinement_level must be in [0, 20]");
    std::lock_guard<std::mutex> lk(L->ds->mutex);
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    auto* T = new adeco_octree_s();
    T->L = L;
    adeco_status st = build_tree(T, roi6, max_tets_per_node, max_level);
    if (st) {
        delete T;
        return st;
    }
    *out = T;
    return ADECO_OK;
}

adeco_status adeco_octree_destroy(adeco_octree T) {
    if (!T) return ADECO_OK;
    std::lock_guard<std::mutex> lk(T->L->ds->mutex);
    (void)hipSetDevice(T->L->device);
    delete T;
    return ADECO_OK;
}

int64_t adeco_octree_num_nodes(adeco_octree T) { return T ? int64_t(T->nodes.size()) : -1; }
int64_t adeco_octree_num_leaves(adeco_octree T) { return T ? int64_t(T->leaves_dfs.size()) : -1; }
int64_t adeco_octree_num_meta_cells(adeco_octree T) { return T ? int64_t(T->n_meta) : -1; }

adeco_status adeco_octree_bbox(adeco_octree T, float* out6) {
    if (!T || !out6) return fail(ADECO_EINVAL, "octree_bbox(): null argument");
    std::memcpy(out6, T->root_bbox, sizeof(T->root_bbox));
    return ADECO_OK;
}

adeco_status adeco_octree_nodes(adeco_octree T, float* ll, float* ext, int32_t* level, int32_t* parent,
                                int32_t* first_child, uint64_t* count) {
    if (!T) return fail(ADECO_EINVAL, "octree_nodes(): null handle");
    for (size_t i = 0; i < T->nodes.size(); ++i) {
        const ONode& n = T->nodes[i];
        if (ll) for (int a = 0; a < 3; ++a) ll[3 * i + a] = n.ll[a];
        if (ext) ext[i] = n.ext;
        if (level) level[i] = n.level;
        if (parent) parent[i] = n.parent;
        if (first_child) first_child[i] = n.first_child;
        if (count) count[i] = n.count;
    }
    return ADECO_OK;
}

int64_t adeco_octree_leaves(adeco_octree T, int32_t* dfs, int32_t* order, int64_t* mc_offsets, int32_t* mc_ids) {
    if (!T) return -1;
    const size_t nl = T->leaves_dfs.size();
    if (dfs) std::memcpy(dfs, T->leaves_dfs.data(), nl * 4);
    if (order) std::memcpy(order, T->leaves_order.data(), nl * 4);
    if (mc_offsets) std::memcpy(mc_offsets, T->leaf_mc_offset.data(), (nl + 1) * 8);
    if (mc_ids) std::memcpy(mc_ids, T->leaf_mc_ids.data(), T->leaf_mc_ids.size() * 4);
    return int64_t(T->leaf_mc_ids.size());
}

adeco_status adeco_octree_meta_bboxes(adeco_octree T, float* out) {
    if (!T || !out) return fail(ADECO_EINVAL, "octree_meta_bboxes(): null argument");
    std::lock_guard<std::mutex> lk(T->L->ds->mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e == hipSuccess) e = hipMemcpy(out, T->d_mc_box.ptr, size_t(T->n_meta) * 24, hipMemcpyDeviceToHost);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "meta bbox download");
}

adeco_status adeco_octree_bin(adeco_octree T, int decomposition, int64_t* leaf_tet_counts) {
    if (!T) return fail(ADECO_EINVAL, "octree_bin(): null handle");
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "octree_bin(): unknown decomposition");
    std::lock_guard<std::mutex> lk(T->L->ds->mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    adeco_status st = octree_bin_tets(T, decomposition);
    if (st) return st;
    if (leaf_tet_counts)
        for (size_t l = 0; l + 1 < T->leaf_tet_offset.size(); ++l)
            leaf_tet_counts[l] = T->leaf_tet_offset[l + 1] - T->leaf_tet_offset[l];
    return ADECO_OK;
}

adeco_status adeco_deposit_leaves(adeco_octree T, int dims, int decomposition, double tet_mass, int mode, int precision) {
    g_last_kernel_ms = 0.0;
    if (!T) return fail(ADECO_EINVAL, "deposit_leaves(): null handle");
    if (dims < 1 || dims > 1024) return fail(ADECO_EINVAL, "deposit_leaves(): dims must be in [1, 1024]");
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "deposit_leaves(): unknown decomposition");
    if (mode != ADECO_MODE_PLAIN && mode != ADECO_MODE_BATCHED) return fail(ADECO_EINVAL, "deposit_leaves(): unknown mode");
    if (precision < 0 || precision > 2) return fail(ADECO_EINVAL, "deposit_leaves(): unknown precision");
    DeviceState& ds = *T->L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    adeco_status st;
    TraceRange phase("adeco:bin_tets");
    if (T->binned_decomposition != decomposition && (st = octree_bin_tets(T, decomposition))) return st;
    phase.next("adeco:deposit_leaves");
    const int64_t nl = int64_t(T->leaves_dfs.size());
    const size_t bytes = size_t(nl) * size_t(dims) * dims * dims * sizeof(double);
    const size_t budget = memory_budget(ds);
    if (bytes > budget) return fail(ADECO_EOVERFLOW, "deposit_leaves(): leaf grids exceed the device memory budget");
    if ((st = T->d_grids.reserve(bytes))) return st;
    T->grid_dims = dims;
    e = launch_fill_f64(T->d_grids.as<double>(), int64_t(bytes / 8), 0.0, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "leaf grid clear");
    g_timings = LastTimings();
    if ((st = timed_begin(ds))) return st;
    const bool plain = mode == ADECO_MODE_PLAIN && precision == ADECO_PREC_FP64;
    if (plain)
        st = deposit_leaves_plain(T, dims, decomposition, tet_mass);
    else
        st = deposit_leaves_batched(T, dims, decomposition, tet_mass, precision);
    if (st) return st;
    if ((st = timed_end(ds))) return st;
    g_timings.total_ms = g_last_kernel_ms;
    g_timings.main_ms = g_last_kernel_ms;
    if (!plain && T->d_fallback) {
        float ms = 0.f;
        unsigned long long cnt[2] = {0, 0};
        e = hipEventElapsedTime(&ms, ds.evk0, ds.evk1);
        if (e == hipSuccess) e = hipMemcpy(cnt, T->d_fallback, 16, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "batched timings");
        g_timings.main_ms = ms;
        g_timings.fallbacks = double(cnt[0]);
        g_timings.partials = double(cnt[1]);
        g_timings.candidates = double(T->last_candidates);
    }
    return ADECO_OK;
}

adeco_status adeco_deposit_leaves_batched(adeco_octree T, int dims, int decomposition, double tet_mass, int precision) {
    return adeco_deposit_leaves(T, dims, decomposition, tet_mass, ADECO_MODE_BATCHED, precision);
}

adeco_status adeco_set_launch_shape(int chunk, int blocks_per_cu) {
    if (chunk < 1 || chunk > 1024 || blocks_per_cu < 1 || blocks_per_cu > 16)
        return fail(ADECO_EINVAL, "set_launch_shape(): chunk must be in 1..1024 and blocks_per_cu in 1..16");
    launch_shape().chunk = chunk;
    launch_shape().blocks_per_cu = blocks_per_cu;
    return ADECO_OK;
}

adeco_status adeco_get_launch_shape(int* chunk, int* blocks_per_cu) {
    if (chunk) *chunk = launch_shape().chunk;
    if (blocks_per_cu) *blocks_per_cu = launch_shape().blocks_per_cu;
    return ADECO_OK;
}

// Times the batched leaf deposit over every (tiles per work-queue pop, resident workgroups per CU)
// of a small grid, keeps the fastest (minimum kernel time over `reps` runs) as the library's launch
// shape and reports it.  The leaf grids hold the result of the last run.
adeco_status adeco_autotune_batched(adeco_octree T, int dims, int decomposition, double tet_mass, int precision, int reps,
                                    int* best_chunk, int* best_blocks_per_cu, double* best_ms) {
    if (!T || reps < 1) return fail(ADECO_EINVAL, "autotune_batched(): null handle or reps < 1");
    static const int chunks[] = {1, 2, 4, 8, 16, 32};
    // workgroups per CU up to what the kernel's resources let a CU hold (at most 4)
    int max_blocks = 1;
    {
        std::lock_guard<std::mutex> lk(T->L->ds->mutex);
        if (hipSetDevice(T->L->device) != hipSuccess) return fail(ADECO_EDEVICE, "autotune_batched(): hipSetDevice failed");
        max_blocks = std::min(4, deposit_occupancy(precision, false));
    }
    const LaunchShape saved = launch_shape();
    double best = 1e300;
    LaunchShape pick = saved;
    for (int c : chunks)
        for (int b = 1; b <= max_blocks; ++b) {
            launch_shape().chunk = c;
            launch_shape().blocks_per_cu = b;
            double t = 1e300;
            for (int r = 0; r < reps; ++r) {
                const adeco_status st = adeco_deposit_leaves(T, dims, decomposition, tet_mass, ADECO_MODE_BATCHED, precision);
                if (st) {
                    launch_shape() = saved;
                    return st;
                }
                t = std::min(t, g_timings.main_ms);
            }
            if (t < best) {
                best = t;
                pick = launch_shape();
            }
        }
    launch_shape() = pick;
    if (best_chunk) *best_chunk = pick.chunk;
    if (best_blocks_per_cu) *best_blocks_per_cu = pick.blocks_per_cu;
    if (best_ms) *best_ms = best;
    return ADECO_OK;
}

// Leaf boxes (DFS order, x0,x1,y0,y1,z0,z1 per leaf) as the deposit uses them: the float
// reference boxes (exact = 0) or the exact dyadic FP64 boxes (exact = 1).
adeco_status adeco_octree_leaf_boxes(adeco_octree T, int exact, double* out) {
    if (!T || !out) return fail(ADECO_EINVAL, "octree_leaf_boxes(): null argument");
    for (size_t l = 0; l < T->leaves_dfs.size(); ++l) {
        if (exact) {
            T->node_box_exact(T->leaves_dfs[l], out + 6 * l);
        } else {
            float b[6];
            T->node_box(T->leaves_dfs[l], b);
            for (int a = 0; a < 6; ++a) out[6 * l + a] = b[a];
        }
    }
    return ADECO_OK;
}

// Binned lattice tet ids of one leaf (DFS index).  Returns the count; out may be null.
int64_t adeco_octree_leaf_tets(adeco_octree T, int64_t leaf, int32_t* out) {
    if (!T || T->binned_decomposition < 0 || leaf < 0 || leaf + 1 >= int64_t(T->leaf_tet_offset.size())) return -1;
    const int64_t b = T->leaf_tet_offset[leaf], n = T->leaf_tet_offset[leaf + 1] - b;
    if (out && n) {
        std::lock_guard<std::mutex> lk(T->L->ds->mutex);
        if (hipSetDevice(T->L->device) != hipSuccess ||
            hipMemcpy(out, T->d_pairs.as<int32_t>() + b, size_t(n) * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return -1;
    }
    return n;
}

// Leaf grids of the last deposit, leaf-major in DFS leaf order, FP64 mass per cell.
adeco_status adeco_leaf_grids(adeco_octree T, int64_t first_leaf, int64_t n_leaves, double* out) {
    if (!T || !out || T->grid_dims <= 0) return fail(ADECO_EINVAL, "leaf_grids(): nothing deposited");
    const int64_t nl = int64_t(T->leaves_dfs.size());
    if (first_leaf < 0 || n_leaves < 0 || first_leaf + n_leaves > nl) return fail(ADECO_EINVAL, "leaf_grids(): leaf range");
    const int64_t cells = int64_t(T->grid_dims) * T->grid_dims * T->grid_dims;
    std::lock_guard<std::mutex> lk(T->L->ds->mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e == hipSuccess)
        e = hipMemcpy(out, T->d_grids.as<double>() + first_leaf * cells, size_t(n_leaves * cells) * 8, hipMemcpyDeviceToHost);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "leaf grid download");
}

}  // extern "C"
