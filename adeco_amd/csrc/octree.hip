// Octree refinement counts and per-leaf tet binning on the GPU (C ABI part 3).
//
// The reference builds its octree depth-first on the host, summing per node a float estimate of
// the tets of each meta-cell it overlaps (AdOctreeGeneration.h:76-258, MetaMesh
// AdTypeDefinitions.h:715-1022).  Here the float work runs on the GPU, one level at a time:
//   * k_meta_bbox: float AABB of each meta-cell's <= 8^3 vertices (min/max: exact in any order);
//   * k_estimate_root / k_refine_level: for every (meta-cell, node) pair the strict-overlap test
//     and est = uint64(float(6*cells) * vol(node ∩ bbox_m) / vol(bbox_m)) with the reference's
//     float operation order (no fused multiply-add can form: the expression has no additions
//     after products), summed per node with 64-bit integer atomics, which is exact in any order.
// Which nodes refine does not depend on the traversal order, so the host only replays the
// reference's DFS stack (children pushed 0..7, popped 7..0) to number the nodes and list leaves
// exactly as the reference does.  Binning then gives every leaf the lattice tets whose AABB
// strictly overlaps it (a superset of nothing: exactly the tets that can deposit mass there).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <list>
#include <stack>
#include <vector>

#include "../../include/adeco.h"
#include "batched_impl.hpp"
#include "deposit_tet.hpp"
#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

namespace adeco {

namespace {

__device__ __forceinline__ bool fstrict(const float* a, const float* b) {  // boxes as x0,x1,y0,y1,z0,z1
    return a[0] < b[1] && b[0] < a[1] && a[2] < b[3] && b[2] < a[3] && a[4] < b[5] && b[4] < a[5];
}

// AdTypeDefinitions.h:812-825 (AABBox::getVolume is ((x1-x0)*(y1-y0))*(z1-z0) in float)
__device__ __forceinline__ unsigned long long estimate(const float* box, const float* mb, unsigned cells) {
    float it[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (fstrict(box, mb)) {
        for (int a = 0; a < 3; ++a) {
            it[2 * a] = box[2 * a] < mb[2 * a] ? mb[2 * a] : box[2 * a];
            it[2 * a + 1] = box[2 * a + 1] < mb[2 * a + 1] ? box[2 * a + 1] : mb[2 * a + 1];
        }
    }
    const float vi = __fmul_rn(__fmul_rn(__fsub_rn(it[1], it[0]), __fsub_rn(it[3], it[2])), __fsub_rn(it[5], it[4]));
    const float vm = __fmul_rn(__fmul_rn(__fsub_rn(mb[1], mb[0]), __fsub_rn(mb[3], mb[2])), __fsub_rn(mb[5], mb[4]));
    const float e = __fdiv_rn(__fmul_rn(float(6ull * cells), vi), vm);
    // float -> uint64 truncation; NaN (flat meta-cell) and negatives map to 0
    return e >= 1.f ? (unsigned long long)(e) : 0ull;
}

// one wave (64 lanes) per meta-cell
__global__ void __launch_bounds__(64) k_meta_bbox(const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, int npx, int npy, int npz,
                                                  int mdx, int mdy, int mdz, float* __restrict__ out_box,
                                                  unsigned* __restrict__ out_cells) {
    const int m = blockIdx.x;
    const int cx = m % mdx, cy = (m / mdx) % mdy, cz = m / (mdx * mdy);
    const int lo[3] = {7 * cx, 7 * cy, 7 * cz};
    const int d[3] = {cx + 1 == mdx ? npx - lo[0] : 8, cy + 1 == mdy ? npy - lo[1] : 8, cz + 1 == mdz ? npz - lo[2] : 8};
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    const int nv = d[0] * d[1] * d[2];
    for (int v = threadIdx.x; v < nv; v += 64) {
        const int i = v % d[0], j = (v / d[0]) % d[1], k = v / (d[0] * d[1]);
        const int64_t g = (lo[0] + i) + int64_t(npx) * ((lo[1] + j) + int64_t(npy) * (lo[2] + k));
        const float p[3] = {x[g], y[g], z[g]};
        for (int a = 0; a < 3; ++a) {
            mn[a] = p[a] < mn[a] ? p[a] : mn[a];
            mx[a] = p[a] > mx[a] ? p[a] : mx[a];
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a) {
            const float o0 = __shfl_down(mn[a], off, 64), o1 = __shfl_down(mx[a], off, 64);
            mn[a] = o0 < mn[a] ? o0 : mn[a];
            mx[a] = o1 > mx[a] ? o1 : mx[a];
        }
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; ++a) {
            out_box[6 * m + 2 * a] = mn[a];
            out_box[6 * m + 2 * a + 1] = mx[a];
        }
        out_cells[m] = unsigned(d[0] - 1) * unsigned(d[1] - 1) * unsigned(d[2] - 1);
    }
}

__global__ void k_estimate_root(const float* __restrict__ mbox, const unsigned* __restrict__ mcells, int M,
                                RootBox rb, unsigned long long* count) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long e = 0;
    if (m < M) e = estimate(rb.b, mbox + 6 * m, mcells[m]);
    // wave reduction, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    if ((threadIdx.x & 63) == 0 && e) atomicAdd(count, e);
}

// pairs of a refined level: (refined-parent slot, meta-cell); for each of the 8 children test
// strict overlap, set a membership bit and add the child's estimate.
__global__ void k_refine_level(const float* __restrict__ mbox, const unsigned* __restrict__ mcells,
                               const int32_t* __restrict__ pair_slot, const int32_t* __restrict__ pair_mc,
                               int64_t n_pairs, const float* __restrict__ child_boxes,
                               unsigned char* __restrict__ flags, unsigned long long* __restrict__ child_count) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int s = pair_slot[p];
    const int m = pair_mc[p];
    const float* mb = mbox + 6 * m;
    unsigned char f = 0;
    for (int c = 0; c < 8; ++c) {
        const float* cb = child_boxes + 6 * (8 * s + c);
        if (fstrict(mb, cb)) {  // distribute_tets_to_children: intersection(meta_cell_bbox, child_bbox)
            f |= (unsigned char)(1u << c);
            const unsigned long long e = estimate(cb, mb, mcells[m]);
            if (e) atomicAdd(child_count + 8 * s + c, e);
        }
    }
    flags[p] = f;
}

// ---- per-tet binning into leaves (tree descent with a bounded explicit stack) ----
constexpr int kMaxStack = 8 * kMaxOctreeDepth + 8;

template <bool kFill>
__device__ __forceinline__ void bin_one_by_descent(const LatticeView& lat, const PackedCorners& ct, int64_t t,
                                                   const float* __restrict__ node_box,
                                                   const int32_t* __restrict__ first_child,
                                                   const int32_t* __restrict__ leaf_index, int* __restrict__ leaf_count,
                                                   const int64_t* __restrict__ leaf_offset, int32_t* __restrict__ pairs,
                                                   int64_t pair_cap) {
    V3<double> v[4];
    lattice_tet_vertices(lat, ct, t, v);
    // the tet AABB in float is exact (vertices are float)
    float tb[6];
    tb[0] = tb[1] = float(v[0].x);
    tb[2] = tb[3] = float(v[0].y);
    tb[4] = tb[5] = float(v[0].z);
    for (int q = 1; q < 4; ++q) {
        const float p[3] = {float(v[q].x), float(v[q].y), float(v[q].z)};
        for (int a = 0; a < 3; ++a) {
            tb[2 * a] = p[a] < tb[2 * a] ? p[a] : tb[2 * a];
            tb[2 * a + 1] = p[a] > tb[2 * a + 1] ? p[a] : tb[2 * a + 1];
        }
    }
    if (!fstrict(tb, node_box)) return;
    int stack[kMaxStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int nd = stack[--sp];
        const int fc = first_child[nd];
        if (fc < 0) {
            const int leaf = leaf_index[nd];
            const int pos = atomicAdd(leaf_count + leaf, 1);
            if (kFill && leaf_offset[leaf] + pos < pair_cap) pairs[leaf_offset[leaf] + pos] = int32_t(t);
            continue;
        }
        for (int c = 0; c < 8 && sp < kMaxStack; ++c)
            if (fstrict(tb, node_box + 6 * (fc + c))) stack[sp++] = fc + c;
    }
}

// (the tets of `ids[0..n)`, the tets the leaf-map pass could not take; n = n_tets, or *n_dev when
// given, grid-stride; fill writes only pairs below pair_cap)
template <bool kFill>
__global__ void __launch_bounds__(256) k_bin_tets(LatticeView lat, PackedCorners ct, int64_t n_tets,
                                                  const int* __restrict__ n_dev, const int32_t* __restrict__ ids,
                                                  const float* __restrict__ node_box, const int32_t* __restrict__ first_child,
                                                  const int32_t* __restrict__ leaf_index, int* __restrict__ leaf_count,
                                                  const int64_t* __restrict__ leaf_offset, int32_t* __restrict__ pairs,
                                                  int64_t pair_cap) {
    const int64_t n = n_dev ? int64_t(*n_dev) : n_tets;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        bin_one_by_descent<kFill>(lat, ct, ids[i], node_box, first_child, leaf_index, leaf_count, leaf_offset, pairs,
                                  pair_cap);
}

__global__ void k_iota(int32_t* __restrict__ out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = int32_t(i);
}

// ---- per-tet binning through a leaf map ----
// The leaves tile the root; on the grid of the finest leaf level (F^3 cells of edge h = root
// extent / F) every cell belongs to one leaf.  A tet's candidate leaves are those of the fine cells
// that its float AABB covers, widened by 1e-4 of a cell each way (a superset of the leaves whose
// float or exact box it strictly overlaps: those boxes differ from the fine-cell faces by at most
// an ulp of the coordinates); each candidate is then kept only if the AABB strictly overlaps the
// leaf's box widened by two float ulps per face, the descent kernel's test.  At most kMapCells
// fine cells per tet; larger tets go to an overflow list for the descent kernel.  Counter atomics
// are aggregated per wave: one per distinct leaf among the wave's lanes.
constexpr int kMapCells = 8;


__device__ __forceinline__ unsigned wave_rank(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u));
}

// tleaf: per tet, written by the count pass and read by the fill pass: its only leaf, kNoLeaf, or
// kRecompute (several leaves, or the descent's), so that the fill pass recomputes only those.
constexpr int32_t kNoLeaf = -2, kRecompute = -1;

template <bool kFill>
__global__ void __launch_bounds__(256) k_bin_map(LatticeView lat, PackedCorners ct, int64_t n_tets, LeafMap lm,
                                                 const float* __restrict__ root_box, int* __restrict__ leaf_count,
                                                 const int64_t* __restrict__ leaf_offset, int32_t* __restrict__ pairs,
                                                 int64_t pair_cap, int32_t* __restrict__ overflow,
                                                 int* __restrict__ n_overflow, int32_t* __restrict__ tleaf) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int leaf[kMapCells];
#pragma unroll
    for (int c = 0; c < kMapCells; ++c) leaf[c] = -1;
    bool over = false;
    const int32_t known = (kFill && t < n_tets) ? tleaf[t] : kRecompute;
    if (kFill && known >= 0) leaf[0] = known;
    if (t < n_tets && known == kRecompute) {
        V3<double> v[4];
        lattice_tet_vertices(lat, ct, t, v);
        float tb[6];
        tb[0] = tb[1] = float(v[0].x);
        tb[2] = tb[3] = float(v[0].y);
        tb[4] = tb[5] = float(v[0].z);
        for (int q = 1; q < 4; ++q) {
            const float p[3] = {float(v[q].x), float(v[q].y), float(v[q].z)};
            for (int a = 0; a < 3; ++a) {
                tb[2 * a] = p[a] < tb[2 * a] ? p[a] : tb[2 * a];
                tb[2 * a + 1] = p[a] > tb[2 * a + 1] ? p[a] : tb[2 * a + 1];
            }
        }
        if (fstrict(tb, root_box)) {
            int lo[3], n[3];
            for (int a = 0; a < 3; ++a) {
                const double u0 = (double(tb[2 * a]) - lm.ll[a]) * lm.inv_h - 1e-4;
                const double u1 = (double(tb[2 * a + 1]) - lm.ll[a]) * lm.inv_h + 1e-4;
                const int i0 = u0 <= 0.0 ? 0 : (u0 >= double(lm.F - 1) ? lm.F - 1 : int(u0));
                const int i1 = u1 <= 0.0 ? 0 : (u1 >= double(lm.F - 1) ? lm.F - 1 : int(u1));
                lo[a] = i0;
                n[a] = i1 - i0 + 1;
            }
            const int cells = n[0] * n[1] * n[2];
            if (cells > kMapCells) {
                over = true;
            } else {
                // the fine cells' leaves, each distinct leaf once (fixed slots: registers only).  Slot c
                // -> (c mod n0, (c / n0) mod n1, c / (n0 n1)) without integer division: for c < 8 and
                // 1 <= n <= 8, floor(c / n) = (c * ceil(64 / n)) >> 6 exactly (ceil(64 / n) packed in 7
                // bits per n)
                constexpr unsigned long long kCeil64Over = 0x102858d2059040ull;
                const int m0 = int((kCeil64Over >> (7 * (n[0] - 1))) & 127u);
                const int m1 = int((kCeil64Over >> (7 * (n[1] - 1))) & 127u);
#pragma unroll
                for (int c = 0; c < kMapCells; ++c) {
                    if (c < cells) {
                        const int q0 = (c * m0) >> 6, q1 = (q0 * m1) >> 6;
                        const int cx = c - q0 * n[0], cy = q0 - q1 * n[1], cz = q1;
                        leaf[c] = lm.map[(lo[0] + cx) + int64_t(lm.F) * ((lo[1] + cy) + int64_t(lm.F) * (lo[2] + cz))];
                    }
                }
#pragma unroll
                for (int c = 1; c < kMapCells; ++c)
#pragma unroll
                    for (int d = 0; d < c; ++d)
                        if (leaf[c] == leaf[d]) leaf[c] = -1;
#pragma unroll
                for (int c = 0; c < kMapCells; ++c)
                    if (leaf[c] >= 0 && !fstrict(tb, lm.leaf_box + 6 * leaf[c])) leaf[c] = -1;
            }
        }
    }
    if (!kFill && t < n_tets) {
        int n = 0, only = kNoLeaf;
#pragma unroll
        for (int c = 0; c < kMapCells; ++c)
            if (leaf[c] >= 0) {
                ++n;
                only = leaf[c];
            }
        tleaf[t] = over || n > 1 ? kRecompute : only;
    }
    // overflow tets: one list append per wave
    if (!kFill) {
        const unsigned long long om = __ballot(over);
        if (om) {
            const int src = __builtin_ctzll(om);
            int base = 0;
            if (lane == src) base = atomicAdd(n_overflow, __popcll(om));
            base = __shfl(base, src, 64);
            if (over) overflow[base + int(wave_rank(om))] = int32_t(t);
        }
    }
#pragma unroll
    for (int c = 0; c < kMapCells; ++c) {
        unsigned long long pending = __ballot(leaf[c] >= 0);
        while (pending) {
            const int src = __builtin_ctzll(pending);
            const int L = __shfl(leaf[c], src, 64);
            const unsigned long long m = __ballot(leaf[c] == L);
            int base = 0;
            if (lane == src) base = atomicAdd(leaf_count + L, __popcll(m));
            base = __shfl(base, src, 64);
            if (kFill && leaf[c] == L) {
                const int64_t at = leaf_offset[L] + base + int(wave_rank(m));
                if (at < pair_cap) pairs[at] = int32_t(t);
            }
            pending &= ~m;
        }
    }
}

}  // namespace

adeco_status octree_meta_bbox(adeco_lattice_s* L, const int md[3], float* d_box, unsigned* d_cells) {
    DeviceState& ds = *L->ds;
    const int M = md[0] * md[1] * md[2];
    hipLaunchKernelGGL(k_meta_bbox, dim3(M), dim3(64), 0, ds.stream, L->d_x, L->d_y, L->d_z, L->np[0], L->np[1],
                       L->np[2], md[0], md[1], md[2], d_box, d_cells);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "meta-cell bbox kernel");
}

adeco_status octree_root_count(DeviceState& ds, const float* d_box, const unsigned* d_cells, int M, const float box[6],
                               unsigned long long* d_count, unsigned long long* h_count) {
    RootBox rb;
    std::memcpy(rb.b, box, sizeof(rb.b));
    hipError_t e = hipMemsetAsync(d_count, 0, sizeof(unsigned long long), ds.stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_estimate_root, dim3((M + 255) / 256), dim3(256), 0, ds.stream, d_box, d_cells, M, rb, d_count);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h_count, d_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "root estimate");
}

adeco_status octree_refine_level(DeviceState& ds, const float* d_box, const unsigned* d_cells,
                                 const std::vector<int32_t>& slot, const std::vector<int32_t>& mc,
                                 const std::vector<float>& child_boxes, std::vector<unsigned char>& flags,
                                 std::vector<unsigned long long>& child_count) {
    const int64_t n = int64_t(slot.size());
    const int64_t nchild = int64_t(child_boxes.size() / 6);
    flags.assign(size_t(n), 0);
    child_count.assign(size_t(nchild), 0);
    if (n == 0) return ADECO_OK;
    DevBuf d_slot, d_mc, d_cb, d_flags, d_cnt;
    adeco_status st;
    if ((st = d_slot.reserve(n * 4)) || (st = d_mc.reserve(n * 4)) || (st = d_cb.reserve(child_boxes.size() * 4)) ||
        (st = d_flags.reserve(n)) || (st = d_cnt.reserve(nchild * 8)))
        return st;
    hipError_t e = hipMemcpyAsync(d_slot.ptr, slot.data(), n * 4, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_mc.ptr, mc.data(), n * 4, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_cb.ptr, child_boxes.data(), child_boxes.size() * 4, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt.ptr, 0, nchild * 8, ds.stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_refine_level, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ds.stream, d_box, d_cells,
                           d_slot.as<int32_t>(), d_mc.as<int32_t>(), n, d_cb.as<float>(), d_flags.as<unsigned char>(),
                           d_cnt.as<unsigned long long>());
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(flags.data(), d_flags.ptr, n, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(child_count.data(), d_cnt.ptr, nchild * 8, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "octree level refinement");
}

namespace {
__global__ void k_set_int(int* __restrict__ p, int v) { *p = v; }
}  // namespace

adeco_status bin_plan(adeco_octree_s* T, int decomposition, BinPlan& p) {
    adeco_lattice_s* L = T->L;
    DeviceState& ds = *L->ds;
    const int64_t n_nodes = int64_t(T->nodes.size());
    const int64_t n_leaves = int64_t(T->leaves_dfs.size());
    std::vector<float> boxes(6 * n_nodes);
    std::vector<int32_t> first(n_nodes), leaf_index(n_nodes, -1);
    for (int64_t i = 0; i < n_nodes; ++i) {
        // widened by two float ulps per face: a conservative superset of the tets that strictly
        // overlap either the float box or the exact dyadic box of the node.  Extra tets deposit
        // nothing (the per-tet deposit re-tests strict overlap with its grid).
        float* b = &boxes[6 * i];
        T->node_box(int(i), b);
        for (int a = 0; a < 3; ++a) {
            b[2 * a] = std::nextafter(std::nextafter(b[2 * a], -FLT_MAX), -FLT_MAX);
            b[2 * a + 1] = std::nextafter(std::nextafter(b[2 * a + 1], FLT_MAX), FLT_MAX);
        }
        first[i] = T->nodes[i].first_child;
    }
    for (int64_t l = 0; l < n_leaves; ++l) leaf_index[T->leaves_dfs[l]] = int32_t(l);
    const int64_t n_tets = 6 * int64_t(L->np[0] - 1) * (L->np[1] - 1) * (L->np[2] - 1);
    adeco_status st;
    if ((st = T->d_node_box.reserve(boxes.size() * 4)) || (st = T->d_first.reserve(n_nodes * 4)) ||
        (st = T->d_leaf_index.reserve(n_nodes * 4)) || (st = T->d_leaf_count.reserve(n_leaves * 4 + 4)) ||
        (st = T->d_leaf_offset.reserve((n_leaves + 1) * 8)) || (st = T->d_overflow.reserve(size_t(n_tets) * 4 + 8)))
        return st;
    hipError_t e = hipMemcpyAsync(T->d_node_box.ptr, boxes.data(), boxes.size() * 4, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T->d_first.ptr, first.data(), n_nodes * 4, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(T->d_leaf_index.ptr, leaf_index.data(), n_nodes * 4, hipMemcpyHostToDevice, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "binning setup");
    // the leaf map on the finest leaf level, when that level is at most 7 (F <= 128)
    int lmax = 0;
    for (int64_t l = 0; l < n_leaves; ++l) lmax = std::max(lmax, T->nodes[T->leaves_dfs[l]].level);
    p = BinPlan();
    p.use_map = lmax <= 7;
    std::vector<int32_t> map;
    std::vector<float> leaf_box;
    if (p.use_map) {
        const int F = 1 << lmax;
        map.assign(size_t(F) * F * F, -1);
        for (int64_t l = 0; l < n_leaves; ++l) {
            const ONode& n = T->nodes[T->leaves_dfs[l]];
            const int w = 1 << (lmax - n.level);
            for (int k = n.ix[2] * w; k < (n.ix[2] + 1) * w; ++k)
                for (int j = n.ix[1] * w; j < (n.ix[1] + 1) * w; ++j)
                    for (int i = n.ix[0] * w; i < (n.ix[0] + 1) * w; ++i) map[i + size_t(F) * (j + size_t(F) * k)] = int32_t(l);
        }
        for (int32_t m : map)
            if (m < 0) return fail(ADECO_EINTERNAL, "binning: the leaves do not tile the root");
        leaf_box.resize(6 * size_t(n_leaves));
        for (int64_t l = 0; l < n_leaves; ++l)
            std::memcpy(&leaf_box[6 * l], &boxes[6 * size_t(T->leaves_dfs[l])], 6 * sizeof(float));
        if ((st = T->d_leaf_map.reserve(map.size() * 4 + leaf_box.size() * 4)) || (st = T->d_tleaf.reserve(size_t(n_tets) * 4)))
            return st;
        e = hipMemcpyAsync(T->d_leaf_map.ptr, map.data(), map.size() * 4, hipMemcpyHostToDevice, ds.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(T->d_leaf_map.as<char>() + map.size() * 4, leaf_box.data(), leaf_box.size() * 4,
                               hipMemcpyHostToDevice, ds.stream);
        if (e != hipSuccess) return hip_fail(e, "leaf map upload");
        p.lm.map = T->d_leaf_map.as<int32_t>();
        p.lm.leaf_box = reinterpret_cast<const float*>(T->d_leaf_map.as<char>() + map.size() * 4);
        p.lm.F = F;
        const double h = double(T->nodes[0].ext) / double(F);
        for (int a = 0; a < 3; ++a) p.lm.ll[a] = double(T->nodes[0].ll[a]);
        p.lm.inv_h = 1.0 / h;
        p.tleaf = T->d_tleaf.as<int32_t>();
    }
    // the host vectors above are the sources of the copies
    e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "binning setup");
    p.lv = LatticeView{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    p.ct = pack_corners(corner_table(decomposition));
    p.n_tets = n_tets;
    p.n_leaves = int(n_leaves);
    p.node_box = T->d_node_box.as<float>();
    p.first = T->d_first.as<int32_t>();
    p.leaf_index = T->d_leaf_index.as<int32_t>();
    p.leaf_count = T->d_leaf_count.as<int>();
    p.n_over = T->d_overflow.as<int>();
    p.over = reinterpret_cast<int32_t*>(T->d_overflow.as<char>() + 8);
    return ADECO_OK;
}

// grid of the descent passes: grid-stride over a count that only the device knows
static unsigned descent_blocks(int64_t n_tets) { return unsigned(std::min<int64_t>((n_tets + 255) / 256, 8192)); }

hipError_t bin_count_async(const BinPlan& p, hipStream_t stream) {
    const unsigned nb = unsigned((p.n_tets + 255) / 256);
    hipError_t e = launch_zero_u32(reinterpret_cast<unsigned*>(p.leaf_count), p.n_leaves, stream);
    if (e != hipSuccess) return e;
    if (p.use_map) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, stream, p.n_over, 0);
        hipLaunchKernelGGL(k_bin_map<false>, dim3(nb), dim3(256), 0, stream, p.lv, p.ct, p.n_tets, p.lm, p.node_box,
                           p.leaf_count, (const int64_t*)nullptr, (int32_t*)nullptr, int64_t(0), p.over, p.n_over, p.tleaf);
    } else {
        hipLaunchKernelGGL(k_iota, dim3(nb), dim3(256), 0, stream, p.over, p.n_tets);
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, stream, p.n_over, int(p.n_tets));
    }
    hipLaunchKernelGGL(k_bin_tets<false>, dim3(descent_blocks(p.n_tets)), dim3(256), 0, stream, p.lv, p.ct, p.n_tets,
                       (const int*)p.n_over, (const int32_t*)p.over, p.node_box, p.first, p.leaf_index, p.leaf_count,
                       (const int64_t*)nullptr, (int32_t*)nullptr, int64_t(0));
    return hipGetLastError();
}

hipError_t bin_fill_async(const BinPlan& p, const int64_t* d_leaf_offset, int32_t* d_pairs, int64_t pair_cap,
                          hipStream_t stream) {
    const unsigned nb = unsigned((p.n_tets + 255) / 256);
    hipError_t e = launch_zero_u32(reinterpret_cast<unsigned*>(p.leaf_count), p.n_leaves, stream);
    if (e != hipSuccess) return e;
    if (p.use_map)
        hipLaunchKernelGGL(k_bin_map<true>, dim3(nb), dim3(256), 0, stream, p.lv, p.ct, p.n_tets, p.lm, p.node_box,
                           p.leaf_count, d_leaf_offset, d_pairs, pair_cap, (int32_t*)nullptr, (int*)nullptr, p.tleaf);
    hipLaunchKernelGGL(k_bin_tets<true>, dim3(descent_blocks(p.n_tets)), dim3(256), 0, stream, p.lv, p.ct, p.n_tets,
                       (const int*)p.n_over, (const int32_t*)p.over, p.node_box, p.first, p.leaf_index, p.leaf_count,
                       d_leaf_offset, d_pairs, pair_cap);
    return hipGetLastError();
}

adeco_status octree_bin_tets(adeco_octree_s* T, int decomposition) {
    DeviceState& ds = *T->L->ds;
    BinPlan p;
    adeco_status st = bin_plan(T, decomposition, p);
    if (st) return st;
    const int64_t n_leaves = p.n_leaves;
    hipError_t e = bin_count_async(p, ds.stream);
    std::vector<int> cnt(n_leaves);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), T->d_leaf_count.ptr, n_leaves * 4, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "binning count pass");
    T->leaf_tet_offset.assign(n_leaves + 1, 0);
    for (int64_t l = 0; l < n_leaves; ++l) T->leaf_tet_offset[l + 1] = T->leaf_tet_offset[l] + cnt[l];
    const int64_t total = T->leaf_tet_offset[n_leaves];
    if ((st = T->d_pairs.reserve(size_t(total) * 4 + 4))) return st;
    e = hipMemcpyAsync(T->d_leaf_offset.ptr, T->leaf_tet_offset.data(), (n_leaves + 1) * 8, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = bin_fill_async(p, T->d_leaf_offset.as<int64_t>(), T->d_pairs.as<int32_t>(), total, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "binning fill pass");
    T->binned_decomposition = decomposition;
    return ADECO_OK;
}

}  // namespace adeco
