// Batched persistent deposit over all octree leaves in one launch.
//
// The plain form gives every lane one tet and lets it walk its covered cells, so a wave runs as
// long as its largest tet and spends most lanes idle on heavy-tailed leaves.  Here the unit of
// work is one (tet, candidate cell) pair:
//   1. k_pair_setup     one lane per binned (leaf, tet) pair: covered cell range in its leaf
//                       (AdDepositTets2Grid.h:514-539) -> candidate count;
//   2. scan             candidate counts -> offsets (64-bit), total C;
//   3. k_tile_map       first pair of every tile of kTile consecutive candidates;
//   4. k_deposit_batched persistent workgroups pull tiles from an atomic work counter.  A tile's
//                       pairs (<= kTile + 1) are staged in LDS once (vertices scaled to the leaf's
//                       cell units in FP64, mass factor, covered range); each lane binary-searches
//                       LDS for its candidate's pair, culls by AABB and classifies the cell with
//                       the 4 face planes evaluated at the extreme corners (conservative margins:
//                       an uncertain cell is clipped, never skipped); trivial cells are deposited
//                       at once and partial cells are compacted into an LDS queue that all lanes
//                       then clip, so the expensive exact clip runs with full waves.
// Masses go to the leaf grids with FP64 global atomics.  Precision modes for the clip:
//   FP64  all FP64;
//   MIXED FP32 clip on cell-local coordinates formed in FP64, accepted only when its error bound
//         16 * eps32 * max(1, M)^2 (M = largest |coordinate| of the cell-local tet; the constant is
//         twice the worst ratio measured over 2e5 random tets against FP64) is below 2.5e-7 of
//         the clipped volume, otherwise recomputed in FP64;
//   FP32  FP32 clip only (precision study).
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/adeco.h"
#include "deposit_tet.hpp"
#include "kernels.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

namespace adeco {

namespace {

constexpr int kTile = 256;  // candidates per tile == threads per workgroup

struct LeafGeomD {
    double gmin[3];
    double delta[3];
    double r[3];      // delta_a * scale_a: the cell edge in scaled units per axis
    double scale[3];  // per-axis scale (all 1/dx under the reference's cubic-cell formula)
    double cell_vol;
};

struct PairLds {
    double S[12];  // scale * (src_q - gmin), oriented vertex order
    double mcell;  // tet_mass * cell_vol / |V|
    int leaf;
    unsigned short lo[3], ext[3];  // covered cells: lo .. lo + ext - 1 (ext = 0: dead pair)
};

__device__ __forceinline__ int upper_bound_i64(const int64_t* a, int n, int64_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// per pair: leaf, covered range, candidate count (dead pairs keep one dummy candidate so every
// tile spans at most kTile + 1 pairs)
__global__ void __launch_bounds__(256) k_pair_setup(LatticeView lat, PackedCorners ct, const int32_t* __restrict__ pairs,
                                                    int64_t n_pairs, const int64_t* __restrict__ leaf_off, int n_leaves,
                                                    const LeafGeomD* __restrict__ lg, int dims,
                                                    unsigned* __restrict__ cand, unsigned long long* __restrict__ range,
                                                    int32_t* __restrict__ pair_leaf) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int leaf = upper_bound_i64(leaf_off, n_leaves + 1, p) - 1;
    V3<double> v[4];
    lattice_tet_vertices(lat, ct, pairs[p], v);
    const LeafGeomD& g = lg[leaf];
    double mn[3] = {v[0].x, v[0].y, v[0].z}, mx[3] = {v[0].x, v[0].y, v[0].z};
    for (int q = 1; q < 4; ++q) {
        const double c[3] = {v[q].x, v[q].y, v[q].z};
        for (int a = 0; a < 3; ++a) {
            mn[a] = c[a] < mn[a] ? c[a] : mn[a];
            mx[a] = c[a] > mx[a] ? c[a] : mx[a];
        }
    }
    unsigned long long pack = 0;
    unsigned count = 1;
    const double vol = vol6(v[0], v[1], v[2], v[3]);
    bool live = vol != 0.0;
    const int dm1 = dims - 1;
    unsigned n = 1;
    for (int a = 0; a < 3; ++a) {
        // covered range of the tet AABB (clamping the overlap to the grid first changes nothing
        // after the clamp of the indices)
        int l = int(floor((mn[a] - g.gmin[a]) / g.delta[a]));
        int h = int(ceil((mx[a] - g.gmin[a]) / g.delta[a]));
        l = l <= 0 ? 0 : (l >= dm1 ? dm1 : l);
        h = h <= 0 ? 0 : (h >= dm1 ? dm1 : h);
        const unsigned e = unsigned(h - l + 1);
        n *= e;
        pack |= (unsigned long long)l << (10 * a);
        pack |= (unsigned long long)e << (30 + 10 * a);
    }
    if (live) count = n;
    else pack = 0;  // ext = 0 marks a dead pair
    cand[p] = count;
    range[p] = pack;
    pair_leaf[p] = leaf;
}

// exclusive scan of uint32 counts into int64 offsets: block-local pass, block-sum pass, add pass
constexpr int kScanBlock = 1024;

__global__ void __launch_bounds__(256) k_scan_local(const unsigned* __restrict__ in, int64_t n, int64_t* __restrict__ out,
                                                    int64_t* __restrict__ block_sum) {
    __shared__ int64_t s[kScanBlock];
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    for (int i = threadIdx.x; i < kScanBlock; i += blockDim.x) s[i] = (base + i < n) ? int64_t(in[base + i]) : 0;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int64_t add[kScanBlock / 256];
        for (int k = 0; k < kScanBlock / 256; ++k) {
            const int i = threadIdx.x + 256 * k;
            add[k] = i >= off ? s[i - off] : 0;
        }
        __syncthreads();
        for (int k = 0; k < kScanBlock / 256; ++k) s[threadIdx.x + 256 * k] += add[k];
        __syncthreads();
    }
    for (int i = threadIdx.x; i < kScanBlock; i += blockDim.x)
        if (base + i < n) out[base + i] = s[i] - ((base + i < n) ? int64_t(in[base + i]) : 0);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = s[kScanBlock - 1];
}

// one workgroup of 256 lanes: exclusive scan of the block sums in chunks of 256 with a carry
__global__ void __launch_bounds__(256) k_scan_sums(int64_t* __restrict__ bs, int64_t nb, int64_t* __restrict__ total) {
    __shared__ int64_t s[256];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += 256) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < nb ? bs[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int64_t a = int(threadIdx.x) >= off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < nb) bs[i] = carry + s[threadIdx.x] - v;
        carry += s[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) k_scan_add(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ bs,
                                                  const int64_t* __restrict__ total) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bs[i / kScanBlock];
    if (i == n) out[n] = *total;
}

__global__ void __launch_bounds__(256) k_tile_map(const int64_t* __restrict__ off, int64_t n_pairs,
                                                  int32_t* __restrict__ tile_first, int64_t n_tiles) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p > n_pairs) return;
    if (p == n_pairs) {
        tile_first[n_tiles] = int32_t(n_pairs - 1);
        return;
    }
    const int64_t t0 = (off[p] + kTile - 1) / kTile, t1 = (off[p + 1] + kTile - 1) / kTile;
    for (int64_t t = t0; t < t1 && t < n_tiles; ++t) tile_first[t] = int32_t(p);
}

template <class R>
__device__ __forceinline__ void load_s(const PairLds& pr, const LeafGeomD& g, int i, int j, int k, V3<R> s[4]) {
    const double ox = double(i) * g.r[0], oy = double(j) * g.r[1], oz = double(k) * g.r[2];
    for (int q = 0; q < 4; ++q) s[q] = mk(R(pr.S[3 * q] - ox), R(pr.S[3 * q + 1] - oy), R(pr.S[3 * q + 2] - oz));
}

// Conservative classification of the cell-local tet s against [0,1]^3 (FP64).
__device__ __forceinline__ int classify_margin(const V3<double> s[4]) {
    const V3<double> e01 = sub(s[1], s[0]), e02 = sub(s[2], s[0]), e03 = sub(s[3], s[0]);
    const V3<double> e13 = sub(s[3], s[1]), e12 = sub(s[2], s[1]);
    const V3<double> n[4] = {cross(e01, e02), cross(e02, e03), cross(e03, e01), cross(e13, e12)};
    bool all_in = true;
    for (int f = 0; f < 4; ++f) {
        const V3<double> o = f == 3 ? s[1] : s[0];
        const double v0 = o.x * n[f].x + o.y * n[f].y + o.z * n[f].z;
        const double pos = fmax(n[f].x, 0.0) + fmax(n[f].y, 0.0) + fmax(n[f].z, 0.0);
        const double neg = fmin(n[f].x, 0.0) + fmin(n[f].y, 0.0) + fmin(n[f].z, 0.0);
        const double scl = (fabs(n[f].x) + fabs(n[f].y) + fabs(n[f].z)) *
                           (1.0 + fmax(fmax(fabs(o.x), fabs(o.y)), fabs(o.z)));
        const double tol = 1e-12 * scl;
        const double vmax = v0 - neg;  // plane value (o - p).n at the corner maximising it
        const double vmin = v0 - pos;
        if (vmax < -tol) return CLS_NONE;
        all_in = all_in && (vmin > tol);
    }
    if (all_in) return CLS_CELL_IN_TET;
    bool tin = true;
    for (int q = 0; q < 4; ++q)
        tin = tin && s[q].x >= 0.0 && s[q].x <= 1.0 && s[q].y >= 0.0 && s[q].y <= 1.0 && s[q].z >= 0.0 && s[q].z <= 1.0;
    return tin ? CLS_TET_IN_CELL : CLS_PARTIAL;
}

template <int kPrec>
__global__ void __launch_bounds__(kTile) k_deposit_batched(const int32_t* __restrict__ pairs, const int32_t* __restrict__ pair_leaf,
                                                           const unsigned long long* __restrict__ range,
                                                           const int64_t* __restrict__ cand_off, const int32_t* __restrict__ tile_first,
                                                           int64_t n_tiles, int64_t n_cand, unsigned* __restrict__ tile_counter,
                                                           LatticeView lat, PackedCorners ct, const LeafGeomD* __restrict__ lg,
                                                           int dims, double tet_mass, double* __restrict__ grids,
                                                           unsigned long long* __restrict__ fallback) {
    __shared__ PairLds s_pair[kTile + 1];
    __shared__ int64_t s_off[kTile + 2];
    __shared__ unsigned short s_qslot[kTile];
    __shared__ unsigned s_qr[kTile];
    __shared__ int s_qn, s_np;
    __shared__ int64_t s_tile;
    const int tid = threadIdx.x;
    const int64_t cells = int64_t(dims) * dims * dims;
    unsigned long long my_fallback = 0;
    for (;;) {
        if (tid == 0) {
            s_tile = int64_t(atomicAdd(tile_counter, 1u));
            s_qn = 0;
        }
        __syncthreads();
        const int64_t t = s_tile;
        if (t >= n_tiles) break;
        const int p0 = tile_first[t];
        const int p1 = tile_first[t + 1];
        const int np = p1 - p0 + 1;  // <= kTile + 1 (every pair holds >= 1 candidate)
        // stage the tile's pairs
        for (int q = tid; q < np; q += kTile) {
            const int p = p0 + q;
            PairLds pr;
            const unsigned long long pk = range[p];
            pr.leaf = pair_leaf[p];
            for (int a = 0; a < 3; ++a) {
                pr.lo[a] = (unsigned short)((pk >> (10 * a)) & 1023u);
                pr.ext[a] = (unsigned short)((pk >> (30 + 10 * a)) & 1023u);
            }
            if (pr.ext[0]) {
                V3<double> v[4];
                lattice_tet_vertices(lat, ct, pairs[p], v);
                const double vol = vol6(v[0], v[1], v[2], v[3]) / 6.0;
                if (vol < 0.0) {
                    const V3<double> tmp = v[0];
                    v[0] = v[3];
                    v[3] = tmp;
                }
                const LeafGeomD& g = lg[pr.leaf];
                for (int w = 0; w < 4; ++w) {
                    pr.S[3 * w] = g.scale[0] * (v[w].x - g.gmin[0]);
                    pr.S[3 * w + 1] = g.scale[1] * (v[w].y - g.gmin[1]);
                    pr.S[3 * w + 2] = g.scale[2] * (v[w].z - g.gmin[2]);
                }
                pr.mcell = tet_mass * (g.cell_vol / fabs(vol));
            } else {
                for (int w = 0; w < 12; ++w) pr.S[w] = 0.0;
                pr.mcell = 0.0;
            }
            s_pair[q] = pr;
            s_off[q] = cand_off[p];
        }
        if (tid == 0) {
            s_np = np;
            s_off[np] = cand_off[p0 + np];
        }
        __syncthreads();
        // classify one candidate per lane
        const int64_t c = t * kTile + tid;
        if (c < n_cand) {
            const int slot = upper_bound_i64(s_off, s_np, c) - 1;
            const PairLds& pr = s_pair[slot];
            if (pr.ext[0]) {
                const unsigned r = unsigned(c - s_off[slot]);
                const unsigned ex = pr.ext[0], ey = pr.ext[1];
                const int i = pr.lo[0] + int(r % ex);
                const int j = pr.lo[1] + int((r / ex) % ey);
                const int k = pr.lo[2] + int(r / (ex * ey));
                const LeafGeomD& g = lg[pr.leaf];
                V3<double> s[4];
                load_s(pr, g, i, j, k, s);
                // strict AABB cull against the cell [0, rx] x [0, ry] x [0, rz] (cell units)
                double mn[3] = {s[0].x, s[0].y, s[0].z}, mx[3] = {s[0].x, s[0].y, s[0].z};
                for (int q = 1; q < 4; ++q) {
                    const double cc[3] = {s[q].x, s[q].y, s[q].z};
                    for (int a = 0; a < 3; ++a) {
                        mn[a] = cc[a] < mn[a] ? cc[a] : mn[a];
                        mx[a] = cc[a] > mx[a] ? cc[a] : mx[a];
                    }
                }
                const bool hit = mn[0] < g.r[0] && 0.0 < mx[0] && mn[1] < g.r[1] && 0.0 < mx[1] && mn[2] < g.r[2] && 0.0 < mx[2];
                if (hit) {
                    const int cls = classify_margin(s);
                    const int64_t idx = int64_t(pr.leaf) * cells + i + int64_t(dims) * (j + int64_t(dims) * k);
                    if (cls == CLS_TET_IN_CELL) {
                        unsafeAtomicAdd(grids + idx, tet_mass);
                    } else if (cls == CLS_CELL_IN_TET) {
                        unsafeAtomicAdd(grids + idx, pr.mcell);
                    } else if (cls == CLS_PARTIAL) {
                        const int qi = atomicAdd(&s_qn, 1);
                        s_qslot[qi] = (unsigned short)slot;
                        s_qr[qi] = r;
                    }
                }
            }
        }
        __syncthreads();
        // clip the compacted partial cells with all lanes
        const int nq = s_qn;
        if (tid == 0 && nq) atomicAdd(fallback + 1, (unsigned long long)nq);
        for (int qi = tid; qi < nq; qi += kTile) {
            const PairLds& pr = s_pair[s_qslot[qi]];
            const unsigned r = s_qr[qi];
            const unsigned ex = pr.ext[0], ey = pr.ext[1];
            const int i = pr.lo[0] + int(r % ex);
            const int j = pr.lo[1] + int((r / ex) % ey);
            const int k = pr.lo[2] + int(r / (ex * ey));
            const LeafGeomD& g = lg[pr.leaf];
            double w6;
            if (kPrec == ADECO_PREC_FP64) {
                V3<double> s[4];
                load_s(pr, g, i, j, k, s);
                w6 = clip_unit_cube_vol6(s);
            } else {
                V3<float> s[4];
                load_s(pr, g, i, j, k, s);
                const float w32 = clip_unit_cube_vol6(s);
                w6 = double(w32);
                if (kPrec == ADECO_PREC_MIXED) {
                    float M = 1.f;
                    for (int q = 0; q < 4; ++q) M = fmaxf(M, fmaxf(fmaxf(fabsf(s[q].x), fabsf(s[q].y)), fabsf(s[q].z)));
                    const float bound = 16.f * 5.9604645e-8f * M * M;
                    if (!(bound <= 2.5e-7f * w32)) {
                        V3<double> sd[4];
                        load_s(pr, g, i, j, k, sd);
                        w6 = clip_unit_cube_vol6(sd);
                        ++my_fallback;
                    }
                }
            }
            const double dm = pr.mcell * (w6 * (1.0 / 6.0));
            if (dm > 0.0) unsafeAtomicAdd(grids + int64_t(pr.leaf) * cells + i + int64_t(dims) * (j + int64_t(dims) * k), dm);
        }
        __syncthreads();
    }
    if (kPrec == ADECO_PREC_MIXED && fallback && my_fallback) atomicAdd(fallback, my_fallback);
}

}  // namespace

// Builds the candidate space for the current binning (pairs x covered cells) and runs the
// persistent deposit.  Scratch buffers live in the octree handle and are reused across calls.
adeco_status deposit_leaves_batched(adeco_octree_s* T, int dims, int decomposition, double tet_mass, int precision) {
    adeco_lattice_s* L = T->L;
    DeviceState& ds = *L->ds;
    if (dims > 1023) return fail(ADECO_EINVAL, "batched deposit: dims must be < 1024");
    const int nl = int(T->leaves_dfs.size());
    const int64_t P = T->leaf_tet_offset[nl];
    if (P == 0) return ADECO_OK;
    if (P >= (int64_t(1) << 31) - 2) return fail(ADECO_EOVERFLOW, "batched deposit: too many (leaf, tet) pairs");
    std::vector<LeafGeomD> lg(nl);
    const int scaling = cell_scaling();
    for (int l = 0; l < nl; ++l) {
        const GridGeom g = T->leaf_geom(l, dims, scaling);
        for (int a = 0; a < 3; ++a) {
            lg[l].gmin[a] = g.gmin[a];
            lg[l].delta[a] = g.delta[a];
            lg[l].r[a] = g.delta[a] * g.scale[a];
            lg[l].scale[a] = g.scale[a];
        }
        lg[l].cell_vol = g.cell_vol;
    }
    // layout of the work buffer
    const int64_t nblk = (P + kScanBlock - 1) / kScanBlock;
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_cand = take(P * 4), o_range = take(P * 8), o_leaf = take(P * 4), o_off = take((P + 1) * 8),
                 o_bs = take(nblk * 8 + 8), o_tot = take(8), o_ctr = take(16), o_fb = take(16);
    adeco_status st;
    if ((st = T->d_work.reserve(off)) || (st = T->d_leaf_geom.reserve(lg.size() * sizeof(LeafGeomD)))) return st;
    char* w = T->d_work.as<char>();
    unsigned* cand = reinterpret_cast<unsigned*>(w + o_cand);
    unsigned long long* range = reinterpret_cast<unsigned long long*>(w + o_range);
    int32_t* pleaf = reinterpret_cast<int32_t*>(w + o_leaf);
    int64_t* coff = reinterpret_cast<int64_t*>(w + o_off);
    int64_t* bs = reinterpret_cast<int64_t*>(w + o_bs);
    int64_t* tot = reinterpret_cast<int64_t*>(w + o_tot);
    unsigned* ctr = reinterpret_cast<unsigned*>(w + o_ctr);
    unsigned long long* fb = reinterpret_cast<unsigned long long*>(w + o_fb);
    hipError_t e = hipMemcpyAsync(T->d_leaf_geom.ptr, lg.data(), lg.size() * sizeof(LeafGeomD), hipMemcpyHostToDevice, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "leaf geometry upload");
    const LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    const PackedCorners ct = pack_corners(corner_table(decomposition));
    const LeafGeomD* dlg = T->d_leaf_geom.as<LeafGeomD>();
    hipLaunchKernelGGL(k_pair_setup, dim3(unsigned((P + 255) / 256)), dim3(256), 0, ds.stream, lv, ct, T->d_pairs.as<int32_t>(), P,
                       T->d_leaf_offset.as<int64_t>(), nl, dlg, dims, cand, range, pleaf);
    hipLaunchKernelGGL(k_scan_local, dim3(unsigned(nblk)), dim3(256), 0, ds.stream, cand, P, coff, bs);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, ds.stream, bs, nblk, tot);
    hipLaunchKernelGGL(k_scan_add, dim3(unsigned((P + 1 + 255) / 256)), dim3(256), 0, ds.stream, coff, P, bs, tot);
    int64_t C = 0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&C, tot, 8, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched setup");
    const int64_t n_tiles = (C + kTile - 1) / kTile;
    if ((st = T->d_tiles.reserve(size_t(n_tiles + 1) * 4))) return st;
    hipLaunchKernelGGL(k_tile_map, dim3(unsigned((P + 1 + 255) / 256)), dim3(256), 0, ds.stream, coff, P,
                       T->d_tiles.as<int32_t>(), n_tiles);
    e = hipMemsetAsync(ctr, 0, 16, ds.stream);
    if (e == hipSuccess) e = hipMemsetAsync(fb, 0, 16, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched counters");
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, L->device);
    const unsigned grid = unsigned(std::min<int64_t>(n_tiles, int64_t(cus) * 4));
#define ADECO_LAUNCH_BATCHED(P_)                                                                                      \
    hipLaunchKernelGGL(k_deposit_batched<P_>, dim3(grid), dim3(kTile), 0, ds.stream, T->d_pairs.as<int32_t>(), pleaf,   \
                       range, coff, T->d_tiles.as<int32_t>(), n_tiles, C, ctr, lv, ct, dlg, dims, tet_mass,             \
                       T->d_grids.as<double>(), fb)
    e = hipEventRecord(ds.evk0, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    if (precision == ADECO_PREC_FP64) ADECO_LAUNCH_BATCHED(ADECO_PREC_FP64);
    else if (precision == ADECO_PREC_MIXED) ADECO_LAUNCH_BATCHED(ADECO_PREC_MIXED);
    else ADECO_LAUNCH_BATCHED(ADECO_PREC_FP32);
#undef ADECO_LAUNCH_BATCHED
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ds.evk1, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched deposit launch");
    T->last_candidates = C;
    T->d_fallback = fb;
    return ADECO_OK;
}

}  // namespace adeco
