// Batched persistent deposit: one launch over every (grid, tet, candidate) of a candidate space.
//
// The plain form gives every lane one tet and lets it walk its covered cells, so a wave runs as
// long as its largest tet and leaves most lanes idle on heavy-tailed leaves.  Here the unit of
// work is one (tet, candidate) pair, and every wavefront works on its own, with no block barriers.
// A candidate is a column segment: up to kMaxSeg consecutive covered cells along the tet's longest
// axis (FP64 on grids of unit scaled cells), else a single cell.
//   1. k_pair_setup     one lane per binned (grid, tet) pair: signed volume and orientation
//                       (v0 <-> v3 when negative, AdDepositTets2Grid.h:547-558), vertices scaled to
//                       the grid's cell units in FP64, mass factor, covered cell range
//                       (AdDepositTets2Grid.h:514-539), column axis -> one 128-byte PairData line
//                       and a candidate count;
//   2. scan             candidate counts -> 64-bit offsets, total C;
//   3. k_tile_map       first pair of every tile of 64 consecutive candidates;
//   4. k_deposit_waves  persistent wavefronts pop chunks of tiles from a work-queue counter.  Per
//                       tile each lane finds its candidate's pair (a marker per pair start and a
//                       wave prefix-max), culls the segment by AABB and classifies it with the
//                       tet's 4 face planes at the extreme corners (a relative margin of 1e-12: an
//                       uncertain segment is clipped, never skipped or filled).  Segments inside the
//                       tet and cells holding the whole tet are deposited at once; partial ones go
//                       to one of 22 per-wave LDS queues (11 in a server step) keyed by their clip
//                       work class (number of x/y faces that cut the tet once the column axis - for
//                       a single cell the axis with the most cutting faces - is taken as the
//                       closed-form z axis, and how many vertices lie inside each of the clip tree's
//                       first two faces).  A queue that reaches 64 entries is clipped by the whole
//                       wave, so the lanes of one clip follow nearly the same path through the
//                       nested clip.
// One x/y clip serves all cells of a segment: every leaf piece is integrated over the z planes
// between the cells in closed form, each plane once for the two cells it separates, into per-lane
// LDS sums.  The clip uses the signed tree (a cut leaving 3 vertices inside gives the tet minus the
// corner beyond the plane: fewer pieces); a cell whose sum is below 1e-6 of the magnitudes it was
// summed from may be cancellation noise, so it goes alone to one more per-wave queue whose waves
// clip it with the exact tree (non-overlapping pieces, all terms positive).  Masses go to the grids
// with FP64 global atomics.  Precision modes for the clip:
//   FP64  all FP64;
//   MIXED FP32 clip on cell-local coordinates formed in FP64 (single cells), accepted only when its
//         error bound 16 * eps32 * max(1, M)^2 (M = largest |coordinate| of the cell-local tet;
//         twice the worst ratio measured over 2e5 random tets against FP64) is below 2.5e-7 of the
//         clipped volume and it is not noise, otherwise recomputed by the exact FP64 tree;
//   FP32  FP32 signed clip only (precision study).
// counters[0] counts the exact recomputes (cells), counters[1] the partial candidates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "../../include/adeco.h"
#include "batched_impl.hpp"
#include "deposit_tet.hpp"
#include "kernels.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

namespace adeco {

GridDesc make_grid_desc(const GridGeom& g, long long base) {
    GridDesc d;
    for (int a = 0; a < 3; ++a) {
        d.gmin[a] = g.gmin[a];
        d.delta[a] = g.delta[a];
        // a cell edge of exactly 1 lets neighbouring cells share their z planes (column segments);
        // it moves cell faces by at most an ulp of the index
        d.r[a] = g.scale[a] == 1.0 / g.delta[a] ? 1.0 : g.delta[a] * g.scale[a];
        d.scale[a] = g.scale[a];
    }
    d.cell_vol = g.cell_vol;
    d.base = base;
    d.dims = g.dims;
    d.pad_ = 0;
    return d;
}

LaunchShape& launch_shape() {
    static LaunchShape s;
    return s;
}

namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = 64 * kWavesPerBlock;
constexpr int kBuckets = 23;       // 22 clip work classes + the exact-recompute queue
constexpr int kExactQueue = 22;
// launches with fewer tiles than this per resident wave drain their queues in merged passes
constexpr int64_t kMergeDrainTiles = 16;
// how a wave drains its queues once the work items run out: a pass per queue (a large launch or a
// step of the request server), or passes filled from all the classes (a small launch)
constexpr int kDrainPerClass = 0, kDrainAll = 1;

// Per-wavefront LDS: the bucketed partial-candidate queues (64 entries each), the entries a tile
// could not place because their queue was full (at most one per lane, written back once that queue
// has been clipped), and each lane's per-cell sums of the column it clips; the pair-start markers
// of the tile being classified share the sums' storage (classification and clipping alternate).
// A queue entry is (pair, first cell | closed-form axis << 30) plus the candidate space it belongs
// to.  qn[b] counts queue b's entries plus its carried ones.
struct WaveLds {
    int qn[24];
    int2 q[kBuckets][kWaveTile];
    unsigned char qs[kBuckets][kWaveTile];
    int2 carry[kWaveTile];            // per lane: its carried entry,
    unsigned char carry_slot[kWaveTile];
    signed char carry_key[kWaveTile]; // its queue (-1: none)
    unsigned char carry_pos[kWaveTile];  // and its position once that queue has been clipped
    union {
        double acc[kMaxSeg][kWaveTile];   // signed mass sums (vol6 units) of the column's cells
        int mark[kWaveTile];
    };
    float sab[kMaxSeg][kWaveTile];    // sums of their magnitudes (the noise threshold's scale)
};
static_assert(kBuckets <= 24 && kBuckets < 32, "queue counts and the drain's prefix sum cover 24 lanes");

__device__ __forceinline__ int upper_bound_i64(const int64_t* a, int n, int64_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// per pair: PairData and candidate count (dead pairs keep one dummy candidate so every pair owns
// at least one candidate and a tile of 64 candidates spans at most 64 pairs).  With columns
// (max_seg > 1 and unit cells) a candidate is a segment of up to max_seg cells along the axis of
// the largest extent (ties -> z, then y); otherwise a single cell.
// (pair_grid: the grid of every pair, or null to search grid_off)
__device__ __forceinline__ void pair_setup_one(LatticeView lat, PackedCorners ct, const int32_t* __restrict__ pairs,
                                               int64_t p, const int64_t* __restrict__ grid_off, int n_grids,
                                               const int32_t* __restrict__ pair_grid, const GridDesc* __restrict__ grids,
                                               double tet_mass, int max_seg, unsigned* __restrict__ cand, PairData& o) {
    const int gi = pair_grid ? pair_grid[p] : upper_bound_i64(grid_off, n_grids + 1, p) - 1;
    V3<double> v[4];
    lattice_tet_vertices(lat, ct, pairs[p], v);
    const GridDesc& g = grids[gi];
    double mn[3] = {v[0].x, v[0].y, v[0].z}, mx[3] = {v[0].x, v[0].y, v[0].z};
    for (int q = 1; q < 4; ++q) {
        const double c[3] = {v[q].x, v[q].y, v[q].z};
        for (int a = 0; a < 3; ++a) {
            mn[a] = c[a] < mn[a] ? c[a] : mn[a];
            mx[a] = c[a] > mx[a] ? c[a] : mx[a];
        }
    }
    unsigned long long pack = 0;
    const double vol = vol6(v[0], v[1], v[2], v[3]) / 6.0;
    const int dm1 = g.dims - 1;
    unsigned ext[3];
    for (int a = 0; a < 3; ++a) {
        // cells whose box can strictly overlap the tet AABB.  The reference covers [floor, ceil]
        // (AdDepositTets2Grid.h:530-539), one cell more per axis than can pass its strict test
        // (AdDepositTets2Grid.h:598); the range here stops at floor, widened by 1e-9 of a cell each
        // way so that rounding in the division cannot drop a touching cell.  Clamping the overlap to
        // the grid first changes nothing after the clamp of the indices.
        int l = int(floor((mn[a] - g.gmin[a]) / g.delta[a] - 1e-9));
        int h = int(floor((mx[a] - g.gmin[a]) / g.delta[a] + 1e-9));
        l = l <= 0 ? 0 : (l >= dm1 ? dm1 : l);
        h = h <= 0 ? 0 : (h >= dm1 ? dm1 : h);
        const unsigned e = unsigned(h - l + 1);
        ext[a] = e;
        pack |= (unsigned long long)l << (10 * a);
        pack |= (unsigned long long)e << (30 + 10 * a);
    }
    const bool col = max_seg > 1 && g.r[0] == 1.0 && g.r[1] == 1.0 && g.r[2] == 1.0;
    const int axis = !col ? 2 : (ext[2] >= ext[1] && ext[2] >= ext[0]) ? 2 : (ext[1] >= ext[0] ? 1 : 0);
    const unsigned ea = axis == 0 ? ext[0] : (axis == 1 ? ext[1] : ext[2]);
    const unsigned seg = col ? unsigned(max_seg) : 1u;
    const unsigned n = (ext[0] * ext[1] * ext[2] / ea) * ((ea + seg - 1) / seg);
    pack |= (unsigned long long)axis << 60;
    pack |= (unsigned long long)(col ? 1 : 0) << 62;
    const bool live = vol != 0.0;
    if (vol < 0.0) {
        const V3<double> t = v[0];
        v[0] = v[3];
        v[3] = t;
    }
    for (int w = 0; w < 4; ++w) {
        o.S[3 * w] = g.scale[0] * (v[w].x - g.gmin[0]);
        o.S[3 * w + 1] = g.scale[1] * (v[w].y - g.gmin[1]);
        o.S[3 * w + 2] = g.scale[2] * (v[w].z - g.gmin[2]);
    }
    o.mcell = live ? tet_mass * (g.cell_vol / fabs(vol)) : 0.0;
    o.range = live ? pack : 0ull;  // extent 0 marks a dead pair
    o.grid = gi;
    o.pad_[0] = o.pad_[1] = o.pad_[2] = 0;
    cand[p] = live ? n : 1u;
}

__global__ void __launch_bounds__(256) k_pair_setup(LatticeView lat, PackedCorners ct, const int32_t* __restrict__ pairs,
                                                    int64_t n_pairs, const int64_t* __restrict__ n_dev,
                                                    const int64_t* __restrict__ grid_off, int n_grids,
                                                    const int32_t* __restrict__ pair_grid,
                                                    const GridDesc* __restrict__ grids, double tet_mass, int max_seg,
                                                    unsigned* __restrict__ cand, PairData* __restrict__ pd) {
    // the 128-byte lines are assembled in LDS (rows of 17 words: the odd stride spreads a wave's
    // row writes over the banks) and stored as whole contiguous rows of the block, non-temporal
    // (the 16 GB of lines of a large launch pass through no cache usefully)
    constexpr int kWords = int(sizeof(PairData) / 8);
    __shared__ unsigned long long rows[256 * (kWords + 1)];
    // the pair count: n_pairs, or the device's count capped at n_pairs (a uniform exit per block)
    if (n_dev) n_pairs = min(*n_dev, n_pairs);
    const int64_t base = int64_t(blockIdx.x) * blockDim.x;
    if (base >= n_pairs) return;
    const int64_t p = base + threadIdx.x;
    if (p < n_pairs) {
        PairData o;
        pair_setup_one(lat, ct, pairs, p, grid_off, n_grids, pair_grid, grids, tet_mass, max_seg, cand, o);
        const unsigned long long* w = reinterpret_cast<const unsigned long long*>(&o);
        for (int k = 0; k < kWords; ++k) rows[threadIdx.x * (kWords + 1) + k] = w[k];
    }
    __syncthreads();
    const int rows_here = int(min(int64_t(256), n_pairs - base));
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(pd + base);
    for (int k = threadIdx.x; k < rows_here * kWords; k += 256)
        __builtin_nontemporal_store(rows[(k / kWords) * (kWords + 1) + k % kWords], dst + k);
}

// grid index of every pair of grid-major pairs below cap: one workgroup per grid writes its range
__global__ void __launch_bounds__(256) k_pair_grid(const int64_t* __restrict__ grid_off, int64_t cap,
                                                   int32_t* __restrict__ pair_grid) {
    const int g = int(blockIdx.x);
    const int64_t end = min(grid_off[g + 1], cap);
    for (int64_t i = grid_off[g] + threadIdx.x; i < end; i += blockDim.x) pair_grid[i] = g;
}

// Exclusive scan of uint32 counts into int64 offsets out[0..n] (out[n] = the total), by
// reduce-then-scan: the counts' sums per block of kScanBlock (k_scan_reduce), one workgroup
// scanning those sums (k_scan_sums), then each block's local scan offset by its block's prefix
// (k_scan_apply): the counts are read twice and the offsets written once.  A lane takes 4
// consecutive counts; a block's 4 waves scan by shuffles and one exchange of wave totals.
// (n_dev: the count on the device, capped at n; blocks past it sum to zero and write nothing, except
// the one holding out[n_dev].)
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int64_t shfl_up_i64(int64_t v, int d) {
    const int lo = __shfl_up(int(uint64_t(v) & 0xffffffffu), d, 64);
    const int hi = __shfl_up(int(uint64_t(v) >> 32), d, 64);
    return int64_t((uint64_t(unsigned(hi)) << 32) | unsigned(lo));
}

__device__ __forceinline__ int64_t shfl_xor_i64(int64_t v, int m) {
    const int lo = __shfl_xor(int(uint64_t(v) & 0xffffffffu), m, 64);
    const int hi = __shfl_xor(int(uint64_t(v) >> 32), m, 64);
    return int64_t((uint64_t(unsigned(hi)) << 32) | unsigned(lo));
}

__global__ void __launch_bounds__(256) k_scan_reduce(const unsigned* __restrict__ in, int64_t n,
                                                     const int64_t* __restrict__ n_dev, int64_t* __restrict__ block_sum) {
    __shared__ int64_t wsum[4];
    if (n_dev) n = min(*n_dev, n);
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    const int t = int(threadIdx.x);
    if (base >= n) {
        if (t == 0) block_sum[blockIdx.x] = 0;
        return;
    }
    const int64_t i0 = base + 4 * t;
    int64_t v = 0;
    for (int k = 0; k < 4; ++k) v += i0 + k < n ? int64_t(in[i0 + k]) : 0;
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_i64(v, m);
    if ((t & 63) == 0) wsum[t >> 6] = v;
    __syncthreads();
    if (t == 0) block_sum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// (launched with n / kScanBlock + 1 blocks, so that the block holding out[n] exists)
__global__ void __launch_bounds__(256) k_scan_apply(const unsigned* __restrict__ in, int64_t n,
                                                    const int64_t* __restrict__ n_dev,
                                                    const int64_t* __restrict__ block_sum,
                                                    const int64_t* __restrict__ total, int64_t* __restrict__ out) {
    __shared__ int64_t wsum[4];
    if (n_dev) n = min(*n_dev, n);
    const int64_t base = int64_t(blockIdx.x) * kScanBlock;
    const int t = int(threadIdx.x), lane = t & 63, w = t >> 6;
    if (base >= n) {
        if (base == n && t == 0) out[n] = *total;
        return;
    }
    const int64_t i0 = base + 4 * t;
    int64_t c[4];
    for (int k = 0; k < 4; ++k) c[k] = i0 + k < n ? int64_t(in[i0 + k]) : 0;
    const int64_t own = c[0] + c[1] + c[2] + c[3];
    int64_t v = own;
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t o = shfl_up_i64(v, d);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    int64_t run = block_sum[blockIdx.x] + v - own;
    for (int k = 0; k < w; ++k) run += wsum[k];
    for (int k = 0; k < 4; ++k) {
        const int64_t i = i0 + k;
        if (i < n) out[i] = run;
        else if (i == n) out[n] = *total;
        run += c[k];
    }
}

// one workgroup of 1024 lanes: exclusive scan of the nb block sums.  Each lane owns a contiguous
// segment of ceil(nb / 1024) sums: it adds them up, the lanes' totals are scanned across the
// workgroup (wave shuffles, then the 16 wave totals), and each lane writes its segment's
// exclusive prefixes from its lane offset.  (Blocks past a device-side count hold zero, so
// scanning all nb of them is the same.)
constexpr int kScanSumsBlock = 1024;

__global__ void __launch_bounds__(kScanSumsBlock) k_scan_sums(int64_t* __restrict__ bs, int64_t nb,
                                                              int64_t* __restrict__ total) {
    __shared__ int64_t wsum[kScanSumsBlock / 64];
    const int t = int(threadIdx.x), lane = t & 63, w = t >> 6;
    const int64_t per = (nb + kScanSumsBlock - 1) / kScanSumsBlock;
    const int64_t lo = min(nb, int64_t(t) * per), hi = min(nb, lo + per);
    // (8 loads in flight per lane: the lanes' segments lie far apart, so each load waits on memory)
    int64_t own = 0;
    int64_t i = lo;
    for (; i + 8 <= hi; i += 8) {
        int64_t x[8];
        for (int k = 0; k < 8; ++k) x[k] = bs[i + k];
        for (int k = 0; k < 8; ++k) own += x[k];
    }
    for (; i < hi; ++i) own += bs[i];
    // inclusive scan of the lanes' totals: within the wave, then over the wave totals
    int64_t v = own;
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t o = shfl_up_i64(v, d);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    if (w == 0) {
        int64_t x = lane < kScanSumsBlock / 64 ? wsum[lane] : 0;
        for (int d = 1; d < kScanSumsBlock / 64; d <<= 1) {
            const int64_t o = shfl_up_i64(x, d);
            if (lane >= d) x += o;
        }
        if (lane < kScanSumsBlock / 64) wsum[lane] = x;
    }
    __syncthreads();
    int64_t run = (w > 0 ? wsum[w - 1] : 0) + v - own;
    for (i = lo; i + 8 <= hi; i += 8) {
        int64_t x[8];
        for (int k = 0; k < 8; ++k) x[k] = bs[i + k];
        for (int k = 0; k < 8; ++k) {
            bs[i + k] = run;
            run += x[k];
        }
    }
    for (; i < hi; ++i) {
        const int64_t x = bs[i];
        bs[i] = run;
        run += x;
    }
    if (t == kScanSumsBlock - 1) *total = wsum[kScanSumsBlock / 64 - 1];
}

// (sizes: when given, the pair count capped at n_pairs and the tile count sizes[1] come from the
// device; with no tiles the map is left alone)
__global__ void __launch_bounds__(256) k_tile_map(const int64_t* __restrict__ off, int64_t n_pairs,
                                                  const int64_t* __restrict__ n_dev, const int64_t* __restrict__ sizes,
                                                  int32_t* __restrict__ tile_first, int64_t n_tiles) {
    if (n_dev) n_pairs = min(*n_dev, n_pairs);
    if (sizes) n_tiles = sizes[1];
    if (n_tiles <= 0) return;
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p > n_pairs) return;
    if (p == n_pairs) {
        tile_first[n_tiles] = int32_t(n_pairs - 1);
        return;
    }
    const int64_t t0 = (off[p] + kWaveTile - 1) / kWaveTile, t1 = (off[p + 1] + kWaveTile - 1) / kWaveTile;
    for (int64_t t = t0; t < t1 && t < n_tiles; ++t) tile_first[t] = int32_t(p);
}

// Device-side sizes of a candidate space whose pair count only the device knows: sizes[0] = the
// candidates, sizes[1] = the tiles; a pair count above pair_cap or a tile count above tile_cap
// sets *overflow and leaves no tiles (the deposit then does nothing).
__global__ void k_space_sizes(const int64_t* __restrict__ off, const int64_t* __restrict__ n_dev, int64_t pair_cap,
                              int64_t tile_cap, int64_t* __restrict__ sizes, unsigned* __restrict__ overflow) {
    const int64_t n = *n_dev;
    const int64_t nc = n <= pair_cap ? off[n] : 0;
    const int64_t nt = (nc + kWaveTile - 1) / kWaveTile;
    const bool over = n > pair_cap || nt > tile_cap;
    sizes[0] = over ? 0 : nc;
    sizes[1] = over ? 0 : nt;
    if (over) *overflow = 1u;
}

template <class R>
__device__ __forceinline__ void cell_local(const PairData& pd, const GridDesc& g, int i, int j, int k, V3<R> s[4]) {
    const double ox = double(i) * g.r[0], oy = double(j) * g.r[1], oz = double(k) * g.r[2];
    for (int q = 0; q < 4; ++q) s[q] = mk(R(pd.S[3 * q] - ox), R(pd.S[3 * q + 1] - oy), R(pd.S[3 * q + 2] - oz));
}

// the axis given to the closed-form slab integration becomes z
template <class R>
__device__ __forceinline__ void rotate_axes(V3<R> s[4], int axis) {
    for (int q = 0; q < 4; ++q) {
        const V3<R> u = s[q];
        s[q] = axis == 0 ? mk(u.y, u.z, u.x) : (axis == 1 ? mk(u.z, u.x, u.y) : u);
    }
}

// Cells [c, c+1] of a column segment (in the rotated frame: x, y in [0,1], c < ls) that one face
// plane of the positively oriented tet t separates from it: bit c.  The face-plane part of
// classify_margin per cell; the plane offset moves by c * n.z from cell to cell, and the margin
// takes the largest coordinate over the segment.
__device__ __forceinline__ unsigned separated_cells(const V3<double> t[4], int ls) {
    const V3<double> e01 = sub(t[1], t[0]), e02 = sub(t[2], t[0]), e03 = sub(t[3], t[0]);
    const V3<double> e13 = sub(t[3], t[1]), e12 = sub(t[2], t[1]);
    unsigned sep = 0u;
    for (int f = 0; f < 4; ++f) {
        const V3<double> n = f == 0 ? cross(e01, e02) : (f == 1 ? cross(e02, e03) : (f == 2 ? cross(e03, e01) : cross(e13, e12)));
        const V3<double> o = f == 3 ? t[1] : t[0];
        const double v0 = o.x * n.x + o.y * n.y + o.z * n.z;
        // the cell corner p with the largest plane value (o - p).n: p.x, p.y in {0, 1}, p.z = c or c + 1
        const double neg = fmin(n.x, 0.0) + fmin(n.y, 0.0) + fmin(n.z, 0.0);
        const double tol = 1e-12 * (fabs(n.x) + fabs(n.y) + fabs(n.z)) *
                           (1.0 + fmax(fmax(fabs(o.x), fabs(o.y)), fabs(o.z) + double(ls)));
        for (int c = 0; c < ls; ++c) sep |= unsigned(v0 - double(c) * n.z - neg < -tol) << c;
    }
    return sep;
}

// Clip work class of a partial cell: the axis with the most faces cutting the tet is integrated in
// closed form (ties -> z); the class counts the remaining x/y faces that cut (from the cell-local
// AABB), whether one face of each axis cuts or both of one, and how many vertices lie inside the
// first two faces of the clip tree (22 classes; kFine false: 11 classes by the number of faces and
// the first face's count, for the request server's steps, which drain all their queues every step,
// so that fewer classes leave fewer partly filled drain passes).  The class only groups cells of
// similar clip work (lanes of a pass then take the same branches of the clip tree more often); the
// clip itself decides every face again.
// A column segment (force >= 0) keeps its axis: force.
template <bool kFine>
__device__ __forceinline__ int work_class(const V3<double> s[4], const double mn[3], const double mx[3], int force,
                                          int& axis) {
    const bool lx = mn[0] < 0.0, hx = mx[0] > 1.0, ly = mn[1] < 0.0, hy = mx[1] > 1.0, lz = mn[2] < 0.0, hz = mx[2] > 1.0;
    const int ax_ = int(lx) + int(hx), ay = int(ly) + int(hy), az = int(lz) + int(hz);
    // closed-form axis: most cutting faces, ties -> z, then x
    axis = force >= 0 ? force : (ax_ > az && ax_ >= ay) ? 0 : ((ay > az && ay > ax_) ? 1 : 2);
    // the clip's x and y after rotation: (y, z) for axis 0, (z, x) for axis 1, (x, y) for axis 2
    const bool lo0 = axis == 0 ? ly : (axis == 1 ? lz : lx), hi0 = axis == 0 ? hy : (axis == 1 ? hz : hx);
    const bool lo1 = axis == 0 ? lz : (axis == 1 ? lx : ly), hi1 = axis == 0 ? hz : (axis == 1 ? hx : hy);
    const int nx = int(lo0) + int(hi0), ny = int(lo1) + int(hi1);
    const int nact = nx + ny;
    if (nact == 0) return 0;
    if (nact == 1) return 1;
    // the first two faces of the clip tree (clip_pieces' canonical frame): X is the clip axis with
    // more cutting faces (ties -> x), Y the other; first X >= 0 (X <= 1 when that is X's only
    // cutting face), then Y's cutting face when each axis has one, else X <= 1
    const bool swap = ny > nx;
    // the clip's x and y as axes of the cell-local frame (see rotate_axes)
    const int xp = axis == 0 ? 1 : (axis == 1 ? 2 : 0), yp = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
    const int xax = swap ? yp : xp, yax = swap ? xp : yp;
    const bool xlo = swap ? lo1 : lo0, ylo = swap ? lo0 : lo1;
    const bool xy = nx == 1 && ny == 1;
    const int ax1 = xax, ax2 = xy ? yax : xax;
    const bool ge1 = xlo, ge2 = xy && ylo;
    // select comparison results, not coordinates: a select of coordinates becomes a select of
    // addresses into a stack copy of the tet
    int n1 = 0, n2 = 0;
    for (int q = 0; q < 4; ++q) {
        const bool gx = s[q].x > 0.0, gy = s[q].y > 0.0, gz = s[q].z > 0.0;
        const bool bx = s[q].x < 1.0, by = s[q].y < 1.0, bz = s[q].z < 1.0;
        n1 += int(ax1 == 0 ? (ge1 ? gx : bx) : (ax1 == 1 ? (ge1 ? gy : by) : (ge1 ? gz : bz)));
        n2 += int(ax2 == 0 ? (ge2 ? gx : bx) : (ax2 == 1 ? (ge2 ? gy : by) : (ge2 ? gz : bz)));
    }
    // vertices inside the first and second face (1-3 each; with both X faces cutting, every vertex
    // is inside one of them, so n1 + n2 >= 4)
    const int t1 = n1 <= 1 ? 0 : (n1 == 2 ? 1 : 2), t2 = n2 <= 1 ? 0 : (n2 == 2 ? 1 : 2);
    // coarse classes (2-10): the number of cutting faces and the first face's count only
    if (!kFine) return 3 * nact - 4 + t1;
    if (xy) return 2 + 3 * t1 + t2;                                           // 2-10
    if (nact == 2) return 11;                                                 // both X faces
    if (nact == 3) return 12 + (t1 == 0 ? 0 : (t1 == 1 ? (t2 <= 1 ? 1 : 2) : 3 + t2));  // 12-17
    return 18 + (t1 <= 1 ? 0 : 1 + t2);                                       // 18-21
}

// n / d for n < 2^30 and 1 <= d < 2^30 without the integer division sequence: the FP64 reciprocal
// refined once is accurate to ~2^-50, so the truncated product is the quotient or one less (exact
// multiples), fixed by one integer correction each way
__device__ __forceinline__ unsigned udiv_small(unsigned n, unsigned d) {
    const double dd = double(d);
    double r = __builtin_amdgcn_rcp(dd);
    r = fma(fma(-dd, r, 1.0), r, r);
    unsigned q = unsigned(double(n) * r);
    q -= unsigned(q * d > n);
    q += unsigned((q + 1) * d <= n);
    return q;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int64_t shfl_i64(int64_t v, int src) {
    const int lo = __shfl(int(uint64_t(v) & 0xffffffffu), src, 64);
    const int hi = __shfl(int(uint64_t(v) >> 32), src, 64);
    return int64_t((uint64_t(unsigned(hi)) << 32) | unsigned(lo));
}

// What a tile's classification needs first: its source slot and index, its first pair, the number
// of pairs it touches and, per lane, the candidate offset of pair p0 + lane.  Loaded one tile
// ahead, so that these dependent loads overlap the previous tile's work.
struct TileHead {
    int slot;
    int p0, np;
    int64_t t;
    int64_t my_off;
};

template <class Sources>
__device__ __forceinline__ TileHead load_head(const Sources& srcs, const int2* __restrict__ entries, int64_t item,
                                              int lane) {
    TileHead h;
    h.slot = 0;
    h.t = item;
    if (entries) {
        const int2 en = entries[item];
        h.slot = __builtin_amdgcn_readfirstlane(en.x);
        h.t = int64_t(__builtin_amdgcn_readfirstlane(en.y));
    }
    const TileSource& src = srcs.get(h.slot);
    h.p0 = src.tile_first[h.t];
    h.np = min(src.tile_first[h.t + 1] - h.p0 + 1, kWaveTile);
    h.my_off = lane < h.np ? src.cand_off[h.p0 + lane] : int64_t(0);
    return h;
}

// Classifies the 64 candidates of tile h.t of `src` (slot h.slot among the launch's sources):
// segments inside the tet and tets inside a single cell are deposited, partial candidates are
// appended to the wave's queues.
template <bool kFine>
__device__ __forceinline__ void classify_tile(const TileSource& src, const TileHead& h, WaveLds& L, int lane,
                                              unsigned long long& n_partial) {
    const int slot = h.slot, p0 = h.p0, np = h.np;
    const int64_t t = h.t;
    const int64_t c0 = t * kWaveTile;
    const int64_t my_off = h.my_off;
    L.mark[lane] = 0;
    wave_sync();
    // every pair owns >= 1 candidate, so pairs p0+1.. start at distinct positions 1..63
    if (lane >= 1 && lane < np && my_off - c0 < kWaveTile) L.mark[int(my_off - c0)] = lane;
    wave_sync();
    int slot_in_tile = L.mark[lane];
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(slot_in_tile, d, 64);
        if (lane >= d) slot_in_tile = max(slot_in_tile, o);
    }
    const int64_t off = shfl_i64(my_off, slot_in_tile);
    const int64_t c = c0 + lane;
    const int p = p0 + slot_in_tile;
    int key = -1, cell = 0;
    if (c < src.n_cand) {
        const PairData& pd = src.pd[p];
        const unsigned long long rg = pd.range;
        const unsigned ex = unsigned(rg >> 30) & 1023u;
        if (ex) {
            const unsigned ey = unsigned(rg >> 40) & 1023u, ez = unsigned(rg >> 50) & 1023u;
            const int pa = int(rg >> 60) & 3;
            const bool col = (rg >> 62) & 1ull;
            // candidate r -> (u fastest, v, segment) over the extents across and along the column
            // axis pa (u, v = the axes after pa, cyclically)
            const unsigned eu = pa == 0 ? ey : (pa == 1 ? ez : ex), ev = pa == 0 ? ez : (pa == 1 ? ex : ey);
            const unsigned ea = pa == 0 ? ex : (pa == 1 ? ey : ez);
            const unsigned lmax = col ? unsigned(kMaxSeg) : 1u;
            const unsigned r = unsigned(c - off);
            const unsigned rest = udiv_small(r, eu);
            const unsigned su = r - rest * eu;
            const unsigned sg = udiv_small(rest, ev);
            const unsigned sv = rest - sg * ev;
            const int ls = int(min(lmax, ea - sg * lmax));
            const int li = int(rg & 1023u), lj = int((rg >> 10) & 1023u), lk = int((rg >> 20) & 1023u);
            const int cu = (pa == 0 ? lj : (pa == 1 ? lk : li)) + int(su);
            const int cv = (pa == 0 ? lk : (pa == 1 ? li : lj)) + int(sv);
            const int ca = (pa == 0 ? li : (pa == 1 ? lj : lk)) + int(sg * lmax);
            const int i = pa == 0 ? ca : (pa == 1 ? cv : cu);
            const int j = pa == 0 ? cu : (pa == 1 ? ca : cv);
            const int k = pa == 0 ? cv : (pa == 1 ? cu : ca);
            const GridDesc& g = src.grids[pd.grid];
            V3<double> s[4];
            cell_local(pd, g, i, j, k, s);
            const double mn[3] = {fmin(fmin(s[0].x, s[1].x), fmin(s[2].x, s[3].x)),
                                  fmin(fmin(s[0].y, s[1].y), fmin(s[2].y, s[3].y)),
                                  fmin(fmin(s[0].z, s[1].z), fmin(s[2].z, s[3].z))};
            const double mx[3] = {fmax(fmax(s[0].x, s[1].x), fmax(s[2].x, s[3].x)),
                                  fmax(fmax(s[0].y, s[1].y), fmax(s[2].y, s[3].y)),
                                  fmax(fmax(s[0].z, s[1].z), fmax(s[2].z, s[3].z))};
            // strict AABB cull against the cell box [0, r] (a segment: ls cells of edge 1 along pa);
            // the reference then classifies and clips against the unit cube in scaled coordinates
            // (AdDepositTets2Grid.h:587-634)
            const double len = double(ls);
            const double bx = pa == 0 ? len * g.r[0] : g.r[0], by = pa == 1 ? len * g.r[1] : g.r[1],
                         bz = pa == 2 ? len * g.r[2] : g.r[2];
            const bool hit = mn[0] < bx && 0.0 < mx[0] && mn[1] < by && 0.0 < mx[1] && mn[2] < bz && 0.0 < mx[2];
            if (hit) {
                const double box[3] = {pa == 0 ? len : 1.0, pa == 1 ? len : 1.0, pa == 2 ? len : 1.0};
                const int cls = classify_margin(s, box);
                const int64_t idx = g.base + i + int64_t(g.dims) * (j + int64_t(g.dims) * k);
                if (cls == CLS_TET_IN_CELL && ls == 1) {
                    unsafeAtomicAdd(src.grid_buf + idx, src.tet_mass);
                } else if (cls == CLS_CELL_IN_TET) {
                    const int64_t stride = pa == 0 ? 1 : (pa == 1 ? int64_t(g.dims) : int64_t(g.dims) * g.dims);
                    for (int q = 0; q < ls; ++q) unsafeAtomicAdd(src.grid_buf + idx + q * stride, pd.mcell);
                } else if (cls != CLS_NONE) {
                    // a segment keeps the pair's axis (its length is measured along it); a single
                    // cell takes the axis with the most cutting faces
                    int axis;
                    key = work_class<kFine>(s, mn, mx, col && ea > 1 ? pa : -1, axis);
                    cell = int(unsigned(i) | (unsigned(j) << 10) | (unsigned(k) << 20) | (unsigned(axis) << 30));
                }
            }
        }
    }
    // append the partial cells to their queues.  Every queue holds < 64 entries here, so a count
    // reaches at most 127; an entry beyond 64 is carried (position - 64 once its queue is clipped).
    int ckey = -1;
    if (key >= 0) {
        const int pos = atomicAdd(&L.qn[key], 1);
        if (pos < kWaveTile) {
            L.q[key][pos] = make_int2(p, cell);
            L.qs[key][pos] = (unsigned char)slot;
        } else {
            ckey = key;
            L.carry[lane] = make_int2(p, cell);
            L.carry_slot[lane] = (unsigned char)slot;
            L.carry_pos[lane] = (unsigned char)(pos - kWaveTile);
        }
    }
    L.carry_key[lane] = (signed char)ckey;
    n_partial += __popcll(__ballot(key >= 0));
    wave_sync();
}

// Clips one queued partial candidate (a cell, or a column segment from its first cell along its
// axis) and deposits its cells, except those whose result may be wrong: a signed-tree cell sum that
// may be cancellation noise, or a mixed-precision FP32 result outside its error bound.  Returns
// the mask of those cells (bit c = c-th cell along the axis); each is recomputed alone by the
// exact tree from the exact queue, as a cell entry with axis code 3, unless the face-plane test
// shows that the tet misses it.  acc/sab: this lane's LDS sums.
template <int kPrec, class Sources>
__device__ __forceinline__ unsigned clip_entry(const Sources& srcs, int2 e, int slot, bool exact, double* acc,
                                               float* sab) {
    const TileSource& src = srcs.get(slot);
    const PairData& pd = src.pd[e.x];
    const unsigned u = unsigned(e.y);
    const int i = int(u & 1023u), j = int((u >> 10) & 1023u), k = int((u >> 20) & 1023u), code = int(u >> 30);
    const GridDesc& g = src.grids[pd.grid];
    // segment length: the covered cells from the first one on along the axis, at most kMaxSeg
    const unsigned long long rg = pd.range;
    int ls = 1;
    if (code != 3 && ((rg >> 62) & 1ull)) {
        const int lo = int((rg >> (10 * code)) & 1023u), ext = int((rg >> (30 + 10 * code)) & 1023u);
        const int ca = code == 0 ? i : (code == 1 ? j : k);
        ls = max(1, min(kMaxSeg, lo + ext - ca));
    }
    for (int c = 0; c < ls; ++c) {
        acc[c * kWaveTile] = 0.0;
        sab[c * kWaveTile] = 0.f;
    }
    unsigned redo = 0u, empty = 0u;
    int axis = code;
    if (kPrec == ADECO_PREC_FP64 || (kPrec == ADECO_PREC_MIXED && exact)) {
        V3<double> s[4];
        cell_local(pd, g, i, j, k, s);
        if (code == 3) {
            // a single cell: the axis with the most cutting faces goes to the closed form
            int cut[3];
            for (int a = 0; a < 3; ++a) {
                double lo = comp(s[0], a), hi = lo;
                for (int q = 1; q < 4; ++q) {
                    lo = fmin(lo, comp(s[q], a));
                    hi = fmax(hi, comp(s[q], a));
                }
                cut[a] = int(lo < 0.0) + int(hi > 1.0);
            }
            axis = (cut[0] > cut[2] && cut[0] >= cut[1]) ? 0 : ((cut[1] > cut[2] && cut[1] > cut[0]) ? 1 : 2);
        }
        rotate_axes(s, axis);
        // most cells whose signed sum is noise lie outside the tet (pieces that cancel there): a
        // cell that a face plane of the tet separates from it holds nothing
        const unsigned sep = exact ? 0u : separated_cells(s, ls);
        ColumnLeaf<double, kWaveTile, float> leaf{acc, sab, ls};
        if (kPrec == ADECO_PREC_FP64) clip_pieces<double, 7>(s, !exact, leaf);
        else clip_pieces<double, 6>(s, false, leaf);
        if (!exact)
            for (int c = 0; c < ls; ++c)
                redo |= unsigned(signed_is_noisy(acc[c * kWaveTile], double(sab[c * kWaveTile]))) << c;
        empty = redo & sep;
        redo &= ~empty;
    } else {
        V3<float> s[4];
        cell_local(pd, g, i, j, k, s);
        rotate_axes(s, axis);
        ColumnLeaf<float, kWaveTile, float> leaf{acc, sab, ls};
        clip_pieces<float, 7>(s, true, leaf);
        if (kPrec == ADECO_PREC_MIXED) {
            float M = 1.f;
            for (int q = 0; q < 4; ++q) M = fmaxf(M, fmaxf(fmaxf(fabsf(s[q].x), fabsf(s[q].y)), fabsf(s[q].z)));
            const float bound = 16.f * 5.9604645e-8f * M * M;
            for (int c = 0; c < ls; ++c) {
                const float w32 = float(acc[c * kWaveTile]);
                const bool bad = !(bound <= 2.5e-7f * w32) || signed_is_noisy(w32, float(sab[c * kWaveTile]));
                redo |= unsigned(bad) << c;
            }
        }
    }
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? int64_t(g.dims) : int64_t(g.dims) * g.dims);
    double* cell0 = src.grid_buf + g.base + i + int64_t(g.dims) * (j + int64_t(g.dims) * k);
    for (int c = 0; c < ls; ++c) {
        const double dm = pd.mcell * (acc[c * kWaveTile] * (1.0 / 6.0));
        if (!(((redo | empty) >> c) & 1u) && dm > 0.0) unsafeAtomicAdd(cell0 + c * stride, dm);
    }
    return redo;
}

struct OneSource {
    TileSource s;
    __device__ const TileSource& get(int) const { return s; }
};
struct ManySources {
    const TileSource* s;
    __device__ const TileSource& get(int i) const { return s[i]; }
};

// The persistent loop of one wavefront over work items [0, n_items): item -> (source slot, tile)
// directly (one source) or through `entries`.  The wave pops `chunk` items per work-queue access;
// a queue reaching 64 entries is clipped at once; when the items run out the queues are drained.
// Every wave leaves the loop once the counter passes n_items and its queues are empty.
template <int kPrec, int kDrain, bool kFine, class Sources>
__device__ __forceinline__ void wave_loop(const Sources& srcs, const int2* __restrict__ entries, int64_t n_items,
                                          int chunk, unsigned* __restrict__ queue_counter,
                                          unsigned long long* __restrict__ counters) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const int lane = threadIdx.x & 63;
    WaveLds& L = lds[threadIdx.x >> 6];
    if (lane < 24) L.qn[lane] = 0;
    L.carry_key[lane] = -1;
    wave_sync();
    unsigned long long n_partial = 0, fb = 0;
    int64_t it = 0, it_end = 0;
    TileHead head;  // of item `it` while it < it_end
    head.slot = head.p0 = head.np = 0;
    head.t = head.my_off = 0;
    bool draining = false;
    // cells of the last clip pass left to the exact queue: this lane's entry and the cells' mask
    unsigned pend = 0u;
    int2 pe = make_int2(0, 0);
    int pslot = 0;
    for (;;) {
        // pending cells go to the exact queue, one per lane and round, up to its 64 entries; a full
        // exact queue is clipped first (exact clips leave no cells behind)
        const unsigned long long pm = __ballot(pend != 0u);
        const int nexact = pm ? __builtin_amdgcn_readfirstlane(L.qn[kExactQueue]) : kWaveTile;
        if (nexact < kWaveTile) {
            const int room = kWaveTile - nexact;
            const int rank =
                int(__builtin_amdgcn_mbcnt_hi(unsigned(pm >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(pm), 0u)));
            if (pend && rank < room) {
                const int c = __builtin_ctz(pend);
                pend &= pend - 1u;
                const unsigned u = unsigned(pe.y);
                const int ax = int(u >> 30);
                const unsigned cell = (u & 0x3fffffffu) + (unsigned(c) << (10 * ax));
                L.q[kExactQueue][nexact + rank] = make_int2(pe.x, int(cell | (3u << 30)));
                L.qs[kExactQueue][nexact + rank] = (unsigned char)pslot;
                ++fb;
            }
            wave_sync();
            if (lane == 0) L.qn[kExactQueue] = nexact + min(int(__popcll(pm)), room);
            wave_sync();
            continue;
        }
        // a queue to clip: full (64) while tiles remain, any non-empty one when draining (or when
        // cells are pending); the exact queue first
        const int need = draining ? 1 : kWaveTile;
        const int qv = lane < kBuckets ? L.qn[lane] : 0;
        const unsigned long long ready = __ballot(qv >= need);
        if (ready) {
            const int b = (ready >> kExactQueue) & 1ull ? kExactQueue : __builtin_ctzll(ready);
            const int cnt = __builtin_amdgcn_readfirstlane(__shfl(qv, b, 64));
            // this pass's entry of each lane: from queue b, or when draining a small launch with no
            // full queue, up to 64 entries across all the work classes (they only group similar
            // work; the clip does not depend on them), from the top of each queue, lowest class
            // first.  Either way a pass takes >= 1 entry, so the drain ends.
            int2 e = make_int2(0, 0);
            int slot = 0;
            bool have = false;
            if (kDrain == kDrainAll && draining && b != kExactQueue && cnt < kWaveTile) {
                const bool merge = lane < kExactQueue;
                const int qc = merge ? qv : 0;
                int pre = qc;
                for (int d = 1; d < 32; d <<= 1) {
                    const int o = __shfl_up(pre, d, 64);
                    if (lane >= d) pre += o;
                }
                const int excl = pre - qc;
                const int tq = merge ? max(0, min(qc, kWaveTile - excl)) : 0;
                for (int qi = 0; qi < kExactQueue; ++qi) {
                    const int ex = __shfl(excl, qi, 64), tk = __shfl(tq, qi, 64), cq = __shfl(qc, qi, 64);
                    if (lane >= ex && lane < ex + tk) {
                        e = L.q[qi][cq - tk + (lane - ex)];
                        slot = L.qs[qi][cq - tk + (lane - ex)];
                        have = true;
                    }
                }
                wave_sync();
                if (merge) L.qn[lane] = qc - tq;
            } else {
                // the queue's entries; what the count holds beyond 64 are carried entries, which
                // take the freed positions
                const int take = min(cnt, kWaveTile);
                const int rest = cnt - take;
                if (lane < take) {
                    e = L.q[b][lane];
                    slot = L.qs[b][lane];
                    have = true;
                }
                wave_sync();
                if (lane == 0) L.qn[b] = rest;
                if (rest > 0 && L.carry_key[lane] == b) {
                    const int at = L.carry_pos[lane];
                    L.q[b][at] = L.carry[lane];
                    L.qs[b][at] = L.carry_slot[lane];
                    L.carry_key[lane] = -1;
                }
            }
            wave_sync();
            // (a pass over another queue starts only with no cells pending; exact clips leave none)
            if (have) {
                const unsigned redo =
                    clip_entry<kPrec>(srcs, e, slot, b == kExactQueue, &L.acc[0][lane], &L.sab[0][lane]);
                if (redo) {
                    pend = redo;
                    pe = e;
                    pslot = slot;
                }
            }
            wave_sync();
            continue;
        }
        if (draining) break;
        if (it == it_end) {
            unsigned base = 0;
            if (lane == 0) base = atomicAdd(queue_counter, 1u);
            base = __builtin_amdgcn_readfirstlane(base);
            it = int64_t(base) * chunk;
            it_end = min(it + int64_t(chunk), n_items);
            if (it >= n_items) {
                draining = true;
                continue;
            }
            head = load_head(srcs, entries, it, lane);
        }
        const TileHead cur = head;
        ++it;
        if (it < it_end) head = load_head(srcs, entries, it, lane);
        classify_tile<kFine>(srcs.get(cur.slot), cur, L, lane, n_partial);
    }
    if (lane == 0 && n_partial) atomicAdd(counters + 1, n_partial);
    if (fb) atomicAdd(counters, fb);
}

template <int kPrec, bool kMergeDrain>
__global__ void __launch_bounds__(kBlock, 2) k_deposit_waves(TileSource src, int chunk, unsigned* __restrict__ queue_counter,
                                                          unsigned long long* __restrict__ counters) {
    if (src.dev_sizes) {
        src.n_cand = src.dev_sizes[0];
        src.n_tiles = src.dev_sizes[1];
    }
    wave_loop<kPrec, kMergeDrain ? kDrainAll : kDrainPerClass, true>(OneSource{src}, nullptr, src.n_tiles, chunk,
                                                                queue_counter, counters);
}

template <int kPrec, bool kMergeDrain>
__global__ void __launch_bounds__(kBlock, 2) k_deposit_step(const TileSource* __restrict__ sources,
                                                         const int2* __restrict__ entries, int n_entries, int chunk,
                                                         unsigned* __restrict__ queue_counter,
                                                         unsigned long long* __restrict__ counters) {
    wave_loop<kPrec, kMergeDrain ? kDrainAll : kDrainPerClass, false>(ManySources{sources}, entries, n_entries, chunk,
                                                             queue_counter, counters);
}

__global__ void k_zero_words(unsigned* p, int n) {
    const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Resident workgroups of a launch: blocks_per_cu per CU, capped at what the kernel's registers
// and LDS let a CU hold (a surplus workgroup would only start once a resident one had drained the
// work queue).
int resident_blocks(int device, int precision, bool merge) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus * std::max(1, std::min(launch_shape().blocks_per_cu, deposit_occupancy(precision, merge)));
}

}  // namespace

int deposit_occupancy(int precision, bool merge) {
    // per (precision, merge) instantiation of k_deposit_waves; k_deposit_step shares its resources
    static std::atomic<int> cache[3][2];
    const int pi = precision == ADECO_PREC_FP64 ? 0 : (precision == ADECO_PREC_MIXED ? 1 : 2);
    std::atomic<int>& n = cache[pi][merge ? 1 : 0];
    const int have = n.load(std::memory_order_relaxed);
    if (have > 0) return have;
    int got = 0;
    hipError_t e = hipErrorInvalidValue;
    const void* f = nullptr;
    if (pi == 0) f = merge ? reinterpret_cast<const void*>(k_deposit_waves<ADECO_PREC_FP64, true>)
                           : reinterpret_cast<const void*>(k_deposit_waves<ADECO_PREC_FP64, false>);
    else if (pi == 1) f = merge ? reinterpret_cast<const void*>(k_deposit_waves<ADECO_PREC_MIXED, true>)
                                : reinterpret_cast<const void*>(k_deposit_waves<ADECO_PREC_MIXED, false>);
    else f = merge ? reinterpret_cast<const void*>(k_deposit_waves<ADECO_PREC_FP32, true>)
                   : reinterpret_cast<const void*>(k_deposit_waves<ADECO_PREC_FP32, false>);
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&got, f, kBlock, 0);
    const int v = (e == hipSuccess && got > 0) ? got : 1;
    n.store(v, std::memory_order_relaxed);  // concurrent first calls store the same value
    return v;
}

size_t candidate_space_bytes(int64_t n_pairs) {
    const int64_t nblk = (n_pairs + kScanBlock - 1) / kScanBlock;
    return align256(size_t(n_pairs) * sizeof(PairData)) + align256(n_pairs * 4) + align256((n_pairs + 1) * 8) +
           align256(nblk * 8 + 8) + align256(8) + align256(n_pairs * 4);
}

size_t tile_map_bytes(int64_t n_cand) { return align256(size_t((n_cand + kWaveTile - 1) / kWaveTile + 1) * 4); }

size_t scan_block_words(int64_t n) { return size_t((n + kScanBlock - 1) / kScanBlock + 1); }

hipError_t launch_scan_u32(const unsigned* in, int64_t n, int64_t* out, int64_t* block_sum, int64_t* total,
                           hipStream_t stream) {
    const int64_t nblk = (n + kScanBlock - 1) / kScanBlock;
    if (nblk > 0)
        hipLaunchKernelGGL(k_scan_reduce, dim3(unsigned(nblk)), dim3(256), 0, stream, in, n, (const int64_t*)nullptr,
                           block_sum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanSumsBlock), 0, stream, block_sum, nblk, total);  // (nblk 0: total 0)
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(n / kScanBlock + 1)), dim3(256), 0, stream, in, n,
                       (const int64_t*)nullptr, block_sum, total, out);
    return hipGetLastError();
}

// Lays out the slab of a candidate space for up to n_pairs pairs and enqueues pair setup and the
// candidate scan; n_dev (nullable) is the pair count on the device, capped at n_pairs.
static void enqueue_candidate_space(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                    const int32_t* pairs, int64_t n_pairs, const int64_t* n_dev, const int64_t* grid_off,
                                    int n_grids, const GridDesc* d_grids, double tet_mass, int precision, void* slab,
                                    CandidateSpace& cs) {
    char* w = static_cast<char*>(slab);
    const int64_t nblk = (n_pairs + kScanBlock - 1) / kScanBlock;
    cs.pd = reinterpret_cast<PairData*>(w);
    w += align256(size_t(n_pairs) * sizeof(PairData));
    cs.cand = reinterpret_cast<unsigned*>(w);
    w += align256(n_pairs * 4);
    cs.cand_off = reinterpret_cast<int64_t*>(w);
    w += align256((n_pairs + 1) * 8);
    cs.block_sum = reinterpret_cast<int64_t*>(w);
    w += align256(nblk * 8 + 8);
    cs.total = reinterpret_cast<int64_t*>(w);
    w += align256(8);
    int32_t* pair_grid = n_grids > 1 ? reinterpret_cast<int32_t*>(w) : nullptr;
    cs.n_pairs = n_pairs;
    cs.n_cand = 0;
    cs.n_tiles = 0;
    if (n_pairs == 0) return;
    if (pair_grid)
        hipLaunchKernelGGL(k_pair_grid, dim3(unsigned(n_grids)), dim3(256), 0, stream, grid_off, n_pairs, pair_grid);
    hipLaunchKernelGGL(k_pair_setup, dim3(unsigned((n_pairs + 255) / 256)), dim3(256), 0, stream, lat, ct, pairs, n_pairs,
                       n_dev, grid_off, n_grids, pair_grid, d_grids, tet_mass, precision == ADECO_PREC_FP64 ? kMaxSeg : 1,
                       cs.cand, cs.pd);
    hipLaunchKernelGGL(k_scan_reduce, dim3(unsigned(nblk)), dim3(256), 0, stream, cs.cand, n_pairs, n_dev, cs.block_sum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanSumsBlock), 0, stream, cs.block_sum, nblk, cs.total);
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(n_pairs / kScanBlock + 1)), dim3(256), 0, stream, cs.cand, n_pairs,
                       n_dev, cs.block_sum, cs.total, cs.cand_off);
}

adeco_status build_candidate_space(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                   const int32_t* pairs, int64_t n_pairs, const int64_t* grid_off, int n_grids,
                                   const GridDesc* d_grids, double tet_mass, int precision, void* slab,
                                   CandidateSpace& cs) {
    enqueue_candidate_space(stream, lat, ct, pairs, n_pairs, nullptr, grid_off, n_grids, d_grids, tet_mass, precision,
                            slab, cs);
    if (n_pairs == 0) return ADECO_OK;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&cs.n_cand, cs.total, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "candidate space setup");
    cs.n_tiles = (cs.n_cand + kWaveTile - 1) / kWaveTile;
    return ADECO_OK;
}

adeco_status build_tile_map(hipStream_t stream, CandidateSpace& cs, int32_t* tiles) {
    cs.tile_first = tiles;
    if (cs.n_pairs == 0) return ADECO_OK;
    hipLaunchKernelGGL(k_tile_map, dim3(unsigned((cs.n_pairs + 1 + 255) / 256)), dim3(256), 0, stream, cs.cand_off,
                       cs.n_pairs, (const int64_t*)nullptr, (const int64_t*)nullptr, tiles, cs.n_tiles);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "tile map");
}

hipError_t enqueue_candidate_space_dev(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                       const int32_t* pairs, int64_t pair_cap, const int64_t* n_dev,
                                       const int64_t* grid_off, int n_grids, const GridDesc* d_grids, double tet_mass,
                                       int precision, void* slab, int32_t* tiles, int64_t tile_cap, int64_t* sizes,
                                       unsigned* overflow, CandidateSpace& cs) {
    enqueue_candidate_space(stream, lat, ct, pairs, pair_cap, n_dev, grid_off, n_grids, d_grids, tet_mass, precision, slab,
                            cs);
    hipLaunchKernelGGL(k_space_sizes, dim3(1), dim3(1), 0, stream, cs.cand_off, n_dev, pair_cap, tile_cap, sizes, overflow);
    cs.tile_first = tiles;
    hipLaunchKernelGGL(k_tile_map, dim3(unsigned((pair_cap + 1 + 255) / 256)), dim3(256), 0, stream, cs.cand_off, pair_cap,
                       n_dev, (const int64_t*)sizes, tiles, tile_cap);
    return hipGetLastError();
}

hipError_t launch_deposit_tiles(hipStream_t stream, int precision, const TileSource& src, unsigned* queue_counter,
                                unsigned long long* counters, int device) {
    if (src.n_tiles <= 0) return hipSuccess;
    // small launches: fewer tiles per pop when there are too few to give every resident wave
    // `chunk` of them, and drain passes that merge the work classes
    // (the merged-drain instantiation decides with the large one's occupancy; it is the same or
    // lower, and resident_blocks caps the grid with the chosen instantiation's own)
    const int64_t resident_waves = int64_t(resident_blocks(device, precision, false)) * kWavesPerBlock;
    const int chunk = int(std::max<int64_t>(1, std::min<int64_t>(launch_shape().chunk, src.n_tiles / resident_waves)));
    const bool merge = src.n_tiles < kMergeDrainTiles * resident_waves;
    const int64_t waves_needed = (src.n_tiles + chunk - 1) / chunk;
    const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>((waves_needed + kWavesPerBlock - 1) / kWavesPerBlock,
                                                                          resident_blocks(device, precision, merge))));
    const dim3 g3(grid), b3(kBlock);
    if (precision == ADECO_PREC_FP64) {
        if (merge)
            hipLaunchKernelGGL((k_deposit_waves<ADECO_PREC_FP64, true>), g3, b3, 0, stream, src, chunk, queue_counter, counters);
        else
            hipLaunchKernelGGL((k_deposit_waves<ADECO_PREC_FP64, false>), g3, b3, 0, stream, src, chunk, queue_counter, counters);
    } else if (precision == ADECO_PREC_MIXED) {
        if (merge)
            hipLaunchKernelGGL((k_deposit_waves<ADECO_PREC_MIXED, true>), g3, b3, 0, stream, src, chunk, queue_counter, counters);
        else
            hipLaunchKernelGGL((k_deposit_waves<ADECO_PREC_MIXED, false>), g3, b3, 0, stream, src, chunk, queue_counter, counters);
    } else {
        if (merge)
            hipLaunchKernelGGL((k_deposit_waves<ADECO_PREC_FP32, true>), g3, b3, 0, stream, src, chunk, queue_counter, counters);
        else
            hipLaunchKernelGGL((k_deposit_waves<ADECO_PREC_FP32, false>), g3, b3, 0, stream, src, chunk, queue_counter, counters);
    }
    return hipGetLastError();
}

hipError_t launch_deposit_step(hipStream_t stream, int precision, const TileSource* d_sources, const int2* d_entries,
                               int n_entries, unsigned* queue_counter, unsigned long long* counters, int device) {
    if (n_entries <= 0) return hipSuccess;
    // small steps as small launches (launch_deposit_tiles)
    const int64_t resident_waves = int64_t(resident_blocks(device, precision, false)) * kWavesPerBlock;
    const int chunk = int(std::max<int64_t>(1, std::min<int64_t>(launch_shape().chunk, n_entries / resident_waves)));
    const bool merge = n_entries < kMergeDrainTiles * resident_waves;
    const int64_t waves_needed = (int64_t(n_entries) + chunk - 1) / chunk;
    const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>((waves_needed + kWavesPerBlock - 1) / kWavesPerBlock,
                                                                          resident_blocks(device, precision, merge))));
    const dim3 g3(grid), b3(kBlock);
    if (precision == ADECO_PREC_FP64) {
        if (merge)
            hipLaunchKernelGGL((k_deposit_step<ADECO_PREC_FP64, true>), g3, b3, 0, stream, d_sources, d_entries, n_entries,
                               chunk, queue_counter, counters);
        else
            hipLaunchKernelGGL((k_deposit_step<ADECO_PREC_FP64, false>), g3, b3, 0, stream, d_sources, d_entries, n_entries,
                               chunk, queue_counter, counters);
    } else if (precision == ADECO_PREC_MIXED) {
        if (merge)
            hipLaunchKernelGGL((k_deposit_step<ADECO_PREC_MIXED, true>), g3, b3, 0, stream, d_sources, d_entries, n_entries,
                               chunk, queue_counter, counters);
        else
            hipLaunchKernelGGL((k_deposit_step<ADECO_PREC_MIXED, false>), g3, b3, 0, stream, d_sources, d_entries, n_entries,
                               chunk, queue_counter, counters);
    } else {
        if (merge)
            hipLaunchKernelGGL((k_deposit_step<ADECO_PREC_FP32, true>), g3, b3, 0, stream, d_sources, d_entries, n_entries,
                               chunk, queue_counter, counters);
        else
            hipLaunchKernelGGL((k_deposit_step<ADECO_PREC_FP32, false>), g3, b3, 0, stream, d_sources, d_entries, n_entries,
                               chunk, queue_counter, counters);
    }
    return hipGetLastError();
}

// Candidate space of an octree's leaf deposit in caller-owned buffers: leaf geometry (geom),
// pair data and scan (work), tile map (tiles); grids are the leaf grids (dims^3 FP64 per leaf, DFS
// leaf order).  Pair setup, scan and tile map run on `stream`; n_cand is read back.
adeco_status prepare_leaves_batched(adeco_octree_s* T, hipStream_t stream, int dims, int decomposition, double tet_mass,
                                    int precision, DevBuf& geom, DevBuf& work, DevBuf& tiles, double* grids,
                                    TileSource& src) {
    adeco_lattice_s* L = T->L;
    if (dims > 1023) return fail(ADECO_EINVAL, "batched deposit: dims must be < 1024");
    const int nl = int(T->leaves_dfs.size());
    const int64_t P = T->leaf_tet_offset[nl];
    src = TileSource{nullptr, nullptr, nullptr, 0, 0, nullptr, grids, tet_mass};
    if (P == 0) return ADECO_OK;
    if (P >= (int64_t(1) << 31) - 2) return fail(ADECO_EOVERFLOW, "batched deposit: too many (leaf, tet) pairs");
    const int64_t cells = int64_t(dims) * dims * dims;
    std::vector<GridDesc> gd(nl);
    const int scaling = cell_scaling();
    for (int l = 0; l < nl; ++l) gd[l] = make_grid_desc(T->leaf_geom(l, dims, scaling), (long long)(l * cells));
    adeco_status st;
    if ((st = work.reserve(candidate_space_bytes(P) + 256)) || (st = geom.reserve(gd.size() * sizeof(GridDesc)))) return st;
    hipError_t e = hipMemcpyAsync(geom.ptr, gd.data(), gd.size() * sizeof(GridDesc), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "leaf geometry upload");
    const LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    const PackedCorners ct = pack_corners(corner_table(decomposition));
    CandidateSpace cs;
    // (synchronises the stream, so gd may go out of scope afterwards)
    if ((st = build_candidate_space(stream, lv, ct, T->d_pairs.as<int32_t>(), P, T->d_leaf_offset.as<int64_t>(), nl,
                                    geom.as<GridDesc>(), tet_mass, precision, work.ptr, cs)))
        return st;
    if ((st = tiles.reserve(tile_map_bytes(cs.n_cand)))) return st;
    if ((st = build_tile_map(stream, cs, tiles.as<int32_t>()))) return st;
    src = TileSource{cs.pd, cs.cand_off, cs.tile_first, cs.n_tiles, cs.n_cand, geom.as<GridDesc>(), grids, tet_mass};
    return ADECO_OK;
}

// Leaf deposit of an octree: one candidate space over all leaves, one persistent launch.
adeco_status deposit_leaves_batched(adeco_octree_s* T, int dims, int decomposition, double tet_mass, int precision) {
    adeco_lattice_s* L = T->L;
    DeviceState& ds = *L->ds;
    TileSource src;
    adeco_status st = prepare_leaves_batched(T, ds.stream, dims, decomposition, tet_mass, precision, T->d_leaf_geom, T->d_work,
                                             T->d_tiles, T->d_grids.as<double>(), src);
    if (st) return st;
    if (src.n_tiles == 0) return ADECO_OK;
    if ((st = T->d_counters.reserve(64))) return st;
    unsigned* ctr = T->d_counters.as<unsigned>();
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(T->d_counters.as<char>() + 16);
    hipError_t e = hipMemsetAsync(T->d_counters.ptr, 0, 64, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched counters");
    e = hipEventRecord(ds.evk0, ds.stream);
    if (e == hipSuccess) e = launch_deposit_tiles(ds.stream, precision, src, ctr, counters, L->device);
    if (e == hipSuccess) e = hipEventRecord(ds.evk1, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched deposit launch");
    T->last_candidates = src.n_cand;
    T->d_fallback = counters;
    return ADECO_OK;
}

// Clears n 32-bit words with a kernel (a memset node in a captured graph is avoided).
hipError_t launch_zero_u32(unsigned* p, int n, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_words, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace adeco
