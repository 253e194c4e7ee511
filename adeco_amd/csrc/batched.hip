// Batched persistent deposit: one launch over every (grid, tet, candidate cell) of a candidate space.
//
// The plain form gives every lane one tet and lets it walk its covered cells, so a wave runs as
// long as its largest tet and leaves most lanes idle on heavy-tailed leaves.  Here the unit of
// work is one (tet, candidate cell) pair:
//   1. k_pair_setup     one lane per binned (grid, tet) pair: covered cell range in its grid
//                       (AdDepositTets2Grid.h:514-539) -> candidate count;
//   2. scan             candidate counts -> 64-bit offsets, total C;
//   3. k_tile_map       first pair of every tile of kTile consecutive candidates;
//   4. process_tile     per tile: the tile's pairs (<= kTile + 1) are staged in LDS once (vertices
//                       scaled to the grid's cell units in FP64, mass factor, covered range); each
//                       lane binary-searches LDS for its candidate's pair, culls by AABB and
//                       classifies the cell with the 4 face planes evaluated at the extreme corners
//                       (a relative margin of 1e-12: an uncertain cell is clipped, never skipped or
//                       filled); trivial cells are deposited at once and partial cells are
//                       compacted into an LDS queue that all lanes then clip, so the exact clip runs
//                       with full waves.
// k_deposit_tiles (octree leaves) pulls the tiles of one candidate space from an atomic work
// counter; k_deposit_step (the request server) pulls (space, tile) entries spanning several
// candidate spaces, so requests of different sizes share one launch.  Masses go to the grids with
// FP64 global atomics.  Precision modes for the clip:
//   FP64  all FP64;
//   MIXED FP32 clip on cell-local coordinates formed in FP64, accepted only when its error bound
//         16 * eps32 * max(1, M)^2 (M = largest |coordinate| of the cell-local tet; twice the worst
//         ratio measured over 2e5 random tets against FP64) is below 2.5e-7 of the clipped
//         volume, otherwise recomputed in FP64;
//   FP32  FP32 clip only (precision study).
#include <hip/hip_runtime.h>

#include <algorithm>
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
        d.r[a] = g.delta[a] * g.scale[a];
        d.scale[a] = g.scale[a];
    }
    d.cell_vol = g.cell_vol;
    d.base = base;
    d.dims = g.dims;
    d.pad_ = 0;
    return d;
}

namespace {

struct PairLds {
    double S[12];  // scale_a * (src_q - gmin), oriented vertex order
    double mcell;  // tet_mass * cell_vol / |V|
    int grid;
    unsigned short lo[3], ext[3];  // covered cells: lo .. lo + ext - 1 (ext = 0: dead pair)
};

struct TileShared {
    PairLds pair[kTile + 1];
    int64_t off[kTile + 2];
    unsigned short qslot[kTile];
    unsigned qr[kTile];
    int qn, np;
    int64_t tile;
    int slot;
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

// per pair: grid, covered range, candidate count (dead pairs keep one dummy candidate so every
// tile spans at most kTile + 1 pairs)
__global__ void __launch_bounds__(256) k_pair_setup(LatticeView lat, PackedCorners ct, const int32_t* __restrict__ pairs,
                                                    int64_t n_pairs, const int64_t* __restrict__ grid_off, int n_grids,
                                                    const GridDesc* __restrict__ grids, unsigned* __restrict__ cand,
                                                    unsigned long long* __restrict__ range, int32_t* __restrict__ pair_grid) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int gi = upper_bound_i64(grid_off, n_grids + 1, p) - 1;
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
    const double vol = vol6(v[0], v[1], v[2], v[3]);
    const int dm1 = g.dims - 1;
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
    const bool live = vol != 0.0;
    cand[p] = live ? n : 1u;
    range[p] = live ? pack : 0ull;  // ext = 0 marks a dead pair
    pair_grid[p] = gi;
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
        if (base + i < n) out[base + i] = s[i] - int64_t(in[base + i]);
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
__device__ __forceinline__ void load_s(const PairLds& pr, const GridDesc& g, int i, int j, int k, V3<R> s[4]) {
    const double ox = double(i) * g.r[0], oy = double(j) * g.r[1], oz = double(k) * g.r[2];
    for (int q = 0; q < 4; ++q) s[q] = mk(R(pr.S[3 * q] - ox), R(pr.S[3 * q + 1] - oy), R(pr.S[3 * q + 2] - oz));
}

// Conservative classification of the cell-local tet s against [0,r] (FP64).
__device__ __forceinline__ int classify_margin(const V3<double> s[4], const double* r) {
    const V3<double> e01 = sub(s[1], s[0]), e02 = sub(s[2], s[0]), e03 = sub(s[3], s[0]);
    const V3<double> e13 = sub(s[3], s[1]), e12 = sub(s[2], s[1]);
    const V3<double> n[4] = {cross(e01, e02), cross(e02, e03), cross(e03, e01), cross(e13, e12)};
    bool all_in = true;
    for (int f = 0; f < 4; ++f) {
        const V3<double> o = f == 3 ? s[1] : s[0];
        const double v0 = o.x * n[f].x + o.y * n[f].y + o.z * n[f].z;
        // plane value (o - p).n at the cube corners p in {0, r}: extremes from the normal's signs
        const double px = n[f].x * r[0], py = n[f].y * r[1], pz = n[f].z * r[2];
        const double pos = fmax(px, 0.0) + fmax(py, 0.0) + fmax(pz, 0.0);
        const double neg = fmin(px, 0.0) + fmin(py, 0.0) + fmin(pz, 0.0);
        const double scl = (fabs(n[f].x) + fabs(n[f].y) + fabs(n[f].z)) *
                           (1.0 + fmax(fmax(fabs(o.x), fabs(o.y)), fabs(o.z)));
        const double tol = 1e-12 * scl;
        const double vmax = v0 - neg;
        const double vmin = v0 - pos;
        if (vmax < -tol) return CLS_NONE;
        all_in = all_in && (vmin > tol);
    }
    if (all_in) return CLS_CELL_IN_TET;
    bool tin = true;
    for (int q = 0; q < 4; ++q)
        tin = tin && s[q].x >= 0.0 && s[q].x <= r[0] && s[q].y >= 0.0 && s[q].y <= r[1] && s[q].z >= 0.0 &&
              s[q].z <= r[2];
    return tin ? CLS_TET_IN_CELL : CLS_PARTIAL;
}

template <int kPrec>
__device__ __forceinline__ void process_tile(const TileSource& src, int64_t t, const LatticeView& lat,
                                             const PackedCorners& ct, TileShared& sh,
                                             unsigned long long& my_fallback, unsigned long long* counters) {
    const int tid = threadIdx.x;
    const int p0 = src.tile_first[t];
    const int p1 = src.tile_first[t + 1];
    const int np = p1 - p0 + 1;  // <= kTile + 1 (every pair holds >= 1 candidate)
    for (int q = tid; q < np; q += kTile) {
        const int p = p0 + q;
        PairLds pr;
        const unsigned long long pk = src.range[p];
        pr.grid = src.pair_grid[p];
        for (int a = 0; a < 3; ++a) {
            pr.lo[a] = (unsigned short)((pk >> (10 * a)) & 1023u);
            pr.ext[a] = (unsigned short)((pk >> (30 + 10 * a)) & 1023u);
        }
        if (pr.ext[0]) {
            V3<double> v[4];
            lattice_tet_vertices(lat, ct, src.pairs[p], v);
            const double vol = vol6(v[0], v[1], v[2], v[3]) / 6.0;
            if (vol < 0.0) {
                const V3<double> tmp = v[0];
                v[0] = v[3];
                v[3] = tmp;
            }
            const GridDesc& g = src.grids[pr.grid];
            for (int w = 0; w < 4; ++w) {
                pr.S[3 * w] = g.scale[0] * (v[w].x - g.gmin[0]);
                pr.S[3 * w + 1] = g.scale[1] * (v[w].y - g.gmin[1]);
                pr.S[3 * w + 2] = g.scale[2] * (v[w].z - g.gmin[2]);
            }
            pr.mcell = src.tet_mass * (g.cell_vol / fabs(vol));
        } else {
            for (int w = 0; w < 12; ++w) pr.S[w] = 0.0;
            pr.mcell = 0.0;
        }
        sh.pair[q] = pr;
        sh.off[q] = src.cand_off[p];
    }
    if (tid == 0) {
        sh.np = np;
        sh.off[np] = src.cand_off[p0 + np];
    }
    __syncthreads();
    // classify one candidate per lane
    const int64_t c = t * kTile + tid;
    if (c < src.n_cand) {
        const int slot = upper_bound_i64(sh.off, sh.np, c) - 1;
        const PairLds& pr = sh.pair[slot];
        if (pr.ext[0]) {
            const unsigned r = unsigned(c - sh.off[slot]);
            const unsigned ex = pr.ext[0], ey = pr.ext[1];
            const int i = pr.lo[0] + int(r % ex);
            const int j = pr.lo[1] + int((r / ex) % ey);
            const int k = pr.lo[2] + int(r / (ex * ey));
            const GridDesc& g = src.grids[pr.grid];
            V3<double> s[4];
            load_s(pr, g, i, j, k, s);
            // strict AABB cull against the cell [0, r] (cell units)
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
                // the reference culls against the cell box but classifies and clips against the
                // unit cube in scaled coordinates (AdDepositTets2Grid.h:587-634)
                const double unit[3] = {1.0, 1.0, 1.0};
                const int cls = classify_margin(s, unit);
                const int64_t idx = g.base + i + int64_t(g.dims) * (j + int64_t(g.dims) * k);
                if (cls == CLS_TET_IN_CELL) {
                    unsafeAtomicAdd(src.grid_buf + idx, src.tet_mass);
                } else if (cls == CLS_CELL_IN_TET) {
                    unsafeAtomicAdd(src.grid_buf + idx, pr.mcell);
                } else if (cls == CLS_PARTIAL) {
                    const int qi = atomicAdd(&sh.qn, 1);
                    sh.qslot[qi] = (unsigned short)slot;
                    sh.qr[qi] = r;
                }
            }
        }
    }
    __syncthreads();
    // clip the compacted partial cells with all lanes
    const int nq = sh.qn;
    if (tid == 0 && nq) atomicAdd(counters + 1, (unsigned long long)nq);
    for (int qi = tid; qi < nq; qi += kTile) {
        const PairLds& pr = sh.pair[sh.qslot[qi]];
        const unsigned r = sh.qr[qi];
        const unsigned ex = pr.ext[0], ey = pr.ext[1];
        const int i = pr.lo[0] + int(r % ex);
        const int j = pr.lo[1] + int((r / ex) % ey);
        const int k = pr.lo[2] + int(r / (ex * ey));
        const GridDesc& g = src.grids[pr.grid];
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
        if (dm > 0.0) unsafeAtomicAdd(src.grid_buf + g.base + i + int64_t(g.dims) * (j + int64_t(g.dims) * k), dm);
    }
    __syncthreads();
}

template <int kPrec>
__global__ void __launch_bounds__(kTile) k_deposit_tiles(TileSource src, LatticeView lat, PackedCorners ct,
                                                         unsigned* __restrict__ tile_counter,
                                                         unsigned long long* __restrict__ counters) {
    __shared__ TileShared sh;
    unsigned long long my_fallback = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            sh.tile = int64_t(atomicAdd(tile_counter, 1u));
            sh.qn = 0;
        }
        __syncthreads();
        const int64_t t = sh.tile;
        if (t >= src.n_tiles) break;
        process_tile<kPrec>(src, t, lat, ct, sh, my_fallback, counters);
    }
    if (kPrec == ADECO_PREC_MIXED && my_fallback) atomicAdd(counters, my_fallback);
}

template <int kPrec>
__global__ void __launch_bounds__(kTile) k_deposit_step(const TileSource* __restrict__ sources, const int2* __restrict__ entries,
                                                        int n_entries, LatticeView lat, PackedCorners ct,
                                                        unsigned* __restrict__ entry_counter,
                                                        unsigned long long* __restrict__ counters) {
    __shared__ TileShared sh;
    unsigned long long my_fallback = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            sh.tile = int64_t(atomicAdd(entry_counter, 1u));
            sh.qn = 0;
        }
        __syncthreads();
        const int64_t e = sh.tile;
        if (e >= n_entries) break;
        const int2 en = entries[e];
        process_tile<kPrec>(sources[en.x], int64_t(en.y), lat, ct, sh, my_fallback, counters);
    }
    if (kPrec == ADECO_PREC_MIXED && my_fallback) atomicAdd(counters, my_fallback);
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

size_t candidate_space_bytes(int64_t n_pairs) {
    const int64_t nblk = (n_pairs + kScanBlock - 1) / kScanBlock;
    return align256(n_pairs * 4) + align256(n_pairs * 4) + align256(n_pairs * 8) + align256((n_pairs + 1) * 8) +
           align256(nblk * 8 + 8) + align256(8);
}

size_t tile_map_bytes(int64_t n_cand) { return align256(size_t((n_cand + kTile - 1) / kTile + 1) * 4); }

adeco_status build_candidate_space(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                   const int32_t* pairs, int64_t n_pairs, const int64_t* grid_off, int n_grids,
                                   const GridDesc* d_grids, void* slab, CandidateSpace& cs) {
    char* w = static_cast<char*>(slab);
    const int64_t nblk = (n_pairs + kScanBlock - 1) / kScanBlock;
    cs.pair_grid = reinterpret_cast<int32_t*>(w);
    w += align256(n_pairs * 4);
    cs.cand = reinterpret_cast<unsigned*>(w);
    w += align256(n_pairs * 4);
    cs.range = reinterpret_cast<unsigned long long*>(w);
    w += align256(n_pairs * 8);
    cs.cand_off = reinterpret_cast<int64_t*>(w);
    w += align256((n_pairs + 1) * 8);
    cs.block_sum = reinterpret_cast<int64_t*>(w);
    w += align256(nblk * 8 + 8);
    cs.total = reinterpret_cast<int64_t*>(w);
    cs.n_pairs = n_pairs;
    cs.n_cand = 0;
    cs.n_tiles = 0;
    if (n_pairs == 0) return ADECO_OK;
    hipLaunchKernelGGL(k_pair_setup, dim3(unsigned((n_pairs + 255) / 256)), dim3(256), 0, stream, lat, ct, pairs, n_pairs,
                       grid_off, n_grids, d_grids, cs.cand, cs.range, cs.pair_grid);
    hipLaunchKernelGGL(k_scan_local, dim3(unsigned(nblk)), dim3(256), 0, stream, cs.cand, n_pairs, cs.cand_off, cs.block_sum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, stream, cs.block_sum, nblk, cs.total);
    hipLaunchKernelGGL(k_scan_add, dim3(unsigned((n_pairs + 1 + 255) / 256)), dim3(256), 0, stream, cs.cand_off, n_pairs,
                       cs.block_sum, cs.total);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&cs.n_cand, cs.total, 8, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "candidate space setup");
    cs.n_tiles = (cs.n_cand + kTile - 1) / kTile;
    return ADECO_OK;
}

adeco_status build_tile_map(hipStream_t stream, CandidateSpace& cs, int32_t* tiles) {
    cs.tile_first = tiles;
    if (cs.n_pairs == 0) return ADECO_OK;
    hipLaunchKernelGGL(k_tile_map, dim3(unsigned((cs.n_pairs + 1 + 255) / 256)), dim3(256), 0, stream, cs.cand_off,
                       cs.n_pairs, tiles, cs.n_tiles);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "tile map");
}

hipError_t launch_deposit_tiles(hipStream_t stream, int precision, const TileSource& src, const LatticeView& lat,
                                const PackedCorners& ct, unsigned* tile_counter, unsigned long long* counters,
                                int grid_blocks) {
    if (src.n_tiles <= 0) return hipSuccess;
    const unsigned grid = unsigned(std::min<int64_t>(src.n_tiles, grid_blocks));
    if (precision == ADECO_PREC_FP64)
        hipLaunchKernelGGL(k_deposit_tiles<ADECO_PREC_FP64>, dim3(grid), dim3(kTile), 0, stream, src, lat, ct, tile_counter, counters);
    else if (precision == ADECO_PREC_MIXED)
        hipLaunchKernelGGL(k_deposit_tiles<ADECO_PREC_MIXED>, dim3(grid), dim3(kTile), 0, stream, src, lat, ct, tile_counter, counters);
    else
        hipLaunchKernelGGL(k_deposit_tiles<ADECO_PREC_FP32>, dim3(grid), dim3(kTile), 0, stream, src, lat, ct, tile_counter, counters);
    return hipGetLastError();
}

hipError_t launch_deposit_step(hipStream_t stream, int precision, const TileSource* d_sources, const int2* d_entries,
                               int n_entries, const LatticeView& lat, const PackedCorners& ct, unsigned* entry_counter,
                               unsigned long long* counters, int grid_blocks) {
    if (n_entries <= 0) return hipSuccess;
    const unsigned grid = unsigned(std::min<int64_t>(n_entries, grid_blocks));
    if (precision == ADECO_PREC_FP64)
        hipLaunchKernelGGL(k_deposit_step<ADECO_PREC_FP64>, dim3(grid), dim3(kTile), 0, stream, d_sources, d_entries, n_entries,
                           lat, ct, entry_counter, counters);
    else if (precision == ADECO_PREC_MIXED)
        hipLaunchKernelGGL(k_deposit_step<ADECO_PREC_MIXED>, dim3(grid), dim3(kTile), 0, stream, d_sources, d_entries, n_entries,
                           lat, ct, entry_counter, counters);
    else
        hipLaunchKernelGGL(k_deposit_step<ADECO_PREC_FP32>, dim3(grid), dim3(kTile), 0, stream, d_sources, d_entries, n_entries,
                           lat, ct, entry_counter, counters);
    return hipGetLastError();
}

int persistent_blocks(int device) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus * 4;
}

// Leaf deposit of an octree: one candidate space over all leaves, one persistent launch.
adeco_status deposit_leaves_batched(adeco_octree_s* T, int dims, int decomposition, double tet_mass, int precision) {
    adeco_lattice_s* L = T->L;
    DeviceState& ds = *L->ds;
    if (dims > 1023) return fail(ADECO_EINVAL, "batched deposit: dims must be < 1024");
    const int nl = int(T->leaves_dfs.size());
    const int64_t P = T->leaf_tet_offset[nl];
    if (P == 0) return ADECO_OK;
    if (P >= (int64_t(1) << 31) - 2) return fail(ADECO_EOVERFLOW, "batched deposit: too many (leaf, tet) pairs");
    const int64_t cells = int64_t(dims) * dims * dims;
    std::vector<GridDesc> gd(nl);
    const int scaling = cell_scaling();
    for (int l = 0; l < nl; ++l) gd[l] = make_grid_desc(T->leaf_geom(l, dims, scaling), (long long)(l * cells));
    adeco_status st;
    if ((st = T->d_work.reserve(candidate_space_bytes(P) + 256)) || (st = T->d_leaf_geom.reserve(gd.size() * sizeof(GridDesc))) ||
        (st = T->d_counters.reserve(64)))
        return st;
    hipError_t e = hipMemcpyAsync(T->d_leaf_geom.ptr, gd.data(), gd.size() * sizeof(GridDesc), hipMemcpyHostToDevice, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "leaf geometry upload");
    const LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    const PackedCorners ct = pack_corners(corner_table(decomposition));
    CandidateSpace cs;
    if ((st = build_candidate_space(ds.stream, lv, ct, T->d_pairs.as<int32_t>(), P, T->d_leaf_offset.as<int64_t>(), nl,
                                    T->d_leaf_geom.as<GridDesc>(), T->d_work.ptr, cs)))
        return st;
    if ((st = T->d_tiles.reserve(tile_map_bytes(cs.n_cand)))) return st;
    if ((st = build_tile_map(ds.stream, cs, T->d_tiles.as<int32_t>()))) return st;
    unsigned* ctr = T->d_counters.as<unsigned>();
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(T->d_counters.as<char>() + 16);
    e = hipMemsetAsync(T->d_counters.ptr, 0, 64, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched counters");
    const TileSource src{T->d_pairs.as<int32_t>(), cs.pair_grid, cs.range, cs.cand_off, cs.tile_first, cs.n_tiles,
                         cs.n_cand, T->d_leaf_geom.as<GridDesc>(), T->d_grids.as<double>(), tet_mass};
    e = hipEventRecord(ds.evk0, ds.stream);
    if (e == hipSuccess) e = launch_deposit_tiles(ds.stream, precision, src, lv, ct, ctr, counters, persistent_blocks(L->device));
    if (e == hipSuccess) e = hipEventRecord(ds.evk1, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "batched deposit launch");
    T->last_candidates = cs.n_cand;
    T->d_fallback = counters;
    return ADECO_OK;
}

}  // namespace adeco
