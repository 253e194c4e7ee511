// Candidate-space machinery of the batched deposit, shared by the octree leaf deposit
// (batched.hip) and the request server (serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "runtime.hpp"

namespace adeco {

constexpr int kTile = 256;  // candidates per tile == lanes per workgroup

// One output grid: a leaf or a request ROI.  gmin/delta/scale in FP64; base = first cell of the
// grid in the grid buffer; cells i + dims*(j + dims*k).
struct GridDesc {
    double gmin[3];
    double delta[3];
    double r[3];  // delta_a * scale_a
    double scale[3];
    double cell_vol;
    long long base;
    int dims;
    int pad_;
};

GridDesc make_grid_desc(const GridGeom& g, long long base);

// A candidate space: binned (grid, tet) pairs x covered cells, cut into tiles of kTile candidates.
struct TileSource {
    const int32_t* pairs;        // lattice tet id per pair
    const int32_t* pair_grid;    // grid index per pair
    const unsigned long long* range;  // packed covered range per pair (10-bit lo / extent per axis)
    const int64_t* cand_off;     // candidate offsets, n_pairs + 1
    const int32_t* tile_first;   // first pair of each tile, n_tiles + 1
    int64_t n_tiles, n_cand;
    const GridDesc* grids;
    double* grid_buf;
    double tet_mass;
};

// Device work arrays of one candidate space (allocated from one slab).
struct CandidateSpace {
    int32_t* pair_grid = nullptr;
    unsigned* cand = nullptr;
    unsigned long long* range = nullptr;
    int64_t* cand_off = nullptr;
    int64_t* block_sum = nullptr;
    int64_t* total = nullptr;
    int32_t* tile_first = nullptr;
    int64_t n_pairs = 0, n_cand = 0, n_tiles = 0;
};

// Bytes of the work slab for n_pairs pairs (tile map excluded: it needs n_cand).
size_t candidate_space_bytes(int64_t n_pairs);
size_t tile_map_bytes(int64_t n_cand);

// Lays out `cs` in `slab` and runs pair setup + scan (asynchronous; n_cand is read back with a
// synchronous copy).  pair_grid_in: nullptr -> grid = leaf from leaf_off (CSR over grids).
adeco_status build_candidate_space(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                   const int32_t* pairs, int64_t n_pairs, const int64_t* leaf_off, int n_grids,
                                   const GridDesc* d_grids, void* slab, CandidateSpace& cs);
// Writes the tile map into `tiles` (tile_map_bytes(cs.n_cand) bytes).
adeco_status build_tile_map(hipStream_t stream, CandidateSpace& cs, int32_t* tiles);

// Persistent deposit over one candidate space; counters[0] += FP64 recomputes, counters[1] +=
// partial cells.  tile_counter must be zero.
hipError_t launch_deposit_tiles(hipStream_t stream, int precision, const TileSource& src, const LatticeView& lat,
                                const PackedCorners& ct, unsigned* tile_counter, unsigned long long* counters,
                                int grid_blocks);

// One serving step: entries (source slot, tile) over several candidate spaces in one launch.
hipError_t launch_deposit_step(hipStream_t stream, int precision, const TileSource* d_sources,
                               const int2* d_entries, int n_entries, const LatticeView& lat, const PackedCorners& ct,
                               unsigned* entry_counter, unsigned long long* counters, int grid_blocks);

}  // namespace adeco
