// Candidate-space machinery of the batched deposit, shared by the octree leaf deposit and the
// request server (both in batched.hip / serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "runtime.hpp"

struct adeco_octree_s;

namespace adeco {

constexpr int kWaveTile = 64;  // candidates per tile == lanes per wavefront
// Cells per column segment: the unit of work of an FP64 deposit on grids of unit scaled cells is a
// run of up to kMaxSeg cells along the tet's longest axis (one x/y clip for all of them); other
// deposits use single cells.
constexpr int kMaxSeg = 8;

// One output grid: a leaf or a request ROI.  gmin/delta/scale in FP64; base = first cell of the
// grid in the grid buffer; cells i + dims*(j + dims*k).
struct GridDesc {
    double gmin[3];
    double delta[3];
    double r[3];  // delta_a * scale_a: the cell edge in the scaled coordinates (1 exactly when scale_a = 1/delta_a)
    double scale[3];
    double cell_vol;
    long long base;
    int dims;
    int pad_;
};

GridDesc make_grid_desc(const GridGeom& g, long long base);

// Everything the deposit needs about one binned (grid, tet) pair, written once by the pair setup:
// one 128-byte line per pair.
struct alignas(128) PairData {
    double S[12];              // scale_a * (v_q - gmin_a), vertices in the oriented order
    double mcell;              // tet_mass * cell_vol / |V|
    unsigned long long range;  // covered cells, 10-bit lo / extent per axis (extent 0: dead pair),
                               // bits 60-61 the column axis, bit 62 set when candidates are columns
    int grid;
    int pad_[3];
};
static_assert(sizeof(PairData) == 128, "PairData is one cache line");

// A candidate space: binned (grid, tet) pairs x their candidates (covered cells, or column segments
// of covered cells), cut into tiles of kWaveTile consecutive candidates.
struct TileSource {
    const PairData* pd;
    const int64_t* cand_off;     // candidate offsets, n_pairs + 1
    const int32_t* tile_first;   // first pair of each tile, n_tiles + 1
    int64_t n_tiles, n_cand;
    const GridDesc* grids;
    double* grid_buf;
    double tet_mass;
    const int64_t* dev_sizes = nullptr;  // when set, n_cand and n_tiles are dev_sizes[0], [1] (device)
};

// Device work arrays of one candidate space (allocated from one slab).
struct CandidateSpace {
    PairData* pd = nullptr;
    unsigned* cand = nullptr;
    int64_t* cand_off = nullptr;
    int64_t* block_sum = nullptr;
    int64_t* total = nullptr;
    int32_t* tile_first = nullptr;
    int64_t n_pairs = 0, n_cand = 0, n_tiles = 0;
};

// Launch shape of the persistent deposit: tiles taken per work-queue pop and resident workgroups
// (of 4 wavefronts) per CU.  Chosen by adeco_autotune_batched.  Defaults from MI355X: one counter
// pop costs ~12 ns of serialized atomic time, so a pop takes 16 tiles (1024 candidates); 4
// workgroups per CU fill the FP32 kernel's occupancy (the FP64 kernel holds 2).
struct LaunchShape {
    int chunk = 16;
    int blocks_per_cu = 4;
};
LaunchShape& launch_shape();
// Workgroups of the persistent deposit kernel (precision, merged drain or not) that one CU holds at
// once (hipOccupancyMaxActiveBlocksPerMultiprocessor on the current device); at least 1.
int deposit_occupancy(int precision, bool merge);

// Bytes of the work slab for n_pairs pairs (tile map excluded: it needs n_cand).
size_t candidate_space_bytes(int64_t n_pairs);
size_t tile_map_bytes(int64_t n_cand);

// Lays out `cs` in `slab` and runs pair setup + scan (asynchronous; n_cand is read back with a
// synchronous copy).  Pairs are grid-major: grid_off is the CSR offset of each grid's pairs.
// precision selects the candidate unit: column segments for ADECO_PREC_FP64 on grids with unit
// scaled cells, single cells otherwise.
adeco_status build_candidate_space(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                   const int32_t* pairs, int64_t n_pairs, const int64_t* grid_off, int n_grids,
                                   const GridDesc* d_grids, double tet_mass, int precision, void* slab,
                                   CandidateSpace& cs);
// Writes the tile map into `tiles` (tile_map_bytes(cs.n_cand) bytes).
adeco_status build_tile_map(hipStream_t stream, CandidateSpace& cs, int32_t* tiles);

// The same without host reads, for graph capture: the pair count is *n_dev (at most pair_cap, the
// slab's capacity: candidate_space_bytes(pair_cap)); pair setup, scan, sizes (sizes[0] = the
// candidates, sizes[1] = the tiles) and the tile map (tile_cap + 1 entries) are enqueued.  A pair
// count above pair_cap or a tile count above tile_cap sets *overflow and leaves no tiles.
hipError_t enqueue_candidate_space_dev(hipStream_t stream, const LatticeView& lat, const PackedCorners& ct,
                                       const int32_t* pairs, int64_t pair_cap, const int64_t* n_dev,
                                       const int64_t* grid_off, int n_grids, const GridDesc* d_grids, double tet_mass,
                                       int precision, void* slab, int32_t* tiles, int64_t tile_cap, int64_t* sizes,
                                       unsigned* overflow, CandidateSpace& cs);

// Exclusive scan of n uint32 counts into int64 offsets out[0..n] (out[n] = the total); block_sum
// holds scan_block_words(n) words, total one.
hipError_t launch_scan_u32(const unsigned* in, int64_t n, int64_t* out, int64_t* block_sum, int64_t* total,
                           hipStream_t stream);
size_t scan_block_words(int64_t n);

// Persistent deposit over one candidate space (built for the same precision); counters[0] += exact
// recomputes, counters[1] += partial candidates.  queue_counter must be zero.
hipError_t launch_deposit_tiles(hipStream_t stream, int precision, const TileSource& src, unsigned* queue_counter,
                                unsigned long long* counters, int device);

// One serving step: entries (source slot, tile) over several candidate spaces in one launch.
hipError_t launch_deposit_step(hipStream_t stream, int precision, const TileSource* d_sources,
                               const int2* d_entries, int n_entries, unsigned* queue_counter,
                               unsigned long long* counters, int device);

// Candidate space of an octree's leaf deposit in caller-owned buffers (leaf geometry, pair data +
// scan, tile map) over the leaf grids `grids`; src describes it for launch_deposit_tiles.
adeco_status prepare_leaves_batched(adeco_octree_s* T, hipStream_t stream, int dims, int decomposition, double tet_mass,
                                    int precision, DevBuf& geom, DevBuf& work, DevBuf& tiles, double* grids,
                                    TileSource& src);

// Clears n 32-bit words with a kernel (a memset node in a captured graph is avoided).
hipError_t launch_zero_u32(unsigned* p, int n, hipStream_t stream);

}  // namespace adeco
