// Internal interface between the host runtime (runtime.hip) and the HIP kernels.  Not part of
// the C-ABI: the kernels take device pointers and plain structs, never torch types.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "adeco_geom.hpp"

namespace adeco {

// One uniform output grid (an octree leaf, a request ROI or the boundary call's bbox).
// Built on the host from the caller's float bbox exactly as the reference does in FP32, but in
// FP64: delta = (max - min) / dims, cell_vol = dx*dy*dz, scale = 1/dx (AdDepositTets2Grid.h:527-569).
struct GridGeom {
    double gmin[3], gmax[3];
    double delta[3];
    double scale;
    double cell_vol;
    int dims;
};

inline GridGeom make_grid_geom(const float bbox[6], int dims) {
    GridGeom g;
    for (int a = 0; a < 3; ++a) {
        g.gmin[a] = double(bbox[2 * a]);
        g.gmax[a] = double(bbox[2 * a + 1]);
        g.delta[a] = (g.gmax[a] - g.gmin[a]) / double(dims);
    }
    g.cell_vol = g.delta[0] * g.delta[1] * g.delta[2];
    g.scale = 1.0 / g.delta[0];
    g.dims = dims;
    return g;
}

// Resident lattice: SoA float vertex coordinates, np[0]*np[1]*np[2] vertices, i fastest.
struct LatticeView {
    const float* x;
    const float* y;
    const float* z;
    int np[3];
};

// Per-launch work statistics (device counters, filled only by the counting variants).
// 0 tets, 1 tets_active, 2 candidates, 3 aabb_pass, 4 none, 5 tet_in_cell, 6 cell_in_tet,
// 7 partial, 8 nonzero, 9 clip stages, 10 clip sub-tets
constexpr int kNumStats = 11;

// ---- deposit.hip launchers (all asynchronous on `stream`) ----
hipError_t launch_deposit_patches_plain(const float* d_positions, int64_t num_patches, int m,
                                        const GridGeom& g, double tet_mass, double* d_grid,
                                        unsigned long long* d_stats, hipStream_t stream);

hipError_t launch_deposit_lattice_plain(const LatticeView& lat, int decomp, const int32_t* d_tet_ids,
                                        int64_t n_tets, const GridGeom& g, double tet_mass,
                                        double* d_grid, unsigned long long* d_stats, hipStream_t stream);

hipError_t launch_deposit_soup_plain(const float* d_tets, int64_t n_tets, const GridGeom& g,
                                     double tet_mass, double* d_grid, unsigned long long* d_stats,
                                     hipStream_t stream);

hipError_t launch_add_f64_to_f32(const double* d_src, float* d_dst, int64_t n, hipStream_t stream);
hipError_t launch_fill_f64(double* d, int64_t n, double v, hipStream_t stream);

}  // namespace adeco
