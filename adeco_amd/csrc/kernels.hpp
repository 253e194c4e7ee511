// Internal interface between the host runtime (runtime.hip) and the HIP kernels.  Not part of
// the C-ABI: the kernels take device pointers and plain structs, never torch types.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "adeco_geom.hpp"

namespace adeco {

// One uniform output grid (an octree leaf, a request ROI or the boundary call's bbox).
// Built on the host from the caller's float bbox as the reference does in FP32, but in FP64:
// delta = (max - min) / dims, cell_vol = dx*dy*dz (AdDepositTets2Grid.h:527-569).  The reference
// then scales every axis by 1/dx (AdDepositTets2Grid.h:569), which assumes cubic cells; a float
// box (ll, ll + ext) rarely has three bit-identical extents, so that formula moves the cell faces
// by up to ~1e-7 of a cell and breaks mass conservation at that level.  kScalingExact (the
// default) scales each axis by its own 1/delta, which partitions the box exactly;
// kScalingReference keeps the reference's formula.
enum CellScaling : int { kScalingExact = 0, kScalingReference = 1 };

struct GridGeom {
    double gmin[3], gmax[3];
    double delta[3];
    double scale[3];
    double cell_vol;
    int dims;
};

inline GridGeom make_grid_geom_d(const double bbox[6], int dims, int scaling = kScalingExact) {
    GridGeom g;
    for (int a = 0; a < 3; ++a) {
        g.gmin[a] = bbox[2 * a];
        g.gmax[a] = bbox[2 * a + 1];
        g.delta[a] = (g.gmax[a] - g.gmin[a]) / double(dims);
    }
    g.cell_vol = g.delta[0] * g.delta[1] * g.delta[2];
    for (int a = 0; a < 3; ++a) g.scale[a] = 1.0 / g.delta[scaling == kScalingReference ? 0 : a];
    g.dims = dims;
    return g;
}

inline GridGeom make_grid_geom(const float bbox[6], int dims, int scaling = kScalingExact) {
    const double b[6] = {bbox[0], bbox[1], bbox[2], bbox[3], bbox[4], bbox[5]};
    return make_grid_geom_d(b, dims, scaling);
}

// Library-wide cell scaling (adeco_set_cell_scaling); read when a deposit starts.
int cell_scaling();

// Resident lattice: SoA float vertex coordinates, np[0]*np[1]*np[2] vertices, i fastest.  With
// `soup` set, tet t is instead the 4 vertices soup[12 t .. 12 t + 11] of a tet soup (AoS xyz).
struct LatticeView {
    const float* x;
    const float* y;
    const float* z;
    int np[3];
    const float* soup = nullptr;
};

// Per-launch work statistics (device counters, filled only by the counting variants).
// 0 tets, 1 tets_active, 2 candidates, 3 aabb_pass, 4 none, 5 tet_in_cell, 6 cell_in_tet,
// 7 partial, 8 nonzero, 9 clip stages, 10 clip sub-tets
constexpr int kNumStats = 11;

// ---- deposit.hip launchers (all asynchronous on `stream`) ----
hipError_t launch_deposit_patches_plain(const float* d_positions, int64_t num_patches, int m, int decomp,
                                        const GridGeom& g, double tet_mass, double* d_grid,
                                        unsigned long long* d_stats, hipStream_t stream);
// The 6 m^3 tets of every meta-patch as a soup (float [n][4][3]); cells with a sentinel vertex
// give degenerate tets (all four vertices at the origin), which deposit nothing.
hipError_t launch_patches_to_soup(const float* d_positions, int64_t num_patches, int m, int decomp, float* d_soup,
                                  hipStream_t stream);

hipError_t launch_deposit_lattice_plain(const LatticeView& lat, int decomp, const int32_t* d_tet_ids,
                                        int64_t n_tets, const GridGeom& g, double tet_mass,
                                        double* d_grid, unsigned long long* d_stats, hipStream_t stream);

hipError_t launch_deposit_soup_plain(const float* d_tets, int64_t n_tets, const GridGeom& g,
                                     double tet_mass, double* d_grid, unsigned long long* d_stats,
                                     hipStream_t stream);

hipError_t launch_add_f64_to_f32(const double* d_src, float* d_dst, int64_t n, hipStream_t stream);
hipError_t launch_fill_f64(double* d, int64_t n, double v, hipStream_t stream);

}  // namespace adeco
