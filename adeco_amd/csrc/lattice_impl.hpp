// The resident-lattice handle behind adeco_lattice.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

struct adeco_lattice_s {
    int device = -1;
    adeco::DeviceState* ds = nullptr;
    int np[3] = {0, 0, 0};
    int64_t n_vertices = 0;
    float* d_x = nullptr;
    float* d_y = nullptr;
    float* d_z = nullptr;
    // float AABBs of the meta-cells (7^3 lattice cubes), [M][6], filled by the first ROI selection
    std::vector<float> meta_box;
    // The global bbox octrees are rooted at, when it is not the vertices' own bbox: the reference's
    // periodic correction moves vertices but keeps the bbox computed before it
    // (AdDepositWorkFlow.h:101-112).
    bool has_domain = false;
    float domain[6] = {0, 0, 0, 0, 0, 0};
    int meta_dims[3] = {0, 0, 0};
    ~adeco_lattice_s() {
        if (d_x) (void)hipFree(d_x);
        if (d_y) (void)hipFree(d_y);
        if (d_z) (void)hipFree(d_z);
    }
};

namespace adeco {
// tet ids are int32 in the binned work lists: at most 2^31 / 6 vertices
constexpr int64_t kMaxLatticeVertices = (int64_t(1) << 31) / 6;
adeco_status lattice_bbox_device(adeco_lattice_s* L, float out6[6]);
adeco_status lattice_alloc(int device_id, int npx, int npy, int npz, adeco_lattice_s*& out);
// The batched persistent deposit of a tet soup (float [n][4][3], on the device) onto one grid,
// accumulating; kernel time, partial cells and FP64 recomputes go to the thread's timings.
adeco_status deposit_soup_batched(DeviceState& ds, const float* d_tets, int64_t n_tets, const GridGeom& g,
                                  double tet_mass, int precision, double* d_grid);
}
