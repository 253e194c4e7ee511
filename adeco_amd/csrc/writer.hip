// Leaf files (C ABI part 6): AdOctreeGridWriter::write_patch (AdOctreeGridWriter.h:67-144) and
// the write step of the single-rank leaf loop (deposit_leaf_nodes, AdDynamicLoadBalancing.h:998,
// 1133, 1303-1312).
//
// Format: one raw float32 file per leaf (host byte order, little-endian here), dims^3 densities
// (mass / cell volume), i fastest, named <output_path>/node_<level>_<ix>_<iy>_<iz>, (ix,iy,iz) the
// leaf's offset from the root corner in cells of its own grid.  The arithmetic follows the
// reference operation by operation: the node box is formed in float (ll, ll + ext) and widened to
// double, the cell size and inverse cell volume are double, each density is
// float(double(mass) * inv_cell_vol), and the returned mass is the double sum of the float masses
// in index order.  Given the same float mass grid, a file is the reference's file bit for bit.
//
// adeco_write_leaves turns the FP64 leaf grids of the last adeco_deposit_leaves into float mass
// grids on the GPU (one rounding, in place of the reference's float accumulation), copies them to
// the host in bounded chunks and writes every leaf.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/adeco.h"
#include "octree_impl.hpp"
#include "runtime.hpp"

using namespace adeco;

namespace {

__global__ void __launch_bounds__(256) k_f64_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] = float(src[i]);
}

// The leaf-file name of a node (AdOctreeGridWriter.h:103-124): offsets in cells of the node's
// grid, delta = ext / dims computed in float, rounded to nearest (rint) in double.
std::string patch_name(const std::string& output_path, int level, const float ll[3], float ext,
                       const float root_ll[3], int dims) {
#pragma clang fp contract(off)
    const double delta = double(ext / float(dims));
    std::string name = output_path + "/node_" + std::to_string(level);
    for (int a = 0; a < 3; ++a) {
        const long long io = static_cast<long long>(std::rint((double(ll[a]) - double(root_ll[a])) / delta));
        name += "_" + std::to_string(io);
    }
    return name;
}

// Densities of one leaf and its mass (AdOctreeGridWriter.h:75-98).
double patch_density(int dims, const float ll[3], float ext, const float* mass, float* density) {
#pragma clang fp contract(off)
    double cell[3];
    for (int a = 0; a < 3; ++a) {
        const double lo = double(ll[a]);
        const double hi = double(ll[a] + ext);  // float sum, as the reference's bbox initialiser
        cell[a] = (hi - lo) / double(dims);
    }
    const double inv_cell_vol = 1. / (cell[0] * cell[1] * cell[2]);
    const size_t n = size_t(dims) * dims * dims;
    double m = 0.;
    for (size_t i = 0; i < n; ++i) {
        m += double(mass[i]);
        density[i] = float(double(mass[i]) * inv_cell_vol);
    }
    return m;
}

adeco_status write_file(const std::string& name, const float* data, size_t n) {
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return fail(ADECO_EIO, "write_patch(): failed to write node data: " + name + ": " + std::strerror(errno));
    const size_t w = std::fwrite(data, sizeof(float), n, f);
    const int c = std::fclose(f);
    if (w != n || c != 0) return fail(ADECO_EIO, "write_patch(): failed to write node data: " + name);
    return ADECO_OK;
}

adeco_status write_patch_impl(const char* output_path, int level, const float* ll, float ext, const float* root_ll,
                              int dims, const float* mass, double* mass_out, std::vector<float>& buf) {
    const size_t n = size_t(dims) * dims * dims;
    buf.resize(n);
    const double m = patch_density(dims, ll, ext, mass, buf.data());
    adeco_status st = write_file(patch_name(output_path, level, ll, ext, root_ll, dims), buf.data(), n);
    if (st) return st;
    if (mass_out) *mass_out = m;
    return ADECO_OK;
}

}  // namespace

extern "C" {

adeco_status adeco_write_patch(const char* output_path, int level, const float* node_ll3, float node_ext,
                               const float* root_ll3, int grid_dims, const float* mass_grid, double* mass_out) {
    if (!output_path || !node_ll3 || !root_ll3 || !mass_grid)
        return fail(ADECO_EINVAL, "write_patch(): null argument");
    if (grid_dims < 1 || grid_dims > 1024) return fail(ADECO_EINVAL, "write_patch(): grid_dims must be in [1, 1024]");
    if (level < 0) return fail(ADECO_EINVAL, "write_patch(): negative level");
    if (!(node_ext > 0.f)) return fail(ADECO_EINVAL, "write_patch(): invalid node extension");
    std::vector<float> buf;
    return write_patch_impl(output_path, level, node_ll3, node_ext, root_ll3, grid_dims, mass_grid, mass_out, buf);
}

adeco_status adeco_write_leaves(adeco_octree T, const char* output_path, double* leaf_mass, double* total_mass) {
    if (!T || !output_path) return fail(ADECO_EINVAL, "write_leaves(): null argument");
    const int dims = T->grid_dims;
    if (dims < 1 || !T->d_grids.ptr) return fail(ADECO_EINVAL, "write_leaves(): no leaf deposit to write");
    DeviceState& ds = *T->L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int64_t nl = int64_t(T->leaves_dfs.size());
    const size_t cells = size_t(dims) * dims * dims;
    // leaves per chunk: at most 64 MiB of float grids staged at once (at least one leaf)
    const int64_t per = std::max<int64_t>(1, int64_t((size_t(64) << 20) / (cells * sizeof(float))));
    DevBuf d_f32;
    adeco_status st = d_f32.reserve(size_t(std::min(per, nl)) * cells * sizeof(float));
    if (st) return st;
    std::vector<float> host(size_t(std::min(per, nl)) * cells), buf;
    std::vector<double> mass(size_t(nl), 0.0);
    const float* root_ll = T->nodes[0].ll;
    for (int64_t first = 0; first < nl; first += per) {
        const int64_t cnt = std::min(per, nl - first);
        const int64_t n = cnt * int64_t(cells);
        const unsigned nb = unsigned(std::min<int64_t>(n / 256 + 1, 4096));
        hipLaunchKernelGGL(k_f64_to_f32, dim3(nb), dim3(256), 0, ds.stream,
                           T->d_grids.as<double>() + size_t(first) * cells, d_f32.as<float>(), n);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(host.data(), d_f32.ptr, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, ds.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
        if (e != hipSuccess) return hip_fail(e, "leaf grid download");
        for (int64_t l = 0; l < cnt; ++l) {
            const ONode& nd = T->nodes[T->leaves_dfs[size_t(first + l)]];
            st = write_patch_impl(output_path, nd.level, nd.ll, nd.ext, root_ll, dims, host.data() + size_t(l) * cells,
                                  &mass[size_t(first + l)], buf);
            if (st) return st;
        }
    }
    if (leaf_mass) std::memcpy(leaf_mass, mass.data(), mass.size() * sizeof(double));
    if (total_mass) {
        // summed in the leaf processing order, as the reference's total_mass
        std::vector<int64_t> dfs_of(T->nodes.size(), -1);
        for (int64_t l = 0; l < nl; ++l) dfs_of[size_t(T->leaves_dfs[size_t(l)])] = l;
        double t = 0.;
        for (int32_t id : T->leaves_order) t += mass[size_t(dfs_of[size_t(id)])];
        *total_mass = t;
    }
    return ADECO_OK;
}

}  // extern "C"
