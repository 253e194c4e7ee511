// Plain deposit kernels: one tet per thread, FP64 geometry, FP64 global atomics.
//
// These are the "plain" GPU form of the ADECO deposit (the reference's CUDA kernel
// resample_without_connectivity_GPU, AdDeposit2Grid.cu:139-211, runs one warp per Lagrangian cell
// and one lane per covered cell).  Here every lane owns one tet end to end: it loads the 4
// vertices (coalesced SoA loads for the resident lattice), computes the covered range and walks
// it.  The batched form (batched.hip) is the load-balanced alternative.
#include <hip/hip_runtime.h>

#include "deposit_tet.hpp"
#include "kernels.hpp"

namespace adeco {

namespace {

struct AtomicEmit {
    double* grid;
    __device__ __forceinline__ void operator()(int64_t idx, double dm) { unsafeAtomicAdd(grid + idx, dm); }
};

__device__ __forceinline__ void flush_counters(const TetCounters& c, unsigned long long* d_stats) {
    for (int s = 0; s < kNumStats; ++s)
        if (c.v[s]) atomicAdd(d_stats + s, c.v[s]);
}

// AoS meta-patches (3*(m+1)^3 floats each), one thread per (cell, tet) of num_patches*m^3 cells.
// Cells with a 666 sentinel vertex are skipped whole (AdDepositTets2Grid.h:1029-1038); the 6 tets
// follow the corner table pc (the reference's slot order is the REF_EFFECTIVE decomposition).
template <bool kCount>
__global__ void __launch_bounds__(256) k_deposit_patches_plain(const float* __restrict__ pos,
                                                               int64_t n_tets, int m, GridGeom g,
                                                               double tet_mass, double* grid,
                                                               unsigned long long* d_stats, PackedCorners pc) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_tets) return;
    const int64_t cell = t / 6;
    const int lt = int(t - cell * 6);
    const int cpp = m * m * m;
    const int64_t p = cell / cpp;
    const int c = int(cell - p * cpp);
    const int ck = c / (m * m), rest = c - ck * m * m, cj = rest / m, ci = rest - cj * m;
    const int64_t fpp = 3ll * (m + 1) * (m + 1) * (m + 1);
    const float* base = pos + p * fpp + 3 * (ci + (m + 1) * (cj + (m + 1) * ck));
    const int ax = 3, ay = 3 * (m + 1), az = 3 * (m + 1) * (m + 1);
    // slot offsets P0..P7 = 000,100,010,110,011,111,001,101 (AdDepositTets2Grid.h:947-972)
    const int slot_off[8] = {0, ax, ay, ay + ax, ay + az, ay + az + ax, az, az + ax};
    for (int s = 0; s < 8; ++s)
        if (is_sentinel(base[slot_off[s]], base[slot_off[s] + 1], base[slot_off[s] + 2])) return;
    V3<double> v[4];
    for (int q = 0; q < 4; ++q) {
        const int code = pc.code(lt, q);
        const float* pv = base + 3 * (code & 1) + ay * ((code >> 1) & 1) + az * ((code >> 2) & 1);
        v[q] = mk(double(pv[0]), double(pv[1]), double(pv[2]));
    }
    AtomicEmit emit{grid};
    if (kCount) {
        TetCounters cnt = {};
        deposit_tet_f64(v, g, tet_mass, emit, &cnt);
        flush_counters(cnt, d_stats);
    } else {
        deposit_tet_f64(v, g, tet_mass, emit, nullptr);
    }
}

template <bool kCount>
__global__ void __launch_bounds__(256) k_deposit_lattice_plain(LatticeView lat, PackedCorners ct,
                                                               const int32_t* __restrict__ tet_ids,
                                                               int64_t n_tets, GridGeom g, double tet_mass,
                                                               double* grid, unsigned long long* d_stats) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n_tets) return;
    const int64_t t = tet_ids ? int64_t(tet_ids[q]) : q;
    V3<double> v[4];
    lattice_tet_vertices(lat, ct, t, v);
    AtomicEmit emit{grid};
    if (kCount) {
        TetCounters cnt = {};
        deposit_tet_f64(v, g, tet_mass, emit, &cnt);
        flush_counters(cnt, d_stats);
    } else {
        deposit_tet_f64(v, g, tet_mass, emit, nullptr);
    }
}

template <bool kCount>
__global__ void __launch_bounds__(256) k_deposit_soup_plain(const float* __restrict__ tets, int64_t n_tets,
                                                            GridGeom g, double tet_mass, double* grid,
                                                            unsigned long long* d_stats) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_tets) return;
    const float* p = tets + 12 * t;
    V3<double> v[4];
    for (int q = 0; q < 4; ++q) v[q] = mk(double(p[3 * q]), double(p[3 * q + 1]), double(p[3 * q + 2]));
    AtomicEmit emit{grid};
    if (kCount) {
        TetCounters cnt = {};
        deposit_tet_f64(v, g, tet_mass, emit, &cnt);
        flush_counters(cnt, d_stats);
    } else {
        deposit_tet_f64(v, g, tet_mass, emit, nullptr);
    }
}

__global__ void k_add_f64_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        dst[i] += float(src[i]);
}

__global__ void k_fill_f64(double* d, int64_t n, double v) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        d[i] = v;
}

// One thread per tet of the meta-patches: its 4 vertices, in the corner table's order, to the
// soup (float [n][4][3]); a cell with a sentinel vertex gives 6 degenerate tets at the origin.
__global__ void __launch_bounds__(256) k_patches_to_soup(const float* __restrict__ pos, int64_t n_tets, int m,
                                                         PackedCorners pc, float* __restrict__ soup) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_tets) return;
    const int64_t cell = t / 6;
    const int lt = int(t - cell * 6);
    const int cpp = m * m * m;
    const int64_t p = cell / cpp;
    const int c = int(cell - p * cpp);
    const int ck = c / (m * m), rest = c - ck * m * m, cj = rest / m, ci = rest - cj * m;
    const int64_t fpp = 3ll * (m + 1) * (m + 1) * (m + 1);
    const float* base = pos + p * fpp + 3 * (ci + (m + 1) * (cj + (m + 1) * ck));
    const int ax = 3, ay = 3 * (m + 1), az = 3 * (m + 1) * (m + 1);
    const int slot_off[8] = {0, ax, ay, ay + ax, ay + az, ay + az + ax, az, az + ax};
    bool skip = false;
    for (int s = 0; s < 8; ++s) skip = skip || is_sentinel(base[slot_off[s]], base[slot_off[s] + 1], base[slot_off[s] + 2]);
    float* out = soup + 12 * t;
    for (int q = 0; q < 4; ++q) {
        const int code = pc.code(lt, q);
        const float* pv = base + 3 * (code & 1) + ay * ((code >> 1) & 1) + az * ((code >> 2) & 1);
        for (int a = 0; a < 3; ++a) out[3 * q + a] = skip ? 0.f : pv[a];
    }
}

inline unsigned blocks_for(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_patches_to_soup(const float* d_positions, int64_t num_patches, int m, int decomp, float* d_soup,
                                  hipStream_t stream) {
    const int64_t n = num_patches * int64_t(m) * m * m * 6;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_patches_to_soup, dim3(blocks_for(n, 256)), dim3(256), 0, stream, d_positions, n, m,
                       pack_corners(corner_table(decomp)), d_soup);
    return hipGetLastError();
}

hipError_t launch_deposit_patches_plain(const float* d_positions, int64_t num_patches, int m, int decomp,
                                        const GridGeom& g, double tet_mass, double* d_grid,
                                        unsigned long long* d_stats, hipStream_t stream) {
    const int64_t n = num_patches * int64_t(m) * m * m * 6;
    if (n <= 0) return hipSuccess;
    const PackedCorners pc = pack_corners(corner_table(decomp));
    if (d_stats)
        hipLaunchKernelGGL(k_deposit_patches_plain<true>, dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                           d_positions, n, m, g, tet_mass, d_grid, d_stats, pc);
    else
        hipLaunchKernelGGL(k_deposit_patches_plain<false>, dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                           d_positions, n, m, g, tet_mass, d_grid, d_stats, pc);
    return hipGetLastError();
}

hipError_t launch_deposit_lattice_plain(const LatticeView& lat, int decomp, const int32_t* d_tet_ids,
                                        int64_t n_tets, const GridGeom& g, double tet_mass,
                                        double* d_grid, unsigned long long* d_stats, hipStream_t stream) {
    if (n_tets <= 0) return hipSuccess;
    const PackedCorners ct = pack_corners(corner_table(decomp));
    if (d_stats)
        hipLaunchKernelGGL(k_deposit_lattice_plain<true>, dim3(blocks_for(n_tets, 256)), dim3(256), 0, stream,
                           lat, ct, d_tet_ids, n_tets, g, tet_mass, d_grid, d_stats);
    else
        hipLaunchKernelGGL(k_deposit_lattice_plain<false>, dim3(blocks_for(n_tets, 256)), dim3(256), 0, stream,
                           lat, ct, d_tet_ids, n_tets, g, tet_mass, d_grid, d_stats);
    return hipGetLastError();
}

hipError_t launch_deposit_soup_plain(const float* d_tets, int64_t n_tets, const GridGeom& g,
                                     double tet_mass, double* d_grid, unsigned long long* d_stats,
                                     hipStream_t stream) {
    if (n_tets <= 0) return hipSuccess;
    if (d_stats)
        hipLaunchKernelGGL(k_deposit_soup_plain<true>, dim3(blocks_for(n_tets, 256)), dim3(256), 0, stream,
                           d_tets, n_tets, g, tet_mass, d_grid, d_stats);
    else
        hipLaunchKernelGGL(k_deposit_soup_plain<false>, dim3(blocks_for(n_tets, 256)), dim3(256), 0, stream,
                           d_tets, n_tets, g, tet_mass, d_grid, d_stats);
    return hipGetLastError();
}

hipError_t launch_add_f64_to_f32(const double* d_src, float* d_dst, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = unsigned(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
    hipLaunchKernelGGL(k_add_f64_to_f32, dim3(nb), dim3(256), 0, stream, d_src, d_dst, n);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* d, int64_t n, double v, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = unsigned(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
    hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, stream, d, n, v);
    return hipGetLastError();
}

}  // namespace adeco
