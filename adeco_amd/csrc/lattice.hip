// Resident lattice and single-grid deposits (C ABI part 2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/adeco.h"
#include "batched_impl.hpp"
#include "deposit_tet.hpp"
#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

namespace adeco {

namespace {

// Per-block float min/max of the three coordinate arrays; out[6*block + 2a] = min, +1 = max.
__global__ void __launch_bounds__(256) k_minmax(const float* __restrict__ x, const float* __restrict__ y,
                                                const float* __restrict__ z, int64_t n, float* out) {
    __shared__ float sm[6][256];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const float p[3] = {x[i], y[i], z[i]};
        for (int a = 0; a < 3; ++a) {
            mn[a] = p[a] < mn[a] ? p[a] : mn[a];
            mx[a] = p[a] > mx[a] ? p[a] : mx[a];
        }
    }
    for (int a = 0; a < 3; ++a) {
        sm[2 * a][threadIdx.x] = mn[a];
        sm[2 * a + 1][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int a = 0; a < 3; ++a) {
                const float o0 = sm[2 * a][threadIdx.x + s], o1 = sm[2 * a + 1][threadIdx.x + s];
                if (o0 < sm[2 * a][threadIdx.x]) sm[2 * a][threadIdx.x] = o0;
                if (o1 > sm[2 * a + 1][threadIdx.x]) sm[2 * a + 1][threadIdx.x] = o1;
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * blockIdx.x + threadIdx.x] = sm[threadIdx.x][0];
}

// apply_boundary_corrections (AdPointDataReaders.h:156-206) per vertex: along each axis, a vertex
// in the upper half of the lattice (index > dims/2) lying below the box's first quarter mark moves
// up by the shift; one in the lower half lying above the third quarter mark moves down.  The
// thresholds come from the host in float exactly as the reference forms them, so the kernel does
// only compares and one float add or subtract per moved coordinate.
struct UnwrapParams {
    float lo[3], hi[3], half[3], shift[3];
    int np[3];
};

__global__ void __launch_bounds__(256) k_periodic_unwrap(float* __restrict__ x, float* __restrict__ y,
                                                         float* __restrict__ z, int64_t n, UnwrapParams p) {
    for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += int64_t(gridDim.x) * blockDim.x) {
        const int64_t plane = int64_t(p.np[0]) * p.np[1];
        const int idx[3] = {int(v % p.np[0]), int((v / p.np[0]) % p.np[1]), int(v / plane)};
        float* c[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
            const float pos = c[a][v];
            const float fi = float(idx[a]);
            if (fi > p.half[a] && pos < p.lo[a])
                c[a][v] = pos + p.shift[a];
            else if (fi < p.half[a] && pos > p.hi[a])
                c[a][v] = pos - p.shift[a];
        }
    }
}

__global__ void __launch_bounds__(256) k_iota_i32(int32_t* out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = int32_t(i);
}

}  // namespace

// Batched deposit of a tet soup (already on the device) onto one grid: every tet is one pair of
// a one-grid candidate space.  Kernel time, partial cells and FP64 recomputes go to g_timings.
adeco_status deposit_soup_batched(DeviceState& ds, const float* d_tets, int64_t n_tets, const GridGeom& g, double tet_mass,
                                  int precision, double* d_grid) {
    g_timings = LastTimings();
    if (n_tets == 0) return ADECO_OK;
    if (n_tets >= (int64_t(1) << 31) - 2) return fail(ADECO_EOVERFLOW, "deposit_soup(): too many tets for the batched form");
    if (g.dims > 1023) return fail(ADECO_EINVAL, "deposit_soup(): the batched form needs dims < 1024");
    DevBuf &ids = ds.soup_ids, &desc = ds.soup_desc, &work = ds.soup_work, &tiles = ds.soup_tiles, &ctr = ds.soup_ctr;
    adeco_status st;
    if ((st = ids.reserve(size_t(n_tets) * 4)) || (st = desc.reserve(512)) ||
        (st = work.reserve(candidate_space_bytes(n_tets) + 256)) || (st = ctr.reserve(64)))
        return st;
    const GridDesc gd = make_grid_desc(g, 0);
    const int64_t goff[2] = {0, n_tets};
    GridDesc* d_gd = desc.as<GridDesc>();
    int64_t* d_goff = reinterpret_cast<int64_t*>(desc.as<char>() + 256);
    hipLaunchKernelGGL(k_iota_i32, dim3(unsigned((n_tets + 255) / 256)), dim3(256), 0, ds.stream, ids.as<int32_t>(), n_tets);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(d_gd, &gd, sizeof(gd), hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_goff, goff, sizeof(goff), hipMemcpyHostToDevice, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "soup setup");
    LatticeView lv{nullptr, nullptr, nullptr, {0, 0, 0}, d_tets};
    const PackedCorners ct = pack_corners(corner_table(0));
    CandidateSpace cs;
    if ((st = build_candidate_space(ds.stream, lv, ct, ids.as<int32_t>(), n_tets, d_goff, 1, d_gd, tet_mass, precision,
                                    work.ptr, cs)))
        return st;
    if ((st = tiles.reserve(tile_map_bytes(cs.n_cand))) || (st = build_tile_map(ds.stream, cs, tiles.as<int32_t>()))) return st;
    unsigned* q = ctr.as<unsigned>();
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(ctr.as<char>() + 16);
    e = launch_zero_u32(q, 16, ds.stream);
    const TileSource src{cs.pd, cs.cand_off, cs.tile_first, cs.n_tiles, cs.n_cand, d_gd, d_grid, tet_mass};
    if (e == hipSuccess) e = hipEventRecord(ds.evk0, ds.stream);
    if (e == hipSuccess) e = launch_deposit_tiles(ds.stream, precision, src, q, counters, ds.id);
    if (e == hipSuccess) e = hipEventRecord(ds.evk1, ds.stream);
    unsigned long long cnt[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, counters, 16, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ds.evk0, ds.evk1);
    if (e != hipSuccess) return hip_fail(e, "soup batched deposit");
    g_timings.main_ms = ms;
    g_timings.candidates = double(cs.n_cand);
    g_timings.fallbacks = double(cnt[0]);
    g_timings.partials = double(cnt[1]);
    return ADECO_OK;
}

adeco_status lattice_bbox_device(adeco_lattice_s* L, float out6[6]) {
    DeviceState& ds = *L->ds;
    const int nb = 512;
    adeco_status st = ds.scratch.reserve(sizeof(float) * 6 * nb);
    if (st) return st;
    hipLaunchKernelGGL(k_minmax, dim3(nb), dim3(256), 0, ds.stream, L->d_x, L->d_y, L->d_z, L->n_vertices,
                       ds.scratch.as<float>());
    hipError_t e = hipGetLastError();
    std::vector<float> part(6 * nb);
    if (e == hipSuccess) e = hipMemcpyAsync(part.data(), ds.scratch.ptr, sizeof(float) * 6 * nb, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "lattice bbox");
    for (int a = 0; a < 3; ++a) {
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int b = 0; b < nb; ++b) {
            mn = std::min(mn, part[6 * b + 2 * a]);
            mx = std::max(mx, part[6 * b + 2 * a + 1]);
        }
        out6[2 * a] = mn;
        out6[2 * a + 1] = mx;
    }
    return ADECO_OK;
}

// A lattice handle with uninitialised device arrays for npx*npy*npz vertices on device_id.
adeco_status lattice_alloc(int device_id, int npx, int npy, int npz, adeco_lattice_s*& out) {
    out = nullptr;
    adeco_status st = check_device(device_id);
    if (st) return st;
    DeviceState& ds = device_state(device_id);
    std::lock_guard<std::mutex> lk(ds.mutex);
    if ((st = ds.ensure(device_id))) return st;
    hipError_t e = hipSetDevice(device_id);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    auto* L = new adeco_lattice_s();
    L->device = device_id;
    L->ds = &ds;
    L->np[0] = npx;
    L->np[1] = npy;
    L->np[2] = npz;
    L->n_vertices = int64_t(npx) * npy * npz;
    const size_t bytes = size_t(L->n_vertices) * sizeof(float);
    e = hipMalloc(&L->d_x, bytes);
    if (e == hipSuccess) e = hipMalloc(&L->d_y, bytes);
    if (e == hipSuccess) e = hipMalloc(&L->d_z, bytes);
    if (e != hipSuccess) {
        delete L;
        return fail(ADECO_ENOMEM, std::string("lattice_create(): ") + hipGetErrorString(e));
    }
    out = L;
    return ADECO_OK;
}

}  // namespace adeco

using namespace adeco;

extern "C" {

adeco_status adeco_lattice_create(int device_id, const float* x, const float* y, const float* z, int npx,
                                  int npy, int npz, adeco_lattice* out) {
    if (!out || !x || !y || !z) return fail(ADECO_EINVAL, "lattice_create(): null pointer argument");
    *out = nullptr;
    if (npx < 2 || npy < 2 || npz < 2) return fail(ADECO_EINVAL, "lattice_create(): need at least 2 vertices per axis");
    const int64_t n = int64_t(npx) * npy * npz;
    if (n > kMaxLatticeVertices) return fail(ADECO_EOVERFLOW, "lattice_create(): lattice too large for 32-bit tet ids");
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]))
            return fail(ADECO_EINVAL, "lattice_create(): non-finite vertex coordinate");
    adeco_lattice_s* L = nullptr;
    adeco_status st = lattice_alloc(device_id, npx, npy, npz, L);
    if (st) return st;
    DeviceState& ds = *L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    const size_t bytes = size_t(n) * sizeof(float);
    hipError_t e = hipMemcpyAsync(L->d_x, x, bytes, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(L->d_y, y, bytes, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(L->d_z, z, bytes, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) {
        delete L;
        return hip_fail(e, "lattice upload");
    }
    *out = L;
    return ADECO_OK;
}

adeco_status adeco_lattice_update(adeco_lattice L, const float* x, const float* y, const float* z) {
    if (!L || !x || !y || !z) return fail(ADECO_EINVAL, "lattice_update(): null pointer argument");
    const int64_t n = L->n_vertices;
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]))
            return fail(ADECO_EINVAL, "lattice_update(): non-finite vertex coordinate");
    DeviceState& ds = *L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(L->device);
    const size_t bytes = size_t(n) * sizeof(float);
    if (e == hipSuccess) e = hipMemcpyAsync(L->d_x, x, bytes, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(L->d_y, y, bytes, hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(L->d_z, z, bytes, hipMemcpyHostToDevice, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "lattice update");
    if (!L->meta_box.empty()) {
        // the meta-cell boxes that ROI selection culls with follow the new positions
        const size_t M = L->meta_box.size() / 6;
        DevBuf box, cells;
        adeco_status st;
        if ((st = box.reserve(M * 6 * sizeof(float))) || (st = cells.reserve(M * sizeof(unsigned)))) return st;
        if ((st = octree_meta_bbox(L, L->meta_dims, box.as<float>(), cells.as<unsigned>()))) return st;
        e = hipMemcpyAsync(L->meta_box.data(), box.ptr, M * 6 * sizeof(float), hipMemcpyDeviceToHost, ds.stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "lattice update");
}

adeco_status adeco_lattice_destroy(adeco_lattice L) {
    if (!L) return ADECO_OK;
    std::lock_guard<std::mutex> lk(L->ds->mutex);
    (void)hipSetDevice(L->device);
    delete L;
    return ADECO_OK;
}

adeco_status adeco_lattice_bbox(adeco_lattice L, float* bbox6) {
    if (!L || !bbox6) return fail(ADECO_EINVAL, "lattice_bbox(): null argument");
    std::lock_guard<std::mutex> lk(L->ds->mutex);
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return lattice_bbox_device(L, bbox6);
}

adeco_status adeco_lattice_positions(adeco_lattice L, float* x, float* y, float* z) {
    if (!L || !x || !y || !z) return fail(ADECO_EINVAL, "lattice_positions(): null argument");
    std::lock_guard<std::mutex> lk(L->ds->mutex);
    hipError_t e = hipSetDevice(L->device);
    const size_t bytes = size_t(L->n_vertices) * sizeof(float);
    if (e == hipSuccess) e = hipMemcpy(x, L->d_x, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(y, L->d_y, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(z, L->d_z, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "lattice positions download");
}

adeco_status adeco_lattice_periodic_unwrap(adeco_lattice L, const float* bbox6, const float* period3) {
#pragma clang fp contract(off)
    if (!L || !bbox6) return fail(ADECO_EINVAL, "lattice_periodic_unwrap(): null argument");
    for (int a = 0; a < 3; ++a) {
        if (!(bbox6[2 * a] <= bbox6[2 * a + 1])) return fail(ADECO_EINVAL, "lattice_periodic_unwrap(): invalid bbox");
        if (period3 && !(period3[a] > 0.f && std::isfinite(period3[a])))
            return fail(ADECO_EINVAL, "lattice_periodic_unwrap(): period must be positive and finite");
    }
    // thresholds as AdPointDataReaders.h:159-162 forms them, in float
    UnwrapParams p;
    for (int a = 0; a < 3; ++a) {
        const float mn = bbox6[2 * a];
        const float ext = bbox6[2 * a + 1] - mn;
        p.lo[a] = mn + ext / 4.0f;
        p.hi[a] = mn + 0.75f * ext;
        p.half[a] = float(L->np[a]) / 2.0f;
        p.shift[a] = period3 ? period3[a] : ext;
        p.np[a] = L->np[a];
    }
    std::lock_guard<std::mutex> lk(L->ds->mutex);
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    DeviceState& ds = *L->ds;
    const unsigned nb = unsigned(std::min<int64_t>(L->n_vertices / 256 + 1, 4096));
    hipLaunchKernelGGL(k_periodic_unwrap, dim3(nb), dim3(256), 0, ds.stream, L->d_x, L->d_y, L->d_z, L->n_vertices, p);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "periodic unwrap");
    L->meta_box.clear();  // vertices moved
    L->has_domain = !period3;
    if (!period3) std::memcpy(L->domain, bbox6, sizeof(L->domain));
    return ADECO_OK;
}

adeco_status adeco_lattice_deposit_grid(adeco_lattice L, int decomposition, double tet_mass,
                                        const int32_t* tet_ids, int64_t n_tets, const float* bbox6, int dims,
                                        int mode, int precision, double* grid_out, uint64_t* stats) {
    g_last_kernel_ms = 0.0;
    if (!L || !bbox6 || !grid_out) return fail(ADECO_EINVAL, "lattice_deposit_grid(): null argument");
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "lattice_deposit_grid(): unknown decomposition");
    if (dims < 1 || !valid_bbox(bbox6)) return fail(ADECO_EINVAL, "lattice_deposit_grid(): invalid grid");
    if (mode != ADECO_MODE_PLAIN && mode != ADECO_MODE_BATCHED) return fail(ADECO_EINVAL, "lattice_deposit_grid(): unknown mode");
    if (precision < 0 || precision > 2) return fail(ADECO_EINVAL, "lattice_deposit_grid(): unknown precision");
    const int64_t all = 6 * int64_t(L->np[0] - 1) * (L->np[1] - 1) * (L->np[2] - 1);
    if (!tet_ids) n_tets = all;
    if (n_tets < 0) return fail(ADECO_EINVAL, "lattice_deposit_grid(): negative tet count");
    if (tet_ids)
        for (int64_t i = 0; i < n_tets; ++i)
            if (tet_ids[i] < 0 || tet_ids[i] >= all) return fail(ADECO_EINVAL, "lattice_deposit_grid(): tet id out of range");
    DeviceState& ds = *L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t ncell = size_t(dims) * dims * dims;
    adeco_status st = ds.grid.reserve(ncell * sizeof(double));
    if (st) return st;
    if (tet_ids && (st = ds.pos.reserve(size_t(n_tets) * sizeof(int32_t) + 4))) return st;
    if ((st = ds.stats.reserve(kNumStats * sizeof(unsigned long long)))) return st;
    const GridGeom g = make_grid_geom(bbox6, dims, cell_scaling());
    e = launch_fill_f64(ds.grid.as<double>(), int64_t(ncell), 0.0, ds.stream);
    if (e == hipSuccess && tet_ids)
        e = hipMemcpyAsync(ds.pos.ptr, tet_ids, size_t(n_tets) * sizeof(int32_t), hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess && stats) e = hipMemsetAsync(ds.stats.ptr, 0, kNumStats * sizeof(unsigned long long), ds.stream);
    if (e != hipSuccess) return hip_fail(e, "deposit setup");
    LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    if ((st = timed_begin(ds))) return st;
    e = launch_deposit_lattice_plain(lv, decomposition, tet_ids ? ds.pos.as<int32_t>() : nullptr, n_tets, g, tet_mass,
                                     ds.grid.as<double>(), stats ? ds.stats.as<unsigned long long>() : nullptr, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "deposit launch");
    if ((st = timed_end(ds))) return st;
    e = hipMemcpyAsync(grid_out, ds.grid.ptr, ncell * sizeof(double), hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess && stats)
        e = hipMemcpyAsync(stats, ds.stats.ptr, kNumStats * sizeof(unsigned long long), hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "deposit download");
    (void)mode;
    (void)precision;
    return ADECO_OK;
}

adeco_status adeco_deposit_soup(int device_id, const float* tets, int64_t n_tets, const float* bbox6, int dims,
                                double tet_mass, int mode, int precision, double* grid_out, uint64_t* stats) {
    g_last_kernel_ms = 0.0;
    if (!tets || !bbox6 || !grid_out || n_tets < 0) return fail(ADECO_EINVAL, "deposit_soup(): invalid argument");
    if (dims < 1 || !valid_bbox(bbox6)) return fail(ADECO_EINVAL, "deposit_soup(): invalid grid");
    if (mode != ADECO_MODE_PLAIN && mode != ADECO_MODE_BATCHED) return fail(ADECO_EINVAL, "deposit_soup(): unknown mode");
    if (precision < 0 || precision > 2) return fail(ADECO_EINVAL, "deposit_soup(): unknown precision");
    adeco_status st = check_device(device_id);
    if (st) return st;
    DeviceState& ds = device_state(device_id);
    std::lock_guard<std::mutex> lk(ds.mutex);
    if ((st = ds.ensure(device_id))) return st;
    hipError_t e = hipSetDevice(device_id);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t ncell = size_t(dims) * dims * dims;
    if ((st = ds.grid.reserve(ncell * sizeof(double)))) return st;
    if ((st = ds.pos.reserve(size_t(n_tets) * 12 * sizeof(float) + 4))) return st;
    if ((st = ds.stats.reserve(kNumStats * sizeof(unsigned long long)))) return st;
    const GridGeom g = make_grid_geom(bbox6, dims, cell_scaling());
    e = launch_fill_f64(ds.grid.as<double>(), int64_t(ncell), 0.0, ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ds.pos.ptr, tets, size_t(n_tets) * 12 * sizeof(float), hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess && stats) e = hipMemsetAsync(ds.stats.ptr, 0, kNumStats * sizeof(unsigned long long), ds.stream);
    if (e != hipSuccess) return hip_fail(e, "soup setup");
    if ((st = timed_begin(ds))) return st;
    if (mode == ADECO_MODE_BATCHED) {
        if ((st = deposit_soup_batched(ds, ds.pos.as<float>(), n_tets, g, tet_mass, precision, ds.grid.as<double>()))) return st;
    } else {
        e = launch_deposit_soup_plain(ds.pos.as<float>(), n_tets, g, tet_mass, ds.grid.as<double>(),
                                      stats ? ds.stats.as<unsigned long long>() : nullptr, ds.stream);
        if (e != hipSuccess) return hip_fail(e, "soup launch");
    }
    if ((st = timed_end(ds))) return st;
    e = hipMemcpyAsync(grid_out, ds.grid.ptr, ncell * sizeof(double), hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess && stats)
        e = hipMemcpyAsync(stats, ds.stats.ptr, kNumStats * sizeof(unsigned long long), hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "soup download");
    return ADECO_OK;
}

}  // extern "C"
