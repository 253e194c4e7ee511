// Host runtime of libadeco_amd.so: the C ABI declared in include/adeco.h.
//
// Device state is kept per GPU (stream at default priority, reusable buffers, timing events) and
// guarded by a per-device mutex, so the boundary can be called concurrently from several host
// threads, one GPU each, as the reference's worker threads do (AdDynamicLoadBalancing.h:183-229).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/adeco.h"
#include "batched_impl.hpp"
#include "deposit_tet.hpp"
#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "runtime.hpp"

namespace adeco {

thread_local std::string g_last_error;
thread_local double g_last_kernel_ms = 0.0;
thread_local LastTimings g_timings;
static size_t g_memory_budget = 0;  // 0 -> 75% of device memory
static std::mutex g_budget_mutex;
static std::atomic<int> g_cell_scaling{kScalingExact};
// the boundary call's form (adeco_deposit_patches): tets, tet mass, kernel form and precision
static std::atomic<int> g_boundary_decomp{REF_EFFECTIVE};
static std::atomic<double> g_boundary_tet_mass{1.0};  // resample_meta_patch, AdDepositTets2Grid.h:1007-1053
static std::atomic<int> g_boundary_mode{ADECO_MODE_PLAIN};
static std::atomic<int> g_boundary_precision{ADECO_PREC_FP64};

int cell_scaling() { return g_cell_scaling.load(); }

adeco_status fail(adeco_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

adeco_status hip_fail(hipError_t e, const char* what) {
    return fail(ADECO_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

DeviceState& device_state(int dev) {
    static std::mutex m;
    static std::vector<std::unique_ptr<DeviceState>> states;
    std::lock_guard<std::mutex> lk(m);
    if (int(states.size()) <= dev) states.resize(dev + 1);
    if (!states[dev]) states[dev].reset(new DeviceState());
    states[dev]->id = dev;
    return *states[dev];
}

adeco_status DeviceState::ensure(int dev) {
    if (ready) return ADECO_OK;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e == hipSuccess) e = hipEventCreate(&evk0);
    if (e == hipSuccess) e = hipEventCreate(&evk1);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    size_t free_b = 0, total_b = 0;
    e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    total_mem = total_b;
    ready = true;
    return ADECO_OK;
}

adeco_status DevBuf::reserve(size_t nbytes) {
    if (nbytes <= cap) return ADECO_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max(nbytes, size_t(1) << 20);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(ADECO_ENOMEM, std::string("hipMalloc of ") + std::to_string(want) + " bytes: " + hipGetErrorString(e));
    }
    cap = want;
    return ADECO_OK;
}

void DevBuf::free_all() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

DevBuf::~DevBuf() {
    if (ptr) (void)hipFree(ptr);
}

size_t memory_budget(const DeviceState& ds) {
    std::lock_guard<std::mutex> lk(g_budget_mutex);
    return g_memory_budget ? g_memory_budget : size_t(0.75 * double(ds.total_mem));
}

adeco_status check_device(int dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(ADECO_EDEVICE, "no usable GPU (hipGetDeviceCount reports none)");
    if (dev < 0 || dev >= n)
        return fail(ADECO_EINVAL, "selected invalid device number: " + std::to_string(dev) + " of " + std::to_string(n));
    return ADECO_OK;
}

bool valid_bbox(const float* b) {
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(b[2 * a]) || !std::isfinite(b[2 * a + 1]) || !(b[2 * a] < b[2 * a + 1])) return false;
    return true;
}

adeco_status timed_begin(DeviceState& ds) {
    hipError_t e = hipEventRecord(ds.ev0, ds.stream);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "hipEventRecord");
}

adeco_status timed_end(DeviceState& ds) {
    hipError_t e = hipEventRecord(ds.ev1, ds.stream);
    if (e == hipSuccess) e = hipEventSynchronize(ds.ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ds.ev0, ds.ev1);
    if (e != hipSuccess) return hip_fail(e, "kernel timing");
    g_last_kernel_ms = ms;
    return ADECO_OK;
}

}  // namespace adeco

using namespace adeco;

namespace {

// One device's share of a boundary call: passes over its patch range within the budget.  The
// plain form runs one thread per tet over the uploaded patches; the batched form expands each
// pass's patches into a tet soup on the device and runs the batched persistent deposit on it, so
// a pass budgets the soup and the soup's candidate space as well (the tile map, about one 64th of
// the candidates in 4 bytes each, is allocated on top).
adeco_status deposit_patches_on_device(int dev, size_t first_patch, size_t n_patches, const float* positions,
                                       int m, const float* bbox, int dims, std::vector<double>& host_sum) {
    adeco_status st = check_device(dev);
    if (st) return st;
    DeviceState& ds = device_state(dev);
    std::lock_guard<std::mutex> lk(ds.mutex);
    if ((st = ds.ensure(dev))) return st;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const size_t fpp = 3 * size_t(m + 1) * (m + 1) * (m + 1);
    const size_t ncell = size_t(dims) * dims * dims;
    const size_t grid_bytes = ncell * sizeof(double);
    const size_t budget = memory_budget(ds);
    if (grid_bytes >= budget)  // AdDeposit2Grid.cu:314-317
        return fail(ADECO_EOVERFLOW, "grid of " + std::to_string(grid_bytes) + " bytes exceeds the device memory budget");
    const size_t avail = budget - grid_bytes;
    const int decomp = g_boundary_decomp.load();
    const double tet_mass = g_boundary_tet_mass.load();
    const bool batched = g_boundary_mode.load() == ADECO_MODE_BATCHED;
    const int precision = g_boundary_precision.load();
    if (!batched && precision != ADECO_PREC_FP64)
        return fail(ADECO_EINVAL, "deposit_tets(): precision modes other than FP64 need the batched form");
    if (batched && dims > 1023) return fail(ADECO_EINVAL, "deposit_tets(): the batched form needs dims < 1024");
    const size_t patch_bytes = fpp * sizeof(float);
    const size_t tets_per_patch = 6 * size_t(m) * m * m;
    const size_t soup_bytes = tets_per_patch * 12 * sizeof(float);
    const size_t per_patch = batched ? patch_bytes + soup_bytes + tets_per_patch * 4 + candidate_space_bytes(int64_t(tets_per_patch))
                                     : patch_bytes;
    const size_t patches_per_pass = std::max<size_t>(1, std::min(n_patches, avail / per_patch));
    if (patches_per_pass * per_patch > avail)
        return fail(ADECO_EOVERFLOW, "memory budget too small for one meta-patch");
    if ((st = ds.grid.reserve(grid_bytes))) return st;
    if ((st = ds.pos.reserve(patches_per_pass * patch_bytes))) return st;
    DevBuf& soup = ds.soup;
    if (batched && (st = soup.reserve(patches_per_pass * soup_bytes))) return st;
    const GridGeom g = make_grid_geom(bbox, dims, cell_scaling());
    if ((e = launch_fill_f64(ds.grid.as<double>(), int64_t(ncell), 0.0, ds.stream)) != hipSuccess)
        return hip_fail(e, "grid clear");
    double total_ms = 0.0;
    for (size_t done = 0; done < n_patches; done += patches_per_pass) {
        const size_t np = std::min(patches_per_pass, n_patches - done);
        e = hipMemcpyAsync(ds.pos.ptr, positions + (first_patch + done) * fpp, np * patch_bytes,
                           hipMemcpyHostToDevice, ds.stream);
        if (e != hipSuccess) return hip_fail(e, "positions upload");
        if (batched) {
            e = launch_patches_to_soup(ds.pos.as<float>(), int64_t(np), m, decomp, soup.as<float>(), ds.stream);
            if (e != hipSuccess) return hip_fail(e, "patch expansion");
            if ((st = deposit_soup_batched(ds, soup.as<float>(), int64_t(np * tets_per_patch), g, tet_mass, precision,
                                           ds.grid.as<double>())))
                return st;
            total_ms += g_timings.main_ms;
            continue;
        }
        if ((st = timed_begin(ds))) return st;
        e = launch_deposit_patches_plain(ds.pos.as<float>(), int64_t(np), m, decomp, g, tet_mass, ds.grid.as<double>(),
                                         nullptr, ds.stream);
        if (e != hipSuccess) return hip_fail(e, "deposit launch");
        if ((st = timed_end(ds))) return st;
        total_ms += g_last_kernel_ms;
    }
    g_last_kernel_ms = total_ms;
    std::vector<double> tmp(ncell);
    e = hipMemcpyAsync(tmp.data(), ds.grid.ptr, grid_bytes, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "grid download");
    for (size_t c = 0; c < ncell; ++c) host_sum[c] += tmp[c];
    return ADECO_OK;
}

}  // namespace

extern "C" {

const char* adeco_last_error(void) { return g_last_error.c_str(); }
int adeco_version(void) { return 100; }

int adeco_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return -1;
    return n;
}

int adeco_get_grid_size(void) { return 1024; }
int adeco_get_block_size(void) { return 256; }

adeco_status adeco_set_memory_budget(size_t bytes) {
    std::lock_guard<std::mutex> lk(g_budget_mutex);
    g_memory_budget = bytes;
    return ADECO_OK;
}

double adeco_last_kernel_ms(void) { return g_last_kernel_ms; }

adeco_status adeco_set_cell_scaling(int mode) {
    if (mode != ADECO_SCALING_EXACT && mode != ADECO_SCALING_REFERENCE)
        return fail(ADECO_EINVAL, "set_cell_scaling(): unknown mode");
    g_cell_scaling.store(mode);
    return ADECO_OK;
}

int adeco_get_cell_scaling(void) { return g_cell_scaling.load(); }

adeco_status adeco_set_tet_decomposition(int decomposition) {
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "set_tet_decomposition(): unknown decomposition");
    g_boundary_decomp.store(decomposition);
    return ADECO_OK;
}
int adeco_get_tet_decomposition(void) { return g_boundary_decomp.load(); }

adeco_status adeco_set_tet_mass(double tet_mass) {
    if (!(tet_mass > 0.0) || !std::isfinite(tet_mass)) return fail(ADECO_EINVAL, "set_tet_mass(): tet mass must be positive and finite");
    g_boundary_tet_mass.store(tet_mass);
    return ADECO_OK;
}
double adeco_get_tet_mass(void) { return g_boundary_tet_mass.load(); }

adeco_status adeco_set_deposit_mode(int mode) {
    if (mode != ADECO_MODE_PLAIN && mode != ADECO_MODE_BATCHED) return fail(ADECO_EINVAL, "set_deposit_mode(): unknown mode");
    g_boundary_mode.store(mode);
    return ADECO_OK;
}
int adeco_get_deposit_mode(void) { return g_boundary_mode.load(); }

adeco_status adeco_set_precision_mode(int precision) {
    if (precision < ADECO_PREC_FP64 || precision > ADECO_PREC_FP32) return fail(ADECO_EINVAL, "set_precision_mode(): unknown precision");
    g_boundary_precision.store(precision);
    return ADECO_OK;
}
int adeco_get_precision_mode(void) { return g_boundary_precision.load(); }

void adeco_last_timings(double* out5) {
    if (!out5) return;
    out5[0] = g_timings.total_ms;
    out5[1] = g_timings.main_ms;
    out5[2] = g_timings.candidates;
    out5[3] = g_timings.partials;
    out5[4] = g_timings.fallbacks;
}

adeco_status adeco_deposit_patches(int use_specific_gpu, size_t device_id, size_t patch_offset,
                                   size_t num_patches, const float* positions, size_t positions_size,
                                   int m, const float* bbox, int dims, float* grid_accum) {
    g_last_kernel_ms = 0.0;
    if (positions_size == 0) return ADECO_OK;  // AdDeposit2Grid.cpp:86-87
    if (!positions || !bbox || !grid_accum) return fail(ADECO_EINVAL, "deposit_tets(): null pointer argument");
    if (m < 1 || dims < 1) return fail(ADECO_EINVAL, "deposit_tets(): meta_cells_per_patch and dims must be >= 1");
    if (!valid_bbox(bbox)) return fail(ADECO_EINVAL, "deposit_tets(): bbox must be finite with min < max on every axis");
    const size_t fpp = 3 * size_t(m + 1) * (m + 1) * (m + 1);
    if (positions_size % fpp != 0)  // AdDeposit2Grid.cpp:100
        return fail(ADECO_EINVAL, "deposit_tets(): positions_size is not a whole number of meta-patches");
    const size_t total = positions_size / fpp;
    if (patch_offset + num_patches > total || patch_offset + num_patches < patch_offset)  // AdDeposit2Grid.cu:265-268
        return fail(ADECO_EINVAL, "deposit_tets(): inconsistent point array (patch range beyond positions)");
    if (num_patches == 0) return ADECO_OK;
    const size_t ncell = size_t(dims) * dims * dims;
    std::vector<double> sum(ncell, 0.0);
    if (use_specific_gpu) {
        adeco_status st = deposit_patches_on_device(int(device_id), patch_offset, num_patches, positions, m, bbox, dims, sum);
        if (st) return st;
    } else {
        const int n = adeco_device_count();
        if (n <= 0) return fail(ADECO_EDEVICE, "no usable GPU (hipGetDeviceCount reports none)");
        const size_t per = (num_patches + size_t(n) - 1) / size_t(n);
        double ms = 0.0;
        for (int d = 0; d < n; ++d) {
            const size_t b = std::min(num_patches, per * size_t(d));
            const size_t cnt = std::min(per, num_patches - b);
            if (!cnt) continue;
            adeco_status st = deposit_patches_on_device(d, patch_offset + b, cnt, positions, m, bbox, dims, sum);
            if (st) return st;
            ms += g_last_kernel_ms;
        }
        g_last_kernel_ms = ms;
    }
    for (size_t c = 0; c < ncell; ++c) grid_accum[c] += float(sum[c]);  // AdDeposit2Grid.cu:510-513
    return ADECO_OK;
}

double adeco_selftest_clip_vol6_f64(const double* t) {
    V3<double> v[4];
    for (int q = 0; q < 4; ++q) v[q] = mk(t[3 * q], t[3 * q + 1], t[3 * q + 2]);
    return clip_unit_cube_vol6(v);
}

float adeco_selftest_clip_vol6_f32(const float* t) {
    V3<float> v[4];
    for (int q = 0; q < 4; ++q) v[q] = mk(t[3 * q], t[3 * q + 1], t[3 * q + 2]);
    return clip_unit_cube_vol6(v);
}

int adeco_selftest_classify_f64(const double* t) {
    V3<double> v[4];
    for (int q = 0; q < 4; ++q) v[q] = mk(t[3 * q], t[3 * q + 1], t[3 * q + 2]);
    return classify_unit(v);
}

double adeco_selftest_beyond_fraction_f64(const double* d) { return beyond_fraction(d[0], d[1], d[2], d[3]); }

void adeco_selftest_column_f64(const double* t, int ls, int use_signed, double* vol6_out, double* sab_out) {
    if (!t || !vol6_out || !sab_out || ls < 1 || ls > 8) return;
    V3<double> v[4];
    for (int q = 0; q < 4; ++q) v[q] = mk(t[3 * q], t[3 * q + 1], t[3 * q + 2]);
    for (int c = 0; c < ls; ++c) vol6_out[c] = sab_out[c] = 0.0;
    ColumnLeaf<double, 1> leaf{vol6_out, sab_out, ls};
    if (use_signed) clip_pieces<double, 7>(v, true, leaf);
    else clip_pieces<double, 6>(v, false, leaf);
}

adeco_status adeco_selftest_deposit_soup_f64(const float* tets, int64_t n_tets, const float* bbox6, int dims,
                                             double tet_mass, double* grid) {
    if (!tets || !bbox6 || !grid || dims < 1 || !valid_bbox(bbox6)) return fail(ADECO_EINVAL, "selftest: invalid argument");
    const GridGeom g = make_grid_geom(bbox6, dims, cell_scaling());
    struct HostEmit {
        double* grid;
        void operator()(int64_t idx, double dm) { grid[idx] += dm; }
    } emit{grid};
    for (int64_t t = 0; t < n_tets; ++t) {
        V3<double> v[4];
        for (int q = 0; q < 4; ++q)
            v[q] = mk(double(tets[12 * t + 3 * q]), double(tets[12 * t + 3 * q + 1]), double(tets[12 * t + 3 * q + 2]));
        deposit_tet_f64(v, g, tet_mass, emit, nullptr);
    }
    return ADECO_OK;
}

void adeco_selftest_decomposition(int decomposition, int32_t* out) {
    const CornerTable t = corner_table(decomposition);
    for (int i = 0; i < 6; ++i)
        for (int v = 0; v < 4; ++v) out[4 * i + v] = t.c[i][v];
}

}  // extern "C"
