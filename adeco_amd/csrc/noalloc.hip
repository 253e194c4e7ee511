// adeco_deposit_patches_noalloc: the boundary deposit (adeco_deposit_patches in the plain form on
// one device) without any device allocation, for callers that cannot allocate device memory.
//
// Its device memory beyond the caller's buffers lives in one file-scope device array, `scratch`,
// of kScratchBytes: the FP64 accumulation grid (dims^3 doubles), the caller's float grid staged
// for the final add (dims^3 floats), then the meta-patches uploaded pass by pass in the rest.  Each
// GPU holds its own copy of the array from module load to process end.  Host code never uses the
// array's name as an address: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device on the first call there, and only that address goes
// to kernels and copies.  Calls on one device share its copy; they serialise on the device's lock,
// which the other boundary calls on that device take too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/adeco.h"
#include "kernels.hpp"
#include "runtime.hpp"

namespace adeco {

constexpr size_t kScratchBytes = size_t(48) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch copy (current device = dev); null on error.
void* scratch_address(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            fail(ADECO_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

// Largest grid side the scratch array holds with room for one meta-patch (m = 7).
static int max_dims_for(size_t patch_bytes) {
    int d = 1;
    while (align256(size_t(d + 1) * (d + 1) * (d + 1) * 8) + align256(size_t(d + 1) * (d + 1) * (d + 1) * 4) +
               patch_bytes <= kScratchBytes)
        ++d;
    return d;
}

}  // namespace adeco

using namespace adeco;

extern "C" {

size_t adeco_noalloc_scratch_bytes(void) { return kScratchBytes; }

adeco_status adeco_deposit_patches_noalloc(int device_id, size_t patch_offset, size_t num_patches,
                                           const float* positions, size_t positions_size, int m, const float* bbox,
                                           int dims, float* grid_accum) {
    g_last_kernel_ms = 0.0;
    if (positions_size == 0) return ADECO_OK;  // AdDeposit2Grid.cpp:86-87
    if (!positions || !bbox || !grid_accum) return fail(ADECO_EINVAL, "deposit_tets_noalloc(): null pointer argument");
    if (m < 1 || dims < 1) return fail(ADECO_EINVAL, "deposit_tets_noalloc(): meta_cells_per_patch and dims must be >= 1");
    if (!valid_bbox(bbox))
        return fail(ADECO_EINVAL, "deposit_tets_noalloc(): bbox must be finite with min < max on every axis");
    const size_t fpp = 3 * size_t(m + 1) * (m + 1) * (m + 1);
    if (positions_size % fpp != 0)
        return fail(ADECO_EINVAL, "deposit_tets_noalloc(): positions_size is not a whole number of meta-patches");
    const size_t total = positions_size / fpp;
    if (patch_offset + num_patches > total || patch_offset + num_patches < patch_offset)
        return fail(ADECO_EINVAL, "deposit_tets_noalloc(): inconsistent point array (patch range beyond positions)");
    if (num_patches == 0) return ADECO_OK;
    const size_t patch_bytes = fpp * sizeof(float);
    // (dims above 1024 cannot fit; the bound also keeps dims^3 from overflowing)
    const size_t ncell = dims <= 1024 ? size_t(dims) * dims * dims : 0;
    const size_t grid_off = 0, acc_off = align256(ncell * sizeof(double)), pos_off = acc_off + align256(ncell * sizeof(float));
    if (dims > 1024 || pos_off + patch_bytes > kScratchBytes)
        return fail(ADECO_EOVERFLOW, "deposit_tets_noalloc(): a " + std::to_string(dims) + "^3 grid does not fit the " +
                                         std::to_string(kScratchBytes >> 20) + " MiB scratch (at most " +
                                         std::to_string(max_dims_for(patch_bytes)) + "^3 cells with these meta-patches)");
    if (device_id < 0 || device_id >= kMaxDevices)
        return fail(ADECO_EINVAL, "deposit_tets_noalloc(): device_id out of range");
    adeco_status st = check_device(device_id);
    if (st) return st;
    DeviceState& ds = device_state(device_id);
    std::lock_guard<std::mutex> lk(ds.mutex);
    if ((st = ds.ensure(device_id))) return st;
    hipError_t e = hipSetDevice(device_id);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    char* base = static_cast<char*>(scratch_address(device_id));
    if (!base) return ADECO_EDEVICE;
    double* d_grid = reinterpret_cast<double*>(base + grid_off);
    float* d_acc = reinterpret_cast<float*>(base + acc_off);
    float* d_pos = reinterpret_cast<float*>(base + pos_off);
    const size_t patches_per_pass = (kScratchBytes - pos_off) / patch_bytes;
    const GridGeom g = make_grid_geom(bbox, dims, cell_scaling());
    const int decomp = adeco_get_tet_decomposition();
    const double tet_mass = adeco_get_tet_mass();
    // the array keeps the last call's contents: the FP64 grid starts at zero
    if ((e = launch_fill_f64(d_grid, int64_t(ncell), 0.0, ds.stream)) != hipSuccess) return hip_fail(e, "grid clear");
    if ((st = timed_begin(ds))) return st;
    for (size_t done = 0; done < num_patches; done += patches_per_pass) {
        const size_t np = std::min(patches_per_pass, num_patches - done);
        e = hipMemcpyAsync(d_pos, positions + (patch_offset + done) * fpp, np * patch_bytes, hipMemcpyHostToDevice,
                           ds.stream);
        if (e != hipSuccess) return hip_fail(e, "positions upload");
        e = launch_deposit_patches_plain(d_pos, int64_t(np), m, decomp, g, tet_mass, d_grid, nullptr, ds.stream);
        if (e != hipSuccess) return hip_fail(e, "deposit launch");
    }
    if ((st = timed_end(ds))) return st;
    // grid_accum[c] += float(sum[c]), as the host add of the boundary call (AdDeposit2Grid.cu:510-513)
    e = hipMemcpyAsync(d_acc, grid_accum, ncell * sizeof(float), hipMemcpyHostToDevice, ds.stream);
    if (e == hipSuccess) e = launch_add_f64_to_f32(d_grid, d_acc, int64_t(ncell), ds.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(grid_accum, d_acc, ncell * sizeof(float), hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "grid accumulate");
    return ADECO_OK;
}

}  // extern "C"
