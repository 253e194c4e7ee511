// DARK_SKY particle files straight into a resident lattice (C ABI part 8): AdDarkSkyDataReader
// (AdPointDataReaders.h:361-767) on one rank, with the id placement of the one-rank particle
// redistribution (particles sorted by id, AdLagrangianRedistribution.h:255-470) done on the GPU.
//
// File: a text header whose "# SDF-EOH" line ends it, then `sha1_chunks` records of 24 bytes (the
// count from the header line "int sha1_chunks = N;"), then one 32-byte record per particle:
// float x, y, z, vx, vy, vz and an int64 id, host byte order.  The id's low 48 bits are the Morton
// code of the particle's lattice site: x in bits 0, 3, 6, ..., y in 1, 4, ..., z in 2, 5, ....
// With stride s (dividing the lattice size) only sites whose three indices are multiples of s are
// kept, at site / s.
//
// Here the header and the size checks run on the host; the particle records stream through a
// pinned host buffer to the device, where a kernel decodes each id and stores the position at its
// lattice site.  Every kept site must be written exactly once: the kernel counts kept records,
// repeated sites and ids outside the lattice, and the host turns any mismatch into an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "../../include/adeco.h"
#include "adeco_geom.hpp"
#include "lattice_impl.hpp"
#include "runtime.hpp"

using namespace adeco;

namespace {

struct DarkSkyRecord {
    float x, y, z, vx, vy, vz;
    long long id;
};
static_assert(sizeof(DarkSkyRecord) == 32, "a DARK_SKY particle record is 32 bytes");

// every third bit of v (bits 0, 3, 6, ...) packed into the low bits
ADECO_HD unsigned every_third_bit(unsigned long long v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return unsigned(v);
}

// lattice site of a DARK_SKY id (the low 48 bits as a Morton code)
ADECO_HD void site_of(long long id, unsigned s[3]) {
    const unsigned long long key = (unsigned long long)id & ((1ull << 48) - 1);
    s[0] = every_third_bit(key);
    s[1] = every_third_bit(key >> 1);
    s[2] = every_third_bit(key >> 2);
}

// counters: 0 kept records, 1 repeated sites, 2 ids outside the lattice
__global__ void __launch_bounds__(256) k_dark_sky_scatter(const DarkSkyRecord* __restrict__ rec, int64_t n, int full,
                                                          int stride, float* __restrict__ x, float* __restrict__ y,
                                                          float* __restrict__ z, unsigned* __restrict__ seen,
                                                          unsigned long long* __restrict__ counters) {
    const int d = full / stride;
    for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < n; base += int64_t(gridDim.x) * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        bool kept = false, repeat = false, outside = false;
        if (i < n) {
            const DarkSkyRecord r = rec[i];
            unsigned s[3];
            site_of(r.id, s);
            outside = s[0] >= unsigned(full) || s[1] >= unsigned(full) || s[2] >= unsigned(full);
            const bool skip = stride != 1 && (s[0] % stride || s[1] % stride || s[2] % stride);
            if (!outside && !skip) {
                const int64_t site = int64_t(s[0] / stride) + int64_t(d) * (int64_t(s[1] / stride) + int64_t(d) * int64_t(s[2] / stride));
                repeat = atomicExch(seen + site, 1u) != 0u;
                kept = !repeat;
                if (kept) {
                    x[site] = r.x;
                    y[site] = r.y;
                    z[site] = r.z;
                }
            }
        }
        const unsigned long long nk = __popcll(__ballot(kept)), nr = __popcll(__ballot(repeat)),
                                 no = __popcll(__ballot(outside));
        if ((threadIdx.x & 63) == 0) {
            if (nk) atomicAdd(counters, nk);
            if (nr) atomicAdd(counters + 1, nr);
            if (no) atomicAdd(counters + 2, no);
        }
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(unsigned* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = 0u;
}

// Header size as the reference's get_header_size (AdPointDataReaders.h:419-466): lines of at most
// 1023 characters, "int sha1_chunks = N;" gives N, the line starting "# SDF-EOH" ends the header,
// which is followed by N records of (uint32, 20 bytes).
adeco_status header_size(const char* filename, int64_t& out) {
    std::ifstream in(filename, std::ios::in | std::ios::binary);
    if (!in.is_open()) return fail(ADECO_EINVAL, std::string("ERROR: get_header_size(): failed to open file: ") + filename);
    char buf[1024];
    int chunks = -1;
    bool eoh = false;
    size_t lines = 0;
    while (in.good() && lines < 100000) {
        ++lines;
        in.getline(buf, sizeof(buf));
        const std::string line(buf);
        if (line.compare(0, 17, "int sha1_chunks =") == 0) {
            const size_t eq = line.find('='), sc = line.find(';');
            chunks = std::atoi(line.substr(eq + 1, sc - eq - 1).c_str());
        }
        if (line.compare(0, 9, "# SDF-EOH") == 0) {
            eoh = true;
            break;
        }
    }
    const long long pos = eoh ? static_cast<long long>(in.tellg()) : -1;
    if (!eoh || chunks < 0 || lines >= 100000 || pos < 0)
        return fail(ADECO_EINVAL, std::string("Failed to parse header of file: ") + filename);
    out = pos + int64_t(chunks) * int64_t(sizeof(int) + 20);
    return ADECO_OK;
}

adeco_status dark_sky_info(const char* filename, int64_t& header, int64_t& n, int& dims) {
    if (!filename || !*filename) return fail(ADECO_EINVAL, "ERROR: AdPointDataReader::load_data(): missing filename.");
    adeco_status st = header_size(filename, header);
    if (st) return st;
    if (header == 0) return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): invalid header size");
    FILE* f = std::fopen(filename, "rb");
    if (!f) return fail(ADECO_EINVAL, std::string("ERROR: AdDarkSkyDataReader::load_data_(): cannot open ") + filename);
    const int sk = std::fseek(f, 0, SEEK_END);
    const long long bytes = sk == 0 ? static_cast<long long>(std::ftell(f)) : -1;
    std::fclose(f);
    if (bytes < header) return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): inconsistent file size ");
    const uint64_t body = uint64_t(bytes - header);
    if (body % sizeof(DarkSkyRecord) != 0)
        return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): inconsistent file size ");
    n = int64_t(body / sizeof(DarkSkyRecord));
    const long long d = std::llrint(std::cbrt(double(n)));
    if (d * d * d != n) return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): lagrangian grid not cubical ");
    dims = int(d);
    return ADECO_OK;
}

}  // namespace

extern "C" {

adeco_status adeco_dark_sky_info(const char* filename, int64_t* header_bytes, int64_t* num_particles, int* dims) {
    int64_t h = 0, n = 0;
    int d = 0;
    const adeco_status st = dark_sky_info(filename, h, n, d);
    if (st) return st;
    if (header_bytes) *header_bytes = h;
    if (num_particles) *num_particles = n;
    if (dims) *dims = d;
    return ADECO_OK;
}

void adeco_selftest_dark_sky_site(long long id, int* site3) {
    unsigned s[3];
    site_of(id, s);
    for (int a = 0; a < 3; ++a) site3[a] = int(s[a]);
}

adeco_status adeco_lattice_create_dark_sky(int device_id, const char* filename, int stride, adeco_lattice* out) {
    if (!out) return fail(ADECO_EINVAL, "lattice_create_dark_sky(): null argument");
    *out = nullptr;
    if (stride < 1) return fail(ADECO_EINVAL, "ERROR: AdPointDataReader::load_data(): invalid 'stride' parameters.");
    int64_t header = 0, n = 0;
    int full = 0;
    adeco_status st = dark_sky_info(filename, header, n, full);
    if (st) return st;
    if (full % stride != 0)
        return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): invalid stride for data dimensions ");
    const int d = full / stride;
    if (d < 2) return fail(ADECO_EINVAL, "lattice_create_dark_sky(): need at least 2 particles per axis");
    if (int64_t(d) * d * d > kMaxLatticeVertices)
        return fail(ADECO_EOVERFLOW, "lattice_create_dark_sky(): lattice too large for 32-bit tet ids");
    adeco_lattice_s* L = nullptr;
    if ((st = lattice_alloc(device_id, d, d, d, L))) return st;
    struct Owner {
        adeco_lattice_s* L;
        ~Owner() { delete L; }
    } owner{L};
    DeviceState& ds = *L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    const int64_t sites = L->n_vertices;
    // two staging buffers: the file read of one chunk overlaps the copy and scatter of the last
    constexpr int64_t kChunk = int64_t(1) << 18;  // records per transfer (8 MiB)
    constexpr size_t kChunkBytes = size_t(kChunk) * sizeof(DarkSkyRecord);
    DevBuf d_seen, d_rec[2], d_cnt;
    if ((st = d_seen.reserve(size_t(sites) * 4)) || (st = d_rec[0].reserve(kChunkBytes)) ||
        (st = d_rec[1].reserve(kChunkBytes)) || (st = d_cnt.reserve(32)))
        return st;
    struct Staging {
        DarkSkyRecord* h[2] = {nullptr, nullptr};
        hipEvent_t copied[2] = {nullptr, nullptr};
        hipStream_t stream = nullptr;
        ~Staging() {
            // nothing queued may still read or write these once they are released
            if (stream) (void)hipStreamSynchronize(stream);
            for (int b = 0; b < 2; ++b) {
                if (h[b]) (void)hipHostFree(h[b]);
                if (copied[b]) (void)hipEventDestroy(copied[b]);
            }
        }
    } stg;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = hipHostMalloc(reinterpret_cast<void**>(&stg.h[b]), kChunkBytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&stg.copied[b], hipEventDisableTiming);
    }
    if (e != hipSuccess) return hip_fail(e, "pinned staging buffers");
    stg.stream = ds.stream;
    const unsigned nb_fill = unsigned(std::min<int64_t>(sites / 256 + 1, 4096));
    hipLaunchKernelGGL(k_fill_u32, dim3(nb_fill), dim3(256), 0, ds.stream, d_seen.as<unsigned>(), sites);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, ds.stream, d_cnt.as<unsigned>(), int64_t(8));
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "dark sky clear");
    FILE* f = std::fopen(filename, "rb");
    if (!f) return fail(ADECO_EINVAL, std::string("ERROR: AdDarkSkyDataReader::load_data_(): cannot open ") + filename);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    if (std::fseek(f, long(header), SEEK_SET) != 0) return fail(ADECO_EIO, "lattice_create_dark_sky(): seek failed");
    int64_t chunk = 0;
    for (int64_t done = 0; done < n; ++chunk) {
        const int b = int(chunk & 1);
        const int64_t take = std::min(kChunk, n - done);
        // host buffer b is free once the copy of two chunks ago has completed; device buffer b is
        // reused in stream order, after that chunk's scatter
        if (chunk >= 2 && (e = hipEventSynchronize(stg.copied[b])) != hipSuccess) return hip_fail(e, "dark sky staging");
        if (std::fread(stg.h[b], sizeof(DarkSkyRecord), size_t(take), f) != size_t(take))
            return fail(ADECO_EIO, std::string("ERROR: AdDarkSkyDataReader::load_data_(): Failed to read all particles "
                                               "for lagrangian input grid: ") + filename);
        e = hipMemcpyAsync(d_rec[b].ptr, stg.h[b], size_t(take) * sizeof(DarkSkyRecord), hipMemcpyHostToDevice,
                           ds.stream);
        if (e == hipSuccess) e = hipEventRecord(stg.copied[b], ds.stream);
        if (e != hipSuccess) return hip_fail(e, "dark sky upload");
        const unsigned nb = unsigned(std::min<int64_t>(take / 256 + 1, 4096));
        hipLaunchKernelGGL(k_dark_sky_scatter, dim3(nb), dim3(256), 0, ds.stream, d_rec[b].as<DarkSkyRecord>(), take,
                           full, stride, L->d_x, L->d_y, L->d_z, d_seen.as<unsigned>(), d_cnt.as<unsigned long long>());
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "dark sky scatter");
        done += take;
    }
    if ((e = hipStreamSynchronize(ds.stream)) != hipSuccess) return hip_fail(e, "dark sky scatter");
    unsigned long long cnt[3] = {0, 0, 0};
    e = hipMemcpy(cnt, d_cnt.ptr, sizeof(cnt), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "dark sky counters");
    if (cnt[2]) return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): invalid dark sky id");
    if (cnt[1] || int64_t(cnt[0]) != sites)
        return fail(ADECO_EINVAL, "ERROR: AdDarkSkyDataReader::load_data_(): inconsistent global particle number");
    owner.L = nullptr;
    *out = L;
    return ADECO_OK;
}

}  // extern "C"
