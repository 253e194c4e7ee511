// Region-of-interest density server over a resident lattice: continuous batching of requests.
//
// A request asks for the density of the lattice's tets on a uniform dims^3 grid over an
// axis-aligned box (the leaf deposit of the reference's single-rank loop, AdDynamicLoadBalancing.h
// :983-1323, with the box given by the caller instead of an octree leaf).  The server keeps the
// lattice in HBM and runs one worker thread:
//   * admission: a request is admitted when a slot is free and its buffers fit the memory budget
//     (FP64 grid, FP32 output, tet list and candidate space, sized once its tets are selected);
//     otherwise it waits until finishing requests release theirs;
//   * selection: meta-cells (7^3 lattice cubes, the reference's MetaMesh blocks) whose AABB
//     overlaps the box are culled on the host; k_roi_count / k_roi_fill list, in meta-cell order,
//     the tets whose float AABB strictly overlaps the box;
//   * steps: every step is one launch of the batched deposit kernel over (request, tile) entries of
//     all admitted requests, each request taking an equal share of the step's tiles, so a request
//     admitted while others run joins the next step and its tiles share the device with theirs.
//     Steps are queued without waiting for the previous one: the step entries are double-buffered
//     (a buffer is refilled once its upload of two steps ago has run), and admission runs on a
//     stream of its own, so selecting and setting up new requests overlaps the running step;
//   * completion: after a request's last tile its FP64 masses are turned into FP32 density
//     (mass / cell volume, as AdOctreeGridWriter::write_patch) and an event is recorded on the
//     step stream; a copier thread waits for that event, moves the density to the caller's buffer
//     and frees the slot.
// adeco_roi_deposit_plain is the serial baseline: the same selection, then the plain
// one-tet-per-thread kernel, one request at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
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

static_assert(sizeof(adeco_serve_config) == 40 && sizeof(adeco_request_info) == 48 && sizeof(adeco_serve_stats) == 72,
              "C-ABI struct layouts are fixed");

constexpr int kMeta = 7;  // lattice cubes per meta-cell edge (AdTypeDefinitions.h:732-734)
constexpr int kMaxSlots = 64;
constexpr int64_t kMaxStepTiles = int64_t(1) << 22;  // two pinned entry buffers of 8 bytes per tile

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct RoiBox {
    float b[6];
};

// tets of meta-cell `m` (cubes [7 mi, min(7 mi + 7, nc)) per axis, 6 tets per cube, tet id =
// 6 * cube + local tet) whose float AABB strictly overlaps the box (AdTypeDefinitions.h:339-367)
__device__ __forceinline__ bool tet_hits(const LatticeView& lat, const PackedCorners& ct, int64_t t, const RoiBox& r) {
    const int64_t c = t / 6;
    const int lt = int(t - 6 * c);
    const int64_t ncx = lat.np[0] - 1, ncy = lat.np[1] - 1;
    const int64_t cj = c / ncx;
    const int64_t i = c - cj * ncx;
    const int64_t k = cj / ncy;
    const int64_t j = cj - k * ncy;
    float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int q = 0; q < 4; ++q) {
        const int code = ct.code(lt, q);
        const int64_t vi = (i + (code & 1)) +
                           int64_t(lat.np[0]) * ((j + ((code >> 1) & 1)) + int64_t(lat.np[1]) * (k + ((code >> 2) & 1)));
        const float p[3] = {lat.x[vi], lat.y[vi], lat.z[vi]};
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], p[a]);
            mx[a] = fmaxf(mx[a], p[a]);
        }
    }
    return mn[0] < r.b[1] && r.b[0] < mx[0] && mn[1] < r.b[3] && r.b[2] < mx[1] && mn[2] < r.b[5] && r.b[4] < mx[2];
}

__device__ __forceinline__ void meta_range(const LatticeView& lat, int m, const int md[3], int lo[3], int n[3]) {
    const int mi = m % md[0], mj = (m / md[0]) % md[1], mk = m / (md[0] * md[1]);
    const int mm[3] = {mi, mj, mk};
    for (int a = 0; a < 3; ++a) {
        lo[a] = kMeta * mm[a];
        n[a] = min(kMeta, lat.np[a] - 1 - lo[a]);
    }
}

struct MetaDims {
    int d[3];
};

// one workgroup per listed meta-cell: count of hitting tets
__global__ void __launch_bounds__(256) k_roi_count(LatticeView lat, PackedCorners ct, MetaDims md, const int* __restrict__ metas,
                                                   RoiBox roi, unsigned* __restrict__ counts) {
    __shared__ unsigned s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    int lo[3], n[3];
    meta_range(lat, metas[blockIdx.x], md.d, lo, n);
    const int total = 6 * n[0] * n[1] * n[2];
    unsigned mine = 0;
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
        const int cube = q / 6, lt = q - 6 * (q / 6);
        const int ci = cube % n[0], cj = (cube / n[0]) % n[1], ck = cube / (n[0] * n[1]);
        const int64_t c = (lo[0] + ci) + int64_t(lat.np[0] - 1) * ((lo[1] + cj) + int64_t(lat.np[1] - 1) * (lo[2] + ck));
        mine += tet_hits(lat, ct, 6 * c + lt, roi) ? 1u : 0u;
    }
    atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_cnt;
}

// one wavefront per listed meta-cell: the hitting tet ids in local order at offs[meta]
__global__ void __launch_bounds__(64) k_roi_fill(LatticeView lat, PackedCorners ct, MetaDims md, const int* __restrict__ metas,
                                                 RoiBox roi, const int64_t* __restrict__ offs, int32_t* __restrict__ out) {
    int lo[3], n[3];
    meta_range(lat, metas[blockIdx.x], md.d, lo, n);
    const int total = 6 * n[0] * n[1] * n[2];
    const int lane = threadIdx.x;
    int64_t base = offs[blockIdx.x];
    for (int q0 = 0; q0 < total; q0 += 64) {
        const int q = q0 + lane;
        bool hit = false;
        int64_t t = 0;
        if (q < total) {
            const int cube = q / 6, lt = q - 6 * (q / 6);
            const int ci = cube % n[0], cj = (cube / n[0]) % n[1], ck = cube / (n[0] * n[1]);
            const int64_t c = (lo[0] + ci) + int64_t(lat.np[0] - 1) * ((lo[1] + cj) + int64_t(lat.np[1] - 1) * (lo[2] + ck));
            t = 6 * c + lt;
            hit = tet_hits(lat, ct, t, roi);
        }
        const unsigned long long m = __ballot(hit);
        if (hit) {
            const unsigned below = __builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u));
            out[base + below] = int32_t(t);
        }
        base += __popcll(m);
    }
}

// FP64 mass -> FP32 density (AdOctreeGridWriter.h:67-144 writes mass * (1 / cell volume))
__global__ void __launch_bounds__(256) k_density_f32(const double* __restrict__ mass, int64_t n, double inv_cell_vol,
                                                     float* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = float(mass[i] * inv_cell_vol);
}

// Request-side grid geometry: the FP64 box of the request with per-axis (or reference) scaling.
GridGeom roi_geom(const double roi[6], int dims) { return make_grid_geom_d(roi, dims, cell_scaling()); }

// Selection of a request's tets into `pairs` (grown as needed); returns the count.
struct Selector {
    const std::vector<float>* meta_box = nullptr;  // [M][6], cached on the lattice handle
    MetaDims md;
    DevBuf d_metas, d_counts, d_offs;
    std::vector<int> metas;
    std::vector<unsigned> counts;
    std::vector<int64_t> offs;

    // Caches the lattice's meta-cell boxes (once per lattice; takes the device mutex).
    adeco_status init(adeco_lattice_s* L) {
        DeviceState& ds = *L->ds;
        std::lock_guard<std::mutex> lk(ds.mutex);
        if (L->meta_box.empty()) {
            for (int a = 0; a < 3; ++a) L->meta_dims[a] = (L->np[a] - 1 + kMeta - 1) / kMeta;
            const int M = L->meta_dims[0] * L->meta_dims[1] * L->meta_dims[2];
            hipError_t e = hipSetDevice(L->device);
            if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
            DevBuf box, cells;
            adeco_status st;
            if ((st = box.reserve(size_t(M) * 6 * sizeof(float))) || (st = cells.reserve(size_t(M) * sizeof(unsigned))))
                return st;
            if ((st = octree_meta_bbox(L, L->meta_dims, box.as<float>(), cells.as<unsigned>()))) return st;
            std::vector<float> hb(size_t(M) * 6);
            e = hipMemcpyAsync(hb.data(), box.ptr, hb.size() * sizeof(float), hipMemcpyDeviceToHost, ds.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
            if (e != hipSuccess) return hip_fail(e, "meta-cell boxes");
            L->meta_box.swap(hb);
        }
        for (int a = 0; a < 3; ++a) md.d[a] = L->meta_dims[a];
        meta_box = &L->meta_box;
        return ADECO_OK;
    }

    // Lists the meta-cells whose box strictly overlaps the request box.
    void cull(const float r[6]) {
        metas.clear();
        const int M = int(meta_box->size() / 6);
        for (int m = 0; m < M; ++m) {
            const float* b = &(*meta_box)[size_t(m) * 6];
            if (b[0] < r[1] && r[0] < b[1] && b[2] < r[3] && r[2] < b[3] && b[4] < r[5] && r[4] < b[5]) metas.push_back(m);
        }
    }

    // Lists the box's tets into `pairs`.  may_grow(bytes) is asked before `pairs` grows to hold
    // them; a refusal returns ADECO_ENOMEM with n_out set and nothing allocated.
    template <class MayGrow>
    adeco_status select(hipStream_t stream, const LatticeView& lv, const PackedCorners& ct, const float r[6], DevBuf& pairs,
                        int64_t& n_out, MayGrow may_grow) {
        cull(r);
        n_out = 0;
        const int nm = int(metas.size());
        if (nm == 0) return ADECO_OK;
        adeco_status st;
        if ((st = d_metas.reserve(size_t(nm) * 4)) || (st = d_counts.reserve(size_t(nm) * 4)) ||
            (st = d_offs.reserve(size_t(nm) * 8)))
            return st;
        RoiBox rb;
        std::memcpy(rb.b, r, sizeof(rb.b));
        hipError_t e = hipMemcpyAsync(d_metas.ptr, metas.data(), size_t(nm) * 4, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_roi_count, dim3(nm), dim3(256), 0, stream, lv, ct, md, d_metas.as<int>(), rb, d_counts.as<unsigned>());
            e = hipGetLastError();
        }
        counts.resize(size_t(nm));
        if (e == hipSuccess) e = hipMemcpyAsync(counts.data(), d_counts.ptr, size_t(nm) * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "ROI tet count");
        offs.resize(size_t(nm));
        int64_t total = 0;
        for (int m = 0; m < nm; ++m) {
            offs[size_t(m)] = total;
            total += counts[size_t(m)];
        }
        n_out = total;
        if (total == 0) return ADECO_OK;
        if (total >= (int64_t(1) << 31) - 2) return fail(ADECO_EOVERFLOW, "ROI request: too many tets");
        if (!may_grow(size_t(total) * 4)) return ADECO_ENOMEM;
        if ((st = pairs.reserve(size_t(total) * 4))) return st;
        e = hipMemcpyAsync(d_offs.ptr, offs.data(), size_t(nm) * 8, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_roi_fill, dim3(nm), dim3(64), 0, stream, lv, ct, md, d_metas.as<int>(), rb, d_offs.as<int64_t>(),
                               pairs.as<int32_t>());
            e = hipGetLastError();
        }
        // the host vectors above are read by the copies: finish them before they change
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        return e == hipSuccess ? ADECO_OK : hip_fail(e, "ROI tet fill");
    }
};

void float_box(const double roi[6], float r[6]) {
    for (int i = 0; i < 6; ++i) r[i] = float(roi[i]);
}

bool valid_roi(const double* r) {
    for (int a = 0; a < 3; ++a)
        if (!(r[2 * a] < r[2 * a + 1]) || !std::isfinite(r[2 * a]) || !std::isfinite(r[2 * a + 1])) return false;
    return true;
}

}  // namespace

}  // namespace adeco

using namespace adeco;

// ------------------------------------------------------------------------------------------------
struct adeco_server_s {
    struct Slot {
        bool busy = false;
        DevBuf grid, out32, pairs, work, tiles, desc;  // desc: GridDesc + the pair CSR offsets {0, n}
        size_t capacity() const { return grid.cap + out32.cap + pairs.cap + work.cap + tiles.cap + desc.cap; }
        void release() {
            grid.free_all();
            out32.free_all();
            pairs.free_all();
            work.free_all();
            tiles.free_all();
            desc.free_all();
        }
    };
    struct Req {
        uint64_t id = 0;
        double roi[6];
        int dims = 0;
        float* host_out = nullptr;
        int slot = -1;
        int64_t n_tets = 0, n_cand = 0, n_tiles = 0, next_tile = 0;
        double t_submit = 0, t_admit = -1, t_done = -1;
        adeco_status status = ADECO_OK;
        bool done = false;
        GridDesc gd;         // host sources of the setup copies (alive until the request is freed)
        int64_t goff[2] = {0, 0};
    };

    adeco_lattice_s* L = nullptr;
    adeco_serve_config cfg;
    PackedCorners ct;
    Selector sel;
    // steps run on `stream`, request setup on `admit_stream`, downloads on `copy_stream`
    hipStream_t stream = nullptr, admit_stream = nullptr, copy_stream = nullptr;
    std::unique_ptr<Slot[]> slot_store;
    int n_slots = 0;
    Slot* slots = nullptr;
    TileSource h_sources[kMaxSlots];
    hipEvent_t slot_done[kMaxSlots] = {};  // recorded after a request's density kernel
    DevBuf d_sources, d_ctr;
    // double-buffered step entries: pinned host side, device side, and an event recorded after
    // each buffer's upload
    int2* h_entries[2] = {nullptr, nullptr};
    DevBuf d_entries[2];
    hipEvent_t entries_uploaded[2] = {nullptr, nullptr};
    bool entries_in_flight[2] = {false, false};
    std::thread worker, copier;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_copy;
    std::deque<Req*> pending, to_copy;
    std::vector<Req*> active;
    std::map<uint64_t, Req*> reqs;
    uint64_t next_id = 1;
    uint64_t n_released = 0;  // slots freed so far (a request waiting for memory waits for this to move)
    bool stop = false;
    std::string worker_error;
    // statistics
    uint64_t n_steps = 0, n_done = 0;
    int64_t tiles_done = 0, tets_done = 0;
    double busy_s = 0;
    // device time of the step launches: an event pair recorded around each one, added up once the
    // pair has run (under mu); spent pairs are reused
    std::deque<std::pair<hipEvent_t, hipEvent_t>> step_events;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spare_events;
    double step_kernel_s = 0;

    void collect_step_times(bool wait) {
        while (!step_events.empty()) {
            auto ev = step_events.front();
            if (wait) (void)hipEventSynchronize(ev.second);
            else if (hipEventQuery(ev.second) != hipSuccess) break;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) step_kernel_s += 1e-3 * double(ms);
            step_events.pop_front();
            spare_events.push_back(ev);
        }
    }

    size_t capacity_in_use() const {
        size_t b = 0;
        for (int i = 0; i < n_slots; ++i) b += slots[i].capacity();
        return b;
    }

    // bytes a request's buffers need in slot s beyond what s already holds
    size_t growth(const Slot& s, size_t grid_b, size_t out_b, size_t pairs_b, size_t work_b) const {
        auto g = [](size_t cap, size_t need) { return need > cap ? need : 0; };
        return g(s.grid.cap, grid_b) + g(s.out32.cap, out_b) + g(s.pairs.cap, pairs_b) + g(s.work.cap, work_b);
    }

    adeco_status admit(Req* r, int si);
    void run_worker();
    void run_copier();
    // streams, events, pinned and device buffers of the server itself (not the slots')
    void release_device_objects() {
        for (int b = 0; b < 2; ++b) {
            if (h_entries[b]) (void)hipHostFree(h_entries[b]);
            if (entries_uploaded[b]) (void)hipEventDestroy(entries_uploaded[b]);
            h_entries[b] = nullptr;
            entries_uploaded[b] = nullptr;
            d_entries[b].free_all();
        }
        for (int i = 0; i < kMaxSlots; ++i)
            if (slot_done[i]) {
                (void)hipEventDestroy(slot_done[i]);
                slot_done[i] = nullptr;
            }
        for (auto& ev : step_events) spare_events.push_back(ev);
        step_events.clear();
        for (auto& ev : spare_events) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
        spare_events.clear();
        d_sources.free_all();
        d_ctr.free_all();
        if (stream) (void)hipStreamDestroy(stream);
        if (admit_stream) (void)hipStreamDestroy(admit_stream);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        stream = admit_stream = copy_stream = nullptr;
    }
};

adeco_status adeco_server_s::admit(Req* r, int si) {
    Slot& s = slots[size_t(si)];
    const LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    float rf[6];
    float_box(r->roi, rf);
    // Every buffer of the request is held to the budget before it grows, except the tile map
    // (4 bytes per 64 candidates), whose size is known only after the pair setup.
    const int64_t cells = int64_t(r->dims) * r->dims * r->dims;
    const size_t grid_b = size_t(cells) * 8, out_b = size_t(cells) * 4, min_b = size_t(1) << 20;
    auto grow = [](const DevBuf& b, size_t need, size_t floor) { return need > b.cap ? std::max(need, floor) : size_t(0); };
    size_t fixed = grow(s.grid, grid_b, min_b) + grow(s.out32, out_b, min_b) + grow(s.desc, 512, min_b) + grow(s.tiles, 16, min_b);
    auto fits = [&](size_t more) { return !cfg.memory_budget || capacity_in_use() + fixed + more <= cfg.memory_budget; };
    adeco_status st = sel.select(admit_stream, lv, ct, rf, s.pairs, r->n_tets,
                                 [&](size_t pairs_b) { return fits(grow(s.pairs, pairs_b, min_b)); });
    if (st) return st;
    const size_t work_b = candidate_space_bytes(r->n_tets) + 256;
    if (!fits(grow(s.work, work_b, min_b))) return ADECO_ENOMEM;  // wait for frees
    if ((st = s.grid.reserve(grid_b)) || (st = s.out32.reserve(out_b)) || (st = s.work.reserve(work_b)) ||
        (st = s.desc.reserve(512)))
        return st;
    r->gd = make_grid_desc(roi_geom(r->roi, r->dims), 0);
    r->goff[0] = 0;
    r->goff[1] = r->n_tets;
    GridDesc* d_gd = s.desc.as<GridDesc>();
    int64_t* d_goff = reinterpret_cast<int64_t*>(s.desc.as<char>() + 256);
    hipError_t e = hipMemcpyAsync(d_gd, &r->gd, sizeof(GridDesc), hipMemcpyHostToDevice, admit_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_goff, r->goff, sizeof(r->goff), hipMemcpyHostToDevice, admit_stream);
    if (e == hipSuccess) e = launch_fill_f64(s.grid.as<double>(), cells, 0.0, admit_stream);
    if (e != hipSuccess) return hip_fail(e, "request setup");
    r->n_cand = 0;
    r->n_tiles = 0;
    if (r->n_tets > 0) {
        CandidateSpace cs;
        if ((st = build_candidate_space(admit_stream, lv, ct, s.pairs.as<int32_t>(), r->n_tets, d_goff, 1, d_gd, cfg.tet_mass,
                                        cfg.precision, s.work.ptr, cs)))
            return st;
        if ((st = s.tiles.reserve(tile_map_bytes(cs.n_cand)))) return st;
        if ((st = build_tile_map(admit_stream, cs, s.tiles.as<int32_t>()))) return st;
        r->n_cand = cs.n_cand;
        r->n_tiles = cs.n_tiles;
        h_sources[si] = TileSource{cs.pd, cs.cand_off, cs.tile_first, cs.n_tiles, cs.n_cand, d_gd, s.grid.as<double>(),
                                   cfg.tet_mass};
        e = hipMemcpyAsync(d_sources.as<TileSource>() + si, &h_sources[si], sizeof(TileSource), hipMemcpyHostToDevice,
                           admit_stream);
        if (e != hipSuccess) return hip_fail(e, "request setup");
    }
    r->next_tile = 0;
    // the request joins the next step launch on the step stream: its setup must have run
    e = hipStreamSynchronize(admit_stream);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "request setup");
}

void adeco_server_s::run_worker() {
    (void)hipSetDevice(L->device);
    std::unique_lock<std::mutex> lk(mu);
    int buf = 0;  // step-entries buffer of the next step
    for (;;) {
        cv_work.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
        if (stop && pending.empty() && active.empty()) break;
        // admission: in arrival order, while a slot is free and the memory budget allows.  A request
        // that does not fit waits while another slot holds a request (running, or being copied out:
        // its slot keeps its memory until the copier frees it); idle slots' buffers are released
        // before it waits or gives up, and it fails only when nothing else holds memory.
        uint64_t seen_release = n_released;
        while (!pending.empty()) {
            int si = -1;
            for (int i = 0; i < n_slots; ++i)
                if (!slots[i].busy) {
                    si = i;
                    break;
                }
            if (si < 0) {
                seen_release = n_released;  // every slot is taken: wait for one to be freed
                break;
            }
            Req* r = pending.front();
            slots[size_t(si)].busy = true;
            lk.unlock();
            adeco_status st = admit(r, si);
            bool must_wait = false;
            while (st == ADECO_ENOMEM && cfg.memory_budget) {
                lk.lock();
                bool others_busy = false, freed = false;
                for (int i = 0; i < n_slots; ++i) {
                    if (i == si) continue;
                    if (slots[i].busy) {
                        others_busy = true;
                    } else if (slots[i].capacity()) {
                        slots[i].release();
                        freed = true;
                    }
                }
                if (!freed && others_busy) {
                    must_wait = true;
                    seen_release = n_released;  // read under the lock that decided to wait
                }
                lk.unlock();
                if (!freed) break;
                st = admit(r, si);
            }
            lk.lock();
            if (must_wait) {
                slots[size_t(si)].busy = false;
                break;
            }
            pending.pop_front();
            r->slot = si;
            r->t_admit = now_s();
            if (st != ADECO_OK) {
                r->status = st == ADECO_ENOMEM ? ADECO_EOVERFLOW : st;
                r->done = true;
                r->t_done = now_s();
                slots[size_t(si)].busy = false;
                ++n_released;
                ++n_done;
                cv_done.notify_all();
                continue;
            }
            active.push_back(r);
        }
        if (active.empty()) {
            // nothing to step: the head request waits until a slot is freed (the slots it waits
            // for are being copied out; a copy that ended since seen_release was read counts)
            if (!pending.empty()) cv_work.wait(lk, [&] { return n_released != seen_release; });
            continue;
        }
        // one step: an equal share of the step's tiles to every admitted request
        std::vector<Req*> run = active;
        lk.unlock();
        const double t0 = now_s();
        hipError_t e = hipSuccess;
        // this buffer's previous upload (two steps ago) must have run before it is refilled
        if (entries_in_flight[buf]) {
            e = hipEventSynchronize(entries_uploaded[buf]);
            entries_in_flight[buf] = false;
        }
        int2* he = h_entries[buf];
        const int64_t cap = cfg.step_tiles;
        int n = 0;
        int64_t left = cap;
        size_t open = 0;
        for (Req* r : run) open += r->next_tile < r->n_tiles ? 1 : 0;
        while (left > 0 && open > 0) {
            const int64_t quota = std::max<int64_t>(1, left / int64_t(open));
            open = 0;
            for (Req* r : run) {
                const int64_t rem = r->n_tiles - r->next_tile;
                if (rem <= 0 || left <= 0) continue;
                const int64_t take = std::min(std::min(rem, quota), left);
                for (int64_t t = 0; t < take; ++t) he[n++] = make_int2(r->slot, int(r->next_tile + t));
                r->next_tile += take;
                left -= take;
                if (r->next_tile < r->n_tiles) ++open;
            }
        }
        unsigned* ctr = d_ctr.as<unsigned>();
        unsigned long long* counters = reinterpret_cast<unsigned long long*>(d_ctr.as<char>() + 16);
        if (n > 0 && e == hipSuccess) {
            e = hipMemcpyAsync(d_entries[buf].ptr, he, size_t(n) * sizeof(int2), hipMemcpyHostToDevice, stream);
            if (e == hipSuccess) e = hipEventRecord(entries_uploaded[buf], stream);
            if (e == hipSuccess) entries_in_flight[buf] = true;
            if (e == hipSuccess) e = launch_zero_u32(ctr, 16, stream);
            std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
            if (e == hipSuccess) {
                lk.lock();
                if (!spare_events.empty()) {
                    ev = spare_events.back();
                    spare_events.pop_back();
                }
                lk.unlock();
                if (!ev.first) {
                    e = hipEventCreate(&ev.first);
                    if (e == hipSuccess) e = hipEventCreate(&ev.second);
                }
            }
            if (e == hipSuccess) e = hipEventRecord(ev.first, stream);
            if (e == hipSuccess)
                e = launch_deposit_step(stream, cfg.precision, d_sources.as<TileSource>(), d_entries[buf].as<int2>(), n, ctr,
                                        counters, L->device);
            if (e == hipSuccess) e = hipEventRecord(ev.second, stream);
            lk.lock();
            if (e == hipSuccess) step_events.push_back(ev);
            else if (ev.first && ev.second) spare_events.push_back(ev);
            else {
                if (ev.first) (void)hipEventDestroy(ev.first);
                if (ev.second) (void)hipEventDestroy(ev.second);
            }
            collect_step_times(false);
            lk.unlock();
        }
        buf ^= 1;
        // requests whose last tile is in this step: density in FP32, then an event for the copier
        std::vector<Req*> finished;
        for (Req* r : run)
            if (r->next_tile >= r->n_tiles) {
                Slot& s = slots[size_t(r->slot)];
                const int64_t cells = int64_t(r->dims) * r->dims * r->dims;
                const GridGeom g = roi_geom(r->roi, r->dims);
                if (e == hipSuccess) {
                    hipLaunchKernelGGL(k_density_f32, dim3(unsigned(std::min<int64_t>((cells + 255) / 256, 4096))), dim3(256), 0,
                                       stream, s.grid.as<double>(), cells, 1.0 / g.cell_vol, s.out32.as<float>());
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipEventRecord(slot_done[r->slot], stream);
                finished.push_back(r);
            }
        const double t1 = now_s();
        lk.lock();
        busy_s += t1 - t0;
        ++n_steps;
        tiles_done += n;
        if (e != hipSuccess) {
            worker_error = hipGetErrorString(e);
            for (Req* r : run) {
                r->status = ADECO_EDEVICE;
                r->done = true;
                r->t_done = now_s();
                slots[size_t(r->slot)].busy = false;
                ++n_released;
                ++n_done;
            }
            active.clear();
            cv_done.notify_all();
            continue;
        }
        for (Req* r : finished) {
            active.erase(std::find(active.begin(), active.end(), r));
            to_copy.push_back(r);
        }
        if (!finished.empty()) cv_copy.notify_one();
    }
    lk.unlock();
    (void)hipStreamSynchronize(stream);
    lk.lock();
    cv_copy.notify_all();
}

void adeco_server_s::run_copier() {
    (void)hipSetDevice(L->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_copy.wait(lk, [&] { return !to_copy.empty() || (stop && pending.empty() && active.empty()); });
        if (to_copy.empty()) break;
        Req* r = to_copy.front();
        to_copy.pop_front();
        Slot& s = slots[size_t(r->slot)];
        const size_t bytes = size_t(r->dims) * r->dims * r->dims * sizeof(float);
        lk.unlock();
        // the request's last step and density kernel, then the download
        hipError_t e = hipEventSynchronize(slot_done[r->slot]);
        if (e == hipSuccess) e = hipMemcpyAsync(r->host_out, s.out32.ptr, bytes, hipMemcpyDeviceToHost, copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(copy_stream);
        lk.lock();
        if (e != hipSuccess) r->status = hip_fail(e, "request download");
        r->t_done = now_s();
        r->done = true;
        s.busy = false;
        ++n_released;
        ++n_done;
        tets_done += r->n_tets;
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

adeco_status adeco_serve_create(adeco_lattice L, const adeco_serve_config* cfg, adeco_server* out) {
    if (!L || !cfg || !out) return fail(ADECO_EINVAL, "serve_create(): null argument");
    *out = nullptr;
    if (cfg->decomposition < 0 || cfg->decomposition > 2) return fail(ADECO_EINVAL, "serve_create(): unknown decomposition");
    if (cfg->precision < 0 || cfg->precision > 2) return fail(ADECO_EINVAL, "serve_create(): unknown precision");
    if (cfg->max_inflight < 1 || cfg->max_inflight > kMaxSlots)
        return fail(ADECO_EINVAL, "serve_create(): max_inflight must be in 1..64");
    if (cfg->step_tiles < 0 || cfg->step_tiles > kMaxStepTiles)
        return fail(ADECO_EINVAL, "serve_create(): step_tiles must be in 0..4194304");
    if (!(cfg->tet_mass > 0.0)) return fail(ADECO_EINVAL, "serve_create(): tet_mass must be positive");
    auto* S = new adeco_server_s();
    S->L = L;
    S->cfg = *cfg;
    if (S->cfg.step_tiles == 0) S->cfg.step_tiles = int64_t(1) << 17;
    S->ct = pack_corners(corner_table(cfg->decomposition));
    adeco_status st = S->sel.init(L);
    hipError_t e = hipSuccess;
    if (!st) {
        e = hipSetDevice(L->device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->admit_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking);
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            e = hipHostMalloc(reinterpret_cast<void**>(&S->h_entries[b]), size_t(S->cfg.step_tiles) * sizeof(int2));
            if (e == hipSuccess) e = hipEventCreateWithFlags(&S->entries_uploaded[b], hipEventDisableTiming);
        }
        for (int i = 0; i < cfg->max_inflight && e == hipSuccess; ++i)
            e = hipEventCreateWithFlags(&S->slot_done[i], hipEventDisableTiming);
        if (e != hipSuccess) st = hip_fail(e, "serve_create()");
    }
    if (!st) st = S->d_sources.reserve(sizeof(TileSource) * kMaxSlots);
    for (int b = 0; b < 2 && !st; ++b) st = S->d_entries[b].reserve(size_t(S->cfg.step_tiles) * sizeof(int2));
    if (!st) st = S->d_ctr.reserve(64);
    if (st) {
        S->release_device_objects();
        delete S;
        return st;
    }
    std::memset(S->h_sources, 0, sizeof(S->h_sources));
    S->slot_store.reset(new adeco_server_s::Slot[size_t(cfg->max_inflight)]);
    S->slots = S->slot_store.get();
    S->n_slots = cfg->max_inflight;
    S->worker = std::thread([S] { S->run_worker(); });
    S->copier = std::thread([S] { S->run_copier(); });
    *out = S;
    return ADECO_OK;
}

adeco_status adeco_serve_submit(adeco_server S, const double* roi6, int dims, float* density_out, uint64_t* request_id) {
    if (!S || !roi6 || !density_out || !request_id) return fail(ADECO_EINVAL, "serve_submit(): null argument");
    if (dims < 1 || dims > 1023) return fail(ADECO_EINVAL, "serve_submit(): dims must be in 1..1023");
    if (!valid_roi(roi6)) return fail(ADECO_EINVAL, "serve_submit(): invalid box (need min < max per axis)");
    auto* r = new adeco_server_s::Req();
    std::memcpy(r->roi, roi6, sizeof(r->roi));
    r->dims = dims;
    r->host_out = density_out;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete r;
            return fail(ADECO_EINVAL, "serve_submit(): server is shutting down");
        }
        r->id = S->next_id++;
        S->reqs[r->id] = r;
        S->pending.push_back(r);
    }
    S->cv_work.notify_all();
    *request_id = r->id;
    return ADECO_OK;
}

adeco_status adeco_serve_wait(adeco_server S, uint64_t id, double timeout_s, adeco_request_info* info) {
    if (!S) return fail(ADECO_EINVAL, "serve_wait(): null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(ADECO_EINVAL, "serve_wait(): unknown request id");
    adeco_server_s::Req* r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_s < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return fail(ADECO_EOVERFLOW, "serve_wait(): timed out");
    if (info) {
        info->status = r->status;
        info->dims = r->dims;
        info->n_tets = r->n_tets;
        info->n_candidates = r->n_cand;
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
    }
    const adeco_status st = r->status;
    S->reqs.erase(it);
    delete r;
    return st ? fail(st, "serve_wait(): request failed: " + S->worker_error) : ADECO_OK;
}

adeco_status adeco_serve_stats_get(adeco_server S, adeco_serve_stats* out) {
    if (!S || !out) return fail(ADECO_EINVAL, "serve_stats(): null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    out->requests_done = S->n_done;
    out->steps = S->n_steps;
    out->tiles = S->tiles_done;
    out->tets = S->tets_done;
    out->busy_s = S->busy_s;
    out->device_bytes = S->capacity_in_use();
    out->pending = int64_t(S->pending.size());
    out->active = int64_t(S->active.size());
    S->collect_step_times(S->pending.empty() && S->active.empty() && S->to_copy.empty());
    out->step_kernel_s = S->step_kernel_s;
    return ADECO_OK;
}

adeco_status adeco_serve_destroy(adeco_server S) {
    if (!S) return ADECO_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_copy.notify_all();
    if (S->worker.joinable()) S->worker.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
    }
    S->cv_copy.notify_all();
    if (S->copier.joinable()) S->copier.join();
    (void)hipSetDevice(S->L->device);
    for (auto& kv : S->reqs) delete kv.second;
    S->slot_store.reset();
    S->slots = nullptr;
    S->release_device_objects();
    delete S;
    return ADECO_OK;
}

adeco_status adeco_roi_deposit_plain(adeco_lattice L, int decomposition, double tet_mass, const double* roi6, int dims,
                                     float* density_out, int64_t* n_tets_out) {
    if (!L || !roi6 || !density_out) return fail(ADECO_EINVAL, "roi_deposit_plain(): null argument");
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "roi_deposit_plain(): unknown decomposition");
    if (dims < 1 || dims > 1023) return fail(ADECO_EINVAL, "roi_deposit_plain(): dims must be in 1..1023");
    if (!valid_roi(roi6)) return fail(ADECO_EINVAL, "roi_deposit_plain(): invalid box");
    Selector sel;
    adeco_status st = sel.init(L);
    if (st) return st;
    DeviceState& ds = *L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const LatticeView lv{L->d_x, L->d_y, L->d_z, {L->np[0], L->np[1], L->np[2]}};
    const PackedCorners ct = pack_corners(corner_table(decomposition));
    float rf[6];
    float_box(roi6, rf);
    int64_t n = 0;
    if ((st = sel.select(ds.stream, lv, ct, rf, ds.pos, n, [](size_t) { return true; }))) return st;
    const int64_t cells = int64_t(dims) * dims * dims;
    if ((st = ds.grid.reserve(size_t(cells) * 8)) || (st = ds.scratch.reserve(size_t(cells) * 4))) return st;
    const GridGeom g = roi_geom(roi6, dims);
    e = launch_fill_f64(ds.grid.as<double>(), cells, 0.0, ds.stream);
    if (e != hipSuccess) return hip_fail(e, "roi_deposit_plain()");
    if ((st = timed_begin(ds))) return st;
    if (n > 0) {
        e = launch_deposit_lattice_plain(lv, decomposition, ds.pos.as<int32_t>(), n, g, tet_mass, ds.grid.as<double>(), nullptr,
                                         ds.stream);
        if (e != hipSuccess) return hip_fail(e, "roi_deposit_plain() launch");
    }
    if ((st = timed_end(ds))) return st;
    hipLaunchKernelGGL(k_density_f32, dim3(unsigned(std::min<int64_t>((cells + 255) / 256, 4096))), dim3(256), 0, ds.stream,
                       ds.grid.as<double>(), cells, 1.0 / g.cell_vol, ds.scratch.as<float>());
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(density_out, ds.scratch.ptr, size_t(cells) * 4, hipMemcpyDeviceToHost, ds.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds.stream);
    if (e != hipSuccess) return hip_fail(e, "roi_deposit_plain() download");
    if (n_tets_out) *n_tets_out = n;
    return ADECO_OK;
}

}  // extern "C"
