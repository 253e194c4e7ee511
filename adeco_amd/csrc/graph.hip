// hipGraph form of the batched leaf deposit: the grid clear, the work-queue reset and the
// persistent deposit launch captured once and replayed.  The candidate space (pair data, scan,
// tile map) is built when the graph is created, from the lattice positions and leaf binning of
// that moment; each replay deposits them again with one graph launch and no host round trip.
// The graph owns every buffer its kernels touch (leaf geometry, pair data, tile map, leaf grids,
// counters), so deposits through the octree handle cannot move them; the lattice must outlive it.
#include <hip/hip_runtime.h>

#include "../../include/adeco.h"
#include "batched_impl.hpp"
#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

using namespace adeco;

struct adeco_deposit_graph_s {
    adeco_octree_s* T = nullptr;
    int device = 0, dims = 0, precision = 0;
    int64_t n_leaves = 0;
    DevBuf geom, work, tiles, grids, counters;
    TileSource src;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~adeco_deposit_graph_s() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

adeco_status adeco_deposit_graph_create(adeco_octree T, int dims, int decomposition, double tet_mass, int precision,
                                        adeco_deposit_graph* out) {
    if (!T || !out) return fail(ADECO_EINVAL, "deposit_graph_create(): null argument");
    *out = nullptr;
    if (dims < 1 || dims > 1023) return fail(ADECO_EINVAL, "deposit_graph_create(): dims must be in 1..1023");
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "deposit_graph_create(): unknown decomposition");
    if (precision < 0 || precision > 2) return fail(ADECO_EINVAL, "deposit_graph_create(): unknown precision");
    DeviceState& ds = *T->L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    adeco_status st;
    if (T->binned_decomposition != decomposition && (st = octree_bin_tets(T, decomposition))) return st;
    auto* G = new adeco_deposit_graph_s();
    G->T = T;
    G->device = T->L->device;
    G->dims = dims;
    G->precision = precision;
    G->n_leaves = int64_t(T->leaves_dfs.size());
    const int64_t words = G->n_leaves * int64_t(dims) * dims * dims;
    auto bail = [&](adeco_status s) {
        delete G;
        return s;
    };
    if (size_t(words) * 8 > memory_budget(ds)) return bail(fail(ADECO_EOVERFLOW, "deposit_graph_create(): leaf grids exceed the budget"));
    if ((st = G->grids.reserve(size_t(words) * 8)) || (st = G->counters.reserve(64))) return bail(st);
    e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&G->e0);
    if (e == hipSuccess) e = hipEventCreate(&G->e1);
    if (e != hipSuccess) return bail(hip_fail(e, "deposit_graph_create()"));
    if ((st = prepare_leaves_batched(T, G->stream, dims, decomposition, tet_mass, precision, G->geom, G->work, G->tiles,
                                     G->grids.as<double>(), G->src)))
        return bail(st);
    unsigned* ctr = G->counters.as<unsigned>();
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(G->counters.as<char>() + 16);
    e = hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = launch_fill_f64(G->grids.as<double>(), words, 0.0, G->stream);
    if (e == hipSuccess) e = launch_zero_u32(ctr, 16, G->stream);
    if (e == hipSuccess) e = launch_deposit_tiles(G->stream, precision, G->src, ctr, counters, G->device);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(G->stream, &graph);
    G->graph = graph;
    if (e == hipSuccess) e = ec;
    if (e == hipSuccess) e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(hip_fail(e, "deposit graph capture"));
    *out = G;
    return ADECO_OK;
}

adeco_status adeco_deposit_graph_launch(adeco_deposit_graph G) {
    g_last_kernel_ms = 0.0;
    if (!G) return fail(ADECO_EINVAL, "deposit_graph_launch(): null handle");
    hipError_t e = hipSetDevice(G->device);
    if (e == hipSuccess) e = hipEventRecord(G->e0, G->stream);
    if (e == hipSuccess) e = hipGraphLaunch(G->exec, G->stream);
    if (e == hipSuccess) e = hipEventRecord(G->e1, G->stream);
    unsigned long long cnt[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, G->counters.as<char>() + 16, 16, hipMemcpyDeviceToHost, G->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, G->e0, G->e1);
    if (e != hipSuccess) return hip_fail(e, "deposit graph launch");
    g_last_kernel_ms = ms;
    g_timings = LastTimings();
    g_timings.total_ms = ms;
    g_timings.main_ms = ms;
    g_timings.candidates = double(G->src.n_cand);
    g_timings.fallbacks = double(cnt[0]);
    g_timings.partials = double(cnt[1]);
    return ADECO_OK;
}

adeco_status adeco_deposit_graph_grids(adeco_deposit_graph G, int64_t first_leaf, int64_t n_leaves, double* out) {
    if (!G || !out) return fail(ADECO_EINVAL, "deposit_graph_grids(): null argument");
    if (first_leaf < 0 || n_leaves < 0 || first_leaf + n_leaves > G->n_leaves)
        return fail(ADECO_EINVAL, "deposit_graph_grids(): leaf range");
    const int64_t cells = int64_t(G->dims) * G->dims * G->dims;
    hipError_t e = hipSetDevice(G->device);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, G->grids.as<double>() + first_leaf * cells, size_t(n_leaves * cells) * 8,
                           hipMemcpyDeviceToHost, G->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G->stream);
    return e == hipSuccess ? ADECO_OK : hip_fail(e, "deposit graph grids");
}

adeco_status adeco_deposit_graph_destroy(adeco_deposit_graph G) {
    if (!G) return ADECO_OK;
    (void)hipSetDevice(G->device);
    if (G->stream) (void)hipStreamSynchronize(G->stream);
    delete G;
    return ADECO_OK;
}

}  // extern "C"
