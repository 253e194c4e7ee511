// hipGraph forms of the batched leaf deposit, replayed with one graph launch and no host round
// trip:
//   * deposit graph: the grid clear, the work-queue reset and the persistent deposit launch.  The
//     candidate space (pair data, scan, tile map) is built when the graph is created, from the
//     lattice positions and leaf binning of that moment, and each replay deposits it again.
//   * step graph: the whole step - per-leaf binning of the lattice's current positions (count
//     pass, leaf-offset scan, fill pass), pair setup, candidate scan, tile map, grid clear, queue
//     reset and the persistent deposit - with every size kept on the device, so one capture serves
//     later positions of the lattice (adeco_lattice_update).  The leaves stay the captured
//     octree's.  Buffers are sized at capture for `slack` times that moment's pairs and tiles; a
//     replay that needs more sets an overflow flag, deposits nothing and reports ADECO_EOVERFLOW.
// Either graph owns every buffer its kernels write (binning counters, pair lists, leaf geometry,
// pair data, tile map, leaf grids, counters) and its own copies of the octree tables the binning
// reads, so deposits through the octree handle cannot move them; the lattice must outlive it.
#include <hip/hip_runtime.h>

#include "../../include/adeco.h"
#include "batched_impl.hpp"
#include "kernels.hpp"
#include "lattice_impl.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

using namespace adeco;

static size_t align_up(size_t b) { return (b + 255) & ~size_t(255); }

struct adeco_deposit_graph_s {
    adeco_octree_s* T = nullptr;
    int device = 0, dims = 0, precision = 0;
    int64_t n_leaves = 0;
    DevBuf geom, work, tiles, grids, counters;
    TileSource src;
    // step graph: binning plan over the graph's own tables and buffers, capacities, sizes
    bool step = false;
    int64_t pair_cap = 0, tile_cap = 0;
    DevBuf b_tables, b_count, b_over, b_tleaf, b_offset, b_scan, b_pairs, b_sizes;
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
    // counters: [0..3] queue counter, overflow flag; [16..31] recomputes, partials
    unsigned long long cnt[2] = {0, 0};
    unsigned flags[4] = {0, 0, 0, 0};
    int64_t sizes[2] = {G->src.n_cand, G->src.n_tiles};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, G->counters.as<char>() + 16, 16, hipMemcpyDeviceToHost, G->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(flags, G->counters.ptr, 16, hipMemcpyDeviceToHost, G->stream);
    if (e == hipSuccess && G->step) e = hipMemcpyAsync(sizes, G->b_sizes.ptr, 16, hipMemcpyDeviceToHost, G->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, G->e0, G->e1);
    if (e != hipSuccess) return hip_fail(e, "deposit graph launch");
    g_last_kernel_ms = ms;
    g_timings = LastTimings();
    g_timings.total_ms = ms;
    g_timings.main_ms = ms;
    g_timings.candidates = double(sizes[0]);
    g_timings.fallbacks = double(cnt[0]);
    g_timings.partials = double(cnt[1]);
    if (G->step && flags[1])
        return fail(ADECO_EOVERFLOW, "deposit_graph_launch(): the lattice's pairs or tiles exceed the step graph's "
                                     "capacity (nothing deposited); create the graph again");
    return ADECO_OK;
}

adeco_status adeco_deposit_graph_create_step(adeco_octree T, int dims, int decomposition, double tet_mass,
                                             int precision, double slack, adeco_deposit_graph* out) {
    if (!T || !out) return fail(ADECO_EINVAL, "deposit_graph_create_step(): null argument");
    *out = nullptr;
    if (dims < 1 || dims > 1023) return fail(ADECO_EINVAL, "deposit_graph_create_step(): dims must be in 1..1023");
    if (decomposition < 0 || decomposition > 2) return fail(ADECO_EINVAL, "deposit_graph_create_step(): unknown decomposition");
    if (precision < 0 || precision > 2) return fail(ADECO_EINVAL, "deposit_graph_create_step(): unknown precision");
    if (!(slack >= 1.0 && slack <= 16.0)) return fail(ADECO_EINVAL, "deposit_graph_create_step(): slack must be in [1, 16]");
    DeviceState& ds = *T->L->ds;
    std::lock_guard<std::mutex> lk(ds.mutex);
    hipError_t e = hipSetDevice(T->L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    adeco_status st;
    // one eager binning and candidate space: the capacities, and the octree's tables
    if ((st = octree_bin_tets(T, decomposition))) return st;
    auto* G = new adeco_deposit_graph_s();
    G->T = T;
    G->device = T->L->device;
    G->dims = dims;
    G->precision = precision;
    G->step = true;
    G->n_leaves = int64_t(T->leaves_dfs.size());
    const int64_t nl = G->n_leaves;
    const int64_t words = nl * int64_t(dims) * dims * dims;
    auto bail = [&](adeco_status s) {
        delete G;
        return s;
    };
    if (size_t(words) * 8 > memory_budget(ds))
        return bail(fail(ADECO_EOVERFLOW, "deposit_graph_create_step(): leaf grids exceed the budget"));
    if ((st = G->grids.reserve(size_t(words) * 8)) || (st = G->counters.reserve(64))) return bail(st);
    e = hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&G->e0);
    if (e == hipSuccess) e = hipEventCreate(&G->e1);
    if (e != hipSuccess) return bail(hip_fail(e, "deposit_graph_create_step()"));
    TileSource eager;
    if ((st = prepare_leaves_batched(T, G->stream, dims, decomposition, tet_mass, precision, G->geom, G->work, G->tiles,
                                     G->grids.as<double>(), eager)))
        return bail(st);
    const int64_t P = T->leaf_tet_offset[nl];
    G->pair_cap = int64_t(double(P) * slack) + 1024;
    G->tile_cap = int64_t(double(eager.n_tiles) * slack) + 64;
    if (G->pair_cap >= (int64_t(1) << 31) - 2) return bail(fail(ADECO_EOVERFLOW, "deposit_graph_create_step(): too many pairs"));
    // the graph's own binning plan: tables copied from the octree's, buffers of its own
    BinPlan plan;
    if ((st = bin_plan(T, decomposition, plan))) return bail(st);
    const size_t nn = T->nodes.size();
    const size_t map_words = plan.use_map ? size_t(plan.lm.F) * plan.lm.F * plan.lm.F : 0;
    const size_t tb = align_up(nn * 24) + align_up(nn * 4) * 2 + align_up(map_words * 4) + align_up(size_t(nl) * 24);
    const int64_t n_tets = plan.n_tets;
    if ((st = G->b_tables.reserve(tb)) || (st = G->b_count.reserve(size_t(nl) * 4 + 4)) ||
        (st = G->b_over.reserve(size_t(n_tets) * 4 + 8)) || (st = G->b_tleaf.reserve(plan.use_map ? size_t(n_tets) * 4 : 4)) ||
        (st = G->b_offset.reserve(size_t(nl + 1) * 8)) || (st = G->b_scan.reserve(scan_block_words(nl) * 8 + 8)) ||
        (st = G->b_pairs.reserve(size_t(G->pair_cap) * 4 + 4)) || (st = G->b_sizes.reserve(16)))
        return bail(st);
    char* t = G->b_tables.as<char>();
    float* node_box = reinterpret_cast<float*>(t);
    int32_t* first = reinterpret_cast<int32_t*>(t + align_up(nn * 24));
    int32_t* leaf_index = reinterpret_cast<int32_t*>(t + align_up(nn * 24) + align_up(nn * 4));
    int32_t* map = reinterpret_cast<int32_t*>(t + align_up(nn * 24) + 2 * align_up(nn * 4));
    float* leaf_box = reinterpret_cast<float*>(t + align_up(nn * 24) + 2 * align_up(nn * 4) + align_up(map_words * 4));
    e = hipMemcpyAsync(node_box, plan.node_box, nn * 24, hipMemcpyDeviceToDevice, G->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(first, plan.first, nn * 4, hipMemcpyDeviceToDevice, G->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(leaf_index, plan.leaf_index, nn * 4, hipMemcpyDeviceToDevice, G->stream);
    if (e == hipSuccess && plan.use_map) {
        e = hipMemcpyAsync(map, plan.lm.map, map_words * 4, hipMemcpyDeviceToDevice, G->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(leaf_box, plan.lm.leaf_box, size_t(nl) * 24, hipMemcpyDeviceToDevice, G->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(G->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "deposit_graph_create_step(): tables"));
    plan.node_box = node_box;
    plan.first = first;
    plan.leaf_index = leaf_index;
    if (plan.use_map) {
        plan.lm.map = map;
        plan.lm.leaf_box = leaf_box;
        plan.tleaf = G->b_tleaf.as<int32_t>();
    }
    plan.leaf_count = G->b_count.as<int>();
    plan.n_over = G->b_over.as<int>();
    plan.over = reinterpret_cast<int32_t*>(G->b_over.as<char>() + 8);
    int64_t* leaf_offset = G->b_offset.as<int64_t>();
    int64_t* scan_bs = G->b_scan.as<int64_t>();
    int64_t* scan_total = scan_bs + scan_block_words(nl);
    const int64_t* n_pairs_dev = leaf_offset + nl;
    // the candidate space at capacity (the graph's own slab, tile map and sizes)
    if ((st = G->work.reserve(candidate_space_bytes(G->pair_cap) + 256)) || (st = G->tiles.reserve(size_t(G->tile_cap + 1) * 4)))
        return bail(st);
    const LatticeView lv{T->L->d_x, T->L->d_y, T->L->d_z, {T->L->np[0], T->L->np[1], T->L->np[2]}};
    unsigned* ctr = G->counters.as<unsigned>();
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(G->counters.as<char>() + 16);
    int64_t* sizes = G->b_sizes.as<int64_t>();
    G->src = TileSource{nullptr, nullptr, nullptr, G->tile_cap, 0, G->geom.as<GridDesc>(), G->grids.as<double>(), tet_mass};
    e = hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = launch_zero_u32(ctr, 16, G->stream);  // queue counter and overflow flag
    if (e == hipSuccess) e = bin_count_async(plan, G->stream);
    if (e == hipSuccess)
        e = launch_scan_u32(reinterpret_cast<const unsigned*>(plan.leaf_count), nl, leaf_offset, scan_bs, scan_total, G->stream);
    if (e == hipSuccess) e = bin_fill_async(plan, leaf_offset, G->b_pairs.as<int32_t>(), G->pair_cap, G->stream);
    CandidateSpace cs;
    if (e == hipSuccess)
        e = enqueue_candidate_space_dev(G->stream, lv, plan.ct, G->b_pairs.as<int32_t>(), G->pair_cap, n_pairs_dev,
                                        leaf_offset, int(nl), G->geom.as<GridDesc>(), tet_mass, precision, G->work.ptr,
                                        G->tiles.as<int32_t>(), G->tile_cap, sizes, ctr + 1, cs);
    if (e == hipSuccess) {
        G->src.pd = cs.pd;
        G->src.cand_off = cs.cand_off;
        G->src.tile_first = cs.tile_first;
        G->src.n_cand = int64_t(G->tile_cap) * 64;
        G->src.dev_sizes = sizes;
        e = launch_fill_f64(G->grids.as<double>(), words, 0.0, G->stream);
    }
    // (the launch shape is chosen from the capture's tile count; the kernel reads the replay's)
    TileSource shape = G->src;
    shape.n_tiles = eager.n_tiles;
    if (e == hipSuccess) e = launch_deposit_tiles(G->stream, precision, shape, ctr, counters, G->device);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(G->stream, &graph);
    G->graph = graph;
    if (e == hipSuccess) e = ec;
    if (e == hipSuccess) e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(hip_fail(e, "step graph capture"));
    G->src.n_tiles = eager.n_tiles;
    G->src.n_cand = eager.n_cand;
    *out = G;
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
