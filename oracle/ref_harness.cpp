// Test infrastructure only (never shipped, never on the product path).
//
// A thin extern "C" harness around the UNMODIFIED ADECO reference sources, compiled where they lie
// under /root/reference/src by oracle/build_ref.sh into oracle/_ref/libadeco_ref.so.  It is used by
// tests/golden/make_golden.py (in the build container only) to write golden vectors, and by the
// checker tests to validate the oracle restatement.  Nothing here re-implements reference logic:
// every computation is a call into the reference's own functions.
//
//   ref_deposit_tets      -> AdaptiveMassDeposit::DepositWrapper::deposit_tets   (AdDeposit2Grid.cpp:67-125)
//   ref_octree_build      -> MetaMesh (AdTypeDefinitions.h:715-1022) + get_global_bbox_mpi
//                            (AdPointDataReaders.h:90-151) + generate_tree (AdOctreeGeneration.h:76-208)
//                            + Octree::getLeafNodes (AdOctree.h:265-305)
//                            + find_best_node_and_remove_it_from_list (AdDynamicLoadBalancing.h:706-730)
//   ref_octree_deposit_leaves -> gather_local_tets (AdDynamicLoadBalancing.h:894-950) + deposit_tets per
//                            leaf, single worker (the 1-thread scheduler result, AdDynamicLoadBalancing.h:384-540)
//   ref_write_patch_mass  -> AdOctreeGridWriter::write_patch mass sum / density (AdOctreeGridWriter.h:67-144)
//   ref_parse_parameter_file -> parse_parameter_file (main.cpp:83-183)
//   ref_cli_main          -> the reference program's main (main.cpp:187-249): parameter file, MPI
//                            start-up, deposit_workflow (AdDepositWorkFlow.h:60-237), MPI_Finalize
//   ref_apply_boundary_corrections -> apply_boundary_corrections (AdPointDataReaders.h:156-206)
//   ref_sorted_by_id_load -> AdSortedByIdsPointDataReader::load_data (AdPointDataReaders.h:770-911)
//   ref_dark_sky_load     -> AdDarkSkyDataReader::load_data (AdPointDataReaders.h:361-767)
//   ref_dark_sky_header_size -> AdDarkSkyDataReader::get_header_size (AdPointDataReaders.h:419-466)

#include <mpi.h>
#include <fstream>
#include <iomanip>
#include <stack>
#include <cstdint>
#include <cstring>
#include <vector>
#include <list>
#include <memory>
#include <string>

#include "AdMeasureMPIWallClockTime.h"
#include "AdTypeDefinitions.h"
#include "AdPointDataReaders.h"
#include "AdOctreeGeneration.h"
#include "AdDeposit2Grid.h"
#include "AdDynamicLoadBalancing.h"

using namespace AdaptiveMassDeposit;

// Defined in the reference's main.cpp, which build_ref.sh compiles with its main renamed to
// adeco_ref_cli_main.
int adeco_ref_cli_main(int argc, char** argv);
void parse_parameter_file(const std::string& filename, AMD_Parameters& parameters);

namespace {

std::string g_err;

// exposes the reader's protected header parser
struct DarkSkyHeader : AdDarkSkyDataReader {
    static size_t size(const std::string& f) { return get_header_size(f); }
};

void ensure_mpi() {
    int flag = 0;
    MPI_Initialized(&flag);
    if (!flag) MPI_Init(nullptr, nullptr);
}

struct OctreeResult {
    std::shared_ptr<Tetrahedra<ColumnMajorOrder>> tets;
    std::shared_ptr<MetaMesh<ColumnMajorOrder>> mesh;
    TetOctree octree;
    AABBox global_bbox;
    std::vector<node_id_t> leaves_dfs;
    std::vector<node_id_t> leaves_order;
};

OctreeResult* g_oct = nullptr;

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

int ref_deposit_tets(size_t patch_offset, size_t num_patches, const float* positions,
                     size_t positions_size, int meta_cells_per_patch, const float* bbox, int dims,
                     float* data) {
    try {
        DepositWrapper::deposit_tets(false, 0, patch_offset, num_patches, positions, positions_size,
                                     meta_cells_per_patch, bbox, dims, data);
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    return 0;
}

// Builds the reference octree for one rank over an (npx,npy,npz) vertex lattice given as SoA
// floats in ColumnMajorOrder (i fastest).  roi may be null (whole domain, the workflow default).
int ref_octree_build(const float* x, const float* y, const float* z, int npx, int npy, int npz,
                     const float* roi, uint64_t max_tets_param, int max_level) {
    try {
        ensure_mpi();
        delete g_oct;
        g_oct = new OctreeResult();
        const size_t n = size_t(npx) * size_t(npy) * size_t(npz);
        g_oct->tets = std::make_shared<Tetrahedra<ColumnMajorOrder>>();
        Particles& p = g_oct->tets->particles;
        p.set_num_positions(n);
        for (size_t i = 0; i < n; ++i) p.set_position(i, x[i], y[i], z[i]);
        g_oct->tets->set_patch_dims(npx, npy, npz);

        // global bbox exactly as the workflow computes it (AdDepositWorkFlow.h:101)
        std::vector<std::shared_ptr<particles_with_ids_chunk>> per_rank(1);
        per_rank[0] = std::make_shared<Particles>(p);
        g_oct->global_bbox = get_global_bbox_mpi(per_rank);
        AABBox region = g_oct->global_bbox;
        if (roi) region = AABBox(PosVec(roi[0], roi[2], roi[4]), PosVec(roi[1], roi[3], roi[5]));

        g_oct->mesh = std::make_shared<MetaMesh<ColumnMajorOrder>>(*g_oct->tets);
        const unsigned long long global_num_tets = g_oct->tets->get_num_tets();
        const uint64_t max_tets_per_node =
            std::min(size_t(global_num_tets / 1), size_t(max_tets_param));  // AdDepositWorkFlow.h:169
        AdOctreeGeneration<ColumnMajorOrder>::generate_tree(1, 0, g_oct->global_bbox, region,
                                                            max_tets_per_node, max_level,
                                                            *g_oct->mesh, g_oct->octree);
        g_oct->octree.getLeafNodes(g_oct->leaves_dfs);
        std::list<node_id_t> leaf_list(g_oct->leaves_dfs.begin(), g_oct->leaves_dfs.end());
        while (!leaf_list.empty())
            g_oct->leaves_order.push_back(find_best_node_and_remove_it_from_list(0, g_oct->octree, leaf_list));
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    return 0;
}

void ref_octree_bbox(float* out6) {
    const AABBox& b = g_oct->global_bbox;
    out6[0] = b.min[0]; out6[1] = b.max[0]; out6[2] = b.min[1];
    out6[3] = b.max[1]; out6[4] = b.min[2]; out6[5] = b.max[2];
}

int64_t ref_octree_num_nodes() { return g_oct ? int64_t(g_oct->octree.getNodes().size()) : 0; }
int64_t ref_octree_num_leaves() { return g_oct ? int64_t(g_oct->leaves_dfs.size()) : 0; }
int64_t ref_octree_num_meta_cells() {
    std::vector<uint64_t> ids;
    g_oct->mesh->get_cell_ids(ids);
    return int64_t(ids.size());
}

// Node table: lower-left (3 floats), extension, level, parent, first child (-1 when unrefined),
// the refinement count num_tets_covering_node_estimate and num_tets_on_proc_estimate[0].
void ref_octree_nodes(float* ll, float* ext, int32_t* level, int32_t* parent, int32_t* first_child,
                      uint64_t* count, uint64_t* local_count) {
    const TetOctree::Nodes& nodes = g_oct->octree.getNodes();
    for (size_t i = 0; i < nodes.size(); ++i) {
        const TetOctree::Node& nd = nodes[i];
        ll[3 * i + 0] = nd.getLowerLeft()[0];
        ll[3 * i + 1] = nd.getLowerLeft()[1];
        ll[3 * i + 2] = nd.getLowerLeft()[2];
        ext[i] = nd.getExtension();
        level[i] = nd.getLevel();
        parent[i] = nd.getParentId();
        first_child[i] = nd.isRefined() ? nd.getChildId(0) : -1;
        count[i] = nd.getNodeData().num_tets_covering_node_estimate;
        const auto& loc = nd.getNodeData().num_tets_on_proc_estimate;
        local_count[i] = loc.empty() ? 0 : loc[0];
    }
}

// Leaves in DFS order and in the single-rank processing order; per-leaf meta-cell lists (CSR
// over the DFS order).  Call with mc_ids == nullptr to get the total list length.
int64_t ref_octree_leaves(int32_t* dfs, int32_t* order, int64_t* mc_offsets, int64_t* mc_ids) {
    int64_t total = 0;
    for (size_t l = 0; l < g_oct->leaves_dfs.size(); ++l) {
        const auto& ids = g_oct->octree.getNode(g_oct->leaves_dfs[l]).getNodeData().local_meta_cell_ids;
        if (mc_ids) {
            dfs[l] = g_oct->leaves_dfs[l];
            order[l] = g_oct->leaves_order[l];
            mc_offsets[l] = total;
            for (size_t m = 0; m < ids.size(); ++m) mc_ids[total + m] = int64_t(ids[m]);
        }
        total += int64_t(ids.size());
    }
    if (mc_ids) mc_offsets[g_oct->leaves_dfs.size()] = total;
    return total;
}

void ref_meta_cell_bboxes(float* out) {
    std::vector<uint64_t> ids;
    g_oct->mesh->get_cell_ids(ids);
    for (size_t m = 0; m < ids.size(); ++m) {
        const AABBox& b = g_oct->mesh->get_bbox(ids[m]);
        out[6 * m + 0] = b.min[0]; out[6 * m + 1] = b.max[0];
        out[6 * m + 2] = b.min[1]; out[6 * m + 3] = b.max[1];
        out[6 * m + 4] = b.min[2]; out[6 * m + 5] = b.max[2];
    }
}

uint64_t ref_estimate(const float* box6, int64_t meta_cell) {
    const AABBox b(PosVec(box6[0], box6[2], box6[4]), PosVec(box6[1], box6[3], box6[5]));
    return g_oct->mesh->get_estimate_num_tets(b, uint64_t(meta_cell));
}

// Deposits one leaf (by node id) the way a single worker thread does: the leaf's meta-patches are
// gathered with gather_local_tets and deposited into a zeroed float grid by deposit_tets.  When
// out_dir is given, the reference writer's write_patch stores the leaf file there and its double
// mass sum is returned.  grid must hold dims^3 floats.
int ref_octree_deposit_leaf(int32_t node_id, int dims, float* grid, double* mass_out,
                            float* bbox_out, const char* out_dir) {
    try {
        const TetOctree::Node& node = g_oct->octree.getNode(node_id);
        const uint64_t est = node.getNodeData().num_tets_on_proc_estimate[0];
        const int max_tets_per_iteration = 50000000;  // AdDynamicLoadBalancing.h:1002
        const size_t iterations = (est + max_tets_per_iteration) / max_tets_per_iteration;
        std::memset(grid, 0, sizeof(float) * size_t(dims) * dims * dims);
        float bbox[6];
        bbox[0] = node.getLowerLeft()[0]; bbox[1] = node.getLowerLeft()[0] + node.getExtension();
        bbox[2] = node.getLowerLeft()[1]; bbox[3] = node.getLowerLeft()[1] + node.getExtension();
        bbox[4] = node.getLowerLeft()[2]; bbox[5] = node.getLowerLeft()[2] + node.getExtension();
        if (bbox_out) std::memcpy(bbox_out, bbox, sizeof(bbox));
        for (size_t it = 0; it < iterations; ++it) {
            particle_array positions;
            gather_local_tets(0, int(it), int(iterations), *g_oct->mesh, node, positions);
            if (positions.empty()) continue;
            const size_t fpp = 3 * MetaMesh<ColumnMajorOrder>::get_number_of_vertices_per_meta_block();
            DepositWrapper::deposit_tets(true, 0, 0, positions.size() / fpp, positions.data(),
                                         positions.size(),
                                         MetaMesh<ColumnMajorOrder>::get_number_of_cells_per_dimension(),
                                         bbox, dims, grid);
        }
        if (out_dir) {
            AdOctreeGridWriter writer;
            std::vector<float> data(grid, grid + size_t(dims) * dims * dims);
            const double mass = writer.write_patch(out_dir, node, g_oct->octree, dims, data);
            if (mass_out) *mass_out = mass;
        }
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    return 0;
}

// Runs the reference program on a parameter file, as `dm_resampling_cpu <param_file>` would.  MPI
// is finalised at its end and cannot start again, so callers run this in a process of its own.
int ref_cli_main(const char* param_file) {
    std::string prog("dm_resampling_cpu"), arg(param_file);
    std::vector<char> a0(prog.begin(), prog.end()), a1(arg.begin(), arg.end());
    a0.push_back(0);
    a1.push_back(0);
    char* argv[] = {a0.data(), a1.data(), nullptr};
    return adeco_ref_cli_main(2, argv);
}

// The reference's parameter-file parser.  strs receives input_file, output_directory, id_type and
// restart_path, cap bytes each (NUL-terminated, truncated); ints = {use_regions_of_interest,
// max_refinement_level, stride, max_tets_per_octree_node, linear_patch_resolution, num_pthreads,
// num_mpi_ranks_per_node, periodic_boundaries}; roi6 = x0,x1,y0,y1,z0,z1.
int ref_parse_parameter_file(const char* path, char* strs, int cap, int64_t* ints, float* roi6) {
    try {
        AMD_Parameters p;
        parse_parameter_file(path, p);
        const std::string* s[4] = {&p.input_file, &p.output_directory, &p.id_type, &p.restart_path};
        for (int i = 0; i < 4; ++i) {
            const size_t n = std::min(s[i]->size(), size_t(cap - 1));
            std::memcpy(strs + size_t(i) * cap, s[i]->data(), n);
            strs[size_t(i) * cap + n] = 0;
        }
        ints[0] = p.use_regions_of_interest;
        ints[1] = p.max_refinement_level;
        ints[2] = p.stride;
        ints[3] = int64_t(p.max_tets_per_octree_node);
        ints[4] = p.linear_patch_resolution;
        ints[5] = p.num_pthreads;
        ints[6] = p.num_mpi_ranks_per_node;
        ints[7] = p.periodic_boundaries;
        for (int a = 0; a < 3; ++a) {
            roi6[2 * a] = p.region_of_interest.min[a];
            roi6[2 * a + 1] = p.region_of_interest.max[a];
        }
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    return 0;
}

// Periodic-boundary correction of an (npx,npy,npz) lattice (ids = ColumnMajorOrder linear index)
// with the given global bbox; x, y, z are corrected in place.
int ref_apply_boundary_corrections(float* x, float* y, float* z, int npx, int npy, int npz, const float* b) {
    try {
        const size_t n = size_t(npx) * size_t(npy) * size_t(npz);
        Particles p;
        for (size_t i = 0; i < n; ++i) p.push_back(x[i], y[i], z[i], ids_t(i));
        const AABBox box(PosVec(b[0], b[2], b[4]), PosVec(b[1], b[3], b[5]));
        apply_boundary_corrections(box, AvVec3i(npx, npy, npz), ColumnMajorOrder(), p);
        const Particles::pos_t_vec& px = *p.get_positions(0);
        const Particles::pos_t_vec& py = *p.get_positions(1);
        const Particles::pos_t_vec& pz = *p.get_positions(2);
        for (size_t i = 0; i < n; ++i) {
            x[i] = px[i];
            y[i] = py[i];
            z[i] = pz[i];
        }
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
    return 0;
}

// Loads a SORTED_BY_ID file on one rank.  Returns the particle count (-1 on error); with cap at
// least that count, x, y, z and ids receive the particles in load order and dims3 the lattice.
int64_t ref_sorted_by_id_load(const char* path, int64_t cap, float* x, float* y, float* z, int64_t* ids, int* dims3) {
    try {
        ensure_mpi();
        std::shared_ptr<AdPointDataReader> reader = get_reader("SORTED_BY_ID");
        reader->load_data(path, 0, 1, 1);
        const Particles& p = *reader->get_particles_per_rank()[0];
        const int64_t n = int64_t(p.get_num());
        if (n <= cap) {
            for (int64_t i = 0; i < n; ++i) {
                pos_t pos[3];
                p.get_position(size_t(i), pos);
                x[i] = pos[0];
                y[i] = pos[1];
                z[i] = pos[2];
                ids[i] = int64_t(p.get_id(size_t(i)));
            }
            const AvVec3i d = reader->get_global_particle_dims();
            dims3[0] = d[0];
            dims3[1] = d[1];
            dims3[2] = d[2];
        }
        return n;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Loads a DARK_SKY file on one rank with the given stride.  Returns the particle count (-1 on
// error); with cap at least that count, x, y, z and ids receive the particles in load order (ids:
// the linear lattice index after the stride) and dims3 the lattice.
int64_t ref_dark_sky_load(const char* path, int stride, int64_t cap, float* x, float* y, float* z, int64_t* ids,
                          int* dims3) {
    try {
        ensure_mpi();
        std::shared_ptr<AdPointDataReader> reader = get_reader("DARK_SKY");
        reader->load_data(path, 0, 1, stride);
        const Particles& p = *reader->get_particles_per_rank()[0];
        const int64_t n = int64_t(p.get_num());
        if (n <= cap) {
            for (int64_t i = 0; i < n; ++i) {
                pos_t pos[3];
                p.get_position(size_t(i), pos);
                x[i] = pos[0];
                y[i] = pos[1];
                z[i] = pos[2];
                ids[i] = int64_t(p.get_id(size_t(i)));
            }
            const AvVec3i d = reader->get_global_particle_dims();
            dims3[0] = d[0];
            dims3[1] = d[1];
            dims3[2] = d[2];
        }
        return n;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int64_t ref_dark_sky_header_size(const char* path) {
    try {
        return int64_t(DarkSkyHeader::size(path));
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"
