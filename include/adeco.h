/*
 * adeco.h — C ABI of the MI355X-native ADECO mass deposit (libadeco_amd.so).
 *
 * Plain C: pointers, sizes and status codes only.  Every call is synchronous unless stated, keeps
 * no pointer to caller memory after it returns, and reports failures through an adeco_status plus
 * a thread-local message (adeco_last_error); nothing calls exit().  All GPU work runs in
 * hand-written HIP kernels for gfx950; there is no CPU fallback: without a usable GPU the
 * compute entry points return ADECO_EDEVICE.
 *
 * Reference interfaces replaced (kaehlerr/ADECO, src/):
 *   adeco_deposit_patches    DepositWrapper::deposit_tets            AdDeposit2Grid.h:82-91
 *                            (CPU AdDeposit2Grid.cpp:67-125, CUDA AdDeposit2Grid.cu:229-584)
 *   adeco_get_grid_size      DepositWrapper::get_cuda_grid_size      AdDeposit2Grid.h:72-75
 *   adeco_get_block_size     DepositWrapper::get_cuda_block_size     AdDeposit2Grid.h:77-80
 *   adeco_lattice_periodic_unwrap  apply_boundary_corrections      AdPointDataReaders.h:156-206
 *   adeco_octree_*           MetaMesh + AdOctreeGeneration::generate_tree + Octree
 *                            AdTypeDefinitions.h:715-1022, AdOctreeGeneration.h:76-258, AdOctree.h:63-392
 *   adeco_deposit_leaves     deposit_leaf_nodes (one rank)           AdDynamicLoadBalancing.h:983-1323
 *                            + gather_local_tets                     AdDynamicLoadBalancing.h:894-950
 *   adeco_parse_parameter_file  parse_parameter_file                 main.cpp:83-183
 *   adeco_read_sorted_by_id  AdSortedByIdsPointDataReader::load_data AdPointDataReaders.h:770-911
 *   adeco_lattice_create_dark_sky  AdDarkSkyDataReader::load_data  AdPointDataReaders.h:361-767
 *   adeco_deposit_workflow   deposit_workflow (one rank)             AdDepositWorkFlow.h:60-237
 *   adeco_write_patch        AdOctreeGridWriter::write_patch         AdOctreeGridWriter.h:67-144
 *   adeco_write_leaves       the write step of deposit_leaf_nodes    AdDynamicLoadBalancing.h:998,1133,1303-1312
 *   adeco_serve_*            (no counterpart: region-of-interest density service on the resident lattice)
 */
#ifndef ADECO_H
#define ADECO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int adeco_status;
#define ADECO_OK 0
#define ADECO_EINVAL 1    /* invalid argument (the reference throws AdRuntimeException or asserts) */
#define ADECO_EOVERFLOW 2 /* a size does not fit (grid larger than the memory budget, ...) */
#define ADECO_ENOMEM 3    /* device or host allocation failed */
#define ADECO_EDEVICE 4   /* no usable GPU, or a HIP runtime error */
#define ADECO_EINTERNAL 5 /* internal consistency check failed */
#define ADECO_EIO 6       /* a file could not be written (the reference throws AdRuntimeException) */

/* Tet decompositions of a Lagrangian lattice cube (6 tets each). */
#define ADECO_DECOMP_REF_EFFECTIVE 0 /* what the reference deposit applies (AdDepositTets2Grid.h:947-972 x :213-218) */
#define ADECO_DECOMP_REF_INTENDED 1  /* the same table in vert[] order (AdDepositTets2Grid.h:188-192): tiles the cube */
#define ADECO_DECOMP_KUHN 2          /* six tets around the (0,0,0)-(1,1,1) diagonal */

/* Deposit execution forms. */
#define ADECO_MODE_PLAIN 0   /* one tet per thread, one launch per output grid */
#define ADECO_MODE_BATCHED 1 /* persistent work-queue launch over all grids, load-balanced (tet, cell) work */

/* Precision of the clip (accumulation is always FP64 on the device). */
#define ADECO_PREC_FP64 0  /* all geometry in FP64 */
#define ADECO_PREC_MIXED 1 /* FP32 clip with FP64 recompute where the FP32 error bound is not met */
#define ADECO_PREC_FP32 2  /* FP32 clip only (for the precision study) */

/* Cell scaling of the clip (library-wide, read when a deposit starts).  The reference scales all
 * three axes by 1/dx (AdDepositTets2Grid.h:569), assuming cubic cells; float boxes (ll, ll + ext)
 * rarely have bit-identical extents, so that formula shifts cell faces by ~1e-7 of a cell and
 * breaks mass conservation at that level.  EXACT (default) scales each axis by its own 1/delta. */
#define ADECO_SCALING_EXACT 0
#define ADECO_SCALING_REFERENCE 1

const char* adeco_last_error(void);

/* Library-wide form of the boundary call adeco_deposit_patches (set before concurrent use): the 6
 * tets of a lattice cube (default ADECO_DECOMP_REF_EFFECTIVE, the reference's), the mass of one tet
 * (default 1, as resample_meta_patch, AdDepositTets2Grid.h:1007-1053), the kernel form (default
 * ADECO_MODE_PLAIN, one thread per tet like the reference's kernel; ADECO_MODE_BATCHED expands
 * each pass's patches into a tet soup on the device and runs the batched persistent deposit on it)
 * and the precision (default ADECO_PREC_FP64; MIXED and FP32 need the batched form). */
adeco_status adeco_set_tet_decomposition(int decomposition);
int adeco_get_tet_decomposition(void);
adeco_status adeco_set_tet_mass(double tet_mass);
double adeco_get_tet_mass(void);
adeco_status adeco_set_deposit_mode(int mode);
int adeco_get_deposit_mode(void);
adeco_status adeco_set_precision_mode(int precision);
int adeco_get_precision_mode(void);
int adeco_version(void);
adeco_status adeco_set_cell_scaling(int mode);
int adeco_get_cell_scaling(void);
/* Number of visible GPUs, or -1 when the HIP runtime reports an error. */
int adeco_device_count(void);

/* ---- drop-in boundary: DepositWrapper::deposit_tets (AdDeposit2Grid.h:82-91) ----
 * positions: concatenated meta-patches of 3*(m+1)^3 floats (AoS xyz, vertex (i,j,k) at float
 * offset 3*(i+(m+1)*(j+(m+1)*k)), 666.f sentinel marks absent vertices); patches
 * [patch_offset, patch_offset+num_patches) are deposited with tet mass 1 into the dims^3 grid over
 * bbox = x0,x1,y0,y1,z0,z1, ADDED into grid_accum (float, index i+dims*(j+dims*k)).  The device
 * accumulates in FP64 and the sum is rounded to float once when added.  use_specific_gpu selects
 * device_id; otherwise the patches are split over all visible GPUs.  Positions beyond the memory
 * budget are streamed in several passes. */
adeco_status adeco_deposit_patches(int use_specific_gpu, size_t device_id, size_t patch_offset,
                                   size_t num_patches, const float* positions, size_t positions_size,
                                   int meta_cells_per_patch, const float* bbox, int dims,
                                   float* grid_accum);
/* adeco_deposit_patches in the plain form on one device, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Arguments and accumulation as
 * adeco_deposit_patches with use_specific_gpu set (library-wide tet decomposition and tet mass;
 * the kernel is always the plain FP64 one).  Device memory beyond the caller's buffers comes from a
 * fixed 48 MiB array each GPU holds from library load to process end (adeco_noalloc_scratch_bytes):
 * the FP64 grid and the staged float grid take 12 * dims^3 bytes of it, the meta-patches stream
 * through the rest, so dims is at most 161 with 7^3-cell patches; larger grids are ADECO_EOVERFLOW.
 * Calls on one device share its array and serialise on that device's lock. */
adeco_status adeco_deposit_patches_noalloc(int device_id, size_t patch_offset, size_t num_patches,
                                           const float* positions, size_t positions_size,
                                           int meta_cells_per_patch, const float* bbox, int dims,
                                           float* grid_accum);
size_t adeco_noalloc_scratch_bytes(void);
int adeco_get_grid_size(void);
int adeco_get_block_size(void);
/* Device-memory budget per call in bytes (0 = 75% of device memory, as the reference's
 * MAX_GPU_USAGE, AdDeposit2Grid.cu:271).  Lowering it forces multi-pass streaming. */
adeco_status adeco_set_memory_budget(size_t bytes);

/* ---- resident lattice ---- */
typedef struct adeco_lattice_s* adeco_lattice;
/* Uploads an (npx,npy,npz) SoA float vertex lattice (i fastest) to device_id once. */
adeco_status adeco_lattice_create(int device_id, const float* x, const float* y, const float* z,
                                  int npx, int npy, int npz, adeco_lattice* out);
adeco_status adeco_lattice_destroy(adeco_lattice lat);
/* Overwrites the resident vertex positions with a new state of the same lattice (the next snapshot
 * of a simulation); octrees, graphs and servers on the lattice see the new positions from their
 * next deposit.  Not concurrently with other calls on the lattice. */
adeco_status adeco_lattice_update(adeco_lattice lat, const float* x, const float* y, const float* z);
/* Float min/max over all vertices (AdPointDataReaders.h:90-151), computed on the device. */
adeco_status adeco_lattice_bbox(adeco_lattice lat, float* bbox6);

/* Copies the resident vertex coordinates (SoA, npx*npy*npz floats each) back to the host. */
adeco_status adeco_lattice_positions(adeco_lattice lat, float* x, float* y, float* z);
/* Periodic-boundary correction on the device, as apply_boundary_corrections
 * (AdPointDataReaders.h:156-206) with the global bbox bbox6 (normally adeco_lattice_bbox before
 * the call; vertex ids are the ColumnMajorOrder lattice index): along each axis a vertex with index
 * > n/2 lying below min + ext/4 moves up by the shift, one with index < n/2 lying above
 * min + 3/4 ext moves down.  period3 null: the reference's shift, the bbox extent, and bbox6 stays
 * the global bbox that later octrees are rooted at (the reference does not recompute it,
 * AdDepositWorkFlow.h:101-112).  period3 given: shift by the period per axis, and later octrees are
 * rooted at the corrected vertices' bbox.  Call before building octrees or servers on the lattice. */
adeco_status adeco_lattice_periodic_unwrap(adeco_lattice lat, const float* bbox6, const float* period3);

/* Deposits lattice tets (all 6*(npx-1)(npy-1)(npz-1), or the listed tet ids) with tet_mass onto
 * one dims^3 grid over bbox6; grid_out (host, double, dims^3) is overwritten.  stats (nullable,
 * 11 x uint64) receives the work counters of a counting launch. */
adeco_status adeco_lattice_deposit_grid(adeco_lattice lat, int decomposition, double tet_mass,
                                        const int32_t* tet_ids, int64_t n_tets, const float* bbox6,
                                        int dims, int mode, int precision, double* grid_out,
                                        uint64_t* stats);

/* Deposits a tet soup (float [n][4][3]) onto one grid (precision study inputs).  MODE_PLAIN runs
 * the one-tet-per-thread FP64 kernel; MODE_BATCHED the persistent (tet, cell) kernel in the given
 * precision, with its kernel time, partial cells and FP64 recomputes in adeco_last_timings. */
adeco_status adeco_deposit_soup(int device_id, const float* tets, int64_t n_tets, const float* bbox6,
                                int dims, double tet_mass, int mode, int precision, double* grid_out,
                                uint64_t* stats);

/* Kernel time (ms, HIP events on the launch stream) of the calling thread's last deposit. */
double adeco_last_kernel_ms(void);

/* ---- octree on the resident lattice: MetaMesh + generate_tree + Octree, one rank ----
 * Root = (bbox.min, max extent) of the lattice's float bbox; meta-cells of 7^3 cubes; a node is
 * refined iff it strictly meets roi6 (null: the whole domain), its estimate count > 0 and
 * > min(num_tets, max_tets_per_node), and level < max_level.  Node ids, the DFS leaf list, the
 * leaf processing order and every count are those of the reference. */
typedef struct adeco_octree_s* adeco_octree;
adeco_status adeco_octree_build(adeco_lattice lat, const float* roi6, uint64_t max_tets_per_node, int max_level,
                                adeco_octree* out);
adeco_status adeco_octree_destroy(adeco_octree tree);
int64_t adeco_octree_num_nodes(adeco_octree tree);
int64_t adeco_octree_num_leaves(adeco_octree tree);
int64_t adeco_octree_num_meta_cells(adeco_octree tree);
adeco_status adeco_octree_bbox(adeco_octree tree, float* bbox6);
/* Per node (any pointer may be null): lower-left corner (3 floats), extension, level, parent id
 * (-1 for the root), first child id (-1 when unrefined) and num_tets_covering_node_estimate. */
adeco_status adeco_octree_nodes(adeco_octree tree, float* ll, float* ext, int32_t* level, int32_t* parent,
                                int32_t* first_child, uint64_t* count);
/* Leaves in DFS order, in processing order, and their meta-cell id lists (CSR over DFS order,
 * mc_offsets has num_leaves + 1 entries).  Returns the total list length (-1 on a null handle). */
int64_t adeco_octree_leaves(adeco_octree tree, int32_t* dfs, int32_t* order, int64_t* mc_offsets, int32_t* mc_ids);
/* Meta-cell float AABBs, 6 floats (x0,x1,y0,y1,z0,z1) per meta-cell. */
adeco_status adeco_octree_meta_bboxes(adeco_octree tree, float* out);
/* Bins every lattice tet (in the given decomposition) to the leaves its AABB strictly overlaps;
 * leaf_tet_counts (nullable, num_leaves) receives the binned tets per leaf (DFS order). */
adeco_status adeco_octree_bin(adeco_octree tree, int decomposition, int64_t* leaf_tet_counts);

/* ---- leaf deposit: deposit_leaf_nodes for one rank ----
 * Deposits every binned tet into its leaves' dims^3 FP64 grids (HBM-resident, DFS leaf order);
 * mode and precision select the kernel form.  The kernel time is reported by
 * adeco_last_kernel_ms. */
adeco_status adeco_deposit_leaves(adeco_octree tree, int dims, int decomposition, double tet_mass, int mode,
                                  int precision);
/* adeco_deposit_leaves in the batched persistent form (ADECO_MODE_BATCHED). */
adeco_status adeco_deposit_leaves_batched(adeco_octree tree, int dims, int decomposition, double tet_mass,
                                          int precision);
/* Copies leaf grids [first_leaf, first_leaf + n_leaves) of the last deposit (FP64 mass per cell,
 * i fastest, leaf-major in DFS leaf order). */
adeco_status adeco_leaf_grids(adeco_octree tree, int64_t first_leaf, int64_t n_leaves, double* out);
/* Leaf boxes in DFS order (6 doubles per leaf): exact = 0 gives the reference's float boxes
 * (ll, ll + ext) that ADECO_SCALING_REFERENCE deposits into; exact = 1 the dyadic FP64
 * subdivision of the root that ADECO_SCALING_EXACT uses, which tiles the root exactly. */
adeco_status adeco_octree_leaf_boxes(adeco_octree tree, int exact, double* out);
/* Binned lattice tet ids of one leaf (DFS index) after adeco_octree_bin / adeco_deposit_leaves;
 * returns their number (-1 on error); out (nullable) receives them. */
int64_t adeco_octree_leaf_tets(adeco_octree tree, int64_t leaf, int32_t* out);
/* ---- leaf files: AdOctreeGridWriter::write_patch (AdOctreeGridWriter.h:67-144) ----
 * Writes <output_path>/node_<level>_<ix>_<iy>_<iz>: grid_dims^3 float32 densities (mass / cell
 * volume, i fastest, host byte order), (ix,iy,iz) = rint((ll - root_ll) / (ext / grid_dims)).  The
 * node box is (ll, ll + ext) in float; cell size and volume are double; density[i] =
 * float(mass[i] * (1 / cell_vol)); *mass_out (nullable) = double sum of mass_grid in index order.
 * Host-only.  ADECO_EIO when the file cannot be written. */
adeco_status adeco_write_patch(const char* output_path, int level, const float* node_ll3, float node_ext,
                               const float* root_ll3, int grid_dims, const float* mass_grid, double* mass_out);
/* Writes every leaf of the last adeco_deposit_leaves with adeco_write_patch (the FP64 leaf grids
 * are rounded to float on the device first).  leaf_mass (nullable, num_leaves, DFS order) receives
 * each leaf's mass; total_mass (nullable) their sum in the leaf processing order, as the
 * reference's total (AdDynamicLoadBalancing.h:998,1133). */
adeco_status adeco_write_leaves(adeco_octree tree, const char* output_path, double* leaf_mass, double* total_mass);
/* Launch shape of the batched kernel: tiles of 64 (tet, candidate) pairs taken per work-queue pop
 * (1..1024) and resident workgroups of 4 wavefronts per CU (1..16).  Library-wide; set it before
 * concurrent use.  No reference counterpart: the reference's GPU deposit uses a fixed 512 x 64
 * launch (AdDeposit2Grid.h:72-80). */
adeco_status adeco_set_launch_shape(int chunk, int blocks_per_cu);
adeco_status adeco_get_launch_shape(int* chunk, int* blocks_per_cu);
/* Times the batched leaf deposit over chunk in {1,2,4,8,16,32} x blocks_per_cu in 1..B, B the
 * workgroups of the kernel that one CU holds at once (hipOccupancyMaxActiveBlocksPerMultiprocessor,
 * at most 4), keeps the fastest (minimum kernel time of reps runs each) as the launch shape and
 * returns it.  Launches never place more workgroups per CU than B, whatever the shape says. */
adeco_status adeco_autotune_batched(adeco_octree tree, int dims, int decomposition, double tet_mass, int precision,
                                    int reps, int* best_chunk, int* best_blocks_per_cu, double* best_ms);
/* hipGraph form of the batched leaf deposit (no reference counterpart: the reference launches per
 * call, AdDeposit2Grid.cu:229-584).  Create builds the candidate space from the current positions
 * and leaf binning and captures {grid clear, work-queue reset, persistent deposit launch}; each
 * launch replays it (synchronous; the graph's time is reported by adeco_last_kernel_ms and
 * adeco_last_timings).  The graph owns its buffers, including its own leaf grids; destroy it
 * before the lattice. */
typedef struct adeco_deposit_graph_s* adeco_deposit_graph;
adeco_status adeco_deposit_graph_create(adeco_octree tree, int dims, int decomposition, double tet_mass, int precision,
                                        adeco_deposit_graph* out);
/* hipGraph of the whole leaf-deposit step (no reference counterpart): per-leaf binning of the
 * lattice's current positions, pair setup, candidate scan, tile map, grid clear and the persistent
 * deposit, with every size kept on the device, so replays follow later positions of the lattice
 * (adeco_lattice_update) without a new capture; the leaves stay the captured octree's.  Buffers are
 * sized for slack (1..16) times the capture's pairs and tiles; a replay that needs more deposits
 * nothing and returns ADECO_EOVERFLOW (create the graph again).  Launched, read and destroyed like
 * the deposit graph. */
adeco_status adeco_deposit_graph_create_step(adeco_octree tree, int dims, int decomposition, double tet_mass,
                                             int precision, double slack, adeco_deposit_graph* out);
adeco_status adeco_deposit_graph_launch(adeco_deposit_graph graph);
adeco_status adeco_deposit_graph_grids(adeco_deposit_graph graph, int64_t first_leaf, int64_t n_leaves, double* out);
adeco_status adeco_deposit_graph_destroy(adeco_deposit_graph graph);
/* The calling thread's last leaf deposit: out5 = {total ms, dominant-kernel ms, (tet, cell)
 * candidates, partial cells clipped, FP64 recomputes of the mixed-precision clip}. */
void adeco_last_timings(double* out5);

/* ---- the single-rank program: parameter file, SORTED_BY_ID reader, deposit_workflow ----
 * AMD_Parameters (AdTypeDefinitions.h:1218-1247) as a plain struct.  adeco_parameters_init sets
 * the defaults the reference's constructor and parser set (max_refinement_level 5, stride 1,
 * linear_patch_resolution 64, one thread and rank, periodic_boundaries on, no tet limit,
 * region_of_interest (0,1)^3 unused). */
typedef struct {
    char input_file[4096];
    char output_directory[4096];
    char id_type[256];         /* file_format / id_type: SORTED_BY_ID */
    char restart_path[4096];   /* parsed, unused (as in the reference) */
    int use_regions_of_interest;
    int max_refinement_level;
    int stride;                /* parsed; the SORTED_BY_ID reader ignores it (as the reference) */
    int linear_patch_resolution;
    int num_pthreads;          /* parsed; one GPU does the work of the reference's threads */
    int num_mpi_ranks_per_node;
    int periodic_boundaries;
    int pad_;
    uint64_t max_tets_per_octree_node;
    float region_of_interest[6]; /* x0,x1,y0,y1,z0,z1 */
} adeco_parameters;
void adeco_parameters_init(adeco_parameters* p);
/* parse_parameter_file (main.cpp:83-183): "key = value" lines, the key being the text before '='
 * less its last character, matched as a substring in the reference's order; flags are on when the
 * value contains "true"; numbers are read as by istream extraction; at most 10000 lines.
 * ADECO_EINVAL with the reference's message on an unreadable file, a bad number or a
 * region_of_interest without 6 values; ADECO_EOVERFLOW for a string longer than its field. */
adeco_status adeco_parse_parameter_file(const char* filename, adeco_parameters* out);
/* AdSortedByIdsPointDataReader::load_data (AdPointDataReaders.h:770-911) on one rank: raw float32
 * x,y,z per particle, in id order, n^3 particles of a cubic lattice (n = rint(cbrt(count))).
 * *num_particles and dims3 receive the count and lattice; when x, y and z are given (capacity >=
 * count) they receive the coordinates (SoA, ColumnMajorOrder).  stride must be >= 1 and is
 * otherwise ignored, as in the reference. */
adeco_status adeco_read_sorted_by_id(const char* filename, int stride, float* x, float* y, float* z, int64_t capacity,
                                     int64_t* num_particles, int* dims3);
/* DARK_SKY particles (AdDarkSkyDataReader, AdPointDataReaders.h:361-767) on one rank, straight
 * into a resident lattice: a text header ending with a "# SDF-EOH" line, "int sha1_chunks = N;"
 * records of 24 bytes after it, then 32-byte particle records {float x,y,z,vx,vy,vz; int64 id},
 * the id's low 48 bits the Morton code of the lattice site (x in bits 0,3,6,..., y in 1,4,...,
 * z in 2,5,...).  n^3 records; with stride s (dividing n) the sites whose indices are multiples of
 * s, at site / s.  The records stream to the device, which decodes the ids and stores each
 * position at its site; an id outside the lattice, a repeated site or a missing one is
 * ADECO_EINVAL (the reference's checks and assertions). */
adeco_status adeco_lattice_create_dark_sky(int device_id, const char* filename, int stride, adeco_lattice* out);
/* Host-only: the header and size checks of a DARK_SKY file (get_header_size,
 * AdPointDataReaders.h:419-466, and the record count checks): header bytes, particle records and
 * particles per axis before any stride. */
adeco_status adeco_dark_sky_info(const char* filename, int64_t* header_bytes, int64_t* num_particles, int* dims);
typedef struct {
    int device;
    int mode;           /* ADECO_MODE_* (default batched) */
    int precision;      /* ADECO_PREC_* (default FP64) */
    int decomposition;  /* ADECO_DECOMP_* (default REF_EFFECTIVE, the reference's tets) */
    double tet_mass;    /* default 1 (AdDepositTets2Grid.h:1007-1053) */
} adeco_workflow_options;
void adeco_workflow_options_init(adeco_workflow_options* o);
typedef struct {
    int64_t num_particles;
    int particle_dims[3];
    float global_bbox[6];  /* before the periodic correction: the octree's root box */
    int64_t num_nodes, num_leaves;
    double total_mass;     /* sum of the leaf files' masses in the leaf processing order */
    double load_s, octree_s, deposit_s, write_s; /* wall time of the steps */
    double kernel_ms;      /* the leaf deposit's GPU time (HIP events around its kernels) */
} adeco_workflow_result;
/* deposit_workflow (AdDepositWorkFlow.h:60-237) on one rank and one GPU: read the particles
 * (id_type SORTED_BY_ID or DARK_SKY, chosen as get_reader does, AdPointDataReaders.h:914-930)
 * into a resident lattice, global bbox, periodic correction when periodic_boundaries is set,
 * octree over the region of interest (the global bbox unless use_regions_of_interest), leaf
 * deposit at linear_patch_resolution and one write_patch file per leaf into output_directory. */
adeco_status adeco_deposit_workflow(const adeco_parameters* p, const adeco_workflow_options* opt,
                                    adeco_workflow_result* res);

/* ---- region-of-interest density requests over the resident lattice ----
 * The reference has no request server: these entry points apply its single-rank leaf deposit
 * (AdDynamicLoadBalancing.h:983-1323, deposit_tets per leaf) to caller-given boxes.  A request's
 * result is the FP32 density (mass / cell volume, as AdOctreeGridWriter.h:67-144) of every lattice
 * tet whose float AABB strictly overlaps the box, on a dims^3 grid over the box (dims <= 1023). */
typedef struct adeco_server_s* adeco_server;
typedef struct {
    int decomposition;    /* ADECO_DECOMP_* */
    int precision;        /* ADECO_PREC_* */
    double tet_mass;      /* mass of one tet (> 0) */
    int max_inflight;     /* requests resident on the device at once, 1..64 */
    int pad_;
    int64_t step_tiles;   /* tiles of 64 (tet, candidate) pairs per step launch (0: 131072; at most
                           * 4194304: two pinned entry buffers of 8 bytes per tile) */
    uint64_t memory_budget; /* device bytes for request buffers (0: no limit but the device's) */
} adeco_serve_config;
typedef struct {
    adeco_status status;
    int dims;
    int64_t n_tets;       /* tets selected for the box */
    int64_t n_candidates; /* (tet, candidate) pairs of the batched deposit (a candidate is a covered
                           * cell, or for FP64 a run of up to 8 covered cells along the tet's
                           * longest axis) */
    double t_submit, t_admit, t_done; /* seconds on the server's steady clock */
} adeco_request_info;
typedef struct {
    uint64_t requests_done;
    uint64_t steps;       /* step launches */
    int64_t tiles;        /* tiles deposited */
    int64_t tets;         /* tets of the completed requests */
    double busy_s;        /* time the worker spent in steps */
    uint64_t device_bytes; /* request buffers held */
    int64_t pending, active;
    double step_kernel_s; /* device time of the finished step launches (k_deposit_step), from HIP
                           * events recorded around each launch on the step stream */
} adeco_serve_stats;
/* Starts a server (a worker and a copier thread) on the lattice's device.  Requests are admitted
 * in arrival order while a slot is free and the memory budget allows; each step launch deposits an
 * equal share of its tiles from every admitted request (continuous batching), and steps are queued
 * without waiting for the previous one.  A request that does not fit the budget waits while other
 * requests hold memory and fails with ADECO_EOVERFLOW only when it cannot fit alone. */
adeco_status adeco_serve_create(adeco_lattice lat, const adeco_serve_config* cfg, adeco_server* out);
/* Queues a request; density_out (dims^3 floats, i fastest) must stay valid until it completes. */
adeco_status adeco_serve_submit(adeco_server srv, const double* roi6, int dims, float* density_out,
                                uint64_t* request_id);
/* Blocks until the request completes (timeout_s < 0: no limit) and releases its record. */
adeco_status adeco_serve_wait(adeco_server srv, uint64_t request_id, double timeout_s, adeco_request_info* info);
adeco_status adeco_serve_stats_get(adeco_server srv, adeco_serve_stats* out);
/* Finishes every queued request, then stops the threads and frees the device buffers. */
adeco_status adeco_serve_destroy(adeco_server srv);
/* The serial baseline: the same selection and output for one box, deposited by the plain
 * one-tet-per-thread kernel, synchronously.  n_tets_out (nullable) receives the selected tets. */
adeco_status adeco_roi_deposit_plain(adeco_lattice lat, int decomposition, double tet_mass, const double* roi6,
                                     int dims, float* density_out, int64_t* n_tets_out);

/* ---- self-test hooks: the shared host/device geometry evaluated on the host ----
 * These exist so the CPU test suite can check the clip and classification code that the kernels
 * compile; no deposit entry point ever calls them. */
double adeco_selftest_clip_vol6_f64(const double* tet12);
float adeco_selftest_clip_vol6_f32(const float* tet12);
int adeco_selftest_classify_f64(const double* tet12);
double adeco_selftest_beyond_fraction_f64(const double* d4);
void adeco_selftest_decomposition(int decomposition, int32_t* corners24);
/* The batched kernel's column-segment clip on the host: tet12 in the segment frame (x, y across
 * the column, cells [c, c+1] along z for c < ls, 1 <= ls <= 8); vol6_out[c] = 6 x the volume in
 * cell c summed by the signed (use_signed != 0) or the exact tree, sab_out[c] = the magnitudes it
 * was summed from. */
void adeco_selftest_column_f64(const double* tet12, int ls, int use_signed, double* vol6_out, double* sab_out);
/* The DARK_SKY id decoding the reader's kernel uses: lattice site (x, y, z) of an id. */
void adeco_selftest_dark_sky_site(long long id, int* site3);
/* The kernels' per-tet FP64 deposit routine run on the host over a tet soup (float [n][4][3]),
 * accumulating into grid (double, dims^3). */
adeco_status adeco_selftest_deposit_soup_f64(const float* tets, int64_t n_tets, const float* bbox6, int dims,
                                             double tet_mass, double* grid);

#ifdef __cplusplus
}
#endif

#endif /* ADECO_H */
