// The program around the deposit on one rank (C ABI part 7): the parameter file
// (parse_parameter_file, main.cpp:83-183), the SORTED_BY_ID particle reader
// (AdSortedByIdsPointDataReader, AdPointDataReaders.h:770-911, with the checks of
// AdPointDataReader::load_data, :231-276; DARK_SKY files go to darksky.hip) and the workflow (deposit_workflow,
// AdDepositWorkFlow.h:60-237) over the resident lattice: read, global bbox, periodic correction,
// octree, leaf deposit, leaf files.  Host code; the compute runs in the HIP kernels behind the
// other entry points.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/adeco.h"
#include "lattice_impl.hpp"
#include "runtime.hpp"

using namespace adeco;

namespace {

// istream extraction: leading white space skipped, trailing text ignored, failure reported
// (the reference's str2number, AdUtils.h:199-208).
template <class T>
bool to_number(const std::string& s, T& out) {
    std::istringstream in(s);
    T v;
    if (!(in >> v)) return false;
    out = v;
    return true;
}

adeco_status bad_number(const std::string& s) {
    return fail(ADECO_EINVAL, "ERROR: AvGeneralUtils::str2number() failed for string: " + s + "\n");
}

adeco_status set_string(char* dst, size_t cap, const std::string& v, const char* what) {
    if (v.size() >= cap)
        return fail(ADECO_EOVERFLOW, std::string(what) + " is too long (at most " + std::to_string(cap - 1) + " bytes)");
    std::memcpy(dst, v.data(), v.size());
    dst[v.size()] = 0;
    return ADECO_OK;
}

// One "key = value" line.  The key is the text before the '=' less its last character (the
// format puts a blank there); without '=' both key and value are the whole line.  Keys are
// matched as substrings, in this order, so "use_region_of_interest" is tested before
// "region_of_interest".
adeco_status apply_line(const std::string& line, adeco_parameters& p) {
    const size_t eq = line.find('=');
    const std::string key = line.substr(0, eq - 1);  // eq == npos or 0: the whole line
    const std::string value = line.substr(eq + 1);   // eq == npos: the whole line
    auto has = [&](const char* k) { return key.find(k) != std::string::npos; };
    const bool yes = value.find("true") != std::string::npos;
    if (has("input_file")) return set_string(p.input_file, sizeof(p.input_file), value, "input_file");
    if (has("output_directory")) return set_string(p.output_directory, sizeof(p.output_directory), value, "output_directory");
    if (has("id_type") || has("file_format")) return set_string(p.id_type, sizeof(p.id_type), value, "id_type");
    if (has("use_region_of_interest")) {
        p.use_regions_of_interest = yes;
        return ADECO_OK;
    }
    if (has("region_of_interest")) {
        std::istringstream words(value);
        std::vector<std::string> w;
        for (std::string t; words >> t;) w.push_back(t);
        if (w.size() != 6) return fail(ADECO_EINVAL, "ERROR: parse_parameter_file(): 'regions_of_interest' needs 6 parameters");
        double v[6];
        for (int i = 0; i < 6; ++i)
            if (!to_number(w[size_t(i)], v[i])) return bad_number(w[size_t(i)]);
        for (int i = 0; i < 6; ++i) p.region_of_interest[i] = float(v[i]);  // x0 x1 y0 y1 z0 z1
        return ADECO_OK;
    }
    int* ints[] = {&p.max_refinement_level, &p.stride, nullptr, &p.linear_patch_resolution, &p.num_pthreads,
                   &p.num_mpi_ranks_per_node};
    const char* int_keys[] = {"max_refinement_level", "stride", "max_tets_per_octree_node", "linear_patch_resolution",
                              "num_pthreads", "num_mpi_ranks_per_node"};
    for (int i = 0; i < 6; ++i) {
        if (!has(int_keys[i])) continue;
        if (!ints[i]) {
            unsigned long long u;
            if (!to_number(value, u)) return bad_number(value);
            p.max_tets_per_octree_node = uint64_t(u);
        } else if (!to_number(value, *ints[i])) {
            return bad_number(value);
        }
        return ADECO_OK;
    }
    if (has("restart_path")) return set_string(p.restart_path, sizeof(p.restart_path), value, "restart_path");
    if (has("periodic_boundaries")) p.periodic_boundaries = yes;
    return ADECO_OK;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

void adeco_parameters_init(adeco_parameters* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->max_refinement_level = 5;
    p->stride = 1;
    p->linear_patch_resolution = 64;
    p->num_pthreads = 1;
    p->num_mpi_ranks_per_node = 1;
    p->periodic_boundaries = 1;
    p->max_tets_per_octree_node = ~uint64_t(0);
    const float roi[6] = {0.f, 1.f, 0.f, 1.f, 0.f, 1.f};
    std::memcpy(p->region_of_interest, roi, sizeof(roi));
}

adeco_status adeco_parse_parameter_file(const char* filename, adeco_parameters* out) {
    if (!filename || !out) return fail(ADECO_EINVAL, "parse_parameter_file(): null argument");
    std::ifstream in(filename);
    if (!in.is_open()) return fail(ADECO_EINVAL, std::string("failed to open parameter file: ") + filename);
    adeco_parameters p;
    adeco_parameters_init(&p);
    constexpr size_t kMaxLines = 10000;
    size_t lines = 0;
    while (in.good() && lines < kMaxLines) {
        ++lines;
        std::string line;
        std::getline(in, line);
        const adeco_status st = apply_line(line, p);
        if (st) return st;
    }
    if (lines >= kMaxLines) return fail(ADECO_EINVAL, std::string("ERROR: Failed to parse parameter file: ") + filename);
    *out = p;
    return ADECO_OK;
}

adeco_status adeco_read_sorted_by_id(const char* filename, int stride, float* x, float* y, float* z, int64_t capacity,
                                     int64_t* num_particles, int* dims3) {
    if (!filename || !*filename) return fail(ADECO_EINVAL, "ERROR: AdPointDataReader::load_data(): missing filename.");
    if (stride < 1) return fail(ADECO_EINVAL, "ERROR: AdPointDataReader::load_data(): invalid 'stride' parameters.");
    FILE* f = std::fopen(filename, "rb");
    if (!f)
        return fail(ADECO_EINVAL, std::string("ERROR: AdInitialization_IDS::load_particle_data(): failed to open file: ") +
                                      filename);
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    if (std::fseek(f, 0, SEEK_END) != 0) return fail(ADECO_EIO, "read_sorted_by_id(): cannot size the file");
    const long long bytes = std::ftell(f);
    if (bytes < 0 || std::fseek(f, 0, SEEK_SET) != 0) return fail(ADECO_EIO, "read_sorted_by_id(): cannot size the file");
    const uint64_t per = 3 * sizeof(float);
    if (uint64_t(bytes) % per != 0)
        return fail(ADECO_EINVAL, "ERROR: AdInitialization_IDS::load_particle_data(): inconsistent file size ");
    const uint64_t n = uint64_t(bytes) / per;
    const int d = int(std::rint(std::cbrt(double(n))));
    if (uint64_t(d) * uint64_t(d) * uint64_t(d) != n)
        return fail(ADECO_EINVAL, "ERROR: AdInitialization_IDS::load_particle_data(): lagrangian grid not cubical ");
    if (num_particles) *num_particles = int64_t(n);
    if (dims3) dims3[0] = dims3[1] = dims3[2] = d;
    if (!x || !y || !z) return ADECO_OK;
    if (capacity < int64_t(n)) return fail(ADECO_EINVAL, "read_sorted_by_id(): capacity below the particle count");
    // ids are the file order: particle p = i + d*(j + d*k) (ColumnMajorOrder), split into SoA
    std::vector<float> buf(size_t(3) << 20);
    for (uint64_t done = 0; done < n;) {
        const uint64_t take = std::min<uint64_t>(n - done, buf.size() / 3);
        if (std::fread(buf.data(), per, size_t(take), f) != size_t(take))
            return fail(ADECO_EIO, std::string("read_sorted_by_id(): short read from ") + filename);
        for (uint64_t i = 0; i < take; ++i) {
            x[done + i] = buf[3 * i];
            y[done + i] = buf[3 * i + 1];
            z[done + i] = buf[3 * i + 2];
        }
        done += take;
    }
    return ADECO_OK;
}

void adeco_workflow_options_init(adeco_workflow_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->mode = ADECO_MODE_BATCHED;
    o->precision = ADECO_PREC_FP64;
    o->decomposition = ADECO_DECOMP_REF_EFFECTIVE;
    o->tet_mass = 1.0;  // resample_meta_patch deposits every tet with mass 1 (AdDepositTets2Grid.h:1007-1053)
}

adeco_status adeco_deposit_workflow(const adeco_parameters* p, const adeco_workflow_options* opt,
                                    adeco_workflow_result* res) {
    if (!p || !opt) return fail(ADECO_EINVAL, "deposit_workflow(): null argument");
    adeco_workflow_result r;
    std::memset(&r, 0, sizeof(r));
    const std::string format(p->id_type);
    const bool sorted = format.find("SORTED_BY_ID") != std::string::npos;
    if (!sorted && format.find("DARK_SKY") == std::string::npos)
        return fail(ADECO_EINVAL, "ERROR: id_type " + format + " not supported.");
    // load (AdDepositWorkFlow.h:77-97)
    auto t0 = std::chrono::steady_clock::now();
    adeco_lattice L = nullptr;
    adeco_status st = ADECO_OK;
    TraceRange phase("adeco:read_particles");
    if (sorted) {
        int64_t n = 0;
        int d[3];
        if ((st = adeco_read_sorted_by_id(p->input_file, p->stride, nullptr, nullptr, nullptr, 0, &n, d))) return st;
        const size_t count = static_cast<size_t>(n);
        std::vector<float> x(count), y(count), z(count);
        if ((st = adeco_read_sorted_by_id(p->input_file, p->stride, x.data(), y.data(), z.data(), n, &n, d))) return st;
        if ((st = adeco_lattice_create(opt->device, x.data(), y.data(), z.data(), d[0], d[1], d[2], &L))) return st;
    } else if ((st = adeco_lattice_create_dark_sky(opt->device, p->input_file, p->stride, &L))) {
        return st;
    }
    struct LatticeGuard {
        adeco_lattice L;
        ~LatticeGuard() { adeco_lattice_destroy(L); }
    } lg{L};
    for (int a = 0; a < 3; ++a) r.particle_dims[a] = L->np[a];
    r.num_particles = L->n_vertices;
    r.load_s = seconds_since(t0);
    phase.next("adeco:bbox_periodic_octree");
    // global bbox, periodic correction, region of interest (AdDepositWorkFlow.h:101-115)
    t0 = std::chrono::steady_clock::now();
    if ((st = adeco_lattice_bbox(L, r.global_bbox))) return st;
    if (p->periodic_boundaries && (st = adeco_lattice_periodic_unwrap(L, r.global_bbox, nullptr))) return st;
    const float* roi = p->use_regions_of_interest ? p->region_of_interest : r.global_bbox;
    // octree (AdDepositWorkFlow.h:160-178; max_tets = min(num_tets, parameter) on one rank)
    adeco_octree T = nullptr;
    if ((st = adeco_octree_build(L, roi, p->max_tets_per_octree_node, p->max_refinement_level, &T))) return st;
    struct TreeGuard {
        adeco_octree T;
        ~TreeGuard() { adeco_octree_destroy(T); }
    } tg{T};
    r.num_nodes = adeco_octree_num_nodes(T);
    r.num_leaves = adeco_octree_num_leaves(T);
    r.octree_s = seconds_since(t0);
    // leaf deposit and leaf files (deposit_leaf_nodes, AdDynamicLoadBalancing.h:983-1323)
    phase.next("adeco:leaf_deposit");
    t0 = std::chrono::steady_clock::now();
    if ((st = adeco_deposit_leaves(T, p->linear_patch_resolution, opt->decomposition, opt->tet_mass, opt->mode,
                                   opt->precision)))
        return st;
    r.kernel_ms = adeco_last_kernel_ms();
    r.deposit_s = seconds_since(t0);
    phase.next("adeco:leaf_files");
    t0 = std::chrono::steady_clock::now();
    if ((st = adeco_write_leaves(T, p->output_directory, nullptr, &r.total_mass))) return st;
    r.write_s = seconds_since(t0);
    if (res) *res = r;
    return ADECO_OK;
}

}  // extern "C"
