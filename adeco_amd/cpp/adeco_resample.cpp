// adeco_resample: the reference's command-line program (main.cpp:187-249, `dm_resampling_cpu
// <parameter file>`) on one GPU: parse the parameter file, run deposit_workflow on one rank
// (AdDepositWorkFlow.h:60-237) and write one density file per octree leaf into
// output_directory, in the reference's file format.
//
//   adeco_resample PARAMETER_FILE [--device N] [--plain] [--mixed] [--reference-geometry]
//
// --plain runs the one-tet-per-thread kernel instead of the batched one, --mixed the
// mixed-precision clip, --reference-geometry the reference's float leaf boxes and 1/dx cell
// scaling instead of the exact dyadic leaf boxes.  Exit status 0 on success, 1 on any error (the
// reference returns 0 when the workflow itself fails, AdDepositWorkFlow.h:229-233).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "adeco.h"

namespace {

void print_parameters(const adeco_parameters& p) {
    std::printf("INFO: AMD_Parameters(): PARAMETERS INFO:\n");
    std::printf("    input_file               = %s\n", p.input_file);
    std::printf("    output_directory         = %s\n", p.output_directory);
    std::printf("    id_type                  = %s\n", p.id_type);
    std::printf("    use_region_of_interest   = %s\n", p.use_regions_of_interest ? "true" : "false");
    std::printf("    region_of_interest       = %g %g %g %g %g %g\n", p.region_of_interest[0], p.region_of_interest[1],
                p.region_of_interest[2], p.region_of_interest[3], p.region_of_interest[4], p.region_of_interest[5]);
    std::printf("    max_refinement_level     = %d\n", p.max_refinement_level);
    std::printf("    stride                   = %d\n", p.stride);
    std::printf("    max_tets_per_octree_node = %llu\n", static_cast<unsigned long long>(p.max_tets_per_octree_node));
    std::printf("    linear_patch_resolution  = %d\n", p.linear_patch_resolution);
    std::printf("    num_pthreads             = %d\n", p.num_pthreads);
    std::printf("    num_mpi_ranks_per_node   = %d\n", p.num_mpi_ranks_per_node);
    std::printf("    restart_path             = %s\n", p.restart_path);
    std::printf("    periodic_boundaries      = %d\n", p.periodic_boundaries);
    std::printf("------------------\n");
}

int error(const std::string& what) {
    std::fprintf(stderr, "ERROR: (MPI rank = 0): Caught exception: \n%s.\n", what.c_str());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return error("ERROR: main(): invalid number of arguments. ");
    adeco_workflow_options opt;
    adeco_workflow_options_init(&opt);
    bool reference_geometry = false;
    for (int i = 2; i < argc; ++i) {
        const std::string a(argv[i]);
        if (a == "--plain")
            opt.mode = ADECO_MODE_PLAIN;
        else if (a == "--mixed")
            opt.precision = ADECO_PREC_MIXED;
        else if (a == "--reference-geometry")
            reference_geometry = true;
        else if (a == "--device" && i + 1 < argc)
            opt.device = std::atoi(argv[++i]);
        else
            return error("ERROR: main(): unknown option " + a);
    }
    if (opt.mode == ADECO_MODE_PLAIN && opt.precision != ADECO_PREC_FP64)
        return error("ERROR: main(): --plain runs FP64 only");
    adeco_parameters p;
    if (adeco_parse_parameter_file(argv[1], &p) != ADECO_OK) return error(adeco_last_error());
    print_parameters(p);
    if (p.stride > 1 && std::strstr(p.id_type, "SORTED_BY_ID"))
        std::printf("WARNING: AdSortedByIdsPointDataReader::load__data(): 'stride' parameter currently not supported - "
                    "will ignore it ... \n");
    if (adeco_set_cell_scaling(reference_geometry ? ADECO_SCALING_REFERENCE : ADECO_SCALING_EXACT) != ADECO_OK)
        return error(adeco_last_error());
    adeco_workflow_result r;
    if (adeco_deposit_workflow(&p, &opt, &r) != ADECO_OK) return error(adeco_last_error());
    std::printf("INFO: deposit_workflow(): particle_dims       = %d %d %d (%lld particles)\n", r.particle_dims[0],
                r.particle_dims[1], r.particle_dims[2], static_cast<long long>(r.num_particles));
    std::printf("INFO: deposit_workflow(): global_bbox         = %.9g %.9g %.9g %.9g %.9g %.9g\n", r.global_bbox[0],
                r.global_bbox[1], r.global_bbox[2], r.global_bbox[3], r.global_bbox[4], r.global_bbox[5]);
    std::printf("INFO: deposit_workflow(): octree nodes        = %lld, leaves = %lld\n",
                static_cast<long long>(r.num_nodes), static_cast<long long>(r.num_leaves));
    std::printf("INFO: RESULT: deposit_leaf_nodes: total mass                  =  %.17g\n", r.total_mass);
    std::printf("INFO: RESULTS: time for loading data          == %g\n", r.load_s);
    std::printf("INFO: RESULTS: time for constructing octree   == %g\n", r.octree_s);
    std::printf("INFO: RESULTS: time for mass deposit          == %g (gpu %g ms)\n", r.deposit_s, r.kernel_ms);
    std::printf("INFO: RESULTS: time for writing leaf files    == %g\n", r.write_s);
    std::printf("INFO: (MPI rank = 0): finished successfully. \n\n");
    return 0;
}
