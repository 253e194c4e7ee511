// Drop-in implementation of the reference's DepositWrapper::deposit_tets on libadeco_amd.so.
//
// Compile this file in place of the reference's AdDeposit2Grid.cpp (CPU build) or
// AdDeposit2Grid.cu (CUDA build) and link -ladeco_amd: the reference's own header
// (AdDeposit2Grid.h:59-93) declares the class, so every caller -- AdWorkerThread::run_
// (AdDynamicLoadBalancing.h:200-205) -- is unchanged.  Failures throw the reference's
// AdRuntimeException (AdException.h:87-95) where it throws, instead of exit(EXIT_FAILURE).
#include <string>

#include "AdDeposit2Grid.h"
#include "AdException.h"
#include "adeco.h"

void AdaptiveMassDeposit::DepositWrapper::deposit_tets(const bool use_specific_gpu, const size_t device_id_to_use,
                                                       const size_t tet_offset, const size_t num_tets_2_process,
                                                       const float* positions, const size_t positions_size,
                                                       const int meta_cells_per_patch, const float bbox[6],
                                                       const int dims, h_grid_mass_type* data) {
    const adeco_status st =
        adeco_deposit_patches(use_specific_gpu ? 1 : 0, device_id_to_use, tet_offset, num_tets_2_process, positions,
                              positions_size, meta_cells_per_patch, bbox, dims, data);
    if (st != ADECO_OK)
        throw AdaptiveMassDeposit::AdRuntimeException(std::string("ERROR: deposit_tets(): ") + adeco_last_error());
}
