// adeco_deposit_wrapper.hpp — header-only C++ face of libadeco_amd.so for C++ host code.
//
// AdaptiveMassDeposit::DepositWrapper with the reference's typedefs, static members and the exact
// deposit_tets signature (AdDeposit2Grid.h:59-93), forwarding to the C ABI (adeco.h).  Errors are
// thrown as adeco::DepositError (a std::runtime_error carrying the adeco_status), where the
// reference throws AdRuntimeException (AdException.h:87-95); callers that catch std::exception,
// as the reference's workflow does (AdDepositWorkFlow.h:232-235), need no change.
//
// Inside a reference source tree use adeco_amd/cpp/AdDeposit2Grid_adeco.cpp instead: it
// implements the reference's own class from its own header and throws its AdRuntimeException.
#ifndef ADECO_DEPOSIT_WRAPPER_HPP
#define ADECO_DEPOSIT_WRAPPER_HPP

#include <cstddef>
#include <stdexcept>
#include <string>

#include "adeco.h"

namespace adeco {

class DepositError : public std::runtime_error {
public:
    DepositError(adeco_status st, const std::string& what) : std::runtime_error(what), status_(st) {}
    adeco_status status() const { return status_; }

private:
    adeco_status status_;
};

inline void check(adeco_status st, const char* where) {
    if (st != ADECO_OK) throw DepositError(st, std::string("ERROR: ") + where + ": " + adeco_last_error());
}

}  // namespace adeco

namespace AdaptiveMassDeposit {

class DepositWrapper {
public:
    typedef float h_grid_mass_type;
    typedef float vertex_type;
    typedef float mass_type;

    static inline int get_cuda_grid_size() { return adeco_get_grid_size(); }
    static inline int get_cuda_block_size() { return adeco_get_block_size(); }

    // tet_offset / num_tets_2_process count meta-patches (AdDeposit2Grid.cpp:69-70); data is
    // accumulated into (+=).
    static void deposit_tets(const bool use_specific_gpu, const size_t device_id_to_use, const size_t tet_offset,
                             const size_t num_tets_2_process, const float* positions, const size_t positions_size,
                             const int meta_cells_per_patch, const float bbox[6], const int dims,
                             h_grid_mass_type* data) {
        adeco::check(adeco_deposit_patches(use_specific_gpu ? 1 : 0, device_id_to_use, tet_offset, num_tets_2_process,
                                           positions, positions_size, meta_cells_per_patch, bbox, dims, data),
                     "deposit_tets()");
    }
};

}  // namespace AdaptiveMassDeposit

#endif  // ADECO_DEPOSIT_WRAPPER_HPP
