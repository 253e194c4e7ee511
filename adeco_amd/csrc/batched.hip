// Batched persistent deposit over all octree leaves (placeholder: routes to per-leaf launches
// until the load-balanced kernel lands).
#include <hip/hip_runtime.h>

#include "../../include/adeco.h"
#include "kernels.hpp"
#include "octree_impl.hpp"
#include "runtime.hpp"

namespace adeco {

adeco_status deposit_leaves_batched(adeco_octree_s* T, int dims, int decomposition, double tet_mass, int precision) {
    (void)T; (void)dims; (void)decomposition; (void)tet_mass; (void)precision;
    return fail(ADECO_EINTERNAL, "batched deposit not available yet");
}

}  // namespace adeco
