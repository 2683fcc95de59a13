// policy.hip -- per-iteration choice of direction and frontier
// representation.  The fallback is the reference's threshold rule
// (default_threshold::is_sparse, api.h:150-153).
#include "internal.hpp"

namespace gpr {

int policy_decide(gpr_graph *g, uint32_t nactv, uint64_t nacte, int *dir, int *out_kind) {
    const bool sparse = (uint64_t)nactv + nacte <= g->m / 20;
    *dir = sparse ? GPR_DIR_PUSH : GPR_DIR_PULL;
    if (*out_kind == 0) *out_kind = sparse ? GPR_FRONTIER_LIST : GPR_FRONTIER_BITMAP;
    return GPR_OK;
}

} // namespace gpr
