// kcore.hpp -- the k-core peeling of kcore.hip, shared by the engine and the
// application loop in apps.cpp.
#pragma once
#include <stdint.h>

#include "../../include/gpr.h"

namespace gpr {

struct KcStats {
    uint32_t kmax;     // largest core number
    uint32_t levels;   // distinct core values peeled
    uint32_t rounds;   // count-down rounds over all levels
    uint64_t edges;    // out-edges of peeled vertices examined
};
// core (u32) = the core number of every vertex of the symmetric graph g.
int kcore_peel(gpr_graph *g, gpr_prop *core, KcStats *st);

} // namespace gpr
