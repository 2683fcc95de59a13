// apr.hpp -- the delta PageRank (APR_vary) pieces of apr.hip, shared by the
// engine and by the application loop in apps.cpp.
#pragma once
#include <stdint.h>

#include "../../include/gpr.h"

namespace gpr {

// Delta PageRank (apr.hip): totals of one iteration's vertex maps.
struct AprTotals {
    double l1;        // sum |delta|
    double sum_x;     // sum |x|
    uint64_t nactv;   // vertices of the new frontier
    uint64_t nacte;   // their out-edges
};
int apr_check_fmt(int fmt);
// contrib = enc(delta / (factor*deg)), sources outside `mask` (when given and
// not all-true) storing 0.
int apr_contrib(gpr_graph *g, int fmt, const gpr_prop *delta, gpr_prop *contrib, float factor, gpr_frontier *mask);
// y = 0, then y[d] += the decoded contribution of every out-edge of F.
int apr_push(gpr_graph *g, int fmt, gpr_frontier *F, const gpr_prop *delta, float factor, gpr_prop *y);
// The vertex maps after the edgemap; `out` receives the active bitmap and
// tot the totals (synchronises).
int apr_tail(gpr_graph *g, bool first, const gpr_prop *y, float factor, float eps2, gpr_prop *x, gpr_prop *delta,
             gpr_frontier *out, AprTotals *tot);
// out (f64) = x * s.
int apr_scale(gpr_graph *g, const gpr_prop *x, double s, gpr_prop *out);

} // namespace gpr
