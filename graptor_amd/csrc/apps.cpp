// apps.cpp -- the callers of the hot path, written against gpr::api the way
// the reference benchmarks are written against Graptor's api:: templates.
//
//   BFSLVLv   bench/BFSLVLv.C:110-254  (run loop :154-251)
//   BFSv      bench/BFSv.C:88-203      (run loop :118-199)
//   CCv       bench/CCv.C:221-588      (run :221-471, emap_step :520-588)
//   PRv       bench/PRv.C:115-290      (iteration :201-280)
//   PR_vary   bench/PR_vary.C:247-757  (run_prec :247-447, staged run :652-757)
//   APR_vary  bench/APR_vary.C:378-1074 (run_prec :378-640, run :900-1074)
//   KC_bucket bench/KC_bucket.C:104-330  (run; validate :332-460)
//
// By default each benchmark runs the reference's DEFERRED_UPDATE=0
// configuration of the record operator (api::record(output, api::reduction,
// api::strong), BFSLVLv.C:173 and CCv.C:543-549): a vertex joins the next
// frontier when its relax changed it.  With gpr_run_opts::deferred_update
// set they run the reference's default DEFERRED_UPDATE=1 instead: BFSLVLv
// records the method level[d] != prev_level[d] and maintains prev_level
// over each output (BFSLVLv.C:166-172, 218-220); CCv records
// reduction_or_method with IDs[d] != prevIDs[d] and copies prevIDs = IDs
// after each step (CCv.C:536-549, 578-580).  For BFS levels and min-label
// CC both configurations activate the same vertices.
#include "../../include/gpr/api.hpp"
#include "apr.hpp"
#include "kcore.hpp"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <vector>

struct gpr_graph;
namespace gpr {
bool threshold_sparse(float pct, uint64_t nactv, uint64_t nacte, uint64_t m);
std::vector<gpr_iter_stats> &trace_of(gpr_graph *g);
uint32_t graph_max_degree(const gpr_graph *g);
int ids_download(gpr_graph *g, const gpr_prop *p, uint32_t *host);
void trace_clear(gpr_graph *g);
double stream_time_ms(gpr_graph *g, bool start);
}

using namespace gpr;
namespace A = gpr::api;

namespace {

gpr_run_opts defaults(const gpr_run_opts *o) {
    // threshold_pct 0: each benchmark's own rule (BFSLVLv and BFSv:
    // default_threshold, integer m/20; CCv: frac_threshold(5))
    gpr_run_opts d = {};
    d.direction = GPR_DIR_AUTO;
    d.threshold_pct = 0.f;
    d.max_iter = 100;
    d.tol = 1e-7;
    if (o) {
        d = *o;
        if (!(d.threshold_pct > 0.f)) d.threshold_pct = 0.f;
        if (d.max_iter <= 0) d.max_iter = 100;
    }
    return d;
}

void record_step(gpr_graph *g, const gpr_run_opts &o, const gpr_iter_stats &st, gpr_run_info *info) {
    if (o.trace) trace_of(g).push_back(st);
    info->iters++;
    if (st.direction == GPR_DIR_PUSH) info->push_iters++;
}

template <class F> int guarded(F &&f) {
    try {
        f();
        return GPR_OK;
    } catch (const gpr::error &e) {
        return e.code;
    } catch (const std::exception &e) {
        check(GPR_E_INVAL, e.what());
        return GPR_E_INVAL;
    }
}

// ---- BFSLVLv -----------------------------------------------------------------
struct BFSLVLv {
    Graph GA;
    gpr_run_opts o;
    A::vertexprop<uint32_t> a_level, a_prev_level;
    int iter = 0;
    BFSLVLv(gpr_graph *g, const gpr_run_opts &opts) : GA(g), o(opts), a_level(GA, "current level") {
        if (o.deferred_update) a_prev_level = A::vertexprop<uint32_t>(GA, "previous level");
    }
    ~BFSLVLv() {
        a_level.del();
        a_prev_level.del();
    }

    void run(uint32_t start, gpr_run_info *info) {
        const uint32_t n = GA.numVertices();
        const bool deferred = o.deferred_update != 0;
        // BFSLVLv.C:117-129: ilvl = n+1 everywhere, 0 at the start vertex
        // (the previous level too under DEFERRED_UPDATE)
        A::make_lazy_executor(GA).vertex_map_fill(a_level, n + 1).materialize();
        A::frontier F = A::frontier::create(GA, start);
        A::make_lazy_executor(GA).vertex_map_fill(F, a_level, 0).materialize();
        if (deferred) {
            A::make_lazy_executor(GA).vertex_map_fill(a_prev_level, n + 1).materialize();
            A::make_lazy_executor(GA).vertex_map_fill(F, a_prev_level, 0).materialize();
        }
        check(gpr_policy_reset(GA.get()), "policy reset");
        info->edges += F.nActiveEdges();
        iter = 0;
        // FUSION (BFSLVLv.C:28-40, 143-152, 246-250): off by default; 1 = on
        // for low-degree graphs (cgraph.h:45-49) and after a dense step;
        // 2 = always
        const int fmode = o.fusion == 0 ? -1 : o.fusion;
        bool enable_fusion = fmode == 2 || (fmode == 1 && (uint64_t)graph_max_degree(GA.get()) * (128 * 1024) < n);
        // 3: small-level fusion, decided per step by the engine
        const int small = fmode == 3 ? GPR_FUSION_SMALL : -1;
        bool fused = false;
        while (!F.isEmpty()) {
            A::frontier output;
            // DEFERRED_UPDATE: api::record(output, level[d] != prev_level[d],
            // strong), BFSLVLv.C:166-172; else the reduction, :173
            const A::record_arg rec = deferred ? A::record(output, A::differs(a_level, a_prev_level), A::strong)
                                               : A::record(output, A::reduction, A::strong);
            auto lazy = A::edgemap(GA,
                                   rec,
                                   A::filter(A::weak, A::src, F),
                                   A::filter(A::weak, A::dst, A::gt(a_level, (uint32_t)iter + 1)),
                                   A::relax(A::min_plus1(a_level)),
                                   A::config(o.direction, o.threshold_pct, o.out_kind),
                                   small >= 0 ? A::fusion_mode(small) : A::fusion_select(enable_fusion));
            lazy.materialize();
            record_step(GA.get(), o, lazy.stats(), info);
            fused |= lazy.stats().fused != 0;
            if (deferred) {
                // maintain_copies(part, output, a_prev_level, a_level),
                // BFSLVLv.C:218-220.  A fused step changed levels that its
                // (empty or handed-back) output does not list: copy them all.
                if (lazy.stats().fused)
                    A::make_lazy_executor(GA).vertex_map_copy(a_prev_level, a_level).materialize();
                else
                    A::maintain_copies(GA, output, a_prev_level, a_level);
            }
            info->edges += output.nActiveEdges();
            F.del();
            F = output;
            // a fused step hands back a frontier that many levels on
            iter += lazy.stats().levels > 0 ? lazy.stats().levels : 1;
            // api.h:150-153 default_threshold::is_sparse on the new frontier
            if (fmode == 1 && !threshold_sparse(0.f, F.nActiveVertices(), F.nActiveEdges(), GA.numEdges()))
                enable_fusion = true;
        }
        F.del();
        if (fused) {
            // a fused step leaves no frontiers to count: the edges of the
            // reached vertices, from their levels
            A::frontier all = A::frontier::all_true(GA);
            A::frontier reached;
            A::make_lazy_executor(GA).vertex_filter(all, reached, a_level, n + 1, true).materialize();
            info->edges = reached.nActiveEdges();
            reached.del();
            all.del();
        }
    }
};

// ---- BFSv ----------------------------------------------------------------------
// BFS parents (BFSv.C:88-203, CONVERGENCE=1, DEFERRED_UPDATE=0): parent = ~0,
// parent[start] = start; each step sets the unset parents of the frontier's
// out-neighbours (api::relax(parent[d].setif(s)), the dst filter
// msbset(parent[d]) implied by the relax) and records them.
struct BFSv {
    Graph GA;
    gpr_run_opts o;
    A::vertexprop<uint32_t> parent;
    int iter = 0;
    uint64_t active = 0;
    BFSv(gpr_graph *g, const gpr_run_opts &opts) : GA(g), o(opts), parent(GA, "parent") {}
    ~BFSv() { parent.del(); }

    void run(uint32_t start, gpr_run_info *info) {
        A::make_lazy_executor(GA).vertex_map_fill(parent, 0xffffffffu).materialize();
        uint32_t si = 0;   // the start's own (internal) ID is its parent
        check(gpr_graph_remap_id(GA.get(), start, &si), "remap id");
        A::frontier F = A::frontier::create(GA, start);
        A::make_lazy_executor(GA).vertex_map_fill(F, parent, si).materialize();
        check(gpr_policy_reset(GA.get()), "policy reset");
        info->edges += F.nActiveEdges();
        iter = 0;
        active = 1;
        while (!F.isEmpty()) {
            A::frontier output;
            auto lazy = A::edgemap(GA,
                                   A::record(output, A::reduction, A::strong),
                                   A::filter(A::strong, A::src, F),
                                   A::relax(A::setif(parent)),
                                   A::config(o.direction, o.threshold_pct, o.out_kind));
            lazy.materialize();
            record_step(GA.get(), o, lazy.stats(), info);
            info->edges += output.nActiveEdges();
            active += output.nActiveVertices();
            F.del();
            F = output;
            ++iter;
        }
        F.del();
    }
};

// ---- CCv -----------------------------------------------------------------------
struct CCv {
    Graph GA;
    gpr_run_opts o;
    A::vertexprop<uint32_t> m_IDs, m_prevIDs;
    int iter = 0;
    CCv(gpr_graph *g, const gpr_run_opts &opts) : GA(g), o(opts), m_IDs(GA, "IDs") {
        if (o.deferred_update) m_prevIDs = A::vertexprop<uint32_t>(GA, "previous IDs");
    }
    ~CCv() {
        m_IDs.del();
        m_prevIDs.del();
    }

    // CCv.C:520-588; FUSION (CCv.C:47-54, 562-566): 1 by default, i.e.
    // api::fusion with default_fusion_select; 2 fuses every sparse step
    int fusion_mode() const {
        const int f = o.fusion == 0 ? 1 : o.fusion;
        return f == 3 ? GPR_FUSION_SMALL : f == 2 ? GPR_FUSION_ON : f == 1 ? GPR_FUSION_DEFAULT : GPR_FUSION_OFF;
    }
    A::frontier emap_step(A::frontier &F, float threshold, gpr_run_info *info) {
        A::frontier output;
        int dir = o.direction;
        float pct = threshold;
        // DEFERRED_UPDATE: api::record(output, api::reduction_or_method,
        // IDs[d] != prevIDs[d], strong), CCv.C:536-547; else the reduction
        const A::record_arg rec = o.deferred_update
                                      ? A::record(output, A::reduction_or_method, A::differs(m_IDs, m_prevIDs), A::strong)
                                      : A::record(output, A::reduction, A::strong);
        auto lazy = A::edgemap(GA,
                               A::config(dir, pct, o.out_kind),
                               rec,
                               A::filter(A::weak, A::src, F),
                               A::filter(A::weak, A::dst, A::ne(m_IDs, 0)),
                               A::relax(A::min_label(m_IDs)),
                               A::fusion_mode(fusion_mode()));
        // .vertex_map(prevIDs[v] = IDs[v]) chained to the edgemap, CCv.C:578-580
        if (o.deferred_update) lazy.vertex_map_copy(m_prevIDs, m_IDs);
        lazy.materialize();
        record_step(GA.get(), o, lazy.stats(), info);
        return output;
    }

    A::frontier find_zeros(A::frontier &F) {
        A::frontier z;
        A::make_lazy_executor(GA).vertex_filter(F, z, m_IDs, 0, false).materialize();
        return z;
    }
    A::frontier find_nonzeros(A::frontier &F) {
        A::frontier z;
        A::make_lazy_executor(GA).vertex_filter(F, z, m_IDs, 0, true).materialize();
        return z;
    }

    void run(gpr_run_info *info) {
        const uint32_t n = GA.numVertices();
        // CCv.C:246-248 IDs[v] = v
        A::make_lazy_executor(GA).vertex_map_iota(m_IDs).materialize();
        check(gpr_policy_reset(GA.get()), "policy reset");
        iter = 0;
        uint32_t initial_zeros = 1;
        // PRESET_ZERO, CCv.C:262-291: IDs[N(0)] = 0, vertex 0 of the
        // (relabelled) graph, i.e. the highest-degree vertex after reordering
        if (n > 0) {
            uint32_t v0 = 0;
            check(gpr_graph_original_id(GA.get(), 0, &v0), "original id");
            A::frontier G0 = A::frontier::empty(GA);
            check(gpr_frontier_set_neighbours(G0.get(), v0), "neighbours of 0");
            A::make_lazy_executor(GA).vertex_map_fill(G0, m_IDs, 0).materialize();
            initial_zeros += G0.nActiveVertices();
            G0.del();
            ++iter;
        }
        // the reference leaves prevIDs unset until the first step (a pull,
        // which records the reduction); here it starts as a copy of IDs
        if (o.deferred_update) A::make_lazy_executor(GA).vertex_map_copy(m_prevIDs, m_IDs).materialize();
        bool switched = false, have_checked = false;
        A::frontier zeros;
        A::frontier F = A::frontier::all_true(GA);
        while (!F.isEmpty()) {
            // emap_step(F, 5), CCv.C:309: frac_threshold(5)
            A::frontier output = emap_step(F, o.threshold_pct > 0.f ? o.threshold_pct : 5.f, info);
            // PUSH_ZERO, CCv.C:373-388: after one step, if few vertices carry
            // label 0, push the zero label first and the rest afterwards
            if (!have_checked) {
                zeros = find_zeros(output);
                if ((uint64_t(zeros.nActiveVertices()) + initial_zeros) * 10 < n) {
                    F.del();
                    F = output;
                    ++iter;
                    switched = true;
                    break;
                }
                zeros.del();
                have_checked = true;
            }
            F.del();
            F = output;
            ++iter;
        }
        F.del();
        if (switched) {
            // CCv.C:398-469, frac_threshold(200): always sparse.  The
            // reference filters each step's output for label 0 (and, in the
            // second loop, for labels other than 0).  Here the output is
            // taken as it is: a label only falls, to the minimum over
            // sources of the frontier, so in the first loop (sources all
            // labelled 0) every activated vertex ends at 0, and in the
            // second (run once no vertex labelled 0 has a neighbour left
            // that is not, and destinations labelled 0 filtered out) none
            // does.  The filter would return its input, at the cost of a
            // launch and a read-back per step.
            while (!zeros.isEmpty()) {
                A::frontier Z = zeros;
                zeros = emap_step(Z, 200.f, info);
                ++iter;
                Z.del();
            }
            zeros.del();
            A::frontier all = A::frontier::all_true(GA);
            A::frontier nonzeros = find_nonzeros(all);
            all.del();
            while (!nonzeros.isEmpty()) {
                A::frontier Z = nonzeros;
                nonzeros = emap_step(Z, 200.f, info);
                ++iter;
                Z.del();
            }
            nonzeros.del();
        }
        info->edges = GA.numEdges();
    }
};

// ---- PRv / PR_vary --------------------------------------------------------------
struct PRv {
    Graph GA;
    gpr_run_opts o;
    A::vertexprop<float> x_mem, y_mem;
    PRv(gpr_graph *g, const gpr_run_opts &opts) : GA(g), o(opts), x_mem(GA, "x"), y_mem(GA, "y") {}
    ~PRv() {
        x_mem.del();
        y_mem.del();
    }

    // run_prec: one precision stage, continuing from x (PR_vary.C:247-447)
    template <class CT>
    void run_prec(int fmt, double tol, int &iter, double &sdelta, gpr_run_info *info) {
        A::vertexprop<CT> contrib(GA, "contrib");
        const float dampen = 0.85f;
        check(gpr_pr_contrib(GA.get(), x_mem.get(), contrib.get(), fmt, dampen), "contributions");
        sdelta = 2.0;
        const bool need_delta = tol > 0.0;
        // without a convergence test between iterations, each iteration's
        // tail (but the last) is deferred into the next edgemap's write
        // stage (gpr_pr_iterate_fused), where the layout allows it
        bool fuse_tail = !need_delta && GA.numEdges() > 0;
        bool pending = false;
        while (iter < o.max_iter && sdelta > tol) {
            int rc = GPR_E_INVAL;
            if (pending) {
                // the sources without out-edges get their ranks in the last
                // fused iteration only (nothing reads them before the final tail)
                rc = gpr_pr_iterate_fused(GA.get(), x_mem.get(), y_mem.get(), fmt, dampen,
                                          iter + 1 < o.max_iter ? GPR_PR_SKIP_UNLINKED | GPR_PR_SKIP_RANKS : 0,
                                          nullptr);
                if (rc == GPR_E_INVAL) {
                    // not fusable here: the deferred tail, then the plain edgemap
                    fuse_tail = false;
                    check(gpr_pr_finish(GA.get(), x_mem.get(), y_mem.get(), contrib.get(), fmt, dampen, nullptr,
                                        nullptr),
                          "pr tail");
                } else {
                    check(rc, "pr fused iteration");
                }
                pending = false;
            }
            if (rc != GPR_OK) {
                auto lazy = A::edgemap(GA, A::relax(A::pr_sum(y_mem, contrib, fmt, dampen)));
                lazy.materialize();
            }
            // the vertex_scan / vertex_map tail of PRv.C:214-269, fused
            // (deferred into the next iteration when that one is fused)
            if (fuse_tail && iter + 1 < o.max_iter)
                pending = true;
            else
                check(gpr_pr_finish(GA.get(), x_mem.get(), y_mem.get(), contrib.get(), fmt, dampen,
                                    need_delta ? &sdelta : nullptr, nullptr),
                      "pr tail");
            gpr_iter_stats st = {};
            st.nactv_in = GA.numVertices();
            st.nacte_in = GA.numEdges();
            st.direction = GPR_DIR_PULL;
            st.edges_scanned = GA.numEdges();
            record_step(GA.get(), o, st, info);
            info->edges += GA.numEdges();
            ++iter;
        }
        contrib.del();
    }

    void run_stage(int fmt, double tol, int &iter, double &sdelta, gpr_run_info *info) {
        if (fmt == GPR_CONTRIB_F32)
            run_prec<float>(fmt, tol, iter, sdelta, info);
        else
            run_prec<uint16_t>(fmt, tol, iter, sdelta, info);
    }

    void run(const gpr_pr_stage *stages, int nstages, gpr_run_info *info) {
        const uint32_t n = GA.numVertices();
        // PRv.C:163-170: x = 1/n
        A::make_lazy_executor(GA).vertex_map_fill(x_mem, 1.0f / (float)n).materialize();
        int iter = 0;
        double sdelta = 2.0;
        for (int s = 0; s < nstages; ++s) run_stage(stages[s].contrib, stages[s].tol, iter, sdelta, info);
        double sx = 0.0;
        check(gpr_pr_scalars(GA.get(), nullptr, &sdelta, &sx, nullptr), "pr scalars");
        info->delta = sdelta;
        info->sum = sx;
    }
};

// ---- APR_vary ----------------------------------------------------------------
// Delta PageRank.  x and delta are FP32 throughout (FloatTy), z is the
// edgemap's accumulator; the frontier, x and delta carry from one stage to
// the next, and so does the iteration count.
struct APRv {
    Graph GA;
    gpr_run_opts o;
    A::vertexprop<float> x_mem, delta_mem, z_mem;
    A::vertexprop<double> xn_mem;
    APRv(gpr_graph *g, const gpr_run_opts &opts)
        : GA(g), o(opts), x_mem(GA, "x"), delta_mem(GA, "delta"), z_mem(GA, "z"), xn_mem(GA, "x normalised") {}
    ~APRv() {
        x_mem.del();
        delta_mem.del();
        z_mem.del();
        xn_mem.del();
    }

    // APR_vary.C:450-457 / :507 / :560 and frac_threshold(1) (api.h:136-139):
    // 0 the first iteration, 1 every source, 2 the frontier's sources by
    // pull, 3 by push.
    int branch(int iter, uint64_t fv, uint64_t fe, uint64_t m) const {
        if (iter == 0) return 0;
        if (((double)fv + (double)fe) / (double)m > 0.6) return 1;
        if (o.direction == GPR_DIR_PULL) return 2;
        if (o.direction == GPR_DIR_PUSH) return 3;
        return (double)(float)(fv + fe) <= (double)(float)m / 100.0 ? 3 : 2;
    }

    template <class CT>
    void run_prec(int fmt, double eps, double eps2, float factor, int &iter, A::frontier &F, AprTotals &tot,
                  gpr_run_info *info) {
        A::vertexprop<CT> contrib(GA, "contrib");
        const uint32_t n = GA.numVertices();
        const uint64_t m = GA.numEdges();
        while (iter < o.max_iter) {
            const uint64_t fv = F.nActiveVertices(), fe = F.nActiveEdges();
            const int kind = branch(iter, fv, fe, m);
            gpr_iter_stats st = {};
            st.nactv_in = kind <= 1 ? n : (uint32_t)fv;
            st.nacte_in = kind <= 1 ? m : fe;
            st.edges_scanned = kind <= 1 ? m : fe;
            const auto t0 = std::chrono::steady_clock::now();
            if (kind <= 2) {
                // contributions, then z[d] += contrib[s] * factor over the
                // CSC (the factor is applied once per vertex in the tail)
                check(apr_contrib(GA.get(), fmt, delta_mem.get(), contrib.get(), factor,
                                  kind == 2 ? F.get() : nullptr),
                      "apr contributions");
                auto lazy = A::edgemap(GA, A::relax(A::pr_sum(z_mem, contrib, fmt)));
                lazy.materialize();
                st.direction = GPR_DIR_PULL;
            } else {
                check(apr_push(GA.get(), fmt, F.get(), delta_mem.get(), factor, z_mem.get()), "apr push");
                st.direction = GPR_DIR_PUSH;
            }
            A::frontier active = A::frontier::empty(GA);
            check(apr_tail(GA.get(), kind == 0, z_mem.get(), factor, (float)eps2, x_mem.get(), delta_mem.get(),
                           active.get(), &tot),
                  "apr vertex maps");
            st.nactv_out = (uint32_t)tot.nactv;
            st.nacte_out = tot.nacte;
            st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            record_step(GA.get(), o, st, info);
            info->edges += st.edges_scanned;
            ++iter;
            F.del();
            F = active;
            if (tot.l1 < eps) break;
        }
        contrib.del();
    }

    void run(const gpr_pr_stage *stages, int nstages, double eps2, gpr_run_info *info) {
        const uint32_t n = GA.numVertices();
        for (int s = 0; s < nstages; ++s) check(apr_check_fmt(stages[s].contrib), "apr_vary stage");
        // APR_vary.C:436-448: factor = 10^-(decimal digits of n), in FP32
        float factor = 1.0f;
        for (uint32_t nnn = n; nnn > 1; nnn /= 10) factor = (float)((double)factor / 10.0);
        // :917-921 and :426-430: x = 0, delta = 1/n
        A::make_lazy_executor(GA)
            .vertex_map_fill(x_mem, 0.0f)
            .vertex_map_fill(delta_mem, (float)(1.0 / (double)n))
            .materialize();
        A::frontier F = A::frontier::all_true(GA);
        AprTotals tot = {};
        int iter = 0;
        for (int s = 0; s < nstages; ++s) {
            if (stages[s].contrib == GPR_CONTRIB_F32)
                run_prec<float>(stages[s].contrib, stages[s].tol, eps2, factor, iter, F, tot, info);
            else
                run_prec<uint16_t>(stages[s].contrib, stages[s].tol, eps2, factor, iter, F, tot, info);
        }
        F.del();
        // :1060-1072: x scaled to unit L1 norm, in FP64
        info->delta = tot.l1;
        info->sum = tot.sum_x;
        check(apr_scale(GA.get(), x_mem.get(), tot.sum_x > 0.0 ? 1.0 / tot.sum_x : 0.0, xn_mem.get()),
              "apr scale");
    }
};

// ---- KC_bucket -----------------------------------------------------------------
// Coreness by peeling (the count-down relax of KC_bucket.C:251-281 applied
// level by level); the output is the core number of every vertex.
struct KCv {
    Graph GA;
    gpr_run_opts o;
    A::vertexprop<uint32_t> coreness;
    uint32_t largestCore = 0;
    KCv(gpr_graph *g, const gpr_run_opts &opts) : GA(g), o(opts), coreness(GA, "count-down degrees / coreness") {}
    ~KCv() { coreness.del(); }

    void run(gpr_run_info *info) {
        KcStats st = {};
        check(kcore_peel(GA.get(), coreness.get(), &st), "kcore");
        largestCore = st.kmax;
        info->iters = (int)st.levels;
        info->push_iters = (int)st.rounds;
        info->edges = st.edges;
    }
};

} // namespace

extern "C" int gpr_kcore(gpr_graph *g, const gpr_run_opts *opts, uint32_t *coreness, uint32_t *largest,
                         gpr_run_info *info) {
    if (!g) return GPR_E_INVAL;
    gpr_run_opts o = defaults(opts);
    return guarded([&] {
        KCv app(g, o);
        gpr_run_info local = {};
        stream_time_ms(g, true);
        app.run(&local);
        local.ms = stream_time_ms(g, false);
        if (coreness) app.coreness.download(coreness);
        if (largest) *largest = app.largestCore;
        if (info) *info = local;
    });
}

extern "C" int gpr_apr_vary(gpr_graph *g, const gpr_pr_stage *stages, int nstages, double eps2,
                            const gpr_run_opts *opts, double *x, gpr_run_info *info) {
    if (!g || !stages || nstages <= 0 || !(eps2 >= 0.0)) return GPR_E_INVAL;
    gpr_run_opts o = defaults(opts);
    uint32_t n = 0;
    gpr_graph_info(g, &n, nullptr, nullptr);
    if (n == 0) return guarded([&] { check(GPR_E_INVAL, "apr_vary: empty graph"); });
    return guarded([&] {
        APRv app(g, o);
        gpr_run_info local = {};
        if (o.trace) trace_clear(g);
        stream_time_ms(g, true);
        app.run(stages, nstages, eps2, &local);
        local.ms = stream_time_ms(g, false);
        if (x) app.xn_mem.download(x);
        if (info) *info = local;
    });
}

extern "C" int gpr_bfslvl(gpr_graph *g, uint32_t start, const gpr_run_opts *opts, uint32_t *levels,
                          gpr_run_info *info) {
    if (!g) return GPR_E_INVAL;
    gpr_run_opts o = defaults(opts);
    uint32_t n = 0;
    gpr_graph_info(g, &n, nullptr, nullptr);
    if (start >= n) return guarded([&] { check(GPR_E_INVAL, "bfslvl: start vertex out of range"); });
    return guarded([&] {
        BFSLVLv app(g, o);
        gpr_run_info local = {};
        if (o.trace) trace_clear(g);
        stream_time_ms(g, true);
        app.run(start, &local);
        local.ms = stream_time_ms(g, false);
        if (levels) app.a_level.download(levels);
        if (info) *info = local;
    });
}

extern "C" int gpr_bfsv(gpr_graph *g, uint32_t start, const gpr_run_opts *opts, uint32_t *parents,
                        gpr_run_info *info) {
    if (!g) return GPR_E_INVAL;
    gpr_run_opts o = defaults(opts);
    uint32_t n = 0;
    gpr_graph_info(g, &n, nullptr, nullptr);
    if (start >= n) return guarded([&] { check(GPR_E_INVAL, "bfsv: start vertex out of range"); });
    return guarded([&] {
        BFSv app(g, o);
        gpr_run_info local = {};
        if (o.trace) trace_clear(g);
        stream_time_ms(g, true);
        app.run(start, &local);
        local.ms = stream_time_ms(g, false);
        if (parents) check(ids_download(g, app.parent.get(), parents), "parents download");
        if (info) *info = local;
    });
}

extern "C" int gpr_ccv(gpr_graph *g, const gpr_run_opts *opts, uint32_t *labels, gpr_run_info *info) {
    if (!g) return GPR_E_INVAL;
    gpr_run_opts o = defaults(opts);
    return guarded([&] {
        CCv app(g, o);
        gpr_run_info local = {};
        if (o.trace) trace_clear(g);
        stream_time_ms(g, true);
        app.run(&local);
        local.ms = stream_time_ms(g, false);
        if (labels) check(gpr_cc_recode(app.m_IDs.get(), labels), "cc recode");
        if (info) *info = local;
    });
}

extern "C" int gpr_pr_vary(gpr_graph *g, const gpr_pr_stage *stages, int nstages, const gpr_run_opts *opts, float *x,
                           gpr_run_info *info) {
    if (!g || !stages || nstages <= 0) return GPR_E_INVAL;
    gpr_run_opts o = defaults(opts);
    return guarded([&] {
        PRv app(g, o);
        gpr_run_info local = {};
        if (o.trace) trace_clear(g);
        stream_time_ms(g, true);
        app.run(stages, nstages, &local);
        local.ms = stream_time_ms(g, false);
        if (x) app.x_mem.download(x);
        if (info) *info = local;
    });
}

extern "C" int gpr_prv(gpr_graph *g, const gpr_run_opts *opts, float *x, gpr_run_info *info) {
    gpr_run_opts o = defaults(opts);
    gpr_pr_stage st = {GPR_CONTRIB_F32, o.tol};
    return gpr_pr_vary(g, &st, 1, &o, x, info);
}
