/*
 * gpr.h -- C ABI of the MI355X (gfx950) edgemap/vertexmap traversal engine.
 *
 * Graptor's operator API is a set of C++ templates (include/graptor/api.h)
 * with no FFI; its lambdas cannot cross a C boundary.  This ABI therefore
 * exports the fixed operator set that BFSLVLv, CCv, PRv and PR_vary use,
 * with plain pointers and sizes.  The header-only C++ layer
 * include/gpr/api.hpp rebuilds Graptor's api::edgemap / api::filter /
 * api::record / api::relax / api::config / make_lazy_executor names on top
 * of it.  Every entry point returns GPR_OK (0) or a GPR_E_* code; the text of
 * the last error of the calling thread is available from gpr_last_error.
 *
 * Ownership: handles own their device memory and are released with the
 * matching *_destroy call.  Host buffers passed in are copied; host buffers
 * passed out are filled before the call returns.  Calls on one graph handle
 * are serialised on that handle's HIP stream and are not thread-safe per
 * handle; separate handles may be used from separate threads.
 */
#ifndef GPR_H
#define GPR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference aborts instead: GraphCSx.h:2227-2244,
 *      api.h:2376-2377) ---------------------------------------------------- */
enum {
    GPR_OK = 0,
    GPR_E_INVAL = 1,   /* bad argument or unsupported combination */
    GPR_E_FORMAT = 2,  /* malformed graph file or CSR */
    GPR_E_OOM = 3,     /* device or host allocation failed */
    GPR_E_HIP = 4,     /* HIP runtime error */
    GPR_E_IO = 5,      /* file IO error */
    GPR_E_NODEV = 6    /* no gfx950 device */
};

typedef struct gpr_graph gpr_graph;
typedef struct gpr_prop gpr_prop;
typedef struct gpr_frontier gpr_frontier;

/* Copy the calling thread's last error message into buf (NUL-terminated). */
int gpr_last_error(char *buf, size_t len);
/* Library version string, e.g. "gpr 0.1 gfx950". */
const char *gpr_version(void);
/* Select and initialise device `device` (0 = first).  Fails with
 * GPR_E_NODEV when the device is not a gfx950. */
int gpr_init(int device);

/* ---- graphs: GraphCSx (graph/GraphCSx.h:293-2806) ------------------------ */
/* Upload a CSR (index has n+1 entries, nbrs m).  For non-symmetric graphs
 * the CSC is built on the host (a counting sort over the caller's CSR,
 * rows ascending) and uploaded beside it.  Replaces GraphCSx(index, edges)
 * and build_degree (GraphCSx.h:2797-2805). */
int gpr_graph_create(const uint64_t *index, const uint32_t *nbrs, uint32_t n,
                     uint64_t m, int symmetric, gpr_graph **out);
/* Read a b2 file (versions 2 and 3) -- GraphCSx::readFromBinaryFile,
 * GraphCSx.h:2137-2264. */
int gpr_graph_load_b2(const char *path, int symmetric, gpr_graph **out);
/* Write a version-2 b2 file -- GraphCSx::writeToBinaryFile, :2009-2027. */
int gpr_graph_save_b2(const gpr_graph *g, const char *path);
/* Generate a cleaned (no self loops, symmetrised, sorted, deduplicated)
 * RMAT graph on the device: 2^scale vertices, edge_factor*2^scale generated
 * edges, quadrant probabilities a, b, c (d = 1-a-b-c), counter-based stream
 * keyed by seed.  Replaces the CleanGraph/RmDupEdges/CvtToUndir pipeline
 * (tools/CleanGraph.C:10-139, RmDupEdges.C:9-59, CvtToUndir.C:30-121). */
int gpr_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a,
                            double b, double c, uint64_t seed, gpr_graph **out);
int gpr_graph_info(const gpr_graph *g, uint32_t *n, uint64_t *m,
                   int *symmetric);
/* Renumber the vertices internally (GraphVEBOGraptor's VEBO relabelling,
 * graph/CGraphVEBOGraptor.h:60-142, graph/VEBOReorder.h:134-209).
 * GPR_ORDER_DEGREE: decreasing degree, ties by ascending ID (symmetric
 * graphs).  Every vertex ID and per-vertex array that crosses this ABI stays
 * in original IDs (remapID/originalID, CGraphVEBOGraptor.h:441-500); values
 * that are themselves vertex IDs (CC labels) are internal until recoded by
 * gpr_cc_recode. */
enum { GPR_ORDER_NONE = 0, GPR_ORDER_DEGREE = 1 };
int gpr_graph_reorder(gpr_graph *g, int mode);
/* order[i] = original ID of internal vertex i (identity if not reordered). */
int gpr_graph_order(const gpr_graph *g, uint32_t *order);
int gpr_graph_remap_id(gpr_graph *g, uint32_t orig, uint32_t *internal);
int gpr_graph_original_id(gpr_graph *g, uint32_t internal, uint32_t *orig);
/* Copy the CSR back in internal IDs (index: n+1 entries, nbrs: m).  Either
 * may be NULL.  With gpr_graph_order this describes the original graph. */
int gpr_graph_download(const gpr_graph *g, uint64_t *index, uint32_t *nbrs);
/* Highest out-degree vertex, ties to the lowest ID, as an original ID
 * (GraphCSx::findHighestDegreeVertex, GraphCSx.h:2446-2451). */
int gpr_graph_highest_degree_vertex(const gpr_graph *g, uint32_t *v);
int gpr_graph_destroy(gpr_graph *g);

/* ---- vertex properties: api::vertexprop (api/vertexprop.h:28-143) ------- */
enum { GPR_U32 = 0, GPR_F32 = 1, GPR_F64 = 2, GPR_U64 = 3, GPR_U16 = 4 };
int gpr_prop_create(gpr_graph *g, int dtype, gpr_prop **out);
int gpr_prop_upload(gpr_prop *p, const void *host);
int gpr_prop_download(const gpr_prop *p, void *host);
int gpr_prop_destroy(gpr_prop *p);
/* CCv recode (bench/CCv.C:121-169): from a u32 label property in internal
 * IDs, labels[v] = minimum original vertex ID of v's component. */
int gpr_cc_recode(const gpr_prop *p, uint32_t *labels);

/* ---- frontiers: frontier (frontier/frontier.h:58-1179) ------------------ */
enum {
    GPR_FRONTIER_EMPTY = 0,
    GPR_FRONTIER_ALL = 1,     /* frontier::all_true */
    GPR_FRONTIER_LIST = 2,    /* ft_sparse: u32 vertex list */
    GPR_FRONTIER_BITMAP = 3   /* ft_bit: one bit per vertex, u64 words */
};
int gpr_frontier_create(gpr_graph *g, gpr_frontier **out); /* empty */
/* frontier::create(n, v, deg) */
int gpr_frontier_set_single(gpr_frontier *f, uint32_t v);
int gpr_frontier_set_all(gpr_frontier *f);
int gpr_frontier_set_list(gpr_frontier *f, const uint32_t *vs, uint32_t k);
/* frontier::sparse(n, deg(v), &edges[index[v]]): the neighbours of v
 * (CCv.C:262-267, PRESET_ZERO). */
int gpr_frontier_set_neighbours(gpr_frontier *f, uint32_t v);
/* Active counts nActiveVertices / nActiveEdges (frontier.h:694-697) and
 * representation (GPR_FRONTIER_*). */
int gpr_frontier_stats(const gpr_frontier *f, uint32_t *nactv,
                       uint64_t *nacte, int *kind);
/* Convert in place: toSparse (frontier.h:576-634) / to a bitmap. */
int gpr_frontier_convert(gpr_frontier *f, int kind);
/* Copy the active vertices out, ascending.  *k receives the count; out
 * must hold nactv entries. */
int gpr_frontier_download(gpr_frontier *f, uint32_t *out, uint32_t *k);
int gpr_frontier_destroy(gpr_frontier *f);

/* ---- edgemap: api::edgemap (api.h:2143-2415) ----------------------------- */
enum {                      /* relax operator, api::relax (api.h:1670-1687) */
    GPR_RELAX_MIN_PLUS1 = 0, /* p[d].min(p[s]+1)   BFSLVLv.C:208-214 */
    GPR_RELAX_MIN = 1,       /* p[d].min(p[s])     CCv.C:571-577 */
    GPR_RELAX_PR_SUM = 2,    /* y[d] += contrib[s] PRv.C:202-212 */
    GPR_RELAX_SETIF = 3      /* p[d].setif(s): p[d] = s (the source's ID) where
                                p[d] has its top bit set; changed once per
                                vertex (BFSv.C:146, dsl/ast/redop.h:1396-1440) */
};
enum {                      /* destination filter, api::filter(dst) */
    GPR_DST_NONE = 0,
    GPR_DST_GT = 1,          /* p[d] >  arg   BFSLVLv.C:193-196 */
    GPR_DST_NE = 2           /* p[d] != arg   CCv.C:559-564 */
};
enum {                      /* api::record (api.h:1450-1516) */
    GPR_RECORD_NONE = 0,      /* no record argument: no output frontier
                                 (csr_sparse_no_f, api.h:2249-2252) */
    GPR_RECORD_REDUCTION = 1, /* api::reduction: activate d when a relax
                                 changed p[d] */
    GPR_RECORD_METHOD_NE = 2, /* api::method with [&](d){ return p[d] !=
                                 prev[d]; }: pull activates every d with
                                 p[d] != prev[d] after its row; push
                                 activates d when a relax changed p[d] and
                                 p[d] != prev[d] right after that relax
                                 (csr_sparse_with_f: vexpr && uexpr,
                                 dsl/edgemap.h:1497-1507) */
    GPR_RECORD_REDUCTION_OR_METHOD = 3 /* api::reduction_or_method
                                 (api.h:1401-1416): the method in push, the
                                 reduction in pull (CCv.C:536-547) */
};
enum {                      /* traversal direction / threshold */
    GPR_DIR_AUTO = 0,        /* threshold rule, api.h:130-160 */
    GPR_DIR_PUSH = 1,        /* always_sparse */
    GPR_DIR_PULL = 2,        /* always_dense */
    GPR_DIR_POLICY = 3       /* learned policy (gpr_policy_*) */
};
enum {                      /* contribution storage for GPR_RELAX_PR_SUM */
    GPR_CONTRIB_F32 = 0,     /* FP32 store, FP64 accumulate (PRv) */
    GPR_CONTRIB_BF16 = 1,    /* bfloat16 RNE store, FP32 accumulate */
    GPR_CONTRIB_F16 = 2,     /* IEEE half + per-iteration 2^k scale */
    GPR_CONTRIB_CFP16 = 3,   /* customfp<6,10> truncating (customfp.h) */
    GPR_CONTRIB_BF16T = 4,   /* scustomfp<true,8,7>: bf16 truncating */
    GPR_CONTRIB_SCFP16 = 5   /* scustomfp<true,6,9,true,0>: signed, truncating
                                (APR_vary.C:743-750); code 0 reads as 0 */
};

typedef struct {
    int relax;            /* GPR_RELAX_* */
    gpr_prop *prop;       /* relaxed property (u32) */
    gpr_prop *prev;       /* previous copy for GPR_RECORD_METHOD_NE and
                             GPR_RECORD_REDUCTION_OR_METHOD */
    int src_strong;       /* api::filter(strong|weak, src, F) */
    int dst_filter;       /* GPR_DST_* */
    uint32_t dst_arg;
    int record;           /* GPR_RECORD_* */
    int direction;        /* GPR_DIR_* */
    int out_kind;         /* 0 = auto, GPR_FRONTIER_LIST / _BITMAP */
    float threshold_pct;  /* > 0: frac_threshold(t), float(nactv + nacte) <=
                             float(m) / (100.0 / t) (api.h:135-139); 0:
                             default_threshold, the integer rule
                             nactv + nacte <= m / 20 (api.h:150-153) */
    /* GPR_RELAX_PR_SUM only: prop is the f32 accumulator y, contrib the
     * stored contributions in format contrib_fmt (GPR_CONTRIB_*). */
    gpr_prop *contrib;
    int contrib_fmt;
    float damping;        /* 0.85 when 0 */
    /* api::fusion with api::fusion_select (api/fusion.h:17-140, api.h:227-248):
     * a push step processes the vertices it activates at once, in the same
     * launch, until none is left; its output frontier is then empty.  For
     * GPR_RELAX_MIN_PLUS1 / GPR_RELAX_MIN; the dst filter is implied by the
     * relax.  GPR_FUSION_OFF, _ON (fusion_select(true)), _DEFAULT
     * (default_fusion_select: when the frontier has >= 4096 edges; on this
     * engine also only while the frontier and each level it reaches stay
     * within fusion_small_limit, as under GPR_FUSION_SMALL below: the fused
     * kernel runs one workgroup per compute unit, and a larger level costs
     * it more than an unfused step does, so the first such level is handed
     * back as the step's output frontier).  Results do not depend on the
     * fusion setting. */
    int fusion;
} gpr_emap_desc;

/* GPR_FUSION_SMALL (no reference counterpart): a sparse step fuses when its
 * frontier's vertices + out-edges are at most the tunable
 * fusion_small_limit, and its levels run until one exceeds that limit,
 * which it hands back as its output frontier (a list): small levels, whose
 * unfused steps cost more in launches and round trips than in work, run in
 * one launch, large ones as unfused steps. */
enum { GPR_FUSION_OFF = 0, GPR_FUSION_ON = 1, GPR_FUSION_DEFAULT = 2, GPR_FUSION_SMALL = 3 };

typedef struct {
    uint32_t nactv_in;
    uint64_t nacte_in;
    uint32_t nactv_out;
    uint64_t nacte_out;
    int direction;        /* GPR_DIR_PUSH or GPR_DIR_PULL as executed */
    int out_kind;         /* GPR_FRONTIER_* produced */
    uint64_t edges_scanned; /* edges examined (upper bound for pull; every
                               relaxation of a fused step, repeats included) */
    double ms;            /* device time of the step */
    int fused;            /* 1: a fused push step (api::fusion) */
    int levels;           /* levels the step advanced: 1 for an unfused
                             step; a fused step's handed-back frontier lies
                             this many levels past its input */
} gpr_iter_stats;

/* One edgemap step over frontier `in`, recording into `out` (which may be
 * NULL when record is GPR_RECORD_NONE).  `in` may be NULL for all-true.
 * With GPR_RECORD_NONE no vertex is activated and `out`, if given, is left
 * empty. */
int gpr_edgemap(gpr_graph *g, const gpr_emap_desc *desc, gpr_frontier *in,
                gpr_frontier *out, gpr_iter_stats *st);

/* ---- vertexmap: make_lazy_executor().vertex_map/scan/filter
 *      (dsl/vertexmap.h:2098-2456) ----------------------------------------- */
enum {
    GPR_VMAP_FILL = 0,       /* p[v] = value */
    GPR_VMAP_COPY = 1,       /* p[v] = q[v]   (maintain_copies :2515-2535) */
    GPR_VMAP_IOTA = 2,       /* p[v] = v */
    GPR_VMAP_FILL_FRONTIER = 3, /* p[v] = value for v in F */
    GPR_VMAP_COPY_FRONTIER = 4  /* p[v] = q[v] for v in F (maintain_copies
                                   over a sparse change set, :2515-2535) */
};
int gpr_vertex_map(gpr_graph *g, int op, gpr_prop *p, const gpr_prop *q,
                   gpr_frontier *f, uint64_t value_bits);
/* vertex_filter(G, F, out, p[v] == value) or != value (neg=1)
 * (CCv.C:481-509 find_zeros / find_nonzeros). */
int gpr_vertex_filter(gpr_graph *g, gpr_frontier *in, gpr_prop *p,
                      uint32_t value, int neg, gpr_frontier *out);

/* PageRank vertex maps.  gpr_pr_contrib: contrib[v] = damping*x[v]/deg(v)
 * in format fmt (PRv.C:176-187, the MEMO contribution array).
 * gpr_pr_finish: the fused tail of one iteration after a GPR_RELAX_PR_SUM
 * edgemap into y (PRv.C:214-269): y += (1-sum y)/n; y *= 1/sum y;
 * delta = sum |x-y|; x = y; contrib = damping*x/deg.  The sums are FP64 and
 * deterministic.  delta and sum_x are returned when non-NULL (this
 * synchronises); otherwise they stay on the device (gpr_pr_scalars). */
int gpr_pr_contrib(gpr_graph *g, const gpr_prop *x, gpr_prop *contrib,
                   int fmt, float damping);
int gpr_pr_finish(gpr_graph *g, gpr_prop *x, const gpr_prop *y,
                  gpr_prop *contrib, int fmt, float damping, double *delta,
                  double *sum_x);
int gpr_pr_scalars(gpr_graph *g, double *sum_y, double *delta, double *sum_x,
                   float *enc_scale);
/* One PageRank iteration with the previous iteration's vertex-map tail
 * fused into this iteration's edgemap (no reference counterpart: PRv.C
 * runs the maps of :214-269 after the edgemap of :202-212).  Equal to
 * gpr_pr_finish(g, x, y, contrib, fmt, damping, NULL, NULL) followed by
 * the GPR_RELAX_PR_SUM edgemap into y, except that the contributions are
 * computed in the edgemap's write stage and never stored (a contribution
 * property is left as it was) and no residual is computed (the device's
 * delta and sum_x keep their values); x and y come out bit-identical to
 * that sequence's.  For fixed iteration counts (no convergence test
 * between iterations).  Needs every edge in the binned layout (tunable
 * pr_split_bytes 0) and a graph with edges: GPR_E_INVAL otherwise.  st
 * (optional) receives the step's stats.  flags GPR_PR_SKIP_UNLINKED: the
 * ranks of sources without out-edges are not written (their contributions
 * are 0 and no edge reads them): x is then complete only after a later
 * call without the flag, or gpr_pr_finish, updates them; a run of fused
 * iterations passes it to all but its last.  GPR_PR_SKIP_RANKS: no rank is
 * written (the next fused call computes its contributions from y, not
 * from x), which implies GPR_PR_SKIP_UNLINKED: x keeps its old values until
 * a later call without either flag; a run of fused iterations passes both
 * to all but its last. */
enum { GPR_PR_SKIP_UNLINKED = 1, GPR_PR_SKIP_RANKS = 2 };
int gpr_pr_iterate_fused(gpr_graph *g, gpr_prop *x, gpr_prop *y, int fmt,
                         float damping, int flags, gpr_iter_stats *st);

/* ---- applications: the callers of the hot path -------------------------- */
typedef struct {
    int direction;        /* GPR_DIR_* (AUTO = Graptor threshold) */
    int out_kind;         /* 0 = auto */
    float threshold_pct;  /* 0: the benchmark's own rule (BFSLVLv, BFSv and
                             the 64-source BFS: default_threshold; CCv:
                             frac_threshold(5)); > 0: frac_threshold(t) */
    int max_iter;         /* PR: -pr:maxiter (default 100) */
    double tol;           /* PR: 1e-7 */
    int trace;            /* record per-iteration stats (-itimes) */
    /* fused sparse steps (the benchmarks' FUSION setting, BFSLVLv.C:28-40,
     * CCv.C:47-54): 0 = the reference's default for the benchmark (BFSLVLv
     * off, CCv FUSION 1), -1 = off, 1 = FUSION 1 (BFSLVLv: on for
     * low-degree graphs and after a dense step; CCv: default_fusion_select),
     * 2 = FUSION 2 (every sparse step), 3 = small levels only
     * (GPR_FUSION_SMALL) */
    int fusion;
    /* the benchmarks' DEFERRED_UPDATE setting (BFSLVLv.C:10-12, CCv.C:26-28):
     * 0 = api::record(output, api::reduction) (the engine's default);
     * 1 = the reference's default: BFSLVLv records the method
     * level[d] != prev_level[d] and keeps prev_level with maintain_copies
     * (BFSLVLv.C:166-172, 218-220); CCv records reduction_or_method with
     * IDs[d] != prevIDs[d] and copies prevIDs = IDs after every step
     * (CCv.C:536-549, 578-580).  Levels and labels are the same either way. */
    int deferred_update;
} gpr_run_opts;

typedef struct {
    int iters;
    int push_iters;
    uint64_t edges;       /* edges traversed for GTEPS */
    double ms;            /* wall time of run() on the device stream */
    double delta;         /* PR: final L1 residual */
    double sum;           /* PR: sum of ranks */
} gpr_run_info;

/* BFSLVLv (bench/BFSLVLv.C:110-254): levels[n], unreached = n+1. */
int gpr_bfslvl(gpr_graph *g, uint32_t start, const gpr_run_opts *o,
               uint32_t *levels, gpr_run_info *info);
/* BFSv (bench/BFSv.C:88-203): parents[n] (original IDs), the start its own
 * parent, unreached = 0xffffffff.  Which frontier neighbour becomes a
 * vertex's parent depends on the order of the relaxes, as in the reference;
 * every parent is an in-neighbour one BFS level closer to the start. */
int gpr_bfsv(gpr_graph *g, uint32_t start, const gpr_run_opts *o,
             uint32_t *parents, gpr_run_info *info);
/* CCv (bench/CCv.C:221-588): labels[n] = min vertex ID of the component
 * (the recode canonical form, CCv.C:121-169). */
int gpr_ccv(gpr_graph *g, const gpr_run_opts *o, uint32_t *labels,
            gpr_run_info *info);
/* PRv (bench/PRv.C:115-290): FP32 ranks, FP64 accumulation. */
int gpr_prv(gpr_graph *g, const gpr_run_opts *o, float *x,
            gpr_run_info *info);
/* PR_vary (bench/PR_vary.C:247-757): stages of (contribution format,
 * tolerance) run in order, each continuing from the previous ranks, all
 * sharing max_iter.  One stage with GPR_CONTRIB_F32 equals gpr_prv. */
typedef struct { int contrib; double tol; } gpr_pr_stage;
int gpr_pr_vary(gpr_graph *g, const gpr_pr_stage *stages, int nstages,
                const gpr_run_opts *o, float *x, gpr_run_info *info);

/* APR_vary (bench/APR_vary.C:378-640, :900-1074): delta PageRank.  The
 * stages of (contribution format, eps) run in order from x = 0,
 * delta = 1/n, sharing max_iter (-pr:maxiter), the frontier, x and delta;
 * a stage ends when sum |delta| < eps.  eps2 is the relative activity
 * threshold (the reference's 1e-2): a vertex stays active while
 * |delta| > eps2 * x.  x[n] receives the ranks scaled to unit L1 norm
 * (FP64, original IDs).  Formats: GPR_CONTRIB_F32 (variant 20, FP64
 * accumulation), GPR_CONTRIB_SCFP16 (variant 46), GPR_CONTRIB_BF16,
 * GPR_CONTRIB_BF16T; the unsigned formats are rejected.  Variant 2046 is
 * {F32, 1e-1} then {SCFP16, 1e-7}.  o->direction: GPR_DIR_AUTO follows the
 * reference (every source above 60% density, push at or below 1% of the
 * edges); GPR_DIR_PUSH / GPR_DIR_PULL force the frontier iterations.
 * info->delta is the last sum |delta|, info->sum the sum |x| before the
 * scaling. */
int gpr_apr_vary(gpr_graph *g, const gpr_pr_stage *stages, int nstages, double eps2,
                 const gpr_run_opts *o, double *x, gpr_run_info *info);

/* KC_bucket (bench/KC_bucket.C:104-330): coreness[n] = the core number of
 * every vertex (original IDs), *largest = the largest (largestCore).  The
 * graph must be symmetric.  info->iters = distinct core values peeled,
 * info->push_iters = count-down rounds, info->edges = out-edges of peeled
 * vertices.  The reference's -kc:buckets (open buckets) has no counterpart:
 * the peeling takes the next core value from a device-wide minimum. */
int gpr_kcore(gpr_graph *g, const gpr_run_opts *o, uint32_t *coreness, uint32_t *largest,
              gpr_run_info *info);

/* Multi-source BFSLVLv: k <= 64 sources (original IDs) as the bit lanes of
 * one traversal; levels[j*n + v] is the level of v from sources[j] (n+1
 * unreached), each lane equal to gpr_bfslvl from that source.  No
 * counterpart in the reference, which runs one source per call
 * (BFSLVLv.C:65, -start). */
int gpr_msbfs(gpr_graph *g, const uint32_t *sources, int k,
              const gpr_run_opts *o, uint32_t *levels, gpr_run_info *info);

/* ---- batches of small graphs: CC over all of them in one persistent
 *      launch (no counterpart in the reference; labels are CCv's canonical
 *      form, CCv.C:121-169, in the batch's global vertex IDs) ------------- */
typedef struct gpr_batch gpr_batch;
/* count graphs, vertex counts log-uniform in [nmin, nmax], even graphs RMAT
 * (edge factor 8), odd graphs Erdos-Renyi (mean degree 6), from seed. */
int gpr_batch_generate(uint32_t count, uint32_t nmin, uint32_t nmax,
                       uint64_t seed, gpr_batch **out);
/* A CSR over the disjoint union (global IDs) with vertex offsets voff
 * (count+1); every edge must stay inside its graph. */
int gpr_batch_create(const uint64_t *index, const uint32_t *nbrs,
                     const uint32_t *voff, uint32_t count, gpr_batch **out);
int gpr_batch_info(const gpr_batch *b, uint32_t *count, uint32_t *n,
                   uint64_t *m);
int gpr_batch_download(const gpr_batch *b, uint64_t *index, uint32_t *nbrs,
                       uint32_t *voff);
/* labels[v] = minimum global vertex ID of v's component.  Runs as one
 * hipGraph launch (queue reset + persistent kernel). */
int gpr_batch_cc(gpr_batch *b, uint32_t *labels, gpr_run_info *info);
int gpr_batch_destroy(gpr_batch *b);
/* gpr_batch_cc on a batch in caller-owned device memory, allocating nothing
 * (no hipMalloc, hipFree or allocating library call).  No counterpart in the
 * reference: it is for callers that cannot allocate, such as code under HIP
 * graph capture or a serving loop.  d_index (n+1 u64), d_nbrs (m u32),
 * d_voff (count+1 u32, d_voff[count] = n) and d_labels (n u32) are device
 * pointers on the current device, laid out as for gpr_batch_create (every
 * edge inside its graph); stream is a hipStream_t (NULL: the null stream).
 * The kernels are enqueued and the call returns without synchronising;
 * d_labels then receives what gpr_batch_cc returns.  Graphs are taken in
 * index order.  The work-queue counter and the parents of graphs above
 * 65536 vertices live in a static device array of each GPU, so a batch may
 * hold at most 2^21 vertices (GPR_E_INVAL above that).  Calls that overlap
 * on one GPU share that array and race. */
int gpr_batch_cc_noalloc(const uint64_t *d_index, const uint32_t *d_nbrs,
                         const uint32_t *d_voff, uint32_t count, uint32_t n,
                         uint32_t *d_labels, void *stream);

/* ---- learned direction / representation policy (GPR_DIR_POLICY) --------
 * No counterpart in the reference, whose rule is the fixed threshold of
 * api.h:150-153 (the fallback when no network is loaded, and the baseline).
 * A 6-32-32-4 FP32 MLP picks each step's action: 0 = the threshold rule's
 * own step (direction by the rule, the caller's fusion setting and output
 * representation), 1 = an unfused sparse push (where the rule pulls, only
 * while vertices + out-edges of the frontier stay within twice the rule's
 * threshold, else the rule's step), 2 = a push fused while its levels stay
 * small (GPR_FUSION_SMALL), 3 = a pull with a list output where the rule
 * pulls (the rule's step where it pushes).  Features: log2 of the frontier
 * size, of its out-edges and of the unvisited edges (both relative to m),
 * the iteration, log mean degree and log degree skew.  Its forward pass
 * runs on the device in the one-wave reduction of the step's frontier
 * counts (on the host after a fused step).  An action other than 0 is taken only when its logit leads
 * action 0's by at least the margin.
 * Parameter layout: W1[6][32] b1[32] W2[32][32] b2[32] W3[32][4] b3[4]. */
int gpr_mlp_param_count(void);
/* Load (params != NULL) or drop (NULL) the network of this graph handle:
 * count = gpr_mlp_param_count() (margin 0), or one more value, the
 * margin. */
int gpr_policy_set(gpr_graph *g, const float *params, int count);
/* Start of a run: iteration 0, all edges unvisited, no pending choice. */
int gpr_policy_reset(gpr_graph *g);
/* Training traces: algo 0 = BFSLVLv from start, 1 = CCv (min-label
 * propagation from all vertices).  The run is sampled at up to cap
 * iterations (evenly spread); for each sample: the 6 features and, for each
 * of the 4 actions, the measured ms (host clock, best of three runs) from
 * the start of that step to the end of the run when the step takes the
 * action and every other step the threshold rule's
 * (BFSLVLv: no fusion; CCv: default_fusion_select, frac_threshold(5)).
 * With GPR_TRACE_FOLLOW or'ed into algo, every other step is the loaded
 * network's instead (states and costs-to-go under that network, for
 * policy iteration); it needs a network (gpr_policy_set).  With
 * GPR_TRACE_PERSIST, the forced action holds from that step to the end of
 * the run (what a run of departures from the rule costs, which single-step
 * departures do not show when each one pays a change of frontier
 * representation).
 * *k receives the sample count. */
enum { GPR_TRACE_FOLLOW = 0x100, GPR_TRACE_PERSIST = 0x200 };
int gpr_policy_trace(gpr_graph *g, int algo, uint32_t start, int cap,
                     float *features, float *times, int *k);
/* The network on the GPU: logits of count samples; gradient of the
 * weighted soft-target cross-entropy (sum_i w_i CE_i / sum_i w_i); and
 * full-batch Adam training in place. */
int gpr_mlp_forward(const float *params, const float *x, int count,
                    float *logits);
int gpr_mlp_grad(const float *params, const float *x, const float *targets,
                 const float *weights, int count, float *grad, float *loss);
int gpr_mlp_train(float *params, const float *x, const float *targets,
                  const float *weights, int count, int epochs, float lr,
                  float *loss);

/* Per-iteration trace of the last run (-itimes, BFSLVLv.C:90-106). */
int gpr_trace_get(const gpr_graph *g, gpr_iter_stats *out, int cap, int *k);

/* Process-wide runtime tunables (no reference counterpart; chosen by the
 * autotuner, graptor_amd/autotune.py):
 *   pr_split_bytes      PageRank: FP32 bytes of source contributions the
 *                       pull gathers (the sources [0, bytes/4)); edges from
 *                       the other sources are summed by streaming through a
 *                       binned layout.  0 bins every edge; >= 4n pulls every
 *                       edge.  Default 0
 *   pull_serial_max     pull rows up to this in-degree run lane-serially
 *   pull_blocks_per_cu  pull grid size per compute unit
 *   push_blocks_per_cu  push grid cap per compute unit
 *   fusion_spin_limit   fused push: polls of an idle wave without any wave
 *                       claiming work, or of a reserved queue slot, before
 *                       the step fails with GPR_E_HIP (default 2^22)
 *   fusion_blocks       fused push: blocks of 1024 threads in its
 *                       cooperative grid (0: one per compute unit)
 *   fusion_small_limit  GPR_FUSION_SMALL and _DEFAULT: the largest frontier
 *                       (vertices + out-edges) that fuses, and the level
 *                       size that ends a fused step (default 2^17)
 *   pr_write_stage_cells  PageRank binned edgemap: a source tile's run
 *                       offsets are staged in LDS when it has at most this
 *                       many non-empty (tile, bin) cells (capped at 4096,
 *                       the default), else read from HBM; the results are
 *                       the same either way
 *   push_bitmap_frac    out_kind 0: a push writes a bitmap (read back by
 *                       the next step as an ID-sorted list) when its input
 *                       frontier has more than this * n out-edges, else a
 *                       list (default 0.125); a pull always a bitmap */
int gpr_tune_set(const char *name, double value);
int gpr_tune_get(const char *name, double *value);

/* Device-event timing of named kernels over the runs since the last reset,
 * measured on the handle's own stream.  name: "pr_pull", "pr_update",
 * "push", "pull", ... */
int gpr_profile_enable(gpr_graph *g, int on);
int gpr_profile_get(const gpr_graph *g, const char *name, double *ms,
                    uint64_t *launches);
int gpr_profile_reset(gpr_graph *g);

#ifdef __cplusplus
}
#endif
#endif /* GPR_H */
