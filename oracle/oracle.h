/*
 * oracle.h -- CPU restatement of Graptor's traversal hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (graptor_amd/, include/)
 * includes, links or calls this code.  It is used by tests/, by the smoke
 * check and by the cpu_baseline leg of bench.py, always as the checker or as
 * the timed CPU baseline, never as the thing measured on the GPU.
 *
 * Two kinds of functions live here:
 *   - textbook oracles that pin results by their mathematics
 *     (queue BFS, union-find CC, FP64 power iteration), and
 *   - an OpenMP restatement of Graptor's own algorithms (sparse push with
 *     CAS-min relax and a dedup flag array, scalar pull over the CSC, the
 *     nactv + nacte <= m/20 switch), which is the timed CPU baseline.
 *
 * Vertex IDs are uint32 (Graptor VID, include/graptor/itraits.h:81), edge
 * IDs uint64 (EID, itraits.h:84).  Graphs are CSR with n+1 index entries.
 */
#ifndef GRAPTOR_ORACLE_H
#define GRAPTOR_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- input generation (no generator exists in the reference) ---------- */
uint64_t orc_mix64(uint64_t x);
/* Quadrant thresholds floor(p * 2^32) for p = a, a+b, a+b+c (64-bit). */
void orc_rmat_thresholds(double a, double b, double c, uint64_t t[3]);
/* Edges [first, first+count) of an RMAT(scale) stream with given seed. */
void orc_rmat_edges(uint32_t scale, const uint64_t t[3], uint64_t seed,
                    uint64_t first, uint64_t count,
                    uint32_t *src, uint32_t *dst);
/* Erdos-Renyi style stream: count uniform endpoint pairs in [0,n). */
void orc_er_edges(uint32_t n, uint64_t seed, uint64_t count,
                  uint32_t *src, uint32_t *dst);

/* Batches of small graphs: sizes and the local CSR of graph g. */
void orc_batch_sizes(uint32_t count, uint32_t nmin, uint32_t nmax, uint64_t seed, uint32_t *nv);
uint64_t orc_batch_graph(uint32_t g, uint32_t nv, uint64_t seed, uint64_t *index, uint32_t *nbrs);

/* Clean an edge list into a CSR: drop self loops, (optionally) add the
 * reverse of every edge, sort and remove duplicates, keep isolated vertices.
 * Semantics of tools/CleanGraph.C:10-139 (without --rm-isolated),
 * tools/RmDupEdges.C:9-59 and tools/CvtToUndir.C:30-121.
 * index must hold n+1 entries; nbrs must hold 2*count entries if symmetrize
 * (count otherwise).  Returns m. */
uint64_t orc_build_csr(uint32_t n, uint64_t count, const uint32_t *src,
                       const uint32_t *dst, int symmetrize,
                       uint64_t *index, uint32_t *nbrs);

/* b2 file format, include/graptor/graph/GraphCSx.h:2009-2027 (write)
 * and :2222-2260 (read: versions 2 and 3). */
int orc_b2_write(const char *path, uint32_t n, uint64_t m,
                 const uint64_t *index, const uint32_t *nbrs);
/* Reads header only when index == NULL.  Returns 0 on success, negative
 * on format or IO error. */
int orc_b2_read(const char *path, uint32_t *n, uint64_t *m,
                uint64_t *index, uint32_t *nbrs);

/* CheckGraph semantics (tools/CheckGraph.C:10-87): sorted, unique, in
 * range, index monotone, index[n]==m.  Returns 0 when valid. */
int orc_check_csr(uint32_t n, uint64_t m, const uint64_t *index,
                  const uint32_t *nbrs);

/* ---- textbook oracles ------------------------------------------------- */
/* Queue BFS.  level[v] = hop distance, unreached = n+1 as in
 * bench/BFSLVLv.C:117-129. */
void orc_bfs_levels(uint32_t n, const uint64_t *index, const uint32_t *nbrs,
                    uint32_t start, uint32_t *level);
/* Union-find with the smaller root kept: label = min vertex ID of the
 * component, the canonical form of bench/CCv.C:121-169 (recode). */
void orc_cc_labels(uint32_t n, const uint64_t *index, const uint32_t *nbrs,
                   uint32_t *label);
/* FP64 PageRank as in bench/PRv.C:176-269.  (cindex,cnbrs) is the CSC
 * (in-edges), outdeg the out-degrees.  Returns the iteration count; the
 * final L1 residual goes to *delta_out. */
int orc_pr_fp64(uint32_t n, const uint64_t *cindex, const uint32_t *cnbrs,
                const uint32_t *outdeg, int max_iter, double tol,
                double damping, double *x, double *delta_out);

/* ---- restatement of Graptor's engine (timed CPU baseline) -------------- */
typedef struct {
    int iters;          /* edgemap calls */
    int push_iters;     /* calls taken sparse (push) */
    uint64_t edges;     /* edges examined */
} orc_run_info;

/* BFSLVLv (bench/BFSLVLv.C:110-254) with the default_threshold switch
 * (api.h:150-153), csr_sparse_with_f push (dsl/edgemap.h:1397-1740) and
 * csc pull with early exit (dsl/emap/GraphCSx_csc.h:50-140). */
void orc_graptor_bfs(uint32_t n, uint64_t m, const uint64_t *index,
                     const uint32_t *nbrs, uint32_t start, uint32_t *level,
                     int threads, orc_run_info *info);
/* CCv (bench/CCv.C:221-588) with frac_threshold(5) and PRESET_ZERO. */
void orc_graptor_cc(uint32_t n, uint64_t m, const uint64_t *index,
                    const uint32_t *nbrs, uint32_t *label, int threads,
                    orc_run_info *info);
/* PRv (bench/PRv.C:115-290): FP32 storage, FP64 accumulation. */
int orc_graptor_pr(uint32_t n, const uint64_t *cindex, const uint32_t *cnbrs,
                   const uint32_t *outdeg, int max_iter, double tol,
                   float *x, int threads, double *delta_out);

/* ---- custom floating point (include/graptor/customfp.h:20-211) -------- */
/* customfp<6,10>: no sign, low 6 exponent bits, top 10 mantissa bits,
 * truncating.  Used by PR_vary variant 46 (bench/PR_vary.C:509-520). */
uint16_t orc_cfp16_encode(float f);
float orc_cfp16_decode(uint16_t h);
/* scustomfp<true,8,7,false,0>: the top half of a float, truncating
 * (bench/PR_vary.C:554-565, variant 99). */
uint16_t orc_bf16t_encode(float f);
/* bfloat16 with round-to-nearest-even (the GPU's conversion). */
uint16_t orc_bf16_rne_encode(float f);
float orc_bf16_decode(uint16_t h);

/* scustomfp<true,6,9,true,0>: sign, low 6 exponent bits, top 9 mantissa
 * bits, truncating; code 0 decodes to 0 (bench/APR_vary.C:743-750). */
uint16_t orc_scfp16_encode(float f);
float orc_scfp16_decode(uint16_t h);

/* APR_vary (bench/APR_vary.C:378-640, :900-1074): delta PageRank over
 * stages of (contribution format, eps); x_out is scaled to unit L1 norm.
 * Returns the iteration count, -1 for a format without a sign. */
int orc_apr_vary(uint32_t n, uint64_t m, const uint64_t *cindex, const uint32_t *cnbrs,
                 const uint32_t *outdeg, const int *fmts, const double *eps, int nstages,
                 int max_iter, double eps2, double *x_out, int threads, double *l1_out,
                 uint64_t *edges_out, double *trace);

/* Core numbers of a symmetric graph (bench/KC_bucket.C:104-330 computes
 * them by bucketed peeling); returns the largest. */
uint32_t orc_kcore(uint32_t n, const uint64_t *index, const uint32_t *nbrs, uint32_t *core);

#ifdef __cplusplus
}
#endif
#endif
