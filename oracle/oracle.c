/*
 * oracle.c -- CPU restatement of Graptor's traversal hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the
 * cpu_baseline leg of bench.py as the checker or the timed CPU baseline.
 * The product never includes, links or calls it.
 *
 * Each function cites the reference file:line it restates.  Parity pins
 * (how each function is itself checked) are listed at the head of
 * tests/test_oracle.py.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

/* ------------------------------------------------------------------------
 * Counter-based random numbers.  The reference ships no generator
 * (no rmat/kronecker/erdos code in the repository), so the stream below is
 * this project's own specification, shared bit-for-bit with the GPU
 * generator: splitmix64's finaliser applied to (edge index, word index).
 * ---------------------------------------------------------------------- */
uint64_t orc_mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* Thresholds are 64-bit so that a probability of 1 maps to 2^32 and
 * every 32-bit draw falls below it. */
void orc_rmat_thresholds(double a, double b, double c, uint64_t t[3]) {
    t[0] = (uint64_t)(a * 4294967296.0);
    t[1] = (uint64_t)((a + b) * 4294967296.0);
    t[2] = (uint64_t)((a + b + c) * 4294967296.0);
}

/* RMAT: for each of `scale` levels (most significant bit first) pick a
 * quadrant with probabilities a (0,0), b (0,1), c (1,0), d (1,1).  Level l
 * uses half (l & 1) of word h(e, l/2). */
void orc_rmat_edges(uint32_t scale, const uint64_t t[3], uint64_t seed,
                    uint64_t first, uint64_t count,
                    uint32_t *src, uint32_t *dst) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        uint64_t e = first + (uint64_t)i;
        uint32_t u = 0, v = 0;
        uint64_t h = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            if ((l & 1) == 0)
                h = orc_mix64(seed ^ orc_mix64((e << 5) | (l >> 1)));
            uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
            uint32_t q = r < t[0] ? 0u : r < t[1] ? 1u : r < t[2] ? 2u : 3u;
            u = (u << 1) | (q >> 1);
            v = (v << 1) | (q & 1u);
        }
        src[i] = u;
        dst[i] = v;
    }
}

void orc_er_edges(uint32_t n, uint64_t seed, uint64_t count,
                  uint32_t *src, uint32_t *dst) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        uint64_t h = orc_mix64(seed ^ orc_mix64((uint64_t)i));
        src[i] = (uint32_t)(((h & 0xffffffffULL) * (uint64_t)n) >> 32);
        dst[i] = (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
    }
}

/* ---- batches of small graphs (the batched CC workload) -------------------
 * Graph g of a batch: n_g log-uniform in [nmin, nmax] from
 * u = (mix64(seed ^ mix64(0x517e0000 + g)) >> 11) * 2^-53,
 * n_g = floor(exp(log nmin + u (log nmax - log nmin))); graph seed
 * gs = mix64(seed ^ mix64(0x5eed0000 + g)); even g: RMAT (a=.57 b=.19
 * c=.19) at scale ceil(log2 n_g), 8 n_g draws keyed by gs, endpoints >= n_g
 * dropped; odd g: 3 n_g uniform pairs h = mix64(gs ^ mix64(e)). */
void orc_batch_sizes(uint32_t count, uint32_t nmin, uint32_t nmax, uint64_t seed, uint32_t *nv) {
    const double l0 = log((double)nmin), l1 = log((double)nmax);
    for (uint32_t g = 0; g < count; ++g) {
        const double u = (double)(orc_mix64(seed ^ orc_mix64(0x517e0000ULL + g)) >> 11) * (1.0 / 9007199254740992.0);
        double x = floor(exp(l0 + u * (l1 - l0)));
        if (x < nmin) x = nmin;
        if (x > nmax) x = nmax;
        nv[g] = (uint32_t)x;
    }
}

/* Local CSR of graph g (index: nv+1 entries, nbrs: capacity 16*nv). */
uint64_t orc_batch_graph(uint32_t g, uint32_t nv, uint64_t seed, uint64_t *index, uint32_t *nbrs) {
    const uint64_t gs = orc_mix64(seed ^ orc_mix64(0x5eed0000ULL + g));
    const uint64_t cnt = (g & 1) == 0 ? 8ULL * nv : 3ULL * nv;
    uint32_t *src = (uint32_t *)malloc(sizeof(uint32_t) * cnt);
    uint32_t *dst = (uint32_t *)malloc(sizeof(uint32_t) * cnt);
    uint64_t k = 0;
    if ((g & 1) == 0) {
        uint32_t sc = 1;
        while ((1u << sc) < nv) ++sc;
        uint64_t t[3];
        orc_rmat_thresholds(0.57, 0.19, 0.19, t);
        for (uint64_t e = 0; e < cnt; ++e) {
            uint32_t u = 0, v = 0;
            uint64_t h = 0;
            for (uint32_t l = 0; l < sc; ++l) {
                if ((l & 1) == 0) h = orc_mix64(gs ^ orc_mix64((e << 5) | (l >> 1)));
                uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
                uint32_t q = r < t[0] ? 0u : r < t[1] ? 1u : r < t[2] ? 2u : 3u;
                u = (u << 1) | (q >> 1);
                v = (v << 1) | (q & 1u);
            }
            if (u < nv && v < nv) { src[k] = u; dst[k] = v; ++k; }
        }
    } else {
        for (uint64_t e = 0; e < cnt; ++e) {
            uint64_t h = orc_mix64(gs ^ orc_mix64(e));
            src[k] = (uint32_t)(((h & 0xffffffffULL) * (uint64_t)nv) >> 32);
            dst[k] = (uint32_t)(((h >> 32) * (uint64_t)nv) >> 32);
            ++k;
        }
    }
    uint64_t m = orc_build_csr(nv, k, src, dst, 1, index, nbrs);
    free(src);
    free(dst);
    return m;
}

/* LSD radix sort of 64-bit keys, 16 bits per pass, passes with a single
 * populated bucket skipped. */
static void radix_sort_u64(uint64_t *a, uint64_t *tmp, size_t len) {
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * 65536);
    for (int pass = 0; pass < 4; ++pass) {
        int sh = pass * 16;
        memset(cnt, 0, sizeof(size_t) * 65536);
        for (size_t i = 0; i < len; ++i) cnt[(a[i] >> sh) & 0xffff]++;
        int used = 0;
        for (int k = 0; k < 65536; ++k) used += cnt[k] != 0;
        if (used <= 1) continue;
        size_t s = 0;
        for (int k = 0; k < 65536; ++k) { size_t c = cnt[k]; cnt[k] = s; s += c; }
        for (size_t i = 0; i < len; ++i) tmp[cnt[(a[i] >> sh) & 0xffff]++] = a[i];
        memcpy(a, tmp, sizeof(uint64_t) * len);
    }
    free(cnt);
}

uint64_t orc_build_csr(uint32_t n, uint64_t count, const uint32_t *src,
                       const uint32_t *dst, int symmetrize,
                       uint64_t *index, uint32_t *nbrs) {
    size_t cap = symmetrize ? 2 * count : count;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (cap ? cap : 1));
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (cap ? cap : 1));
    size_t k = 0;
    for (uint64_t i = 0; i < count; ++i) {
        uint32_t u = src[i], v = dst[i];
        if (u == v) continue;                     /* CleanGraph --rm-self */
        keys[k++] = ((uint64_t)u << 32) | v;
        if (symmetrize) keys[k++] = ((uint64_t)v << 32) | u; /* CvtToUndir */
    }
    radix_sort_u64(keys, tmp, k);
    uint64_t m = 0;
    for (uint32_t v = 0; v <= n; ++v) index[v] = 0;
    for (size_t i = 0; i < k; ++i) {
        if (i > 0 && keys[i] == keys[i - 1]) continue;   /* RmDupEdges */
        uint32_t u = (uint32_t)(keys[i] >> 32);
        nbrs[m++] = (uint32_t)keys[i];
        index[u + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) index[v + 1] += index[v];
    free(keys);
    free(tmp);
    return m;
}

/* GraphCSx.h:2009-2027: header {2,1,n,m,sizeof(VID),sizeof(EID),0,0},
 * n index entries (index[n] implied), m neighbours. */
int orc_b2_write(const char *path, uint32_t n, uint64_t m,
                 const uint64_t *index, const uint32_t *nbrs) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    uint64_t hdr[8] = {2, 1, n, m, 4, 8, 0, 0};
    int ok = fwrite(hdr, sizeof(hdr), 1, f) == 1
        && fwrite(index, sizeof(uint64_t), n, f) == n
        && fwrite(nbrs, sizeof(uint32_t), m, f) == m;
    fclose(f);
    return ok ? 0 : -1;
}

/* GraphCSx.h:2222-2260: accept versions 2 and 3; version 3 stores the
 * (n+1)-th index entry, which is skipped (:2255-2256); index[n]=m (:2253). */
int orc_b2_read(const char *path, uint32_t *n, uint64_t *m,
                uint64_t *index, uint32_t *nbrs) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    uint64_t hdr[8];
    if (fread(hdr, sizeof(hdr), 1, f) != 1) { fclose(f); return -2; }
    if ((hdr[0] != 2 && hdr[0] != 3) || hdr[4] != 4 || hdr[5] != 8
        || hdr[2] > 0xffffffffULL) { fclose(f); return -3; }
    *n = (uint32_t)hdr[2];
    *m = hdr[3];
    if (!index) { fclose(f); return 0; }
    int ok = fread(index, sizeof(uint64_t), *n, f) == *n;
    index[*n] = *m;
    if (ok && hdr[0] == 3) ok = fseek(f, 8, SEEK_CUR) == 0;
    if (ok) ok = fread(nbrs, sizeof(uint32_t), *m, f) == *m;
    fclose(f);
    return ok ? 0 : -2;
}

int orc_check_csr(uint32_t n, uint64_t m, const uint64_t *index,
                  const uint32_t *nbrs) {
    if (index[0] != 0 || index[n] != m) return 1;
    for (uint32_t v = 0; v < n; ++v) {
        if (index[v + 1] < index[v]) return 2;
        for (uint64_t e = index[v]; e < index[v + 1]; ++e) {
            if (nbrs[e] >= n) return 3;
            if (e > index[v] && nbrs[e] <= nbrs[e - 1]) return 4;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
void orc_bfs_levels(uint32_t n, const uint64_t *index, const uint32_t *nbrs,
                    uint32_t start, uint32_t *level) {
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t v = 0; v < n; ++v) level[v] = n + 1;
    size_t head = 0, tail = 0;
    level[start] = 0;
    q[tail++] = start;
    while (head < tail) {
        uint32_t u = q[head++];
        for (uint64_t e = index[u]; e < index[u + 1]; ++e) {
            uint32_t v = nbrs[e];
            if (level[v] == n + 1) { level[v] = level[u] + 1; q[tail++] = v; }
        }
    }
    free(q);
}

static uint32_t uf_find(uint32_t *p, uint32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

void orc_cc_labels(uint32_t n, const uint64_t *index, const uint32_t *nbrs,
                   uint32_t *label) {
    for (uint32_t v = 0; v < n; ++v) label[v] = v;
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t e = index[u]; e < index[u + 1]; ++e) {
            uint32_t a = uf_find(label, u), b = uf_find(label, nbrs[e]);
            if (a < b) label[b] = a; else if (b < a) label[a] = b;
        }
    for (uint32_t v = 0; v < n; ++v) label[v] = uf_find(label, v);
}

/* PRv.C:176-269 in FP64 throughout. */
int orc_pr_fp64(uint32_t n, const uint64_t *cindex, const uint32_t *cnbrs,
                const uint32_t *outdeg, int max_iter, double tol,
                double damping, double *x, double *delta_out) {
    double *y = (double *)malloc(sizeof(double) * n);
    double *contrib = (double *)malloc(sizeof(double) * n);
    for (uint32_t v = 0; v < n; ++v) {
        x[v] = 1.0 / (double)n;
        contrib[v] = outdeg[v] ? damping * x[v] / (double)outdeg[v] : 0.0;
    }
    int iter = 0;
    double delta = 2.0;
    while (iter < max_iter && delta > tol) {
        double s = 0.0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : s)
        for (int64_t d = 0; d < (int64_t)n; ++d) {
            double acc = 0.0;
            for (uint64_t e = cindex[d]; e < cindex[d + 1]; ++e)
                acc += contrib[cnbrs[e]];
            y[d] = acc;
            s += acc;
        }
        double w = (1.0 - s) / (double)n;
        double s2 = 0.0;
        for (uint32_t v = 0; v < n; ++v) { y[v] += w; s2 += y[v]; }
        double inv = 1.0 / s2;
        delta = 0.0;
        for (uint32_t v = 0; v < n; ++v) {
            y[v] *= inv;
            delta += fabs(x[v] - y[v]);
            x[v] = y[v];
            contrib[v] = outdeg[v] ? damping * x[v] / (double)outdeg[v] : 0.0;
        }
        ++iter;
    }
    free(y);
    free(contrib);
    if (delta_out) *delta_out = delta;
    return iter;
}

/* ------------------------------------------------------------------------
 * Restatement of Graptor's engine (OpenMP).  Sparse push: per-thread
 * output buffers, CAS-loop atomic min (simd/vector_ref.h:205-225), dedup
 * with a byte flag array set by fetch-or (dsl/emap/edgechunk.h:389-392),
 * flags reset afterwards (dsl/edgemap.h:1698-1703).  Dense pull: scalar
 * loop over in-edges with the destination filter re-checked per edge
 * (dsl/emap/GraphCSx_csc.h:110-111).
 * ---------------------------------------------------------------------- */
static int cas_min_u32(uint32_t *p, uint32_t val) {
    uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (val < old) {
        if (__atomic_compare_exchange_n(p, &old, val, 1, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED))
            return 1;
    }
    return 0;
}

typedef struct {
    uint32_t *list;
    uint32_t nactv;
    uint64_t nacte;
} orc_frontier;

/* Push step over a sparse frontier.  kind 0 = BFS level (relax
 * level[d].min(level[s]+1), dst filter level[d] > iter+1), kind 1 = CC
 * (relax IDs[d].min(IDs[s]), dst filter IDs[d] != 0). */
static void push_step(int kind, uint32_t n, const uint64_t *index,
                      const uint32_t *nbrs, uint32_t *val, uint32_t iter,
                      const orc_frontier *F, orc_frontier *out,
                      unsigned char *zf, uint32_t *scratch, int threads,
                      uint64_t *edges) {
    uint32_t total = 0;
    uint64_t nacte = 0, ecount = 0;
    uint32_t *offs = (uint32_t *)calloc((size_t)threads + 1, sizeof(uint32_t));
#pragma omp parallel num_threads(threads) reduction(+ : nacte, ecount)
    {
        int t = omp_get_thread_num();
        int nt = omp_get_num_threads();
        /* private segment of the scratch list, sized by edge share */
        uint32_t *mine = NULL;
        size_t cnt = 0, cap = 1024;
        mine = (uint32_t *)malloc(sizeof(uint32_t) * cap);
#pragma omp for schedule(dynamic, 64)
        for (int64_t j = 0; j < (int64_t)F->nactv; ++j) {
            uint32_t s = F->list[j];
            for (uint64_t e = index[s]; e < index[s + 1]; ++e) {
                uint32_t d = nbrs[e];
                ++ecount;
                uint32_t nv;
                if (kind == 0) {
                    if (!(__atomic_load_n(&val[d], __ATOMIC_RELAXED) > iter + 1)) continue;
                    nv = __atomic_load_n(&val[s], __ATOMIC_RELAXED) + 1;
                } else {
                    if (__atomic_load_n(&val[d], __ATOMIC_RELAXED) == 0) continue;
                    nv = __atomic_load_n(&val[s], __ATOMIC_RELAXED);
                }
                if (cas_min_u32(&val[d], nv)) {
                    if (zf[d] == 0 && __atomic_fetch_or(&zf[d], 1, __ATOMIC_RELAXED) == 0) {
                        if (cnt == cap) { cap *= 2; mine = (uint32_t *)realloc(mine, sizeof(uint32_t) * cap); }
                        mine[cnt++] = d;
                        nacte += index[d + 1] - index[d];
                    }
                }
            }
        }
        offs[t + 1] = (uint32_t)cnt;
#pragma omp barrier
#pragma omp single
        {
            for (int i = 0; i < nt; ++i) offs[i + 1] += offs[i];
            total = offs[nt];
        }
        memcpy(scratch + offs[t], mine, sizeof(uint32_t) * cnt);
        for (size_t i = 0; i < cnt; ++i) zf[mine[i]] = 0;
        free(mine);
    }
    free(offs);
    out->list = scratch;
    out->nactv = total;
    out->nacte = nacte;
    *edges += ecount;
}

/* Pull step over the CSC (symmetric graph: CSC = CSR).  The frontier is
 * a dense byte array.  Records vertices whose value differs from prev. */
static void pull_step(int kind, uint32_t n, const uint64_t *index,
                      const uint32_t *nbrs, uint32_t *val, uint32_t *prev,
                      uint32_t iter, const unsigned char *fb,
                      orc_frontier *out, uint32_t *scratch, int threads,
                      uint64_t *edges) {
    uint64_t nacte = 0, ecount = 0;
    uint32_t total = 0;
    uint32_t *offs = (uint32_t *)calloc((size_t)threads + 1, sizeof(uint32_t));
#pragma omp parallel num_threads(threads) reduction(+ : nacte, ecount)
    {
        int t = omp_get_thread_num();
        int nt = omp_get_num_threads();
        int64_t lo = (int64_t)n * t / nt, hi = (int64_t)n * (t + 1) / nt;
        size_t cnt = 0;
        for (int64_t d = lo; d < hi; ++d) {
            uint32_t cur = __atomic_load_n(&val[d], __ATOMIC_RELAXED);
            int active = kind == 0 ? cur > iter + 1 : cur != 0;
            for (uint64_t e = index[d]; active && e < index[d + 1]; ++e) {
                uint32_t s = nbrs[e];
                ++ecount;
                if (!fb[s]) continue;
                uint32_t nv = __atomic_load_n(&val[s], __ATOMIC_RELAXED) + (kind == 0);
                if (nv < cur) cur = nv;
                active = kind == 0 ? cur > iter + 1 : cur != 0;
            }
            __atomic_store_n(&val[d], cur, __ATOMIC_RELAXED);
        }
#pragma omp barrier
        for (int64_t d = lo; d < hi; ++d)
            if (val[d] != prev[d]) {
                scratch[lo + cnt++] = (uint32_t)d;
                nacte += index[d + 1] - index[d];
                prev[d] = val[d];
            }
        offs[t + 1] = (uint32_t)cnt;
#pragma omp barrier
#pragma omp single
        {
            for (int i = 0; i < nt; ++i) offs[i + 1] += offs[i];
            total = offs[nt];
        }
        /* compact in place: segments move left only, in thread order */
        for (int i = 0; i < nt; ++i) {
#pragma omp barrier
            if (i == t && offs[t] != (uint32_t)lo)
                memmove(scratch + offs[t], scratch + lo, sizeof(uint32_t) * cnt);
        }
    }
    free(offs);
    out->list = scratch;
    out->nactv = total;
    out->nacte = nacte;
    *edges += ecount;
}

static void run_levels(int kind, uint32_t n, uint64_t m, const uint64_t *index,
                       const uint32_t *nbrs, uint32_t *val, orc_frontier F,
                       int threads, int threshold_pct, orc_run_info *info) {
    uint32_t *prev = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *bufA = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *bufB = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    unsigned char *zf = (unsigned char *)calloc(n ? n : 1, 1);
    unsigned char *fb = (unsigned char *)calloc(n ? n : 1, 1);
    memcpy(prev, val, sizeof(uint32_t) * n);
    if (F.list) { memcpy(bufA, F.list, sizeof(uint32_t) * F.nactv); }
    else { for (uint32_t v = 0; v < n; ++v) bufA[v] = v; F.nactv = n; F.nacte = m; }
    F.list = bufA;
    uint32_t iter = 0;
    info->iters = 0; info->push_iters = 0; info->edges = 0;
    while (F.nactv > 0) {
        orc_frontier out;
        uint32_t *dst = F.list == bufA ? bufB : bufA;
        /* frac_threshold(t): (nactv + nacte) <= m / (100/t)   api.h:130-139 */
        int sparse = (float)((uint64_t)F.nactv + F.nacte)
                     <= (float)m / (100.0f / (float)threshold_pct);
        if (sparse) {
            push_step(kind, n, index, nbrs, val, iter, &F, &out, zf, dst,
                      threads, &info->edges);
            for (uint32_t i = 0; i < out.nactv; ++i) prev[out.list[i]] = val[out.list[i]];
            info->push_iters++;
        } else {
            memset(fb, 0, n);
            for (uint32_t i = 0; i < F.nactv; ++i) fb[F.list[i]] = 1;
            pull_step(kind, n, index, nbrs, val, prev, iter, fb, &out, dst,
                      threads, &info->edges);
        }
        F = out;
        ++iter;
        info->iters++;
    }
    free(prev); free(bufA); free(bufB); free(zf); free(fb);
}

void orc_graptor_bfs(uint32_t n, uint64_t m, const uint64_t *index,
                     const uint32_t *nbrs, uint32_t start, uint32_t *level,
                     int threads, orc_run_info *info) {
    /* BFSLVLv.C:117-129 initialisation */
    for (uint32_t v = 0; v < n; ++v) level[v] = n + 1;
    level[start] = 0;
    orc_frontier F;
    uint32_t s = start;
    F.list = &s; F.nactv = 1; F.nacte = index[start + 1] - index[start];
    run_levels(0, n, m, index, nbrs, level, F, threads, 5, info);
}

void orc_graptor_cc(uint32_t n, uint64_t m, const uint64_t *index,
                    const uint32_t *nbrs, uint32_t *label, int threads,
                    orc_run_info *info) {
    /* CCv.C:246-248 IDs[v]=v; PRESET_ZERO CCv.C:262-291 */
    for (uint32_t v = 0; v < n; ++v) label[v] = v;
    if (n > 0)
        for (uint64_t e = index[0]; e < index[1]; ++e) label[nbrs[e]] = 0;
    orc_frontier F;
    F.list = NULL; F.nactv = n; F.nacte = m;
    run_levels(1, n, m, index, nbrs, label, F, threads, 5, info);
}

/* PRv.C:115-290: FP32 storage with FP64 register accumulation (:127). */
int orc_graptor_pr(uint32_t n, const uint64_t *cindex, const uint32_t *cnbrs,
                   const uint32_t *outdeg, int max_iter, double tol,
                   float *x, int threads, double *delta_out) {
    float *y = (float *)malloc(sizeof(float) * n);
    float *contrib = (float *)malloc(sizeof(float) * n);
    const float damp = 0.85f;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        x[v] = 1.0f / (float)n;
        contrib[v] = outdeg[v] ? damp * x[v] / (float)outdeg[v] : 0.0f;
    }
    int iter = 0;
    double sdelta = 2.0;
    while (iter < max_iter && sdelta > tol) {
        double s = 0.0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096) reduction(+ : s)
        for (int64_t d = 0; d < (int64_t)n; ++d) {
            double acc = 0.0;
            for (uint64_t e = cindex[d]; e < cindex[d + 1]; ++e)
                acc += (double)contrib[cnbrs[e]];
            y[d] = (float)acc;
            s += (double)y[d];
        }
        double w = (1.0 - s) / (double)n;
        double s2 = 0.0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : s2)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            y[v] = (float)((double)y[v] + w);
            s2 += (double)y[v];
        }
        float inv = 1.0f / (float)s2;
        double dl = 0.0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : dl)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            float yv = y[v] * inv;
            dl += (double)fabsf(x[v] - yv);
            x[v] = yv;
            contrib[v] = outdeg[v] ? damp * x[v] / (float)outdeg[v] : 0.0f;
        }
        sdelta = dl;
        ++iter;
    }
    free(y);
    free(contrib);
    if (delta_out) *delta_out = sdelta;
    return iter;
}

/* ---------------------------------------------------------------------- */
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* customfp.h:95-117 with S=false, E=6, M=10, B=0: shift the float right by
 * 23-M and keep E+M bits. */
uint16_t orc_cfp16_encode(float f) {
    return (uint16_t)((f2u(f) >> 13) & 0xffffu);
}
/* customfp.h:141-169: OR in the missing exponent bits
 * ((1 << (8-E-1)) - 1) << (E+M) and shift left by 23-M. */
float orc_cfp16_decode(uint16_t h) {
    uint32_t t = (uint32_t)h | (1u << 16);
    return u2f(t << 13);
}
/* scustomfp<true,8,7,false,0>: keep the upper half, no rounding. */
uint16_t orc_bf16t_encode(float f) { return (uint16_t)(f2u(f) >> 16); }
uint16_t orc_bf16_rne_encode(float f) {
    uint32_t u = f2u(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float orc_bf16_decode(uint16_t h) { return u2f((uint32_t)h << 16); }

/* scustomfp<true,6,9,true,0> (bench/APR_vary.C:743-750, variant 46):
 * encode keeps the sign, the low 6 exponent bits and the top 9 mantissa
 * bits, truncating (customfp.h:95-117; the vector narrowing,
 * target/conversion.h:467-497, gives the same bits).  Decode restores the
 * exponent bits 01 above the six (customfp.h:141-169) and, as the vector
 * widening does for a format that may hold zero (target/conversion.h:
 * 150-189), decodes code 0 to 0. */
uint16_t orc_scfp16_encode(float f) {
    uint32_t u = f2u(f);
    return (uint16_t)(((u >> 14) & 0x7fffu) | ((u >> 16) & 0x8000u));
}
float orc_scfp16_decode(uint16_t h) {
    if (h == 0) return 0.0f;
    uint32_t t = ((uint32_t)(h & 0x8000u) << 16) | 0x20000000u | ((uint32_t)(h & 0x7fffu) << 14);
    return u2f(t);
}

/* Contribution formats of the delta PageRank: FP32 (accumulated in FP64),
 * bfloat16 RNE, bfloat16 truncating, scustomfp<true,6,9> (FP32
 * accumulation).  Returns -1 for formats that cannot hold a sign. */
static int apr_fmt_ok(int fmt) { return fmt == 0 || fmt == 1 || fmt == 4 || fmt == 5; }
static uint32_t apr_enc(int fmt, float v) {
    switch (fmt) {
    case 0: return f2u(v);
    case 1: return orc_bf16_rne_encode(v);
    case 4: return orc_bf16t_encode(v);
    default: return orc_scfp16_encode(v);
    }
}
static float apr_dec(int fmt, uint32_t c) {
    switch (fmt) {
    case 0: return u2f(c);
    case 1:
    case 4: return orc_bf16_decode((uint16_t)c);
    default: return orc_scfp16_decode((uint16_t)c);
    }
}

/* APR_vary (bench/APR_vary.C): delta PageRank.  run() (:900-1074) starts
 * from x = 0, delta = 1/n and runs the stages in order (variant 2046 is
 * {FP32 contributions with FP64 accumulation, eps 1e-1} then
 * {scustomfp<true,6,9>, eps 1e-7}), all sharing max_iter, the frontier,
 * x and delta; the result is x scaled to unit L1 norm (:1060-1072).
 * One stage is run_prec (:378-640):
 *   factor = 10^-(digits of n) in FP32 (:436-448), outdeg = 1/(factor*deg);
 *   iteration 0 (:459-505): contrib = delta*outdeg over all vertices,
 *     z = sum over in-edges of contrib*factor, delta = d*z + (1-d)/n,
 *     x += delta, delta -= 1/n, active = |delta| > eps2*x;
 *   later iterations: over every source when the frontier's density
 *     (nactv+nacte)/m exceeds 0.6 (:507-558), over the frontier's sources
 *     only otherwise (:560-612; frac_threshold(1) picks push at <= 1%),
 *     delta = z*d, active = |delta| > eps2*x, x += delta where active;
 *   L1 = sum |delta| (:632-639), frontier = active, stop when L1 < eps.
 * Every vertex operation is FP32 as FloatTy is; z accumulates in FP64 for
 * FP32 contributions (variant 20) and FP32 otherwise (variant 46), in CSC
 * order; L1 is summed in FP64.  trace (4 doubles per iteration, may be
 * NULL): kind (0 first, 1 all sources, 2 frontier sources by pull, 3 by
 * push), nactv, nacte of the new frontier, L1. */
int orc_apr_vary(uint32_t n, uint64_t m, const uint64_t *cindex, const uint32_t *cnbrs,
                 const uint32_t *outdeg, const int *fmts, const double *eps, int nstages,
                 int max_iter, double eps2, double *x_out, int threads, double *l1_out,
                 uint64_t *edges_out, double *trace) {
    for (int s = 0; s < nstages; ++s)
        if (!apr_fmt_ok(fmts[s])) return -1;
    if (n == 0) return 0;
    float *x = (float *)malloc(sizeof(float) * n);
    float *delta = (float *)malloc(sizeof(float) * n);
    float *zf = (float *)malloc(sizeof(float) * n);
    uint32_t *cb = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint8_t *act = (uint8_t *)malloc(n);
    uint8_t *nact = (uint8_t *)malloc(n);
    const float dampen = 0.85f;
    float factor = 1.0f;
    for (uint32_t nnn = n; nnn > 1; nnn /= 10) factor = (float)((double)factor / 10.0);
    const float ninv = 1.0f / (float)n;
    const float cst = (1.0f - dampen) * ninv;
    const float e2 = (float)eps2;
    const float d0 = (float)(1.0 / (double)n);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        x[v] = 0.0f;
        delta[v] = d0;
        act[v] = 1;
    }
    uint64_t fv = n, fe = m, edges = 0;
    int iter = 0;
    double l1 = 0.0;
    for (int st = 0; st < nstages; ++st) {
        const int fmt = fmts[st];
        while (iter < max_iter) {
            int kind;
            if (iter == 0) kind = 0;
            else if (((double)fv + (double)fe) / (double)m > 0.6) kind = 1;
            else kind = (double)(float)(fv + fe) <= (double)(float)m / 100.0 ? 3 : 2;
            const int all = kind <= 1;
            edges += all ? m : fe;
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int64_t v = 0; v < (int64_t)n; ++v)
                if ((all || act[v]) && outdeg[v]) {
                    const float od = 1.0f / (factor * (float)outdeg[v]);
                    cb[v] = apr_enc(fmt, delta[v] * od);
                }
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
            for (int64_t d = 0; d < (int64_t)n; ++d) {
                if (fmt == 0) {
                    double acc = 0.0;
                    for (uint64_t e = cindex[d]; e < cindex[d + 1]; ++e) {
                        const uint32_t s = cnbrs[e];
                        if (all || act[s]) acc += (double)apr_dec(fmt, cb[s]) * (double)factor;
                    }
                    zf[d] = (float)acc;
                } else {
                    float acc = 0.0f;
                    for (uint64_t e = cindex[d]; e < cindex[d + 1]; ++e) {
                        const uint32_t s = cnbrs[e];
                        if (all || act[s]) acc += apr_dec(fmt, cb[s]) * factor;
                    }
                    zf[d] = acc;
                }
            }
            double sum = 0.0;
            uint64_t cv = 0, ce = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : sum, cv, ce)
            for (int64_t v = 0; v < (int64_t)n; ++v) {
                float dl;
                int a;
                if (kind == 0) {
                    dl = dampen * zf[v] + cst;
                    x[v] = x[v] + dl;
                    dl = dl - ninv;
                    a = fabsf(dl) > e2 * x[v];
                } else {
                    dl = zf[v] * dampen;
                    a = fabsf(dl) > e2 * x[v];
                    if (a) x[v] = x[v] + dl;
                }
                delta[v] = dl;
                nact[v] = (uint8_t)a;
                sum += (double)fabsf(dl);
                cv += (uint64_t)a;
                ce += a ? outdeg[v] : 0;
            }
            memcpy(act, nact, n);
            fv = cv;
            fe = ce;
            l1 = sum;
            if (trace) {
                trace[4 * iter] = kind;
                trace[4 * iter + 1] = (double)cv;
                trace[4 * iter + 2] = (double)ce;
                trace[4 * iter + 3] = sum;
            }
            ++iter;
            if (l1 < eps[st]) break;
        }
    }
    double ss = 0.0;
    for (uint32_t v = 0; v < n; ++v) ss += fabs((double)x[v]);
    const double sc = 1.0 / ss;
    for (uint32_t v = 0; v < n; ++v) x_out[v] = (double)x[v] * sc;
    free(x);
    free(delta);
    free(zf);
    free(cb);
    free(act);
    free(nact);
    if (l1_out) *l1_out = l1;
    if (edges_out) *edges_out = edges;
    return iter;
}

/* k-core numbers (KC_bucket, bench/KC_bucket.C:104-330): the output of the
 * reference's bucketed peeling is the core number of every vertex, which is
 * unique; it is computed here by the textbook O(n+m) bucket algorithm
 * (Batagelj and Zaversnik, 2003): vertices sorted by current degree, each
 * taken in order lowers the degree of its higher-degree neighbours by one.
 * Symmetric graphs only.  Returns the largest core number (largestCore,
 * KC_bucket.C:228). */
uint32_t orc_kcore(uint32_t n, const uint64_t *index, const uint32_t *nbrs, uint32_t *core) {
    if (n == 0) return 0;
    uint32_t md = 0;
    for (uint32_t v = 0; v < n; ++v) {
        core[v] = (uint32_t)(index[v + 1] - index[v]);
        if (core[v] > md) md = core[v];
    }
    uint32_t *bin = (uint32_t *)calloc((size_t)md + 2, sizeof(uint32_t));
    uint32_t *pos = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *vert = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t v = 0; v < n; ++v) bin[core[v]]++;
    uint32_t start = 0;
    for (uint32_t d = 0; d <= md; ++d) {
        uint32_t num = bin[d];
        bin[d] = start;
        start += num;
    }
    for (uint32_t v = 0; v < n; ++v) {
        pos[v] = bin[core[v]];
        vert[pos[v]] = v;
        bin[core[v]]++;
    }
    for (uint32_t d = md; d >= 1; --d) bin[d] = bin[d - 1];
    bin[0] = 0;
    uint32_t kmax = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = vert[i];
        if (core[v] > kmax) kmax = core[v];
        for (uint64_t e = index[v]; e < index[v + 1]; ++e) {
            const uint32_t u = nbrs[e];
            if (core[u] > core[v]) {
                const uint32_t du = core[u], pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                }
                bin[du]++;
                core[u]--;
            }
        }
    }
    free(bin);
    free(pos);
    free(vert);
    return kmax;
}
