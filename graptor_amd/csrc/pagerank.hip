// pagerank.hip -- the PageRank edgemap (relax y[d] += contrib[s]) and its
// fused vertex-map tail.
//
// Reference: PRv run loop bench/PRv.C:115-290 (edgemap :202-212, scans and
// maps :214-269), PR_vary run_prec bench/PR_vary.C:247-447 and its encodings
// (customfp.h:20-211, variants :449-648), dense pull over the CSC
// (dsl/emap/GraptorCSCDataParCached.h:60-302 / GraphCSx_csc.h:50-140).
//
// Per iteration, four launches on the handle's stream, no host round trip:
//   k_pr_pull    buckets of <= STREAM_ITEMS rows+edges: neighbour IDs are read
//                coalesced, contributions gathered into LDS, then each row is
//                summed from LDS (lane per short row, wave per row above 64
//                edges).  Rows above LONG_DEG are cut into LONG_CHUNK-edge
//                chunks, one workgroup each, summed by k_pr_long.
//   k_pr_reduce  deterministic FP64 sum of the workgroup partials of y and
//                the maximum of y/deg; the last workgroup derives w, 1/s and
//                the FP16 encode scale.
//   k_pr_update  x' = (y + w) / s, |x - x'| partials, x = x', contrib =
//                d*x'/deg encoded in the contribution format.
//   k_pr_reduce  (again) sums the residual partials into delta.
// FP32 contributions are accumulated in FP64 registers as PRv does (y_acc
// is double over FP32 storage, PRv.C:127); 16-bit contributions are
// accumulated in FP32.
#include "internal.hpp"
#include "device.hpp"
#include "pr_codec.hpp"
#include "grid_sync.hpp"

#include <math.h>

namespace gpr {

constexpr int PR_THREADS = 256;
constexpr int PR_LDS_EDGES = STREAM_ITEMS + (int)LONG_DEG;
constexpr uint32_t PR_MED = 64;
constexpr int PR_UNROLL = PR_LDS_EDGES / PR_THREADS;  // 16

// PR_UNROLL gathers of one lane, issued back to back: every load is
// unconditional (masked lanes read contrib[0]) and decoding follows all of
// them, so no load waits for the previous one.  Sources at or above
// cold_from (and the 0xffffffff padding) contribute 0.
template <int FMT>
__device__ __forceinline__ void gather_all(const typename Codec<FMT>::store_t *__restrict__ c, const uint32_t *s,
                                           uint32_t cold_from, float *v) {
    typename Codec<FMT>::store_t raw[PR_UNROLL];
#pragma unroll
    for (int u = 0; u < PR_UNROLL; ++u) raw[u] = c[s[u] < cold_from ? s[u] : 0u];
#pragma unroll
    for (int u = 0; u < PR_UNROLL; ++u) v[u] = s[u] < cold_from ? Codec<FMT>::dec(raw[u]) : 0.0f;
}

template <int FMT>
__global__ void __launch_bounds__(PR_THREADS)
k_pr_pull(uint32_t nbuckets, const uint32_t *__restrict__ bucket, const uint64_t *__restrict__ cindex,
          const uint32_t *__restrict__ cnbrs, const uint32_t *__restrict__ deg,
          const typename Codec<FMT>::store_t *__restrict__ contrib, float *__restrict__ y,
          const uint32_t *__restrict__ long_rows, const uint32_t *__restrict__ long_base, uint32_t nlong,
          double *__restrict__ long_partial, double *__restrict__ part, float *__restrict__ part_max,
          const PrScalars *__restrict__ prs, uint32_t cold_from) {
    using C = Codec<FMT>;
    using acc_t = typename C::acc_t;
    __shared__ float vals[PR_LDS_EDGES];
    __shared__ uint32_t rs[STREAM_ITEMS + 1];
    __shared__ uint32_t medq[PR_LDS_EDGES / PR_MED + 1];
    __shared__ uint32_t nmed;
    __shared__ double dred[PR_THREADS / 64];
    __shared__ float fred[PR_THREADS / 64];
    const float dscale = FMT == GPR_CONTRIB_F16 ? 1.0f / prs->enc_scale : 1.0f;
    const uint32_t b = blockIdx.x;
    if (b >= nbuckets) {
        // ---- one chunk of a long row
        const uint32_t c = b - nbuckets;
        uint32_t lo = 0, hi = nlong;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (long_base[mid] <= c) lo = mid + 1; else hi = mid;
        }
        const uint32_t li = lo - 1;
        const uint32_t row = long_rows[li];
        const uint64_t rb = cindex[row], re = cindex[row + 1];
        const uint64_t cb = rb + (uint64_t)(c - long_base[li]) * LONG_CHUNK;
        const uint64_t ce = min(re, cb + LONG_CHUNK);
        acc_t acc = 0;
        // PR_UNROLL independent neighbour loads, then as many gathers, in
        // flight per lane
        for (uint64_t base = cb; base < ce; base += (uint64_t)PR_THREADS * PR_UNROLL) {
            uint32_t s[PR_UNROLL];
#pragma unroll
            for (int u = 0; u < PR_UNROLL; ++u) {
                const uint64_t e = base + (uint64_t)u * PR_THREADS + threadIdx.x;
                s[u] = e < ce ? __builtin_nontemporal_load(&cnbrs[e]) : 0xffffffffu;
            }
            float v[PR_UNROLL];
            gather_all<FMT>(contrib, s, cold_from, v);
#pragma unroll
            for (int u = 0; u < PR_UNROLL; ++u) acc += (acc_t)v[u];
        }
        double tot = block_sum((double)acc, dred);
        if (threadIdx.x == 0) {
            long_partial[c] = tot;
            part[b] = 0.0;
            part_max[b] = 0.0f;
        }
        return;
    }
    // ---- a bucket of rows
    const uint32_t r0 = bucket[b], r1 = bucket[b + 1];
    double my_sum = 0.0;
    float my_max = 0.0f;
    if (r0 < r1) {
        uint32_t r1x = r1;
        if (cindex[r1] - cindex[r1 - 1] > LONG_DEG) r1x = r1 - 1;
        const uint64_t e0 = cindex[r0];
        const uint32_t nrows = r1x - r0;
        const uint32_t ne = (uint32_t)(cindex[r1x] - e0);
        if (threadIdx.x == 0) nmed = 0;
        for (uint32_t base = 0; base <= nrows; base += PR_THREADS * 8) {
            uint64_t ix[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t i = base + u * PR_THREADS + threadIdx.x;
                ix[u] = i <= nrows ? __builtin_nontemporal_load(&cindex[r0 + i]) : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t i = base + u * PR_THREADS + threadIdx.x;
                if (i <= nrows) rs[i] = (uint32_t)(ix[u] - e0);
            }
        }
        // a bucket holds < PR_LDS_EDGES edges: one round of PR_UNROLL
        // independent neighbour loads and gathers per lane covers it
        for (uint32_t base = 0; base < ne; base += PR_THREADS * PR_UNROLL) {
            uint32_t s[PR_UNROLL];
#pragma unroll
            for (int u = 0; u < PR_UNROLL; ++u) {
                const uint32_t i = base + u * PR_THREADS + threadIdx.x;
                s[u] = i < ne ? __builtin_nontemporal_load(&cnbrs[e0 + i]) : 0xffffffffu;
            }
            float v[PR_UNROLL];
            gather_all<FMT>(contrib, s, cold_from, v);
#pragma unroll
            for (int u = 0; u < PR_UNROLL; ++u) {
                const uint32_t i = base + u * PR_THREADS + threadIdx.x;
                if (i < ne) vals[i] = v[u];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nrows; i += PR_THREADS) {
            const uint32_t a = rs[i], z = rs[i + 1];
            if (z - a > PR_MED) {
                medq[atomicAdd(&nmed, 1u)] = i;
                continue;
            }
            acc_t acc = 0;
            for (uint32_t j = a; j < z; ++j) acc += (acc_t)vals[j];
            const float yv = (float)acc * dscale;
            y[r0 + i] = yv;
            my_sum += (double)yv;
            if (FMT == GPR_CONTRIB_F16) {
                const uint32_t dg = deg[r0 + i];
                if (dg) my_max = fmaxf(my_max, yv / (float)dg);
            }
        }
        __syncthreads();
        const uint32_t nm = nmed;
        const unsigned w = threadIdx.x / 64, ln = lane();
        for (uint32_t q = w; q < nm; q += PR_THREADS / 64) {
            const uint32_t i = medq[q];
            const uint32_t a = rs[i], z = rs[i + 1];
            acc_t acc = 0;
            for (uint32_t j = a + ln; j < z; j += 64) acc += (acc_t)vals[j];
            acc = wave_sum(acc);
            if (ln == 0) {
                const float yv = (float)acc * dscale;
                y[r0 + i] = yv;
                my_sum += (double)yv;
                if (FMT == GPR_CONTRIB_F16) {
                    const uint32_t dg = deg[r0 + i];
                    if (dg) my_max = fmaxf(my_max, yv / (float)dg);
                }
            }
        }
    }
    const double ts = block_sum(my_sum, dred);
    my_max = wave_max(my_max);
    __syncthreads();
    if (lane() == 0) fred[threadIdx.x / 64] = my_max;
    __syncthreads();
    if (threadIdx.x == 0) {
        float mx = fred[0];
        for (int i = 1; i < PR_THREADS / 64; ++i) mx = fmaxf(mx, fred[i]);
        part[b] = ts;
        part_max[b] = mx;
    }
}

// One thread per long row: fold its chunk partials in chunk order.
__global__ void k_pr_long(uint32_t nlong, const uint32_t *__restrict__ long_rows, const uint32_t *__restrict__ long_base,
                          const double *__restrict__ long_partial, const uint32_t *__restrict__ deg, float *y,
                          double *part, float *part_max, uint32_t slot0, const PrScalars *__restrict__ prs, int f16) {
    const float dscale = f16 ? 1.0f / prs->enc_scale : 1.0f;
    for (uint32_t li = blockIdx.x * blockDim.x + threadIdx.x; li < nlong; li += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (uint32_t c = long_base[li]; c < long_base[li + 1]; ++c) s += long_partial[c];
        const uint32_t row = long_rows[li];
        const float yv = (float)s * dscale;
        y[row] = yv;
        part[slot0 + li] = (double)yv;
        const uint32_t dg = deg[row];
        part_max[slot0 + li] = dg ? yv / (float)dg : 0.0f;
    }
}

// Deterministic reduction of `count` partial sums (and maxima).  mode 0:
// after the pull -- set sum_y, w, inv_s and the encode scale.  mode 1: after
// the update -- set delta and sum_x from the two partial arrays.  mode 2:
// maximum only.  mode 3: the cold-edge mass -- set the fixed-point unit.
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 64;
// Workgroups of a reduction: one for a few thousand partials (a launch of
// 64 and the last-workgroup fold cost more than the sums), RED_BLOCKS above
// that; the sums stay deterministic for a given count
static unsigned red_grid(uint32_t count) { return count <= 16384u ? 1u : (unsigned)RED_BLOCKS; }
__global__ void __launch_bounds__(RED_THREADS)
k_pr_reduce(const double *__restrict__ part, const float *__restrict__ part_max, const double *__restrict__ part2,
            uint32_t count, double *slices, unsigned int *ticket, PrScalars *prs, int mode, uint32_t n,
            float damping, int f16) {
    __shared__ double r1[RED_THREADS], r2[RED_THREADS];
    __shared__ float r3[RED_THREADS];
    __shared__ bool last;
    const uint32_t per = (count + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * per, hi = min(count, lo + per);
    double s1 = 0.0, s2 = 0.0;
    float mx = 0.0f;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += RED_THREADS) {
        s1 += part[i];
        if (mode == 0 || mode == 2) mx = fmaxf(mx, part_max[i]);
        if (mode == 1) s2 += part2[i];
    }
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = s2;
    r3[threadIdx.x] = mx;
    __syncthreads();
    for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) {
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
            r3[threadIdx.x] = fmaxf(r3[threadIdx.x], r3[threadIdx.x + o]);
        }
        __syncthreads();
    }
    // slices: agent-scope stores, acknowledged (s_waitcnt vmcnt(0)) before
    // the ticket; the last workgroup reads them with agent-scope loads (no
    // fence, which would write back the XCD's L2)
    if (threadIdx.x == 0) {
        __hip_atomic_store(&slices[3 * blockIdx.x], r1[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&slices[3 * blockIdx.x + 1], r2[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&slices[3 * blockIdx.x + 2], (double)r3[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    double a = 0.0, c = 0.0, m = 0.0;
    for (unsigned i = 0; i < gridDim.x; ++i) {
        a += __hip_atomic_load(&slices[3 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        c += __hip_atomic_load(&slices[3 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        m = fmax(m, __hip_atomic_load(&slices[3 * i + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    if (mode == 0) {
        // PRv.C:214-240: w = (1 - s)/n; y += w; s = sum y; y *= 1/s.
        // sum(y + w) = s + n*w is formed analytically in FP64.
        const double w = (1.0 - a) / (double)n;
        const double s2t = a + (double)n * w;
        prs->sum_y = a;
        prs->w = w;
        prs->inv_s = 1.0 / s2t;
        prs->max_ratio = (float)m;
        if (f16) {
            // largest next contribution <= d * (max y/deg + w) / s; keep it
            // at or below 2^15 so the FP16 store cannot overflow (max 65504)
            const double u = (double)damping * (m + w) / s2t;
            int k = u > 0.0 ? (int)floor(log2(32768.0 / u)) : 0;
            k = k > 100 ? 100 : k < -100 ? -100 : k;
            prs->enc_scale = (float)ldexp(1.0, k);
        } else {
            prs->enc_scale = 1.0f;
        }
    } else if (mode == 1) {
        prs->delta = a;
        prs->sum_x = c;
    } else if (mode == 3) {
        // fixed-point unit of the cold sums: 2^E with M <= 2^61 units, M the
        // total |contribution| over the cold edges
        int e = 0;
        if (a > 0.0 && a < 1e300) e = ilogb(a) + 1 - 61;
        prs->cold_exp = e;
        prs->cold_unit = ldexp(1.0, e);
    }
    // mode 2 (maximum only): the slices are read by k_pr_set_scale
    *ticket = 0;
}

template <int FMT>
__global__ void __launch_bounds__(256)
k_pr_update(uint32_t n, const float *__restrict__ y, float *__restrict__ x, typename Codec<FMT>::store_t *contrib,
            const uint32_t *__restrict__ deg, float damping, const PrScalars *__restrict__ prs, double *part_delta,
            double *part_sum) {
    __shared__ double r1[4], r2[4];
    const double w = prs->w, inv_s = prs->inv_s;
    const float sc = FMT == GPR_CONTRIB_F16 ? prs->enc_scale : 1.0f;
    double dl = 0.0, sx = 0.0;
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
        const float xn = (float)(((double)y[v] + w) * inv_s);
        dl += (double)fabsf(x[v] - xn);
        sx += (double)xn;
        x[v] = xn;
        const uint32_t dg = deg[v];
        const float cv = dg ? damping * xn / (float)dg : 0.0f;
        contrib[v] = Codec<FMT>::enc(cv * sc);
    }
    dl = block_sum(dl, r1);
    sx = block_sum(sx, r2);
    if (threadIdx.x == 0) {
        part_delta[blockIdx.x] = dl;
        part_sum[blockIdx.x] = sx;
    }
}

// contrib = d*x/deg from the current ranks (start of a run or of a stage).
template <int FMT>
__global__ void __launch_bounds__(256)
k_pr_contrib(uint32_t n, const float *__restrict__ x, typename Codec<FMT>::store_t *contrib,
             const uint32_t *__restrict__ deg, float damping, const PrScalars *__restrict__ prs) {
    const float sc = FMT == GPR_CONTRIB_F16 ? prs->enc_scale : 1.0f;
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
        const uint32_t dg = deg[v];
        const float cv = dg ? damping * x[v] / (float)dg : 0.0f;
        contrib[v] = Codec<FMT>::enc(cv * sc);
    }
}

__global__ void __launch_bounds__(256)
k_pr_xmax(uint32_t n, const float *__restrict__ x, double *part, float *part_max) {
    __shared__ float fr[4];
    float mx = 0.0f;
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) mx = fmaxf(mx, x[v]);
    mx = wave_max(mx);
    if (lane() == 0) fr[threadIdx.x / 64] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = 0.0;
        part_max[blockIdx.x] = fmaxf(fmaxf(fr[0], fr[1]), fmaxf(fr[2], fr[3]));
    }
}

__global__ void k_pr_set_scale(PrScalars *prs, const double *slices, uint32_t nslices, float damping, int f16) {
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (uint32_t i = 0; i < nslices; ++i) m = fmax(m, slices[3 * i + 2]);
    if (f16) {
        const double u = (double)damping * m;
        int k = u > 0.0 ? (int)floor(log2(32768.0 / u)) : 0;
        k = k > 100 ? 100 : k < -100 ? -100 : k;
        prs->enc_scale = (float)ldexp(1.0, k);
    } else {
        prs->enc_scale = 1.0f;
    }
}

// ---- host side -----------------------------------------------------------------
static size_t fmt_bytes(int fmt) { return fmt == GPR_CONTRIB_F32 ? 4 : 2; }

static unsigned update_grid(uint32_t n) { return grid_for(n, 256, 4096); }

// First cold source of the PageRank edgemap: sources at or above it are
// summed through the binned layout (pr_cold.hip), those below by the pull.
// The split is the FP32 bytes of pulled contributions (tunable
// pr_split_bytes): 0 bins every edge, >= 4n pulls every edge.
static uint32_t split_point(gpr_graph *g) {
    const double h = floor(tun().pr_split_bytes / 4.0);
    return h >= (double)g->n ? g->n : (uint32_t)h;
}

// x_tail != nullptr: the previous iteration's tail is fused into the write
// stage (the caller has checked cold_tail_fusable after cold_prepare)
static int pr_edgemap_impl(gpr_graph *g, const gpr_emap_desc *desc, gpr_frontier *in, gpr_iter_stats *st,
                           float *x_tail) {
    GPR_REQUIRE(!in || in->kind == GPR_FRONTIER_ALL, "PR edgemap runs over the all-true frontier only");
    GPR_REQUIRE(desc->prop && desc->prop->dtype == GPR_F32, "PR edgemap: accumulator must be an f32 property");
    const int fmt = desc->contrib_fmt;
    GPR_REQUIRE(fmt >= GPR_CONTRIB_F32 && fmt <= GPR_CONTRIB_SCFP16, "PR edgemap: unknown contribution format %d", fmt);
    GPR_REQUIRE(desc->contrib && desc->contrib->elem == fmt_bytes(fmt),
                "PR edgemap: contribution property has the wrong element size for format %d", fmt);
    uint32_t cold_from = split_point(g);
    if (cold_from < g->n) {
        GPR_TRY(cold_prepare(g, cold_from));
        if (g->cold->ntiles == 0) cold_from = g->n;   // no cold edges
    }
    const bool split = cold_from < g->n;
    const uint32_t nblk = g->nbuckets + g->nlong_chunks;
    size_t need = (size_t)nblk + g->nlong + 4096;
    if (split) need += g->cold->nbins + (size_t)g->cold->nmulti * (COLD_BIN / 1024);
    GPR_TRY(ensure_partials(g, need));
    float *y = (float *)desc->prop->d;
    const void *cb = desc->contrib->d;
    const float damping = desc->damping > 0.f ? desc->damping : 0.85f;
    GPR_HIP(hipEventRecord(g->ev0, g->stream));
    if (split) {
        ProfScope ps(g, "pr_cold_write");
        GPR_TRY(cold_launch_write(g, fmt, cb, y, x_tail, damping));
        hipLaunchKernelGGL(k_pr_reduce, dim3(red_grid(g->cold->nwchunks)), dim3(RED_THREADS), 0, g->stream, g->cold->mass,
                           (const float *)nullptr, (const double *)nullptr, g->cold->nwchunks, g->dsum_slices,
                           &g->ctr->ticket2, g->prs, 3, g->n, damping, 0);
    }
    if (cold_from > 0) {
        // the pull over sources [0, cold_from); every source when not split
        ProfScope ps(g, "pr_pull");
#define PR_PULL(F)                                                                                              \
    hipLaunchKernelGGL(k_pr_pull<F>, dim3(nblk), dim3(PR_THREADS), 0, g->stream, g->nbuckets, g->bucket,        \
                       g->cindex, g->cnbrs, g->deg, (const typename Codec<F>::store_t *)cb, y, g->long_rows,   \
                       g->long_base, g->nlong, g->long_partial, g->partial, g->partial_f, g->prs, cold_from)
        switch (fmt) {
        case GPR_CONTRIB_F32: PR_PULL(GPR_CONTRIB_F32); break;
        case GPR_CONTRIB_BF16: PR_PULL(GPR_CONTRIB_BF16); break;
        case GPR_CONTRIB_F16: PR_PULL(GPR_CONTRIB_F16); break;
        case GPR_CONTRIB_CFP16: PR_PULL(GPR_CONTRIB_CFP16); break;
        case GPR_CONTRIB_BF16T: PR_PULL(GPR_CONTRIB_BF16T); break;
        case GPR_CONTRIB_SCFP16: PR_PULL(GPR_CONTRIB_SCFP16); break;
        }
#undef PR_PULL
    }
    if (cold_from > 0 && g->nlong) {
        ProfScope ps(g, "pr_long");
        hipLaunchKernelGGL(k_pr_long, dim3(grid_for(g->nlong, 256, 1024)), dim3(256), 0, g->stream, g->nlong,
                           g->long_rows, g->long_base, g->long_partial, g->deg, y, g->partial, g->partial_f, nblk,
                           g->prs, fmt == GPR_CONTRIB_F16);
    }
    uint32_t nparts = nblk + g->nlong;
    if (split) {
        // y (+)= the binned sums; their per-bin partials replace the pull's
        ProfScope ps(g, "pr_cold_accum");
        GPR_TRY(cold_launch_accum(g, fmt, y, cold_from > 0 ? 1 : 0, g->partial, g->partial_f, &nparts));
    }
    {
        ProfScope ps(g, "pr_reduce");
        hipLaunchKernelGGL(k_pr_reduce, dim3(red_grid(nparts)), dim3(RED_THREADS), 0, g->stream, g->partial, g->partial_f,
                           (const double *)nullptr, nparts, g->dsum_slices, &g->ctr->ticket2, g->prs, 0, g->n,
                           damping, fmt == GPR_CONTRIB_F16);
    }
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipEventRecord(g->ev1, g->stream));
    if (st) {
        gpr_iter_stats s = {};
        s.nactv_in = g->n;
        s.nacte_in = g->m;
        s.direction = GPR_DIR_PULL;
        s.edges_scanned = g->m;
        GPR_HIP(hipEventSynchronize(g->ev1));
        float ms = 0.f;
        GPR_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        s.ms = ms;
        *st = s;
    }
    return GPR_OK;
}

int pr_edgemap(gpr_graph *g, const gpr_emap_desc *desc, gpr_frontier *in, gpr_iter_stats *st) {
    return pr_edgemap_impl(g, desc, in, st, nullptr);
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_pr_iterate_fused(gpr_graph *g, gpr_prop *x, gpr_prop *y, int fmt, float damping, int flags,
                                    gpr_iter_stats *st) {
    if (!g || !x || !y || x->dtype != GPR_F32 || y->dtype != GPR_F32 || x->g != g || y->g != g) return GPR_E_INVAL;
    GPR_REQUIRE(fmt >= GPR_CONTRIB_F32 && fmt <= GPR_CONTRIB_SCFP16, "pr_iterate_fused: unknown format %d", fmt);
    GPR_REQUIRE(split_point(g) == 0, "pr_iterate_fused: needs every edge binned (pr_split_bytes 0)");
    GPR_TRY(cold_prepare(g, 0));
    GPR_REQUIRE(cold_tail_fusable(g), "pr_iterate_fused: the graph has no binned edges");
    gpr_emap_desc d = {};
    d.relax = GPR_RELAX_PR_SUM;
    d.prop = y;
    d.contrib_fmt = fmt;
    d.damping = damping;
    // the contribution property is never read: the write stage computes
    // the contributions; a stand-in of the right element size passes the
    // checks
    gpr_prop stand_in = *y;
    stand_in.elem = (uint32_t)fmt_bytes(fmt);
    d.contrib = &stand_in;
    g->pr_skip_unlinked = (flags & (GPR_PR_SKIP_UNLINKED | GPR_PR_SKIP_RANKS)) != 0;
    g->pr_skip_ranks = (flags & GPR_PR_SKIP_RANKS) != 0;
    const int r = pr_edgemap_impl(g, &d, nullptr, st, (float *)x->d);
    g->pr_skip_unlinked = false;
    g->pr_skip_ranks = false;
    return r;
}

extern "C" int gpr_pr_contrib(gpr_graph *g, const gpr_prop *x, gpr_prop *contrib, int fmt, float damping) {
    if (!g || !x || !contrib || x->dtype != GPR_F32) return GPR_E_INVAL;
    GPR_REQUIRE(fmt >= GPR_CONTRIB_F32 && fmt <= GPR_CONTRIB_SCFP16 && contrib->elem == fmt_bytes(fmt),
                "pr_contrib: bad contribution format %d / property", fmt);
    const uint32_t n = g->n;
    const unsigned grid = update_grid(n);
    GPR_TRY(ensure_partials(g, grid + 64));
    if (fmt == GPR_CONTRIB_F16) {
        hipLaunchKernelGGL(k_pr_xmax, dim3(grid), dim3(256), 0, g->stream, n, (const float *)x->d, g->partial,
                           g->partial_f);
        hipLaunchKernelGGL(k_pr_reduce, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, g->stream, g->partial, g->partial_f,
                           (const double *)nullptr, grid, g->dsum_slices, &g->ctr->ticket2, g->prs, 2, n, damping, 0);
        hipLaunchKernelGGL(k_pr_set_scale, dim3(1), dim3(64), 0, g->stream, g->prs, g->dsum_slices, RED_BLOCKS,
                           damping, 1);
    } else {
        hipLaunchKernelGGL(k_pr_set_scale, dim3(1), dim3(64), 0, g->stream, g->prs, g->dsum_slices, 0, damping, 0);
    }
#define PR_CONTRIB(F)                                                                                          \
    hipLaunchKernelGGL(k_pr_contrib<F>, dim3(grid), dim3(256), 0, g->stream, n, (const float *)x->d,            \
                       (typename Codec<F>::store_t *)contrib->d, g->deg, damping, g->prs)
    switch (fmt) {
    case GPR_CONTRIB_F32: PR_CONTRIB(GPR_CONTRIB_F32); break;
    case GPR_CONTRIB_BF16: PR_CONTRIB(GPR_CONTRIB_BF16); break;
    case GPR_CONTRIB_F16: PR_CONTRIB(GPR_CONTRIB_F16); break;
    case GPR_CONTRIB_CFP16: PR_CONTRIB(GPR_CONTRIB_CFP16); break;
    case GPR_CONTRIB_BF16T: PR_CONTRIB(GPR_CONTRIB_BF16T); break;
    case GPR_CONTRIB_SCFP16: PR_CONTRIB(GPR_CONTRIB_SCFP16); break;
    }
#undef PR_CONTRIB
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

extern "C" int gpr_pr_finish(gpr_graph *g, gpr_prop *x, const gpr_prop *y, gpr_prop *contrib, int fmt, float damping,
                             double *delta, double *sum_x) {
    if (!g || !x || !y || !contrib || x->dtype != GPR_F32 || y->dtype != GPR_F32) return GPR_E_INVAL;
    GPR_REQUIRE(fmt >= GPR_CONTRIB_F32 && fmt <= GPR_CONTRIB_SCFP16 && contrib->elem == fmt_bytes(fmt),
                "pr_finish: bad contribution format %d / property", fmt);
    const uint32_t n = g->n;
    const unsigned grid = update_grid(n);
    GPR_TRY(ensure_partials(g, 2 * (size_t)grid + 64));
    double *pd = g->partial, *psum = g->partial + grid;
    {
        ProfScope ps(g, "pr_update");
#define PR_UPDATE(F)                                                                                           \
    hipLaunchKernelGGL(k_pr_update<F>, dim3(grid), dim3(256), 0, g->stream, n, (const float *)y->d, (float *)x->d, \
                       (typename Codec<F>::store_t *)contrib->d, g->deg, damping, g->prs, pd, psum)
        switch (fmt) {
        case GPR_CONTRIB_F32: PR_UPDATE(GPR_CONTRIB_F32); break;
        case GPR_CONTRIB_BF16: PR_UPDATE(GPR_CONTRIB_BF16); break;
        case GPR_CONTRIB_F16: PR_UPDATE(GPR_CONTRIB_F16); break;
        case GPR_CONTRIB_CFP16: PR_UPDATE(GPR_CONTRIB_CFP16); break;
        case GPR_CONTRIB_BF16T: PR_UPDATE(GPR_CONTRIB_BF16T); break;
        case GPR_CONTRIB_SCFP16: PR_UPDATE(GPR_CONTRIB_SCFP16); break;
        }
#undef PR_UPDATE
        hipLaunchKernelGGL(k_pr_reduce, dim3(red_grid(grid)), dim3(RED_THREADS), 0, g->stream, pd, g->partial_f, psum, grid,
                           g->dsum_slices, &g->ctr->ticket2, g->prs, 1, n, damping, 0);
    }
    GPR_HIP(hipGetLastError());
    if (delta || sum_x) {
        PrScalars h;
        GPR_HIP(hipMemcpyAsync(&h, g->prs, sizeof(h), hipMemcpyDeviceToHost, g->stream));
        GPR_HIP(hipStreamSynchronize(g->stream));
        if (delta) *delta = h.delta;
        if (sum_x) *sum_x = h.sum_x;
    }
    return GPR_OK;
}

extern "C" int gpr_pr_scalars(gpr_graph *g, double *sum_y, double *delta, double *sum_x, float *enc_scale) {
    if (!g) return GPR_E_INVAL;
    PrScalars h;
    GPR_HIP(hipMemcpyAsync(&h, g->prs, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    if (sum_y) *sum_y = h.sum_y;
    if (delta) *delta = h.delta;
    if (sum_x) *sum_x = h.sum_x;
    if (enc_scale) *enc_scale = h.enc_scale;
    return GPR_OK;
}
