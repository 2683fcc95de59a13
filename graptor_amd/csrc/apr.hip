// apr.hip -- the delta PageRank (APR_vary) vertex maps and its push.
//
// Reference: bench/APR_vary.C run_prec :378-640 (contributions :462-465,
// :512-516, :566-569; the edgemaps :467-480, :518-531, :573-590; the vertex
// maps :484-505, :535-557, :590-612; the L1 scan :632-639) and run()
// :900-1074 (initial values :917-921, final scaling :1060-1072).
//
// One iteration on the handle's stream:
//   k_apr_contrib  contrib[v] = enc(delta[v] / (factor*deg v)); with a
//                  frontier mask, sources outside it store code 0 (read as
//                  0), so the all-sources pull then adds exactly the
//                  frontier's contributions (the strong source filter of
//                  :576 applied where the contributions are written).
//   the pull       the PageRank edgemap (pagerank.hip / pr_cold.hip) sums
//                  y[d] = sum of decoded contributions over in-edges;
//   or k_apr_push  for frontiers at or below 1% of the edges (:574,
//                  frac_threshold(1)): y = 0, then every out-edge of a
//                  frontier source adds its decoded contribution with a
//                  hardware FP32 atomic, edges split evenly over workgroups
//                  by the frontier's degree prefix.
//   k_apr_tail     z = y*factor, delta, x, the active test, the new frontier
//                  as a bitmap (one ballot per 64 vertices), per-workgroup
//                  sums of |delta| and |x| and the frontier counts;
//   k_apr_reduce   the deterministic FP64 totals, read by the host to pick
//                  the next iteration's branch and to stop.
// Vertex arithmetic is FP32 with no contraction, as FloatTy is float in
// every variant; contraction is switched off so each product and sum
// rounds separately.
#include "internal.hpp"
#include "device.hpp"
#include "pr_codec.hpp"
#include "apr.hpp"

#include <algorithm>

namespace gpr {

constexpr int AC_THREADS = 256;
constexpr int AP_THREADS = 256;
constexpr int AP_ITEMS = 4;
constexpr int AP_TILE = AP_THREADS * AP_ITEMS;
constexpr int AP_LDS = AP_TILE + 1;
constexpr int AT_THREADS = 256;
constexpr int AT_WAVES = AT_THREADS / 64;
constexpr unsigned AT_MAX_BLOCKS = 2048;

// The value the pull adds for source s: the stored contribution, decoded.
template <int FMT>
__device__ __forceinline__ float apr_value(float delta, uint32_t deg, float factor) {
#pragma clang fp contract(off)
    const float od = 1.0f / (factor * (float)deg);
    return Codec<FMT>::dec(Codec<FMT>::enc(delta * od));
}

template <int FMT>
__global__ void __launch_bounds__(AC_THREADS)
k_apr_contrib(uint32_t n, const float *__restrict__ delta, const uint32_t *__restrict__ deg, float factor,
              const uint64_t *__restrict__ mask, typename Codec<FMT>::store_t *__restrict__ contrib) {
#pragma clang fp contract(off)
    using S = typename Codec<FMT>::store_t;
    for (uint32_t v = blockIdx.x * AC_THREADS + threadIdx.x; v < n; v += gridDim.x * AC_THREADS) {
        const uint32_t d = deg[v];
        bool on = d != 0;
        if (mask) on = on && bit_test(mask, v);
        S c = (S)0;
        if (on) {
            const float od = 1.0f / (factor * (float)d);
            c = Codec<FMT>::enc(delta[v] * od);
        }
        contrib[v] = c;
    }
}

__global__ void __launch_bounds__(256) k_apr_zero(float *__restrict__ y, uint32_t n) {
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) y[v] = 0.0f;
}

// Edge tiles of AP_TILE over the frontier's degree prefix; the sources of a
// tile and their values are staged in LDS when they fit (a tile spans at
// most AP_TILE sources with edges, more only through sources without
// edges, which then take the per-edge search of the global prefix).
template <int FMT>
__global__ void __launch_bounds__(AP_THREADS)
k_apr_push(const uint32_t *__restrict__ flist, uint32_t k, const uint64_t *__restrict__ pos,
           const uint64_t *__restrict__ btot, uint64_t E, const uint64_t *__restrict__ index,
           const uint32_t *__restrict__ nbrs, const uint32_t *__restrict__ deg, const float *__restrict__ delta,
           float factor, float *y) {
    __shared__ uint64_t s_pos[AP_LDS];
    __shared__ uint32_t s_src[AP_LDS];
    __shared__ float s_val[AP_LDS];
    __shared__ uint32_t s_i0, s_cnt;
    const uint64_t ntiles = (E + AP_TILE - 1) / AP_TILE;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t e0 = t * AP_TILE;
        const uint64_t e1 = min(E, e0 + AP_TILE);
        if (threadIdx.x < 64) {
            // the tile's first and last sources, by one wave's 64-way search
            const uint32_t i0 = wave_ub_pos(pos, btot, k, e0, 0u) - 1;
            const uint32_t i1 = wave_ub_pos(pos, btot, k, e1 - 1, i0 + 1) - 1;
            if (threadIdx.x == 0) {
                s_i0 = i0;
                s_cnt = i1 - i0 + 1;
            }
        }
        __syncthreads();
        const uint32_t i0 = s_i0, cnt = s_cnt;
        const bool in_lds = cnt <= (uint32_t)AP_LDS;
        if (in_lds)
            for (uint32_t j = threadIdx.x; j < cnt; j += AP_THREADS) {
                const uint32_t s = flist[i0 + j];
                s_pos[j] = frontier_pos(pos, btot, i0 + j);
                s_src[j] = s;
                const uint32_t dg = deg[s];
                s_val[j] = dg ? apr_value<FMT>(delta[s], dg, factor) : 0.0f;
            }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < AP_ITEMS; ++it) {
            const uint64_t e = e0 + (uint64_t)it * AP_THREADS + threadIdx.x;
            if (e >= e1) continue;
            uint64_t pj;
            uint32_t s;
            float c;
            if (in_lds) {
                uint32_t lo = 0, hi = cnt;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_pos[mid] <= e) lo = mid + 1; else hi = mid;
                }
                pj = s_pos[lo - 1];
                s = s_src[lo - 1];
                c = s_val[lo - 1];
            } else {
                const uint32_t j = ub_pos(pos, btot, k, e) - 1;
                pj = frontier_pos(pos, btot, j);
                s = flist[j];
                c = apr_value<FMT>(delta[s], deg[s], factor);
            }
            const uint32_t d = nbrs[index[s] + (e - pj)];
            unsafeAtomicAdd(&y[d], c);
        }
        __syncthreads();
    }
}

// Waves take 64-vertex words in a grid stride; per-thread sums run in a
// fixed order for a given n, so the totals are deterministic.
template <bool FIRST>
__global__ void __launch_bounds__(AT_THREADS)
k_apr_tail(uint32_t n, const float *__restrict__ y, float factor, float dampen, float cst, float ninv, float eps2,
           float *__restrict__ x, float *__restrict__ delta, const uint32_t *__restrict__ deg,
           uint64_t *__restrict__ bits, double *__restrict__ part, uint64_t *__restrict__ cpart) {
#pragma clang fp contract(off)
    __shared__ double r1[AT_WAVES], r2[AT_WAVES];
    __shared__ uint64_t r3[AT_WAVES], r4[AT_WAVES];
    const uint32_t words = (uint32_t)(((uint64_t)n + 63) / 64);
    const uint32_t wv = threadIdx.x / 64;
    const uint32_t nwv = gridDim.x * AT_WAVES;
    double sl1 = 0.0, sx = 0.0;
    uint64_t ce = 0, cv = 0;
    for (uint32_t w = blockIdx.x * AT_WAVES + wv; w < words; w += nwv) {
        const uint32_t v = w * 64 + lane();
        bool a = false;
        if (v < n) {
            const float z = y[v] * factor;
            float xv = x[v], dl;
            if (FIRST) {
                // :484-505: delta = d*z + (1-d)/n; x += delta; delta -= 1/n
                dl = dampen * z + cst;
                xv = xv + dl;
                x[v] = xv;
                dl = dl - ninv;
                a = fabsf(dl) > eps2 * xv;
            } else {
                // :535-557 / :590-612: delta = z*d; x += delta where active
                dl = z * dampen;
                a = fabsf(dl) > eps2 * xv;
                if (a) {
                    xv = xv + dl;
                    x[v] = xv;
                }
            }
            delta[v] = dl;
            sl1 += (double)fabsf(dl);
            sx += (double)fabsf(xv);
            if (a) ce += deg[v];
        }
        const uint64_t b = __ballot(a);
        if (lane() == 0) {
            bits[w] = b;
            cv += (uint64_t)__popcll(b);
        }
    }
    sl1 = wave_sum(sl1);
    sx = wave_sum(sx);
    ce = wave_sum(ce);
    cv = wave_sum(cv);
    if (lane() == 0) {
        r1[wv] = sl1;
        r2[wv] = sx;
        r3[wv] = cv;
        r4[wv] = ce;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        uint64_t c = 0, d = 0;
        for (int i = 0; i < AT_WAVES; ++i) {
            a += r1[i];
            b += r2[i];
            c += r3[i];
            d += r4[i];
        }
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
        cpart[2 * blockIdx.x] = c;
        cpart[2 * blockIdx.x + 1] = d;
    }
}

// Totals of the tail's partials in a fixed order: out = {L1, sum |x|,
// nactv, nacte} (the last two as integers in the same 8-byte slots).
__global__ void __launch_bounds__(256)
k_apr_reduce(const double *__restrict__ part, const uint64_t *__restrict__ cpart, uint32_t count,
             AprTotals *out) {
    __shared__ double r1[256], r2[256];
    __shared__ uint64_t r3[256], r4[256];
    double a = 0.0, b = 0.0;
    uint64_t c = 0, d = 0;
    for (uint32_t i = threadIdx.x; i < count; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
        c += cpart[2 * i];
        d += cpart[2 * i + 1];
    }
    r1[threadIdx.x] = a;
    r2[threadIdx.x] = b;
    r3[threadIdx.x] = c;
    r4[threadIdx.x] = d;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) {
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
            r3[threadIdx.x] += r3[threadIdx.x + o];
            r4[threadIdx.x] += r4[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out->l1 = r1[0];
        out->sum_x = r2[0];
        out->nactv = r3[0];
        out->nacte = r4[0];
    }
}

__global__ void __launch_bounds__(256) k_apr_scale(uint32_t n, const float *__restrict__ x, double s,
                                                   double *__restrict__ out) {
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) out[v] = (double)x[v] * s;
}

// ---- host side ------------------------------------------------------------------
static bool apr_fmt_ok(int fmt) {
    return fmt == GPR_CONTRIB_F32 || fmt == GPR_CONTRIB_BF16 || fmt == GPR_CONTRIB_BF16T || fmt == GPR_CONTRIB_SCFP16;
}

int apr_check_fmt(int fmt) {
    GPR_REQUIRE(apr_fmt_ok(fmt), "apr: contribution format %d cannot hold negative deltas "
                                 "(use fp32, bf16, bf16t or scfp16)", fmt);
    return GPR_OK;
}

int apr_contrib(gpr_graph *g, int fmt, const gpr_prop *dp, gpr_prop *cp, float factor, gpr_frontier *mask) {
    GPR_TRY(apr_check_fmt(fmt));
    GPR_REQUIRE(dp && dp->dtype == GPR_F32 && cp && cp->elem == (fmt == GPR_CONTRIB_F32 ? 4u : 2u),
                "apr_contrib: delta must be f32 and contrib sized for format %d", fmt);
    const float *delta = (const float *)dp->d;
    void *contrib = cp->d;
    const uint32_t n = g->n;
    const unsigned grid = grid_for(n, AC_THREADS, 4096);
    const uint64_t *bits = nullptr;
    if (mask && mask->kind != GPR_FRONTIER_ALL) {
        GPR_TRY(frontier_to_bits(mask));
        bits = mask->bits;
    }
    ProfScope ps(g, "apr_contrib");
#define AC(F)                                                                                                  \
    hipLaunchKernelGGL(k_apr_contrib<F>, dim3(grid), dim3(AC_THREADS), 0, g->stream, n, delta, g->deg, factor, \
                       bits, (typename Codec<F>::store_t *)contrib)
    switch (fmt) {
    case GPR_CONTRIB_F32: AC(GPR_CONTRIB_F32); break;
    case GPR_CONTRIB_BF16: AC(GPR_CONTRIB_BF16); break;
    case GPR_CONTRIB_BF16T: AC(GPR_CONTRIB_BF16T); break;
    case GPR_CONTRIB_SCFP16: AC(GPR_CONTRIB_SCFP16); break;
    }
#undef AC
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

int apr_push(gpr_graph *g, int fmt, gpr_frontier *F, const gpr_prop *dp, float factor, gpr_prop *yp) {
    GPR_TRY(apr_check_fmt(fmt));
    GPR_REQUIRE(dp && dp->dtype == GPR_F32 && yp && yp->dtype == GPR_F32 && F && F->g == g,
                "apr_push: delta and y must be f32 properties, F a frontier of this graph");
    const float *delta = (const float *)dp->d;
    float *y = (float *)yp->d;
    const uint32_t n = g->n;
    {
        ProfScope ps(g, "apr_zero");
        hipLaunchKernelGGL(k_apr_zero, dim3(grid_for(n, 256, 4096)), dim3(256), 0, g->stream, y, n);
    }
    if (F->kind == GPR_FRONTIER_EMPTY) return GPR_OK;
    GPR_TRY(frontier_to_list(F));
    const uint32_t k = F->nactv;
    const uint64_t E = F->nacte;
    if (k == 0 || E == 0) return GPR_OK;
    GPR_TRY(scan_frontier_degrees(g, F->list, k));
    const unsigned grid = grid_for(E, AP_TILE, (unsigned)std::max(1.0, tun().push_blocks_per_cu * num_cus()));
    ProfScope ps(g, "apr_push");
#define AP(FM)                                                                                                 \
    hipLaunchKernelGGL(k_apr_push<FM>, dim3(grid), dim3(AP_THREADS), 0, g->stream, F->list, k, g->scan,         \
                       g->scan_blk, E, g->index, g->nbrs, g->deg, delta, factor, y)
    switch (fmt) {
    case GPR_CONTRIB_F32: AP(GPR_CONTRIB_F32); break;
    case GPR_CONTRIB_BF16: AP(GPR_CONTRIB_BF16); break;
    case GPR_CONTRIB_BF16T: AP(GPR_CONTRIB_BF16T); break;
    case GPR_CONTRIB_SCFP16: AP(GPR_CONTRIB_SCFP16); break;
    }
#undef AP
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

int apr_tail(gpr_graph *g, bool first, const gpr_prop *yp, float factor, float eps2, gpr_prop *xp, gpr_prop *dp,
             gpr_frontier *out, AprTotals *tot) {
    GPR_REQUIRE(yp && xp && dp && yp->dtype == GPR_F32 && xp->dtype == GPR_F32 && dp->dtype == GPR_F32 && out &&
                    out->g == g,
                "apr_tail: y, x and delta must be f32 properties, out a frontier of this graph");
    const float *y = (const float *)yp->d;
    float *x = (float *)xp->d, *delta = (float *)dp->d;
    const uint32_t n = g->n;
    GPR_TRY(frontier_ensure_bits(out));
    const unsigned grid = grid_for(nwords(n), AT_WAVES, AT_MAX_BLOCKS);
    GPR_TRY(ensure_partials(g, 4 * (size_t)grid + 8));
    double *part = g->partial;
    uint64_t *cpart = (uint64_t *)(g->partial + 2 * (size_t)grid);
    AprTotals *dtot = (AprTotals *)(g->partial + 4 * (size_t)grid);
    const float dampen = 0.85f;
    const float ninv = 1.0f / (float)n;
    const float cst = (1.0f - dampen) * ninv;
    {
        ProfScope ps(g, "apr_tail");
        if (first)
            hipLaunchKernelGGL(k_apr_tail<true>, dim3(grid), dim3(AT_THREADS), 0, g->stream, n, y, factor, dampen,
                               cst, ninv, eps2, x, delta, g->deg, out->bits, part, cpart);
        else
            hipLaunchKernelGGL(k_apr_tail<false>, dim3(grid), dim3(AT_THREADS), 0, g->stream, n, y, factor, dampen,
                               cst, ninv, eps2, x, delta, g->deg, out->bits, part, cpart);
        hipLaunchKernelGGL(k_apr_reduce, dim3(1), dim3(256), 0, g->stream, part, cpart, grid, dtot);
    }
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(tot, dtot, sizeof(AprTotals), hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    out->nactv = (uint32_t)tot->nactv;
    out->nacte = tot->nacte;
    out->kind = tot->nactv ? GPR_FRONTIER_BITMAP : GPR_FRONTIER_EMPTY;
    return GPR_OK;
}

int apr_scale(gpr_graph *g, const gpr_prop *xp, double s, gpr_prop *op) {
    GPR_REQUIRE(xp && op && xp->dtype == GPR_F32 && op->dtype == GPR_F64, "apr_scale: x f32, out f64");
    hipLaunchKernelGGL(k_apr_scale, dim3(grid_for(g->n, 256, 4096)), dim3(256), 0, g->stream, g->n,
                       (const float *)xp->d, s, (double *)op->d);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

} // namespace gpr
