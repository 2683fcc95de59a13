// frontier.hip -- device frontiers: u32 vertex lists and u64-word bitmaps,
// their conversions and active counts.
//
// Reference: frontier class frontier/frontier.h:58-1179 (toSparse :576-634,
// packDense :923-980, nActiveVertices/nActiveEdges :694-697) and
// calculateActiveCounts frontier/impl.h:233-325.  The reference keeps a
// tagged union of eight dense/sparse encodings; on the GPU two suffice:
// a list for push and a bitmap (one wave64 ballot per word) for pull.
#include "internal.hpp"
#include "device.hpp"

#include <algorithm>

namespace gpr {

__global__ void k_fill_u64(uint64_t *p, size_t words, uint64_t v) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < words; i += stride) p[i] = v;
}

// All-true bitmap over n vertices (bits beyond n stay zero).
__global__ void k_fill_all_bits(uint64_t *p, uint32_t n) {
    size_t words = ((size_t)n + 63) / 64;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < words; i += stride) {
        uint64_t w = ~0ull;
        if ((i + 1) * 64 > n) w = (n % 64) ? ((1ull << (n % 64)) - 1) : ~0ull;
        p[i] = w;
    }
}

// One atomic per run of consecutive entries in the same bitmap word: a
// push's list comes out in tile order, ascending within a tile, so the
// neighbours of one source (a hub's million, on RMAT) share words, and one
// atomic each would queue on those words.  The wave ORs each run's bits
// together (a segmented scan over lanes whose word matches) and the run's
// last lane stores them; a repeated word in another run is ORed twice,
// which changes nothing.
__global__ void k_list_to_bits(const uint32_t *__restrict__ list, uint32_t k, uint64_t *bits) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t ln = lane();
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < k; base += stride) {
        const uint32_t i = base + ln;
        const bool on = i < k;
        const uint32_t v = on ? list[i] : 0u;
        const uint32_t w = on ? v >> 6 : 0xffffffffu - ln;   // idle lanes: a word of their own
        uint64_t b = on ? 1ull << (v & 63) : 0ull;
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t ob = __shfl_up(b, off, 64);
            const uint32_t ow = __shfl_up(w, off, 64);
            if (ln >= (uint32_t)off && ow == w) b |= ob;
        }
        const uint32_t nw = __shfl_down(w, 1, 64);
        if (on && (ln == 63 || nw != w)) atomicOr((unsigned long long *)&bits[w], b);
    }
}

// Bitmap -> ascending vertex list (packDense, frontier.h:923-980), in two
// phases: phase 0 counts the set bits of each workgroup's 256 words (and
// their out-degree sum); after an exclusive scan of those counts, phase 1
// writes each word's vertices at its wave-scanned offset.  The list comes
// out globally ascending and the result is deterministic.
constexpr int B2L_THREADS = 256;
__global__ void __launch_bounds__(B2L_THREADS)
k_bits_to_list(const uint64_t *__restrict__ bits, size_t words, uint32_t *list,
               const uint32_t *__restrict__ deg, unsigned long long *blk_counts, int phase,
               DevCounters *ctr) {
    __shared__ uint32_t red[B2L_THREADS / 64];
    __shared__ uint64_t redl[B2L_THREADS / 64];
    size_t i = blockIdx.x * (size_t)B2L_THREADS + threadIdx.x;
    uint64_t w = i < words ? bits[i] : 0ull;
    uint32_t c = __popcll(w);
    if (phase == 0) {
        // count per workgroup, and edge sum
        uint64_t es = 0;
        uint64_t t = w;
        while (t) {
            int b = __ffsll((long long)t) - 1;
            t &= t - 1;
            es += deg[i * 64 + b];
        }
        uint32_t tot = block_sum(c, red);
        uint64_t etot = block_sum(es, redl);
        if (threadIdx.x == 0) {
            blk_counts[blockIdx.x] = tot;
            atomicAdd(&ctr_stripe(ctr)->nactv, tot);
            atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)etot);
        }
        return;
    }
    // phase 1: blk_counts holds exclusive offsets
    uint32_t incl = wave_incl_scan(c);
    __shared__ uint32_t wtot[B2L_THREADS / 64];
    unsigned wv = threadIdx.x / 64;
    if (lane() == 63) wtot[wv] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (unsigned j = 0; j < wv; ++j) off += wtot[j];
    uint32_t pos = (uint32_t)blk_counts[blockIdx.x] + off + incl - c;
    uint64_t t = w;
    while (t) {
        int b = __ffsll((long long)t) - 1;
        t &= t - 1;
        list[pos++] = (uint32_t)(i * 64 + b);
    }
}

// Exclusive scan of per-workgroup counts in place (one workgroup).
__global__ void __launch_bounds__(1024) k_excl_scan_u64(unsigned long long *a, uint32_t count) {
    __shared__ unsigned long long carry;
    __shared__ unsigned long long wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < count; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        unsigned long long v = i < count ? a[i] : 0ull;
        unsigned long long inc = wave_incl_scan(v);
        unsigned wv = threadIdx.x / 64;
        if (lane() == 63) wsum[wv] = inc;
        __syncthreads();
        unsigned long long off = carry;
        for (unsigned j = 0; j < wv; ++j) off += wsum[j];
        if (i < count) a[i] = off + inc - v;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = off + inc;
        __syncthreads();
    }
}

// Zero every dedup word that holds a listed vertex (all set bits of the
// dedup bitmap belong to listed vertices, so whole words may be cleared).
__global__ void k_clear_words_of_list(const uint32_t *__restrict__ list, const unsigned int *__restrict__ count,
                                      uint64_t *bits) {
    uint32_t k = *count;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    for (; i < k; i += stride) bits[list[i] >> 6] = 0ull;
}

__global__ void k_list_edges(const uint32_t *__restrict__ list, uint32_t k,
                             const uint32_t *__restrict__ deg, DevCounters *ctr) {
    __shared__ uint64_t red[4];
    uint64_t s = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
        s += deg[list[i]];
    s = block_sum(s, red);
    if (threadIdx.x == 0 && s) atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)s);
}

int fill_words(gpr_graph *g, uint64_t *p, size_t words, uint64_t v) {
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(words, 256, 4096)), dim3(256), 0, g->stream, p, words, v);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

int clear_dedup_of_list(gpr_graph *g, const uint32_t *list, uint64_t max_count) {
    hipLaunchKernelGGL(k_clear_words_of_list, dim3(grid_for(max_count, 256, 4096)), dim3(256), 0, g->stream,
                       list, &g->ctr->listpos, g->dedup);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

int frontier_ensure_list(gpr_frontier *f) {
    if (f->list) return GPR_OK;
    auto &pool = f->g->pool_list;
    if (!pool.empty()) {
        f->list = (uint32_t *)pool.back();
        pool.pop_back();
        return GPR_OK;
    }
    return dev_alloc((void **)&f->list, sizeof(uint32_t) * (size_t)std::max<uint32_t>(f->g->n, 1), "frontier list");
}

int frontier_ensure_bits(gpr_frontier *f) {
    if (f->bits) return GPR_OK;
    auto &pool = f->g->pool_bits;
    if (!pool.empty()) {
        f->bits = (uint64_t *)pool.back();
        pool.pop_back();
        return GPR_OK;
    }
    return dev_alloc((void **)&f->bits, sizeof(uint64_t) * nwords(f->g->n), "frontier bitmap");
}

void frontier_release(gpr_frontier *f) {
    // buffers go back to the graph's pool; stream order keeps reuse safe
    if (f->list) f->g->pool_list.push_back(f->list);
    if (f->bits) f->g->pool_bits.push_back(f->bits);
    f->list = nullptr;
    f->bits = nullptr;
}

int frontier_to_bits(gpr_frontier *f) {
    gpr_graph *g = f->g;
    if (f->kind == GPR_FRONTIER_BITMAP) return GPR_OK;
    GPR_TRY(frontier_ensure_bits(f));
    size_t words = nwords(g->n);
    ProfScope ps(g, "frontier_convert");
    if (f->kind == GPR_FRONTIER_ALL) {
        hipLaunchKernelGGL(k_fill_all_bits, dim3(grid_for(words, 256, 4096)), dim3(256), 0, g->stream, f->bits, g->n);
    } else {
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(words, 256, 4096)), dim3(256), 0, g->stream, f->bits, words, 0ull);
        if (f->kind == GPR_FRONTIER_LIST && f->nactv)
            hipLaunchKernelGGL(k_list_to_bits, dim3(grid_for(f->nactv, 256, 8192)), dim3(256), 0, g->stream,
                               f->list, f->nactv, f->bits);
    }
    GPR_HIP(hipGetLastError());
    f->kind = GPR_FRONTIER_BITMAP;
    return GPR_OK;
}

void excl_scan_counts(gpr_graph *g, unsigned long long *counts, uint32_t count) {
    hipLaunchKernelGGL(k_excl_scan_u64, dim3(1), dim3(1024), 0, g->stream, counts, count);
}

// Bitmap/all -> ascending list.  Recomputes nactv and nacte.
int frontier_to_list(gpr_frontier *f) {
    gpr_graph *g = f->g;
    if (f->kind == GPR_FRONTIER_LIST) return GPR_OK;
    GPR_TRY(frontier_ensure_list(f));
    if (f->kind == GPR_FRONTIER_EMPTY) { f->kind = GPR_FRONTIER_LIST; f->nactv = 0; f->nacte = 0; return GPR_OK; }
    if (f->kind == GPR_FRONTIER_ALL) GPR_TRY(frontier_to_bits(f));
    size_t words = nwords(g->n);
    unsigned blocks = grid_for(words, B2L_THREADS);
    GPR_TRY(ensure_partials(g, blocks));
    unsigned long long *cnt = (unsigned long long *)g->partial;
    GPR_TRY(reset_counters(g));
    {
        ProfScope ps(g, "frontier_convert");
        hipLaunchKernelGGL(k_bits_to_list, dim3(blocks), dim3(B2L_THREADS), 0, g->stream, f->bits, words, f->list,
                           g->deg, cnt, 0, g->ctr);
        hipLaunchKernelGGL(k_excl_scan_u64, dim3(1), dim3(1024), 0, g->stream, cnt, blocks);
        hipLaunchKernelGGL(k_bits_to_list, dim3(blocks), dim3(B2L_THREADS), 0, g->stream, f->bits, words, f->list,
                           g->deg, cnt, 1, g->ctr);
    }
    GPR_HIP(hipGetLastError());
    GPR_TRY(read_counters(g));
    f->nactv = g->hctr->nactv;
    f->nacte = g->hctr->nacte;
    f->kind = GPR_FRONTIER_LIST;
    return GPR_OK;
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_frontier_create(gpr_graph *g, gpr_frontier **out) {
    if (!g || !out) return GPR_E_INVAL;
    gpr_frontier *f = new gpr_frontier();
    f->g = g;
    *out = f;
    return GPR_OK;
}

extern "C" int gpr_frontier_destroy(gpr_frontier *f) {
    if (!f) return GPR_E_INVAL;
    frontier_release(f);
    delete f;
    return GPR_OK;
}

// frontier::sparse(n, deg(v), &edges[index[v]]) -- the neighbours of v as a
// list frontier (CCv.C:262-267).
extern "C" int gpr_frontier_set_neighbours(gpr_frontier *f, uint32_t v) {
    if (!f) return GPR_E_INVAL;
    gpr_graph *g = f->g;
    GPR_REQUIRE(v < g->n, "vertex %u out of range", v);
    GPR_TRY(to_internal_id(g, v, &v));
    uint64_t ix[2];
    GPR_HIP(hipMemcpyAsync(ix, g->index + v, sizeof(ix), hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    const uint32_t k = (uint32_t)(ix[1] - ix[0]);
    GPR_TRY(frontier_ensure_list(f));
    if (k) GPR_HIP(hipMemcpyAsync(f->list, g->nbrs + ix[0], sizeof(uint32_t) * k, hipMemcpyDeviceToDevice, g->stream));
    GPR_TRY(reset_counters(g));
    if (k) hipLaunchKernelGGL(k_list_edges, dim3(grid_for(k, 256, 1024)), dim3(256), 0, g->stream, f->list, k, g->deg, g->ctr);
    GPR_HIP(hipGetLastError());
    GPR_TRY(read_counters(g));
    f->kind = k ? GPR_FRONTIER_LIST : GPR_FRONTIER_EMPTY;
    f->nactv = k;
    f->nacte = g->hctr->nacte;
    return GPR_OK;
}

extern "C" int gpr_frontier_set_all(gpr_frontier *f) {
    if (!f) return GPR_E_INVAL;
    f->kind = GPR_FRONTIER_ALL;
    f->nactv = f->g->n;
    f->nacte = f->g->m;
    return GPR_OK;
}

extern "C" int gpr_frontier_set_list(gpr_frontier *f, const uint32_t *vs, uint32_t k) {
    if (!f || (k && !vs)) return GPR_E_INVAL;
    gpr_graph *g = f->g;
    GPR_REQUIRE(k <= g->n, "frontier list of %u entries exceeds n=%u", k, g->n);
    for (uint32_t i = 0; i < k; ++i) GPR_REQUIRE(vs[i] < g->n, "frontier vertex %u out of range", vs[i]);
    GPR_TRY(frontier_ensure_list(f));
    if (k) GPR_HIP(hipMemcpyAsync(f->list, vs, sizeof(uint32_t) * k, hipMemcpyHostToDevice, g->stream));
    GPR_TRY(map_ids(g, f->list, k, true));
    GPR_TRY(reset_counters(g));
    if (k) hipLaunchKernelGGL(k_list_edges, dim3(grid_for(k, 256, 1024)), dim3(256), 0, g->stream, f->list, k, g->deg, g->ctr);
    GPR_HIP(hipGetLastError());
    GPR_TRY(read_counters(g));
    f->kind = GPR_FRONTIER_LIST;
    f->nactv = k;
    f->nacte = g->hctr->nacte;
    return GPR_OK;
}

extern "C" int gpr_frontier_set_single(gpr_frontier *f, uint32_t v) {
    if (!f) return GPR_E_INVAL;
    return gpr_frontier_set_list(f, &v, 1);
}

extern "C" int gpr_frontier_stats(const gpr_frontier *f, uint32_t *nactv, uint64_t *nacte, int *kind) {
    if (!f) return GPR_E_INVAL;
    if (nactv) *nactv = f->nactv;
    if (nacte) *nacte = f->nacte;
    if (kind) *kind = f->kind;
    return GPR_OK;
}

extern "C" int gpr_frontier_convert(gpr_frontier *f, int kind) {
    if (!f) return GPR_E_INVAL;
    if (kind == GPR_FRONTIER_LIST) return frontier_to_list(f);
    if (kind == GPR_FRONTIER_BITMAP) return frontier_to_bits(f);
    set_error("frontier conversion to kind %d not supported", kind);
    return GPR_E_INVAL;
}

extern "C" int gpr_frontier_download(gpr_frontier *f, uint32_t *out, uint32_t *k) {
    if (!f || !k) return GPR_E_INVAL;
    gpr_graph *g = f->g;
    if (f->kind == GPR_FRONTIER_EMPTY || f->nactv == 0) { *k = 0; return GPR_OK; }
    if (f->kind != GPR_FRONTIER_LIST) {
        // convert a copy through the bitmap so the caller's representation
        // is preserved
        GPR_TRY(frontier_to_list(f));
    }
    *k = f->nactv;
    if (out) {
        const uint32_t *src = f->list;
        if (g->order) {
            uint32_t *tmp = (uint32_t *)g->scan;
            GPR_HIP(hipMemcpyAsync(tmp, f->list, sizeof(uint32_t) * f->nactv, hipMemcpyDeviceToDevice, g->stream));
            GPR_TRY(map_ids(g, tmp, f->nactv, false));
            src = tmp;
        }
        GPR_HIP(hipMemcpyAsync(out, src, sizeof(uint32_t) * f->nactv, hipMemcpyDeviceToHost, g->stream));
        GPR_HIP(hipStreamSynchronize(g->stream));
        std::sort(out, out + f->nactv);
    }
    return GPR_OK;
}
