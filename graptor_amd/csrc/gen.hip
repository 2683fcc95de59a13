// gen.hip -- RMAT input generation and cleaning on the device.
//
// The reference ships no generator; its cleaning tools define what a clean
// graph is: no self loops (tools/CleanGraph.C --rm-self), symmetric
// (CvtToUndir.C:30-121), sorted and free of duplicates (RmDupEdges.C:9-59),
// isolated vertices kept.  The edge stream is this project's own
// specification, shared bit-for-bit with the CPU oracle: for edge e and
// level l (most significant bit first), r = half (l & 1) of
// mix64(seed ^ mix64((e << 5) | (l >> 1))), quadrant a/b/c/d chosen by
// comparing r with floor(a*2^32), floor((a+b)*2^32), floor((a+b+c)*2^32);
// mix64 is the splitmix64 finaliser.
//
// Pipeline: one thread per generated edge writes the keys (u<<s)|v and
// (v<<s)|u (self loops become the sentinel (2^2s)-1, which sorts last), an
// LSD radix sort over 2s bits (rocPRIM), a unique pass, then the CSR
// offsets from the row boundaries and the neighbour column.
#include "internal.hpp"
#include "device.hpp"

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <vector>

namespace gpr {

int graph_from_device(gpr_graph *g);
int graph_new_device(uint32_t n, uint64_t m, gpr_graph **out);
void graph_release(gpr_graph *g);

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_rmat_keys(uint32_t scale, uint64_t t0, uint64_t t1, uint64_t t2, uint64_t seed, uint64_t count,
                            uint64_t *keys) {
    const uint64_t sentinel = (scale == 32) ? ~0ull : ((1ull << (2 * scale)) - 1);
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < count;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t u = 0, v = 0;
        uint64_t h = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            if ((l & 1) == 0) h = mix64(seed ^ mix64((e << 5) | (l >> 1)));
            const uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
            const uint32_t q = r < t0 ? 0u : r < t1 ? 1u : r < t2 ? 2u : 3u;
            u = (u << 1) | (q >> 1);
            v = (v << 1) | (q & 1u);
        }
        uint64_t k1 = sentinel, k2 = sentinel;
        if (u != v) {
            k1 = ((uint64_t)u << scale) | v;
            k2 = ((uint64_t)v << scale) | u;
        }
        keys[2 * e] = k1;
        keys[2 * e + 1] = k2;
    }
}

// Sorted unique keys -> neighbour column, plus the end offset of every
// non-empty row at ends[row] (ends zero-filled beforehand).  The row offsets
// follow as index[0] = 0, index[u+1] = max(ends[0..u]) (an inclusive max
// scan), which fills the rows without edges.
__global__ void k_keys_to_rows(const uint64_t *__restrict__ keys, uint64_t m, uint32_t scale, uint32_t *nbrs,
                               uint64_t *ends) {
    const uint64_t mask = (1ull << scale) - 1;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < m; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[p];
        const uint32_t u = (uint32_t)(k >> scale);
        nbrs[p] = (uint32_t)(k & mask);
        if (p + 1 == m || (uint32_t)(keys[p + 1] >> scale) != u) ends[u] = p + 1;
    }
}

__global__ void k_zero_u64(uint64_t *p, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

static int keys_to_csr(const uint64_t *keys, uint64_t m, uint32_t scale, uint32_t n, uint32_t *nbrs, uint64_t *index,
                       hipStream_t st) {
    uint64_t *ends = nullptr;
    void *tmp = nullptr;
    int r = GPR_OK;
    do {
        if ((r = dev_alloc((void **)&ends, 8ull * (n ? n : 1), "row ends"))) break;
        hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(n, 256, 65536)), dim3(256), 0, st, ends, (uint64_t)n);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, index, (uint64_t)1);
        if (m)
            hipLaunchKernelGGL(k_keys_to_rows, dim3(grid_for(m, 256, 65536)), dim3(256), 0, st, keys, m, scale, nbrs, ends);
        size_t tsz = 0;
        if (rocprim::inclusive_scan(nullptr, tsz, ends, index + 1, (size_t)n, rocprim::maximum<uint64_t>(), st) !=
            hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc(&tmp, tsz, "scan workspace"))) break;
        if (rocprim::inclusive_scan(tmp, tsz, ends, index + 1, (size_t)n, rocprim::maximum<uint64_t>(), st) !=
            hipSuccess) { r = GPR_E_HIP; break; }
        if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) { r = GPR_E_HIP; break; }
    } while (0);
    if (r == GPR_E_HIP) set_error("CSR offsets build failed");
    dev_free(ends);
    dev_free(tmp);
    return r;
}

static int rmat_impl(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, gpr_graph **out) {
    GPR_REQUIRE(scale >= 1 && scale <= 31, "rmat scale %u out of range [1,31]", scale);
    GPR_REQUIRE(ef >= 1, "rmat edge factor must be positive");
    GPR_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "rmat probabilities invalid");
    const uint32_t n = 1u << scale;
    const uint64_t count = (uint64_t)ef << scale;
    const uint64_t nkeys = 2 * count;
    // 64-bit so that a probability of 1 maps to 2^32, above every draw
    const uint64_t t0 = (uint64_t)(a * 4294967296.0);
    const uint64_t t1 = (uint64_t)((a + b) * 4294967296.0);
    const uint64_t t2 = (uint64_t)((a + b + c) * 4294967296.0);
    hipStream_t st;
    GPR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    uint64_t *k0 = nullptr, *k1 = nullptr, *ucount = nullptr;
    void *tmp = nullptr;
    int r = GPR_OK;
    gpr_graph *g = nullptr;
    do {
        if ((r = dev_alloc((void **)&k0, nkeys * 8, "rmat keys"))) break;
        if ((r = dev_alloc((void **)&k1, nkeys * 8, "rmat keys"))) break;
        if ((r = dev_alloc((void **)&ucount, 8, "rmat unique count"))) break;
        hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(count, 256, 65536)), dim3(256), 0, st, scale, t0, t1, t2, seed,
                           count, k0);
        rocprim::double_buffer<uint64_t> db(k0, k1);
        size_t tsz = 0;
        if (rocprim::radix_sort_keys(nullptr, tsz, db, (size_t)nkeys, 0, 2 * scale, st) != hipSuccess) {
            set_error("radix sort sizing failed");
            r = GPR_E_HIP;
            break;
        }
        if ((r = dev_alloc(&tmp, tsz, "sort workspace"))) break;
        if (rocprim::radix_sort_keys(tmp, tsz, db, (size_t)nkeys, 0, 2 * scale, st) != hipSuccess) {
            set_error("radix sort failed");
            r = GPR_E_HIP;
            break;
        }
        uint64_t *sorted = db.current();
        uint64_t *other = db.alternate();
        dev_free(tmp);
        tmp = nullptr;
        size_t usz = 0;
        if (rocprim::unique(nullptr, usz, sorted, other, ucount, (size_t)nkeys, rocprim::equal_to<uint64_t>(), st) !=
            hipSuccess) {
            set_error("unique sizing failed");
            r = GPR_E_HIP;
            break;
        }
        if ((r = dev_alloc(&tmp, usz, "unique workspace"))) break;
        if (rocprim::unique(tmp, usz, sorted, other, ucount, (size_t)nkeys, rocprim::equal_to<uint64_t>(), st) !=
            hipSuccess) {
            set_error("unique failed");
            r = GPR_E_HIP;
            break;
        }
        uint64_t nu = 0;
        if (hipMemcpyAsync(&nu, ucount, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("unique count readback failed");
            r = GPR_E_HIP;
            break;
        }
        // the sentinel (if any self loop occurred) is the last unique key
        const uint64_t sentinel = (1ull << (2 * scale)) - 1;
        uint64_t last = 0;
        if (nu) (void)hipMemcpy(&last, other + nu - 1, 8, hipMemcpyDeviceToHost);
        const uint64_t m = (nu && last == sentinel) ? nu - 1 : nu;
        dev_free(tmp);
        tmp = nullptr;
        dev_free(sorted);
        if (sorted == k0) k0 = nullptr; else k1 = nullptr;
        if ((r = graph_new_device(n, m, &g))) break;
        if ((r = dev_alloc((void **)&g->index, sizeof(uint64_t) * ((size_t)n + 1), "CSR index"))) break;
        if ((r = dev_alloc((void **)&g->nbrs, sizeof(uint32_t) * (m ? m : 1), "CSR neighbours"))) break;
        if ((r = keys_to_csr(other, m, scale, n, g->nbrs, g->index, st))) break;
        dev_free(other);
        if (other == k0) k0 = nullptr; else k1 = nullptr;
        r = graph_from_device(g);
    } while (0);
    dev_free(k0);
    dev_free(k1);
    dev_free(ucount);
    dev_free(tmp);
    (void)hipStreamDestroy(st);
    if (r != GPR_OK) {
        if (g) graph_release(g);
        return r;
    }
    *out = g;
    return GPR_OK;
}

int keys_to_csr_ext(const uint64_t *keys, uint64_t m, uint32_t scale, uint32_t n, uint32_t *nbrs, uint64_t *index,
                    hipStream_t st) {
    return keys_to_csr(keys, m, scale, n, nbrs, index, st);
}

// ---- degree-sorted relabelling ------------------------------------------------
// Graptor's default graph type renumbers vertices (VEBO, graph/VEBOReorder.h:
// 134-209) and maps results back with remapID/originalID
// (graph/CGraphVEBOGraptor.h:441-500).  Here vertices are renumbered by
// decreasing degree (ties by ascending ID), so that the few vertices that
// receive most gathers (RMAT hubs, scattered over the whole ID range in the
// generated order) occupy a contiguous, cache-resident prefix of every
// per-vertex array, and a wave's 64 consecutive vertices have similar
// degrees.
__global__ void k_degree_keys(const uint32_t *__restrict__ deg, uint32_t n, uint32_t *key, uint32_t *val) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        key[v] = ~deg[v];
        val[v] = v;
    }
}

__global__ void k_invert_perm(const uint32_t *__restrict__ order, uint32_t n, uint32_t *perm) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) perm[order[i]] = i;
}

// Row of edge e: the last r with index[r] <= e, searched in [lo, hi).
__device__ __forceinline__ uint32_t row_of(const uint64_t *index, uint32_t lo, uint32_t hi, uint64_t e) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (index[mid] <= e) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// key = (perm[v] << s) | perm[u] for every edge (v, u).  Workgroups take
// RL_TILE consecutive edges (balanced whatever the degrees); the tile's row
// offsets are staged in LDS when they fit.
constexpr int RL_TILE = 4096;
__global__ void __launch_bounds__(256)
k_relabel_keys(const uint64_t *__restrict__ index, const uint32_t *__restrict__ nbrs, uint32_t n, uint64_t m,
               const uint32_t *__restrict__ perm, uint32_t s, uint64_t *keys) {
    __shared__ uint64_t sidx[RL_TILE + 1];
    __shared__ uint32_t s_r0, s_cnt;
    const uint64_t ntiles = (m + RL_TILE - 1) / RL_TILE;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t e0 = t * RL_TILE, e1 = min(m, e0 + RL_TILE);
        if (threadIdx.x == 0) {
            const uint32_t r0 = row_of(index, 0, n + 1, e0);
            const uint32_t r1 = row_of(index, r0, n + 1, e1 - 1);
            s_r0 = r0;
            s_cnt = r1 - r0 + 2;
        }
        __syncthreads();
        const uint32_t r0 = s_r0, cnt = s_cnt;
        const bool in_lds = cnt <= RL_TILE + 1;
        if (in_lds)
            for (uint32_t j = threadIdx.x; j < cnt; j += 256) sidx[j] = index[r0 + j];
        __syncthreads();
        for (uint64_t e = e0 + threadIdx.x; e < e1; e += 256) {
            uint32_t r;
            if (in_lds) {
                uint32_t lo = 0, hi = cnt;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (sidx[mid] <= e) lo = mid + 1; else hi = mid;
                }
                r = r0 + lo - 1;
            } else {
                r = row_of(index, r0, r0 + cnt, e);
            }
            keys[e] = ((uint64_t)perm[r] << s) | perm[nbrs[e]];
        }
        __syncthreads();
    }
}

int graph_replace_csr(gpr_graph *g, uint64_t *index, uint32_t *nbrs);

static int reorder_impl(gpr_graph *g) {
    GPR_REQUIRE(g->symmetric, "reordering needs a symmetric graph");
    GPR_REQUIRE(!g->order, "graph is already reordered");
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    uint32_t s = 1;
    while (s < 32 && (1ull << s) < n) ++s;
    hipStream_t st = g->stream;
    uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *perm = nullptr;
    uint64_t *e0 = nullptr, *e1 = nullptr, *nindex = nullptr;
    uint32_t *nnbrs = nullptr;
    void *tmp = nullptr;
    int r = GPR_OK;
    do {
        // 1. order: vertices by (degree desc, id asc); stable radix sort
        if ((r = dev_alloc((void **)&k0, 4ull * n, "reorder keys"))) break;
        if ((r = dev_alloc((void **)&k1, 4ull * n, "reorder keys"))) break;
        if ((r = dev_alloc((void **)&v0, 4ull * n, "reorder values"))) break;
        if ((r = dev_alloc((void **)&v1, 4ull * n, "reorder values"))) break;
        hipLaunchKernelGGL(k_degree_keys, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, g->deg, n, k0, v0);
        rocprim::double_buffer<uint32_t> kb(k0, k1), vb(v0, v1);
        size_t tsz = 0;
        if (rocprim::radix_sort_pairs(nullptr, tsz, kb, vb, (size_t)n, 0, 32, st) != hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc(&tmp, tsz, "sort workspace"))) break;
        if (rocprim::radix_sort_pairs(tmp, tsz, kb, vb, (size_t)n, 0, 32, st) != hipSuccess) { r = GPR_E_HIP; break; }
        dev_free(tmp);
        tmp = nullptr;
        uint32_t *order = vb.current();
        if ((r = dev_alloc((void **)&perm, 4ull * n, "perm"))) break;
        hipLaunchKernelGGL(k_invert_perm, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, order, n, perm);
        // 2. relabelled edge keys, sorted
        if ((r = dev_alloc((void **)&e0, 8ull * (m ? m : 1), "relabel keys"))) break;
        if ((r = dev_alloc((void **)&e1, 8ull * (m ? m : 1), "relabel keys"))) break;
        if (m)
            hipLaunchKernelGGL(k_relabel_keys, dim3(grid_for(m, RL_TILE, 16384)), dim3(256), 0, st, g->index, g->nbrs, n,
                               m, perm, s, e0);
        rocprim::double_buffer<uint64_t> eb(e0, e1);
        tsz = 0;
        if (rocprim::radix_sort_keys(nullptr, tsz, eb, (size_t)m, 0, 2 * s, st) != hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc(&tmp, tsz, "sort workspace"))) break;
        if (rocprim::radix_sort_keys(tmp, tsz, eb, (size_t)m, 0, 2 * s, st) != hipSuccess) { r = GPR_E_HIP; break; }
        dev_free(tmp);
        tmp = nullptr;
        uint64_t *sorted = eb.current();
        // free the alternate key buffer and the old CSR before the new one
        dev_free(eb.alternate());
        if (eb.alternate() == e0) e0 = nullptr; else e1 = nullptr;
        if ((r = dev_alloc((void **)&nindex, 8ull * ((size_t)n + 1), "CSR index"))) break;
        if ((r = dev_alloc((void **)&nnbrs, 4ull * (m ? m : 1), "CSR neighbours"))) break;
        if ((r = keys_to_csr(sorted, m, s, n, nnbrs, nindex, st))) break;
        dev_free(sorted);
        if (sorted == e0) e0 = nullptr; else e1 = nullptr;
        // keep order (new -> original) and perm (original -> new)
        g->order = order;
        g->perm = perm;
        if (order == v0) v0 = nullptr; else v1 = nullptr;
        perm = nullptr;
        r = graph_replace_csr(g, nindex, nnbrs);
        nindex = nullptr;
        nnbrs = nullptr;
    } while (0);
    if (r == GPR_E_HIP && !g->order) set_error("device sort failed while reordering");
    dev_free(k0);
    dev_free(k1);
    dev_free(v0);
    dev_free(v1);
    dev_free(perm);
    dev_free(e0);
    dev_free(e1);
    dev_free(nindex);
    dev_free(nnbrs);
    dev_free(tmp);
    return r;
}

} // namespace gpr

extern "C" int gpr_graph_reorder(gpr_graph *g, int mode) {
    if (!g) return GPR_E_INVAL;
    if (mode == GPR_ORDER_NONE) return GPR_OK;
    GPR_REQUIRE(mode == GPR_ORDER_DEGREE, "unknown reorder mode %d", mode);
    return gpr::reorder_impl(g);
}

extern "C" int gpr_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                       uint64_t seed, gpr_graph **out) {
    if (!out) return GPR_E_INVAL;
    return gpr::rmat_impl(scale, edge_factor, a, b, c, seed, out);
}
