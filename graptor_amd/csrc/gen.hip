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

__global__ void k_rmat_keys(uint32_t scale, uint32_t t0, uint32_t t1, uint32_t t2, uint64_t seed, uint64_t count,
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

// Unique keys -> neighbour column and row offsets.  index[u] = first
// position whose source is >= u.
__global__ void k_keys_to_csr(const uint64_t *__restrict__ keys, uint64_t m, uint32_t scale, uint32_t n,
                              uint32_t *nbrs, uint64_t *index) {
    const uint64_t mask = (1ull << scale) - 1;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p <= m; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t su = p < m ? (uint32_t)(keys[p] >> scale) : n;
        const int64_t sp = p == 0 ? -1 : (int64_t)(keys[p - 1] >> scale);
        if (p < m) nbrs[p] = (uint32_t)(keys[p] & mask);
        for (int64_t u = sp + 1; u <= (int64_t)su; ++u) index[u] = p;
    }
}

static int rmat_impl(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, gpr_graph **out) {
    GPR_REQUIRE(scale >= 1 && scale <= 31, "rmat scale %u out of range [1,31]", scale);
    GPR_REQUIRE(ef >= 1, "rmat edge factor must be positive");
    GPR_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "rmat probabilities invalid");
    const uint32_t n = 1u << scale;
    const uint64_t count = (uint64_t)ef << scale;
    const uint64_t nkeys = 2 * count;
    const uint32_t t0 = (uint32_t)(a * 4294967296.0);
    const uint32_t t1 = (uint32_t)((a + b) * 4294967296.0);
    const uint32_t t2 = (uint32_t)((a + b + c) * 4294967296.0);
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
        hipLaunchKernelGGL(k_keys_to_csr, dim3(grid_for(m + 1, 256, 65536)), dim3(256), 0, st, other, m, scale, n,
                           g->nbrs, g->index);
        if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) {
            set_error("CSR build failed");
            r = GPR_E_HIP;
            break;
        }
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

} // namespace gpr

extern "C" int gpr_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                       uint64_t seed, gpr_graph **out) {
    if (!out) return GPR_E_INVAL;
    return gpr::rmat_impl(scale, edge_factor, a, b, c, seed, out);
}
