// batch.hip -- connected components over thousands of small graphs in one
// persistent launch.
//
// A batch is the disjoint union of G graphs stored as one CSR with global
// vertex IDs plus vertex offsets voff[G+1] (graph g owns vertices
// voff[g]..voff[g+1]).  The labels equal CCv's canonical output
// (bench/CCv.C:121-169: minimum vertex ID of the component), in global IDs.
//
// k_batch_cc is persistent: one 1024-thread workgroup per CU pulls graph
// indices from a queue (graphs sorted by decreasing edge count, so the big
// ones start first) until the queue is empty.  A graph of at most 65536
// vertices is solved entirely in LDS: 16-bit parent pointers (128 KB),
// union-find where a link hooks the larger root under the smaller by a CAS
// on the 32-bit word holding the parent, so every root is the minimum of
// its tree and the final root is the component minimum; then a pointer-
// jumping pass writes the labels.  Larger graphs run the same algorithm on
// a global-memory parent array.  The queue counter is cleared by a one-lane
// kernel captured in the same hipGraph (no memset node), so the pair
// replays without host work.
#include "internal.hpp"
#include "device.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <rocprim/rocprim.hpp>

struct gpr_batch {
    uint32_t count = 0;
    uint32_t n = 0;
    uint64_t m = 0;
    uint64_t *index = nullptr;
    uint32_t *nbrs = nullptr;
    uint32_t *voff = nullptr;      // count+1
    uint32_t *work = nullptr;      // graph order of the queue
    uint32_t *labels = nullptr;    // n
    uint32_t *gparent = nullptr;   // scratch for graphs above LDS capacity
    uint32_t max_nv = 0;
    unsigned int *queue = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<uint32_t> hvoff;
};

namespace gpr {

constexpr int BCC_THREADS = 1024;
constexpr uint32_t BCC_LDS_V = 65536;

template <class P> struct Parents;
// 16-bit parents packed two per 32-bit LDS word
template <> struct Parents<uint16_t> {
    uint32_t *w;
    __device__ uint32_t get(uint32_t x) const {
        const uint32_t v = __hip_atomic_load(&w[x >> 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return (x & 1) ? (v >> 16) : (v & 0xffffu);
    }
    __device__ void set(uint32_t x, uint32_t p) const {
        // only used during initialisation and the final pass (one writer)
        uint16_t *h = reinterpret_cast<uint16_t *>(w);
        h[x] = (uint16_t)p;
    }
    // hook root r under q if r is still a root; true on success
    __device__ bool hook(uint32_t r, uint32_t q) const {
        uint32_t *a = &w[r >> 1];
        const int sh = (r & 1) * 16;
        uint32_t old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        while (((old >> sh) & 0xffffu) == r) {
            const uint32_t nw = (old & ~(0xffffu << sh)) | (q << sh);
            if (__hip_atomic_compare_exchange_strong(a, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return true;
        }
        return false;
    }
};
template <> struct Parents<uint32_t> {
    uint32_t *w;
    __device__ uint32_t get(uint32_t x) const {
        return __hip_atomic_load(&w[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ void set(uint32_t x, uint32_t p) const { w[x] = p; }
    __device__ bool hook(uint32_t r, uint32_t q) const {
        uint32_t exp = r;
        return __hip_atomic_compare_exchange_strong(&w[r], &exp, q, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    }
};

template <class PT>
__device__ __forceinline__ uint32_t uf_find(const PT &P, uint32_t x) {
    uint32_t p = P.get(x);
    while (p != x) {
        x = p;
        p = P.get(x);
    }
    return x;
}

template <class PT>
__device__ __forceinline__ void uf_link(const PT &P, uint32_t u, uint32_t v) {
    while (true) {
        uint32_t ru = uf_find(P, u), rv = uf_find(P, v);
        if (ru == rv) return;
        if (ru > rv) {
            const uint32_t t = ru;
            ru = rv;
            rv = t;
        }
        if (P.hook(rv, ru)) return;
        u = ru;
        v = rv;
    }
}

// Edge offsets are clamped to [0, m] and edges that leave the graph are
// skipped, so a malformed batch in caller memory (gpr_batch_cc_noalloc,
// which cannot check it on the host) yields wrong labels, never an access
// outside the arrays.
template <class PT>
__device__ void solve_graph(const PT &P, uint32_t base, uint32_t nv, const uint64_t *__restrict__ index,
                            const uint32_t *__restrict__ nbrs, uint64_t m, uint32_t *labels) {
    for (uint32_t v = threadIdx.x; v < nv; v += BCC_THREADS) P.set(v, v);
    __syncthreads();
    const unsigned ln = lane();
    // vertex per lane; rows above 32 edges are walked by the whole wave.
    // Each undirected edge is linked once, from its smaller endpoint.
    for (uint32_t v0 = (threadIdx.x & ~63u); v0 < nv; v0 += BCC_THREADS) {
        const uint32_t v = v0 + ln;
        uint64_t b = 0, e = 0;
        if (v < nv) {
            b = min(index[base + v], m);
            e = min(max(index[base + v + 1], b), m);
        }
        const bool heavy = e - b > 32;
        if (!heavy)
            for (uint64_t i = b; i < e; ++i) {
                const uint32_t u = nbrs[i] - base;
                if (u > v && u < nv) uf_link(P, v, u);
            }
        uint64_t hm = __ballot(heavy);
        while (hm) {
            const int L = __ffsll((long long)hm) - 1;
            hm &= hm - 1;
            const uint64_t hb = __shfl(b, L, 64), he = __shfl(e, L, 64);
            const uint32_t hv = v0 + L;
            for (uint64_t i = hb + ln; i < he; i += 64) {
                const uint32_t u = nbrs[i] - base;
                if (u > hv && u < nv) uf_link(P, hv, u);
            }
        }
    }
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < nv; v += BCC_THREADS) labels[base + v] = base + uf_find(P, v);
    __syncthreads();
}

__global__ void k_batch_reset(unsigned int *queue) {
    if (threadIdx.x == 0) *queue = 0;
}

// work: queue order of the graphs (null: index order)
__global__ void __launch_bounds__(BCC_THREADS)
k_batch_cc(uint32_t count, uint32_t n, const uint32_t *__restrict__ work, const uint32_t *__restrict__ voff,
           const uint64_t *__restrict__ index, const uint32_t *__restrict__ nbrs, uint32_t *labels,
           uint32_t *gparent, unsigned int *queue) {
    __shared__ uint32_t words[BCC_LDS_V / 2];
    __shared__ uint32_t s_item;
    const uint64_t m = index[n];
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= count) return;    // every workgroup reaches this exit
        const uint32_t gidx = work ? work[item] : item;
        const uint32_t base = voff[gidx], end = voff[gidx + 1];
        if (end < base || end > n) continue;   // malformed offsets: the graph is skipped
        const uint32_t nv = end - base;
        if (nv <= BCC_LDS_V) {
            Parents<uint16_t> P{words};
            solve_graph(P, base, nv, index, nbrs, m, labels);
        } else {
            // graphs are vertex-disjoint, so each owns its slice of gparent
            Parents<uint32_t> P{gparent + base};
            solve_graph(P, base, nv, index, nbrs, m, labels);
        }
    }
}

// ---- batch generation -----------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t bmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// One thread per generated edge of the whole batch.  Graph g: even g RMAT
// (a=.57 b=.19 c=.19, scale ceil(log2 n), 8n draws, endpoints >= n dropped),
// odd g Erdos-Renyi (3n uniform pairs, mean degree 6 after symmetrising).
__global__ void k_batch_keys(uint32_t count, const uint64_t *__restrict__ ebase, const uint32_t *__restrict__ voff,
                             const uint32_t *__restrict__ gscale, uint64_t seed, uint64_t t0, uint64_t t1, uint64_t t2,
                             uint32_t S, uint64_t total, uint64_t *keys) {
    const uint64_t sentinel = (S >= 32) ? ~0ull : ((1ull << (2 * S)) - 1);
    for (uint64_t E = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; E < total; E += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = count;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ebase[mid] <= E) lo = mid + 1; else hi = mid;
        }
        const uint32_t g = lo - 1;
        const uint64_t e = E - ebase[g];
        const uint32_t nvg = voff[g + 1] - voff[g];
        const uint64_t gs = bmix64(seed ^ bmix64(0x5eed0000ull + g));
        uint32_t u = 0, v = 0;
        if ((g & 1) == 0) {
            uint64_t h = 0;
            const uint32_t sc = gscale[g];
            for (uint32_t l = 0; l < sc; ++l) {
                if ((l & 1) == 0) h = bmix64(gs ^ bmix64((e << 5) | (l >> 1)));
                const uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
                const uint32_t q = r < t0 ? 0u : r < t1 ? 1u : r < t2 ? 2u : 3u;
                u = (u << 1) | (q >> 1);
                v = (v << 1) | (q & 1u);
            }
        } else {
            const uint64_t h = bmix64(gs ^ bmix64(e));
            u = (uint32_t)(((h & 0xffffffffull) * nvg) >> 32);
            v = (uint32_t)(((h >> 32) * nvg) >> 32);
        }
        uint64_t k1 = sentinel, k2 = sentinel;
        if (u != v && u < nvg && v < nvg) {
            const uint64_t gu = voff[g] + u, gv = voff[g] + v;
            k1 = (gu << S) | gv;
            k2 = (gv << S) | gu;
        }
        keys[2 * E] = k1;
        keys[2 * E + 1] = k2;
    }
}

int keys_to_csr_ext(const uint64_t *keys, uint64_t m, uint32_t scale, uint32_t n, uint32_t *nbrs, uint64_t *index,
                    hipStream_t st);

static void batch_free(gpr_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    dev_free(b->index);
    dev_free(b->nbrs);
    dev_free(b->voff);
    dev_free(b->work);
    dev_free(b->labels);
    dev_free(b->gparent);
    dev_free(b->queue);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// queue order (decreasing edge count), scratch, events; index/nbrs/voff set
static int batch_finish(gpr_batch *b) {
    std::vector<uint64_t> idx(b->hvoff.size());
    for (size_t g = 0; g < b->hvoff.size(); ++g)
        GPR_HIP(hipMemcpy(&idx[g], b->index + b->hvoff[g], 8, hipMemcpyDeviceToHost));
    std::vector<uint32_t> work(b->count);
    std::iota(work.begin(), work.end(), 0u);
    std::stable_sort(work.begin(), work.end(), [&](uint32_t x, uint32_t y) {
        return idx[x + 1] - idx[x] > idx[y + 1] - idx[y];
    });
    b->max_nv = 0;
    for (uint32_t g = 0; g < b->count; ++g) b->max_nv = std::max(b->max_nv, b->hvoff[g + 1] - b->hvoff[g]);
    GPR_TRY(dev_alloc((void **)&b->work, 4ull * std::max<uint32_t>(b->count, 1), "batch queue order"));
    GPR_HIP(hipMemcpy(b->work, work.data(), 4ull * b->count, hipMemcpyHostToDevice));
    GPR_TRY(dev_alloc((void **)&b->labels, 4ull * std::max<uint32_t>(b->n, 1), "batch labels"));
    // global parents only for graphs above the LDS capacity
    GPR_TRY(dev_alloc((void **)&b->gparent, b->max_nv > BCC_LDS_V ? 4ull * b->n : 64, "batch parents"));
    GPR_TRY(dev_alloc((void **)&b->queue, 64, "batch queue"));
    GPR_HIP(hipMemset(b->queue, 0, 64));
    GPR_HIP(hipEventCreate(&b->e0));
    GPR_HIP(hipEventCreate(&b->e1));
    return GPR_OK;
}

static int batch_new(gpr_batch **out) {
    gpr_batch *b = new gpr_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return GPR_E_HIP;
    }
    *out = b;
    return GPR_OK;
}

// log-uniform vertex counts in [nmin, nmax] from the seed (host arithmetic)
static void batch_sizes(uint32_t count, uint32_t nmin, uint32_t nmax, uint64_t seed, std::vector<uint32_t> &nv) {
#pragma clang fp contract(off)
    nv.resize(count);
    const double l0 = std::log((double)nmin), l1 = std::log((double)nmax);
    for (uint32_t g = 0; g < count; ++g) {
        const double u = (double)(bmix64(seed ^ bmix64(0x517e0000ull + g)) >> 11) * (1.0 / 9007199254740992.0);
        double x = std::floor(std::exp(l0 + u * (l1 - l0)));
        if (x < nmin) x = nmin;
        if (x > nmax) x = nmax;
        nv[g] = (uint32_t)x;
    }
}

// ---- the entry point that allocates nothing ----------------------------------
// gpr_batch_cc_noalloc's scratch, one copy per GPU for the life of the
// process: word 0 is the work-queue counter (cleared by k_batch_reset at
// every call); from word NOALLOC_PARENTS on, the union-find parents of the
// graphs above the LDS capacity, indexed by global vertex ID, so a batch may
// hold at most NOALLOC_MAX_V vertices.  Host code reaches the array only
// through the address hipGetSymbolAddress returns for the current device.
constexpr uint32_t NOALLOC_MAX_V = 1u << 21;
constexpr uint32_t NOALLOC_PARENTS = 64;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) uint32_t scratch[NOALLOC_PARENTS + NOALLOC_MAX_V];

struct NoallocDevice {
    std::atomic<uint32_t *> ptr{nullptr};
    std::atomic<int> cus{0};
};
static NoallocDevice noalloc_dev[NOALLOC_MAX_DEVICES];

// Once per device, with that device current: the device address of
// `scratch` and the compute-unit count (gpr_init, else the first call).
int batch_noalloc_setup(int dev) {
    GPR_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device %d: no scratch slot", dev);
    NoallocDevice &d = noalloc_dev[dev];
    if (d.ptr.load(std::memory_order_acquire)) return GPR_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        return GPR_E_HIP;
    }
    int cus = 0;
    GPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    d.cus.store(cus > 0 ? cus : 1, std::memory_order_relaxed);
    d.ptr.store((uint32_t *)p, std::memory_order_release);
    return GPR_OK;
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_batch_cc_noalloc(const uint64_t *d_index, const uint32_t *d_nbrs, const uint32_t *d_voff,
                                    uint32_t count, uint32_t n, uint32_t *d_labels, void *stream) {
    GPR_REQUIRE(d_index && d_voff && d_labels && (d_nbrs || n == 0) && count > 0,
                "batch_cc_noalloc: null buffer or empty batch");
    GPR_REQUIRE(n <= NOALLOC_MAX_V, "batch_cc_noalloc: %u vertices exceed the limit of %u", n, NOALLOC_MAX_V);
    int dev = 0;
    GPR_HIP(hipGetDevice(&dev));
    GPR_TRY(batch_noalloc_setup(dev));
    uint32_t *sp = noalloc_dev[dev].ptr.load(std::memory_order_acquire);
    const int cus = noalloc_dev[dev].cus.load(std::memory_order_relaxed);
    hipStream_t st = (hipStream_t)stream;
    const unsigned grid = (unsigned)std::max(1, std::min<int>(cus, (int)std::min<uint32_t>(count, 1u << 30)));
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, (unsigned int *)sp);
    hipLaunchKernelGGL(k_batch_cc, dim3(grid), dim3(BCC_THREADS), 0, st, count, n, (const uint32_t *)nullptr, d_voff,
                       d_index, d_nbrs, d_labels, sp + NOALLOC_PARENTS, (unsigned int *)sp);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

extern "C" int gpr_batch_generate(uint32_t count, uint32_t nmin, uint32_t nmax, uint64_t seed, gpr_batch **out) {
    if (!out || count == 0 || nmin < 2 || nmax < nmin || nmax > (1u << 24)) {
        set_error("batch_generate: need count > 0 and 2 <= nmin <= nmax <= 2^24");
        return GPR_E_INVAL;
    }
    std::vector<uint32_t> nv;
    batch_sizes(count, nmin, nmax, seed, nv);
    std::vector<uint32_t> voff(count + 1, 0), gscale(count);
    std::vector<uint64_t> ebase(count + 1, 0);
    for (uint32_t g = 0; g < count; ++g) {
        voff[g + 1] = voff[g] + nv[g];
        uint32_t sc = 1;
        while ((1u << sc) < nv[g]) ++sc;
        gscale[g] = sc;
        ebase[g + 1] = ebase[g] + ((g & 1) == 0 ? 8ull * nv[g] : 3ull * nv[g]);
    }
    GPR_REQUIRE((uint64_t)voff[count] < 0xfffffffeull, "batch too large");
    const uint32_t N = voff[count];
    uint32_t S = 1;
    while ((1ull << S) < N) ++S;
    const uint64_t total = ebase[count];
    const uint64_t t0 = (uint64_t)(0.57 * 4294967296.0), t1 = (uint64_t)((0.57 + 0.19) * 4294967296.0),
                   t2 = (uint64_t)((0.57 + 0.19 + 0.19) * 4294967296.0);
    gpr_batch *b = nullptr;
    GPR_TRY(batch_new(&b));
    b->count = count;
    b->n = N;
    b->hvoff = voff;
    uint64_t *k0 = nullptr, *k1 = nullptr, *deb = nullptr, *ucount = nullptr;
    uint32_t *dsc = nullptr;
    void *tmp = nullptr;
    int r = GPR_OK;
    hipStream_t st = b->stream;
    do {
        if ((r = dev_alloc((void **)&b->voff, 4ull * (count + 1), "batch offsets"))) break;
        if ((r = dev_alloc((void **)&deb, 8ull * (count + 1), "batch edge offsets"))) break;
        if ((r = dev_alloc((void **)&dsc, 4ull * count, "batch scales"))) break;
        if (hipMemcpy(b->voff, voff.data(), 4ull * (count + 1), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(deb, ebase.data(), 8ull * (count + 1), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dsc, gscale.data(), 4ull * count, hipMemcpyHostToDevice) != hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc((void **)&k0, 16ull * total, "batch keys"))) break;
        if ((r = dev_alloc((void **)&k1, 16ull * total, "batch keys"))) break;
        if ((r = dev_alloc((void **)&ucount, 8, "batch unique count"))) break;
        hipLaunchKernelGGL(k_batch_keys, dim3(grid_for(total, 256, 65536)), dim3(256), 0, st, count, deb, b->voff, dsc,
                           seed, t0, t1, t2, S, total, k0);
        rocprim::double_buffer<uint64_t> db(k0, k1);
        size_t tsz = 0;
        if (rocprim::radix_sort_keys(nullptr, tsz, db, (size_t)(2 * total), 0, 2 * S, st) != hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc(&tmp, tsz, "sort workspace"))) break;
        if (rocprim::radix_sort_keys(tmp, tsz, db, (size_t)(2 * total), 0, 2 * S, st) != hipSuccess) { r = GPR_E_HIP; break; }
        dev_free(tmp);
        tmp = nullptr;
        size_t usz = 0;
        if (rocprim::unique(nullptr, usz, db.current(), db.alternate(), ucount, (size_t)(2 * total),
                            rocprim::equal_to<uint64_t>(), st) != hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc(&tmp, usz, "unique workspace"))) break;
        if (rocprim::unique(tmp, usz, db.current(), db.alternate(), ucount, (size_t)(2 * total),
                            rocprim::equal_to<uint64_t>(), st) != hipSuccess) { r = GPR_E_HIP; break; }
        uint64_t nu = 0, last = 0;
        if (hipMemcpyAsync(&nu, ucount, 8, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) { r = GPR_E_HIP; break; }
        if (nu && hipMemcpy(&last, db.alternate() + nu - 1, 8, hipMemcpyDeviceToHost) != hipSuccess) { r = GPR_E_HIP; break; }
        const uint64_t sentinel = (S >= 32) ? ~0ull : ((1ull << (2 * S)) - 1);
        const uint64_t m = (nu && last == sentinel) ? nu - 1 : nu;
        b->m = m;
        if ((r = dev_alloc((void **)&b->index, 8ull * ((size_t)N + 1), "batch index"))) break;
        if ((r = dev_alloc((void **)&b->nbrs, 4ull * (m ? m : 1), "batch neighbours"))) break;
        if ((r = keys_to_csr_ext(db.alternate(), m, S, N, b->nbrs, b->index, st))) break;
        r = batch_finish(b);
    } while (0);
    if (r == GPR_E_HIP) set_error("batch generation: HIP error %s", hipGetErrorString(hipGetLastError()));
    dev_free(k0);
    dev_free(k1);
    dev_free(deb);
    dev_free(dsc);
    dev_free(ucount);
    dev_free(tmp);
    if (r != GPR_OK) {
        batch_free(b);
        return r;
    }
    *out = b;
    return GPR_OK;
}

extern "C" int gpr_batch_create(const uint64_t *index, const uint32_t *nbrs, const uint32_t *voff, uint32_t count,
                                gpr_batch **out) {
    if (!index || !voff || !out || count == 0) return GPR_E_INVAL;
    const uint32_t N = voff[count];
    const uint64_t m = index[N];
    GPR_REQUIRE(voff[0] == 0 && index[0] == 0, "batch: offsets must start at 0");
    for (uint32_t g = 0; g < count; ++g) {
        GPR_REQUIRE(voff[g + 1] >= voff[g], "batch: vertex offsets decrease at graph %u", g);
        for (uint32_t v = voff[g]; v < voff[g + 1]; ++v) {
            GPR_REQUIRE(index[v + 1] >= index[v], "batch: index decreases at %u", v);
            for (uint64_t e = index[v]; e < index[v + 1]; ++e)
                GPR_REQUIRE(nbrs[e] >= voff[g] && nbrs[e] < voff[g + 1],
                            "batch: edge %llu of graph %u leaves the graph", (unsigned long long)e, g);
        }
    }
    gpr_batch *b = nullptr;
    GPR_TRY(batch_new(&b));
    b->count = count;
    b->n = N;
    b->m = m;
    b->hvoff.assign(voff, voff + count + 1);
    int r = GPR_OK;
    do {
        if ((r = dev_alloc((void **)&b->index, 8ull * ((size_t)N + 1), "batch index"))) break;
        if ((r = dev_alloc((void **)&b->nbrs, 4ull * (m ? m : 1), "batch neighbours"))) break;
        if ((r = dev_alloc((void **)&b->voff, 4ull * (count + 1), "batch offsets"))) break;
        if (hipMemcpy(b->index, index, 8ull * ((size_t)N + 1), hipMemcpyHostToDevice) != hipSuccess ||
            (m && hipMemcpy(b->nbrs, nbrs, 4ull * m, hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(b->voff, voff, 4ull * (count + 1), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("batch upload failed");
            r = GPR_E_HIP;
            break;
        }
        r = batch_finish(b);
    } while (0);
    if (r != GPR_OK) {
        batch_free(b);
        return r;
    }
    *out = b;
    return GPR_OK;
}

extern "C" int gpr_batch_info(const gpr_batch *b, uint32_t *count, uint32_t *n, uint64_t *m) {
    if (!b) return GPR_E_INVAL;
    if (count) *count = b->count;
    if (n) *n = b->n;
    if (m) *m = b->m;
    return GPR_OK;
}

extern "C" int gpr_batch_download(const gpr_batch *b, uint64_t *index, uint32_t *nbrs, uint32_t *voff) {
    if (!b) return GPR_E_INVAL;
    if (index) GPR_HIP(hipMemcpy(index, b->index, 8ull * ((size_t)b->n + 1), hipMemcpyDeviceToHost));
    if (nbrs && b->m) GPR_HIP(hipMemcpy(nbrs, b->nbrs, 4ull * b->m, hipMemcpyDeviceToHost));
    if (voff) std::memcpy(voff, b->hvoff.data(), 4ull * (b->count + 1));
    return GPR_OK;
}

extern "C" int gpr_batch_cc(gpr_batch *b, uint32_t *labels, gpr_run_info *info) {
    if (!b) return GPR_E_INVAL;
    if (!b->exec) {
        // capture: queue reset (a kernel, not a memset node) + the persistent
        // kernel, one workgroup per CU
        int dev = 0, cus = 0;
        GPR_HIP(hipGetDevice(&dev));
        GPR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const unsigned grid = (unsigned)std::max(1, std::min<int>(cus, (int)b->count));
        GPR_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->queue);
        hipLaunchKernelGGL(k_batch_cc, dim3(grid), dim3(BCC_THREADS), 0, b->stream, b->count, b->n, b->work, b->voff,
                           b->index, b->nbrs, b->labels, b->gparent, b->queue);
        hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
        if (ce != hipSuccess) {
            set_error("graph capture: %s", hipGetErrorString(ce));
            return GPR_E_HIP;
        }
        GPR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    GPR_HIP(hipEventRecord(b->e0, b->stream));
    GPR_HIP(hipGraphLaunch(b->exec, b->stream));
    GPR_HIP(hipEventRecord(b->e1, b->stream));
    GPR_HIP(hipEventSynchronize(b->e1));
    float ms = 0.f;
    GPR_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
    if (labels && b->n) GPR_HIP(hipMemcpy(labels, b->labels, 4ull * b->n, hipMemcpyDeviceToHost));
    if (info) {
        gpr_run_info r = {};
        r.iters = 1;
        r.edges = b->m;
        r.ms = ms;
        *info = r;
    }
    return GPR_OK;
}

extern "C" int gpr_batch_destroy(gpr_batch *b) {
    if (!b) return GPR_E_INVAL;
    batch_free(b);
    return GPR_OK;
}
