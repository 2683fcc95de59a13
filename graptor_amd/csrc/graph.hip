// graph.hip -- device graphs: upload, b2 IO, degree and pull partitions,
// vertex properties.
//
// Reference: GraphCSx (graph/GraphCSx.h:293-2806): readFromBinaryFile
// :2137-2264, writeToBinaryFile :2009-2027, build_degree :2797-2805,
// findHighestDegreeVertex :2446-2451; api::vertexprop api/vertexprop.h:28-143.
#include "internal.hpp"
#include "device.hpp"

#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>

namespace gpr {

__global__ void k_degree(const uint64_t *__restrict__ index, uint32_t n, uint32_t *deg) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        deg[v] = (uint32_t)(index[v + 1] - index[v]);
}

// Host-side construction of the pull partitions from the CSC offsets:
//   buckets: rows r with floor((cindex[r] + r) / STREAM_ITEMS) == b.  Both
//            terms are non-decreasing, so a bucket holds at most
//            STREAM_ITEMS rows and its rows start within a window of
//            STREAM_ITEMS edges; a row longer than LONG_DEG can only be the
//            last row of its bucket and is handled by the long-row chunks.
//   long rows: in-degree > LONG_DEG, with a prefix of LONG_CHUNK chunk counts.
static int build_partitions(gpr_graph *g, const uint64_t *cindex) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint64_t nb = (m + n) / STREAM_ITEMS + 1;
    if (nb > 0xffffffffull) {
        set_error("graph too large for the pull partition (%llu buckets)", (unsigned long long)nb);
        return GPR_E_INVAL;
    }
    std::vector<uint32_t> bstart(nb + 1, n);
    uint64_t kprev = 0;
    bool first = true;
    for (uint32_t r = 0; r < n; ++r) {
        uint64_t k = (cindex[r] + r) / STREAM_ITEMS;
        uint64_t lo = first ? 0 : kprev + 1;
        for (uint64_t b = lo; b <= k; ++b) bstart[b] = r;
        kprev = k;
        first = false;
    }
    // buckets after the last row's key stay n (empty)
    std::vector<uint32_t> lrows;
    std::vector<uint32_t> lbase(1, 0);
    uint64_t best = 0;
    uint32_t hdv = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint64_t d = cindex[r + 1] - cindex[r];
        if (d > LONG_DEG) {
            lrows.push_back(r);
            lbase.push_back(lbase.back() + (uint32_t)((d + LONG_CHUNK - 1) / LONG_CHUNK));
        }
    }
    g->nbuckets = (uint32_t)nb;
    GPR_TRY(dev_alloc((void **)&g->bucket, sizeof(uint32_t) * (nb + 1), "pull buckets"));
    GPR_HIP(hipMemcpy(g->bucket, bstart.data(), sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice));
    g->nlong = (uint32_t)lrows.size();
    g->nlong_chunks = lbase.back();
    GPR_TRY(dev_alloc((void **)&g->long_rows, sizeof(uint32_t) * std::max<size_t>(1, lrows.size()), "long rows"));
    GPR_TRY(dev_alloc((void **)&g->long_base, sizeof(uint32_t) * lbase.size(), "long rows"));
    if (!lrows.empty())
        GPR_HIP(hipMemcpy(g->long_rows, lrows.data(), sizeof(uint32_t) * lrows.size(), hipMemcpyHostToDevice));
    GPR_HIP(hipMemcpy(g->long_base, lbase.data(), sizeof(uint32_t) * lbase.size(), hipMemcpyHostToDevice));
    GPR_TRY(dev_alloc((void **)&g->long_partial, sizeof(double) * std::max<uint32_t>(1, g->nlong_chunks), "long partials"));
    (void)best;
    (void)hdv;
    return GPR_OK;
}

// Out-degree maximum, ties to the lowest vertex ID.
uint32_t highest_degree(const uint64_t *index, uint32_t n) {
    uint64_t best = 0;
    uint32_t v = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint64_t d = index[r + 1] - index[r];
        if (d > best) { best = d; v = r; }
    }
    return v;
}

// Everything after the CSR/CSC arrays are on the device.  host_index and
// host_cindex are host copies of the offsets (they may alias).
static int finish_graph(gpr_graph *g, const uint64_t *host_index, const uint64_t *host_cindex) {
    const uint32_t n = g->n;
    GPR_TRY(dev_alloc((void **)&g->deg, sizeof(uint32_t) * std::max<uint32_t>(n, 1), "degrees"));
    hipLaunchKernelGGL(k_degree, dim3(grid_for(n, 256, 8192)), dim3(256), 0, g->stream, g->index, n, g->deg);
    GPR_HIP(hipGetLastError());
    GPR_TRY(build_partitions(g, host_cindex));
    g->hdv = highest_degree(host_index, n);
    g->max_deg = n ? (uint32_t)(host_index[g->hdv + 1] - host_index[g->hdv]) : 0;
    GPR_TRY(dev_alloc((void **)&g->ctr, COUNTER_BYTES, "counters"));
    GPR_HIP(hipMemset(g->ctr, 0, COUNTER_BYTES));
    // the counter read-back target: fine-grained (coherent) pinned host
    // memory that a one-wave kernel writes the sums into, so that a step's
    // counts come back with the stream's completion and no copy
    GPR_HIP(hipHostMalloc((void **)&g->hctr, sizeof(DevCounters), hipHostMallocMapped | hipHostMallocCoherent));
    GPR_HIP(hipHostGetDevicePointer((void **)&g->hctr_dev, g->hctr, 0));
    GPR_TRY(dev_alloc((void **)&g->dedup, sizeof(uint64_t) * nwords(n), "dedup bitmap"));
    GPR_HIP(hipMemset(g->dedup, 0, sizeof(uint64_t) * nwords(n)));
    GPR_TRY(dev_alloc((void **)&g->scan, sizeof(uint64_t) * ((size_t)n + 1), "frontier prefix"));
    GPR_TRY(dev_alloc((void **)&g->scan_blk, sizeof(uint64_t) * ((size_t)n / SCAN_BLOCK + 2), "frontier prefix"));
    GPR_TRY(dev_alloc((void **)&g->dsum_slices, sizeof(double) * 256, "reduction slices"));
    GPR_TRY(dev_alloc((void **)&g->prs, sizeof(PrScalars), "pagerank scalars"));
    GPR_HIP(hipMemset(g->prs, 0, sizeof(PrScalars)));
    GPR_HIP(hipEventCreate(&g->ev0));
    GPR_HIP(hipEventCreate(&g->ev1));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

static void graph_free(gpr_graph *g) {
    if (!g) return;
    if (g->stream) hipStreamSynchronize(g->stream);
    prof_flush(g);
    dev_free(g->cindex32);
    if (g->cindex != g->index) dev_free(g->cindex);
    if (g->cnbrs != g->nbrs) dev_free(g->cnbrs);
    dev_free(g->index);
    dev_free(g->nbrs);
    dev_free(g->deg);
    dev_free(g->order);
    dev_free(g->perm);
    dev_free(g->bucket);
    dev_free(g->long_rows);
    dev_free(g->long_base);
    dev_free(g->long_partial);
    dev_free(g->ctr);
    if (g->hctr) hipHostFree(g->hctr);
    dev_free(g->dedup);
    dev_free(g->scan);
    dev_free(g->scan_blk);
    dev_free(g->partial);
    dev_free(g->partial_f);
    dev_free(g->dsum_slices);
    dev_free(g->prs);
    dev_free(g->pol_params);
    cold_free(g);
    dev_free(g->fbuf);
    dev_free(g->fctl);
    dev_free(g->kbuf);
    dev_free(g->kctl);
    if (g->ev0) hipEventDestroy(g->ev0);
    if (g->ev1) hipEventDestroy(g->ev1);
    if (g->rev0) hipEventDestroy(g->rev0);
    if (g->rev1) hipEventDestroy(g->rev1);
    for (void *p : g->pool_list) dev_free(p);
    for (void *p : g->pool_bits) dev_free(p);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
}

static int new_graph(uint32_t n, uint64_t m, int symmetric, gpr_graph **out) {
    gpr_graph *g = new gpr_graph();
    g->n = n;
    g->m = m;
    g->symmetric = symmetric ? 1 : 0;
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        delete g;
        return GPR_E_HIP;
    }
    *out = g;
    return GPR_OK;
}

// Host CSR validation (CheckGraph semantics, tools/CheckGraph.C:10-87,
// without the sortedness requirement, which the engine does not need).
static int check_csr(const uint64_t *index, const uint32_t *nbrs, uint32_t n, uint64_t m) {
    if (index[0] != 0 || index[n] != m) {
        set_error("CSR index[0]=%llu index[n]=%llu, expected 0 and m=%llu", (unsigned long long)index[0],
                  (unsigned long long)index[n], (unsigned long long)m);
        return GPR_E_FORMAT;
    }
    for (uint32_t v = 0; v < n; ++v)
        if (index[v + 1] < index[v]) {
            set_error("CSR index decreases at vertex %u", v);
            return GPR_E_FORMAT;
        }
    for (uint64_t e = 0; e < m; ++e)
        if (nbrs[e] >= n) {
            set_error("CSR neighbour %u at edge %llu out of range n=%u", nbrs[e], (unsigned long long)e, n);
            return GPR_E_FORMAT;
        }
    return GPR_OK;
}

int graph_from_device(gpr_graph *g) {
    // The CSR arrays are already on the device (generator path): bring the
    // offsets to the host for the partition build.
    std::vector<uint64_t> h(g->n + 1);
    GPR_HIP(hipMemcpy(h.data(), g->index, sizeof(uint64_t) * (g->n + 1), hipMemcpyDeviceToHost));
    g->cindex = g->index;
    g->cnbrs = g->nbrs;
    return finish_graph(g, h.data(), h.data());
}

int graph_new_device(uint32_t n, uint64_t m, gpr_graph **out) {
    GPR_TRY(new_graph(n, m, 1, out));
    return GPR_OK;
}

void graph_release(gpr_graph *g) { graph_free(g); }

uint32_t graph_max_degree(const gpr_graph *g) { return g->max_deg; }

// Install a new (relabelled, symmetric) CSR in place of the current one and
// rebuild the degree array and pull partitions; scratch stays.
int graph_replace_csr(gpr_graph *g, uint64_t *index, uint32_t *nbrs) {
    GPR_HIP(hipStreamSynchronize(g->stream));
    dev_free(g->cindex32);
    g->cindex32 = nullptr;
    if (g->cindex != g->index) dev_free(g->cindex);
    if (g->cnbrs != g->nbrs) dev_free(g->cnbrs);
    dev_free(g->index);
    dev_free(g->nbrs);
    dev_free(g->bucket);
    dev_free(g->long_rows);
    dev_free(g->long_base);
    dev_free(g->long_partial);
    cold_free(g);
    g->bucket = nullptr;
    g->long_rows = g->long_base = nullptr;
    g->long_partial = nullptr;
    g->index = g->cindex = index;
    g->nbrs = g->cnbrs = nbrs;
    hipLaunchKernelGGL(k_degree, dim3(grid_for(g->n, 256, 8192)), dim3(256), 0, g->stream, g->index, g->n, g->deg);
    GPR_HIP(hipGetLastError());
    std::vector<uint64_t> h((size_t)g->n + 1);
    GPR_HIP(hipMemcpyAsync(h.data(), g->index, sizeof(uint64_t) * h.size(), hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return build_partitions(g, h.data());
}

template <class T> __global__ void k_gather_pos(const T *__restrict__ src, const uint32_t *__restrict__ order, uint32_t n, T *dst) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[order[i]];
}
template <class T> __global__ void k_scatter_pos(const T *__restrict__ src, const uint32_t *__restrict__ order, uint32_t n, T *dst) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[order[i]] = src[i];
}
__global__ void k_map_ids(uint32_t *ids, uint32_t k, const uint32_t *__restrict__ map) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) ids[i] = map[ids[i]];
}

int map_ids(gpr_graph *g, uint32_t *d_ids, uint32_t k, bool to_internal) {
    if (!g->order || k == 0) return GPR_OK;
    hipLaunchKernelGGL(k_map_ids, dim3(grid_for(k, 256, 8192)), dim3(256), 0, g->stream, d_ids, k,
                       to_internal ? g->perm : g->order);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

int to_internal_id(gpr_graph *g, uint32_t orig, uint32_t *internal) {
    if (!g->order) {
        *internal = orig;
        return GPR_OK;
    }
    GPR_HIP(hipMemcpyAsync(internal, g->perm + orig, 4, hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

// Move a per-vertex array between original and internal positions through
// the scan scratch (8(n+1) bytes, free between edgemap steps).
static int permute_positions(gpr_graph *g, const void *src, void *dst, size_t elem, bool to_internal) {
    const unsigned grid = grid_for(g->n, 256, 8192);
#define PERM_CASE(T)                                                                                          \
    if (to_internal)                                                                                          \
        hipLaunchKernelGGL(k_gather_pos<T>, dim3(grid), dim3(256), 0, g->stream, (const T *)src, g->order, g->n, (T *)dst); \
    else                                                                                                      \
        hipLaunchKernelGGL(k_scatter_pos<T>, dim3(grid), dim3(256), 0, g->stream, (const T *)src, g->order, g->n, (T *)dst);
    if (elem == 2) { PERM_CASE(uint16_t) }
    else if (elem == 4) { PERM_CASE(uint32_t) }
    else if (elem == 8) { PERM_CASE(uint64_t) }
    else {
        set_error("permute: element size %zu", elem);
        return GPR_E_INVAL;
    }
#undef PERM_CASE
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

// Giant components send most vertices to one address: read first and only
// take the atomic when it can lower the value (the running minimum drops
// quickly, so almost every vertex skips it).
__global__ void k_recode_min(const uint32_t *__restrict__ label, const uint32_t *__restrict__ order, uint32_t n,
                             uint32_t *minorig) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t l = label[i], o = order ? order[i] : i;
        if (o < __hip_atomic_load(&minorig[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&minorig[l], o);
    }
}
__global__ void k_recode_out(const uint32_t *__restrict__ label, const uint32_t *__restrict__ order, uint32_t n,
                             const uint32_t *__restrict__ minorig, uint32_t *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[order ? order[i] : i] = minorig[label[i]];
}
// Values that are vertex IDs (BFSv parents), in original order and as
// original IDs; values with the top bit set (unset entries) are kept.
__global__ void k_ids_out(const uint32_t *__restrict__ v, const uint32_t *__restrict__ order, uint32_t n,
                          uint32_t *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t x = v[i];
        out[order ? order[i] : i] = (x & 0x80000000u) ? x : (order ? order[x] : x);
    }
}
__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_graph_create(const uint64_t *index, const uint32_t *nbrs, uint32_t n, uint64_t m, int symmetric,
                                gpr_graph **out) {
    if (!index || (!nbrs && m) || !out) return GPR_E_INVAL;
    GPR_REQUIRE(n < 0xfffffffeu, "n=%u too large for 32-bit vertex IDs", n);
    GPR_TRY(check_csr(index, nbrs, n, m));
    gpr_graph *g = nullptr;
    GPR_TRY(new_graph(n, m, symmetric, &g));
    int r = GPR_OK;
    std::vector<uint64_t> cidx;
    std::vector<uint32_t> cnb;
    do {
        if ((r = dev_alloc((void **)&g->index, sizeof(uint64_t) * ((size_t)n + 1), "CSR index"))) break;
        if ((r = dev_alloc((void **)&g->nbrs, sizeof(uint32_t) * std::max<uint64_t>(m, 1), "CSR neighbours"))) break;
        if (hipMemcpy(g->index, index, sizeof(uint64_t) * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
            (m && hipMemcpy(g->nbrs, nbrs, sizeof(uint32_t) * m, hipMemcpyHostToDevice) != hipSuccess)) {
            set_error("upload of the CSR failed");
            r = GPR_E_HIP;
            break;
        }
        const uint64_t *hc = index;
        if (symmetric) {
            g->cindex = g->index;
            g->cnbrs = g->nbrs;
        } else {
            // transpose by counting sort; rows come out ascending
            cidx.assign((size_t)n + 1, 0);
            cnb.resize(std::max<uint64_t>(m, 1));
            for (uint64_t e = 0; e < m; ++e) cidx[nbrs[e] + 1]++;
            for (uint32_t v = 0; v < n; ++v) cidx[v + 1] += cidx[v];
            std::vector<uint64_t> fill(cidx.begin(), cidx.end() - 1);
            for (uint32_t s = 0; s < n; ++s)
                for (uint64_t e = index[s]; e < index[s + 1]; ++e) cnb[fill[nbrs[e]]++] = s;
            if ((r = dev_alloc((void **)&g->cindex, sizeof(uint64_t) * ((size_t)n + 1), "CSC index"))) break;
            if ((r = dev_alloc((void **)&g->cnbrs, sizeof(uint32_t) * std::max<uint64_t>(m, 1), "CSC neighbours"))) break;
            if (hipMemcpy(g->cindex, cidx.data(), sizeof(uint64_t) * ((size_t)n + 1), hipMemcpyHostToDevice) !=
                    hipSuccess ||
                (m && hipMemcpy(g->cnbrs, cnb.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice) != hipSuccess)) {
                set_error("upload of the CSC failed");
                r = GPR_E_HIP;
                break;
            }
            hc = cidx.data();
        }
        r = finish_graph(g, index, hc);
    } while (0);
    if (r != GPR_OK) {
        graph_free(g);
        return r;
    }
    *out = g;
    return GPR_OK;
}

extern "C" int gpr_graph_load_b2(const char *path, int symmetric, gpr_graph **out) {
    if (!path || !out) return GPR_E_INVAL;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("cannot open '%s'", path);
        return GPR_E_IO;
    }
    uint64_t hdr[8];
    if (fread(hdr, sizeof(hdr), 1, f) != 1) {
        fclose(f);
        set_error("'%s': short header", path);
        return GPR_E_FORMAT;
    }
    // GraphCSx.h:2240-2248: versions 2 and 3, sizeof(VID)=4, sizeof(EID)=8
    if ((hdr[0] != 2 && hdr[0] != 3) || hdr[4] != 4 || hdr[5] != 8 || hdr[2] >= 0xfffffffeull) {
        fclose(f);
        set_error("'%s': unsupported header {%llu,%llu,%llu,%llu,%llu,%llu}", path, (unsigned long long)hdr[0],
                  (unsigned long long)hdr[1], (unsigned long long)hdr[2], (unsigned long long)hdr[3],
                  (unsigned long long)hdr[4], (unsigned long long)hdr[5]);
        return GPR_E_FORMAT;
    }
    const uint32_t n = (uint32_t)hdr[2];
    const uint64_t m = hdr[3];
    fseek(f, 0, SEEK_END);
    long long len = ftell(f);
    long long want = 64 + 8ll * n + (hdr[0] == 3 ? 8 : 0) + 4ll * (long long)m;
    if (len < want) {
        fclose(f);
        set_error("'%s': file has %lld bytes, header implies %lld", path, len, want);
        return GPR_E_FORMAT;
    }
    fseek(f, 64, SEEK_SET);
    std::vector<uint64_t> index((size_t)n + 1);
    std::vector<uint32_t> nbrs(std::max<uint64_t>(m, 1));
    bool ok = fread(index.data(), sizeof(uint64_t), n, f) == n;
    index[n] = m;   // GraphCSx.h:2253
    if (ok && hdr[0] == 3) ok = fseek(f, 8, SEEK_CUR) == 0;
    if (ok && m) ok = fread(nbrs.data(), sizeof(uint32_t), m, f) == m;
    fclose(f);
    if (!ok) {
        set_error("'%s': read failed", path);
        return GPR_E_IO;
    }
    return gpr_graph_create(index.data(), nbrs.data(), n, m, symmetric, out);
}

extern "C" int gpr_graph_save_b2(const gpr_graph *g, const char *path) {
    if (!g || !path) return GPR_E_INVAL;
    std::vector<uint64_t> index((size_t)g->n + 1);
    std::vector<uint32_t> nbrs(std::max<uint64_t>(g->m, 1));
    GPR_TRY(gpr_graph_download(g, index.data(), nbrs.data()));
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot create '%s'", path);
        return GPR_E_IO;
    }
    uint64_t hdr[8] = {2, 1, g->n, g->m, sizeof(uint32_t), sizeof(uint64_t), 0, 0};
    bool ok = fwrite(hdr, sizeof(hdr), 1, f) == 1 && fwrite(index.data(), sizeof(uint64_t), g->n, f) == g->n &&
              (g->m == 0 || fwrite(nbrs.data(), sizeof(uint32_t), g->m, f) == g->m);
    fclose(f);
    if (!ok) {
        set_error("write of '%s' failed", path);
        return GPR_E_IO;
    }
    return GPR_OK;
}

extern "C" int gpr_graph_info(const gpr_graph *g, uint32_t *n, uint64_t *m, int *symmetric) {
    if (!g) return GPR_E_INVAL;
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (symmetric) *symmetric = g->symmetric;
    return GPR_OK;
}

extern "C" int gpr_graph_download(const gpr_graph *g, uint64_t *index, uint32_t *nbrs) {
    if (!g) return GPR_E_INVAL;
    if (index) GPR_HIP(hipMemcpyAsync(index, g->index, sizeof(uint64_t) * ((size_t)g->n + 1), hipMemcpyDeviceToHost, g->stream));
    if (nbrs && g->m) GPR_HIP(hipMemcpyAsync(nbrs, g->nbrs, sizeof(uint32_t) * g->m, hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

extern "C" int gpr_graph_highest_degree_vertex(const gpr_graph *g, uint32_t *v) {
    if (!g || !v) return GPR_E_INVAL;
    *v = g->hdv;
    return GPR_OK;
}

extern "C" int gpr_graph_destroy(gpr_graph *g) {
    if (!g) return GPR_E_INVAL;
    graph_free(g);
    return GPR_OK;
}

extern "C" int gpr_prop_create(gpr_graph *g, int dtype, gpr_prop **out) {
    if (!g || !out) return GPR_E_INVAL;
    size_t elem = dtype == GPR_U32 || dtype == GPR_F32 ? 4 : dtype == GPR_F64 || dtype == GPR_U64 ? 8
                  : dtype == GPR_U16 ? 2 : 0;
    GPR_REQUIRE(elem, "unknown property dtype %d", dtype);
    gpr_prop *p = new gpr_prop();
    p->g = g;
    p->dtype = dtype;
    p->elem = elem;
    int r = dev_alloc(&p->d, elem * std::max<uint32_t>(g->n, 1), "vertex property");
    if (r) {
        delete p;
        return r;
    }
    GPR_HIP(hipMemsetAsync(p->d, 0, elem * std::max<uint32_t>(g->n, 1), g->stream));
    *out = p;
    return GPR_OK;
}

extern "C" int gpr_prop_upload(gpr_prop *p, const void *host) {
    if (!p || !host) return GPR_E_INVAL;
    gpr_graph *g = p->g;
    if (g->order) {
        GPR_HIP(hipMemcpyAsync(g->scan, host, p->elem * g->n, hipMemcpyHostToDevice, g->stream));
        GPR_TRY(permute_positions(g, g->scan, p->d, p->elem, true));
    } else {
        GPR_HIP(hipMemcpyAsync(p->d, host, p->elem * g->n, hipMemcpyHostToDevice, g->stream));
    }
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

extern "C" int gpr_prop_download(const gpr_prop *p, void *host) {
    if (!p || !host) return GPR_E_INVAL;
    gpr_graph *g = p->g;
    if (g->order) {
        GPR_TRY(permute_positions(g, p->d, g->scan, p->elem, false));
        GPR_HIP(hipMemcpyAsync(host, g->scan, p->elem * g->n, hipMemcpyDeviceToHost, g->stream));
    } else {
        GPR_HIP(hipMemcpyAsync(host, p->d, p->elem * g->n, hipMemcpyDeviceToHost, g->stream));
    }
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

namespace gpr {
int ids_download(gpr_graph *g, const gpr_prop *p, uint32_t *host) {
    GPR_REQUIRE(p && host && p->dtype == GPR_U32 && p->g == g, "ids_download: a u32 property of this graph");
    const uint32_t n = g->n;
    uint32_t *out = (uint32_t *)g->scan;   // 8(n+1) bytes of scratch
    hipLaunchKernelGGL(k_ids_out, dim3(grid_for(n, 256, 8192)), dim3(256), 0, g->stream, (const uint32_t *)p->d,
                       g->order, n, out);
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(host, out, 4ull * n, hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}
} // namespace gpr

// CCv recode (bench/CCv.C:121-169): every vertex gets the minimum original
// vertex ID of its component; labels[] is indexed by original ID.
extern "C" int gpr_cc_recode(const gpr_prop *p, uint32_t *labels) {
    if (!p || !labels || p->dtype != GPR_U32) return GPR_E_INVAL;
    gpr_graph *g = p->g;
    const uint32_t n = g->n;
    uint32_t *minorig = (uint32_t *)g->scan;
    uint32_t *out = minorig + n;
    const unsigned grid = grid_for(n, 256, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, g->stream, minorig, n, 0xffffffffu);
    hipLaunchKernelGGL(k_recode_min, dim3(grid), dim3(256), 0, g->stream, (const uint32_t *)p->d, g->order, n, minorig);
    hipLaunchKernelGGL(k_recode_out, dim3(grid), dim3(256), 0, g->stream, (const uint32_t *)p->d, g->order, n, minorig, out);
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(labels, out, 4ull * n, hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

extern "C" int gpr_graph_order(const gpr_graph *g, uint32_t *order) {
    if (!g || !order) return GPR_E_INVAL;
    if (!g->order) {
        for (uint32_t i = 0; i < g->n; ++i) order[i] = i;
        return GPR_OK;
    }
    GPR_HIP(hipMemcpyAsync(order, g->order, 4ull * g->n, hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

extern "C" int gpr_graph_remap_id(gpr_graph *g, uint32_t orig, uint32_t *internal) {
    if (!g || !internal || orig >= g->n) return GPR_E_INVAL;
    return to_internal_id(g, orig, internal);
}

extern "C" int gpr_graph_original_id(gpr_graph *g, uint32_t internal, uint32_t *orig) {
    if (!g || !orig || internal >= g->n) return GPR_E_INVAL;
    if (!g->order) {
        *orig = internal;
        return GPR_OK;
    }
    GPR_HIP(hipMemcpyAsync(orig, g->order + internal, 4, hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    return GPR_OK;
}

extern "C" int gpr_prop_destroy(gpr_prop *p) {
    if (!p) return GPR_E_INVAL;
    if (p->g && p->g->stream) hipStreamSynchronize(p->g->stream);
    dev_free(p->d);
    delete p;
    return GPR_OK;
}
