// msbfs.hip -- multi-source BFS: up to 64 BFSLVLv runs as the bit lanes of
// one traversal.
//
// Per vertex: seen[v] (bit j: reached from source j), cur[v] (bits that
// reached v in the previous level), nxt[v] (bits arriving this level).  One
// level is
//   push (sparse): for the active list, every edge (u,v) ORs cur[u] & ~seen[v]
//                  into nxt[v] (64-bit atomicOr) -- workgroups take tiles of
//                  consecutive frontier edges as in the single-source push;
//   pull (dense):  every v not yet seen by all sources ORs cur[u] over its
//                  in-edges, stopping once every missing bit has arrived;
// followed by a dense finalize: new = nxt & ~seen, seen |= new, cur = new,
// levels[j][v] = level+1 for every bit j of new, and the next active list
// (wave-aggregated append) with its out-degree sum.  Direction follows the
// reference's rule on the aggregate frontier (nactv + nacte <= m/20 pushes,
// api.h:150-153).  Each lane's result equals a single-source BFSLVLv run
// (bench/BFSLVLv.C:110-254) from that source.
#include "internal.hpp"
#include "device.hpp"

#include <algorithm>
#include <vector>

namespace gpr {

constexpr int MS_THREADS = 256;
constexpr int MS_TILE = 1024;

__global__ void k_ms_init(uint32_t n, uint32_t k, uint32_t *levels, uint64_t *seen, uint64_t *cur, uint64_t *nxt) {
    const uint64_t total = (uint64_t)k * n;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x)
        levels[i] = n + 1;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        seen[v] = 0;
        cur[v] = 0;
        nxt[v] = 0;
    }
}

__global__ void k_ms_sources(const uint32_t *__restrict__ src, uint32_t k, uint32_t n, uint32_t *levels, uint64_t *seen,
                             uint64_t *cur) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t s = src[j];
            seen[s] |= 1ull << j;
            cur[s] |= 1ull << j;
            levels[(uint64_t)j * n + s] = 0;
        }
}

// push over the active list (degree prefix in pos/btot as for the
// single-source push)
__global__ void __launch_bounds__(MS_THREADS)
k_ms_push(const uint32_t *__restrict__ flist, uint32_t kf, const uint64_t *__restrict__ pos,
          const uint64_t *__restrict__ btot, uint64_t E, const uint64_t *__restrict__ index,
          const uint32_t *__restrict__ nbrs, const uint64_t *__restrict__ cur, const uint64_t *__restrict__ seen,
          uint64_t *nxt) {
    __shared__ uint64_t s_pos[MS_TILE + 1];
    __shared__ uint32_t s_v[MS_TILE + 1];
    __shared__ uint32_t s_i0, s_cnt;
    const uint64_t ntiles = (E + MS_TILE - 1) / MS_TILE;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t e0 = t * MS_TILE, e1 = min(E, e0 + MS_TILE);
        if (threadIdx.x < 64) {
            // the tile's first and last sources, by one wave's 64-way search
            const uint32_t i0 = wave_ub_pos(pos, btot, kf, e0, 0u) - 1;
            const uint32_t i1 = wave_ub_pos(pos, btot, kf, e1 - 1, i0 + 1);
            if (threadIdx.x == 0) {
                s_i0 = i0;
                s_cnt = i1 - i0;
            }
        }
        __syncthreads();
        const uint32_t i0 = s_i0, cnt = s_cnt;
        const bool in_lds = cnt <= MS_TILE + 1;
        if (in_lds)
            for (uint32_t j = threadIdx.x; j < cnt; j += MS_THREADS) {
                s_pos[j] = pos[i0 + j] + btot[(i0 + j) / SCAN_BLOCK];
                s_v[j] = flist[i0 + j];
            }
        __syncthreads();
        for (uint64_t e = e0 + threadIdx.x; e < e1; e += MS_THREADS) {
            uint32_t j;
            if (in_lds) {
                uint32_t lo = 0, hi = cnt;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (s_pos[mid] <= e) lo = mid + 1; else hi = mid;
                }
                j = lo - 1;
            } else {
                uint32_t lo = i0, hi = i0 + cnt;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (pos[mid] + btot[mid / SCAN_BLOCK] <= e) lo = mid + 1; else hi = mid;
                }
                j = lo - 1 - i0;
            }
            const uint32_t s = in_lds ? s_v[j] : flist[i0 + j];
            const uint64_t ps = in_lds ? s_pos[j] : pos[i0 + j] + btot[(i0 + j) / SCAN_BLOCK];
            const uint32_t d = nbrs[index[s] + (e - ps)];
            const uint64_t bits = cur[s] & ~seen[d];
            if (bits && (nxt[d] & bits) != bits) atomicOr((unsigned long long *)&nxt[d], bits);
        }
        __syncthreads();
    }
}

// pull: wave per 64 destinations, lane-serial rows up to 32 in-edges, the
// wave together on heavier rows; stop once every missing source arrived
__global__ void __launch_bounds__(MS_THREADS)
k_ms_pull(uint32_t n, const uint64_t *__restrict__ cindex, const uint32_t *__restrict__ cnbrs,
          const uint64_t *__restrict__ cur, const uint64_t *__restrict__ seen, uint64_t full, uint64_t *nxt) {
    const uint32_t ngroups = (n + 63) / 64;
    const uint32_t wid = (blockIdx.x * MS_THREADS + threadIdx.x) / 64, nw = gridDim.x * (MS_THREADS / 64);
    const unsigned ln = lane();
    for (uint32_t grp = wid; grp < ngroups; grp += nw) {
        const uint32_t d = grp * 64 + ln;
        const bool valid = d < n;
        const uint64_t sd = valid ? seen[d] : full;
        const uint64_t miss = full & ~sd;
        uint64_t acc = 0;
        uint64_t beg = 0, end = 0;
        if (miss) {
            beg = cindex[d];
            end = cindex[d + 1];
        }
        const bool heavy = miss && end - beg > 32;
        if (miss && !heavy)
            for (uint64_t e = beg; e < end; ++e) {
                acc |= cur[cnbrs[e]];
                if ((acc & miss) == miss) break;
            }
        uint64_t hm = __ballot(heavy);
        while (hm) {
            const int L = __ffsll((long long)hm) - 1;
            hm &= hm - 1;
            const uint64_t b = __shfl(beg, L, 64), en = __shfl(end, L, 64), ms = __shfl(miss, L, 64);
            uint64_t a = 0;
            for (uint64_t base = b; base < en; base += 64) {
                const uint64_t e = base + ln;
                uint64_t v = e < en ? cur[cnbrs[e]] : 0ull;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
                a |= v;
                if ((a & ms) == ms) break;
            }
            if ((int)ln == L) acc = a;
        }
        if (valid && (acc & miss)) nxt[d] = acc & miss;
    }
}

// finalize a level: levels of the new bits, seen/cur update, next active
// list with its out-degree sum
__global__ void __launch_bounds__(MS_THREADS)
k_ms_finalize(uint32_t n, uint32_t level, uint32_t *levels, uint64_t *seen, uint64_t *cur, uint64_t *nxt,
              const uint32_t *__restrict__ deg, uint32_t *list, DevCounters *ctr) {
    __shared__ uint64_t red[MS_THREADS / 64];
    uint64_t my_e = 0;
    const uint32_t ngroups = (n + 63) / 64;
    const uint32_t wid = (blockIdx.x * MS_THREADS + threadIdx.x) / 64, nw = gridDim.x * (MS_THREADS / 64);
    const unsigned ln = lane();
    for (uint32_t grp = wid; grp < ngroups; grp += nw) {
        const uint32_t v = grp * 64 + ln;
        uint64_t nb = 0;
        if (v < n) {
            const uint64_t x = nxt[v];
            cur[v] = 0;
            if (x) {
                nb = x & ~seen[v];
                nxt[v] = 0;
                if (nb) {
                    seen[v] |= nb;
                    cur[v] = nb;
                    uint64_t t = nb;
                    while (t) {
                        const int j = __ffsll((long long)t) - 1;
                        t &= t - 1;
                        levels[(uint64_t)j * n + v] = level + 1;
                    }
                    my_e += deg[v];
                }
            }
        }
        const uint64_t w = __ballot(nb != 0);
        if (w) {
            uint32_t base = 0;
            if (ln == 0) base = atomicAdd(&ctr->listpos, (unsigned)__popcll(w));
            base = __shfl(base, 0, 64);
            if (nb) list[base + mbcnt(w)] = v;
        }
    }
    const uint64_t te = block_sum(my_e, red);
    if (threadIdx.x == 0 && te) atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)te);
}

// edges traversed: sum over sources of the out-degrees of reached vertices
__global__ void __launch_bounds__(256)
k_ms_edges(uint32_t n, const uint64_t *__restrict__ seen, const uint32_t *__restrict__ deg, DevCounters *ctr) {
    __shared__ uint64_t red[4];
    uint64_t s = 0;
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256)
        s += (uint64_t)__popcll(seen[v]) * deg[v];
    s = block_sum(s, red);
    if (threadIdx.x == 0 && s) atomicAdd(&ctr_stripe(ctr)->edges, (unsigned long long)s);
}

__global__ void k_ms_scatter(const uint32_t *__restrict__ src, const uint32_t *__restrict__ order, uint32_t n,
                             uint32_t k, uint32_t *dst) {
    const uint64_t total = (uint64_t)k * n;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = i / n, v = i % n;
        dst[j * n + order[v]] = src[i];
    }
}

int scan_frontier_degrees(gpr_graph *g, const uint32_t *list, uint32_t k);

} // namespace gpr

using namespace gpr;

extern "C" int gpr_msbfs(gpr_graph *g, const uint32_t *sources, int k, const gpr_run_opts *opts, uint32_t *levels,
                         gpr_run_info *info) {
    if (!g || !sources || k <= 0 || k > 64) {
        set_error("msbfs: need 1..64 sources");
        return GPR_E_INVAL;
    }
    const uint32_t n = g->n;
    for (int j = 0; j < k; ++j) GPR_REQUIRE(sources[j] < n, "msbfs: source %u out of range", sources[j]);
    const int dir_opt = opts ? opts->direction : GPR_DIR_AUTO;
    const float pct = opts ? opts->threshold_pct : 0.f;   // 0: BFSLVLv's default_threshold
    uint32_t *lev = nullptr, *src = nullptr, *list[2] = {nullptr, nullptr};
    uint64_t *seen = nullptr, *cur = nullptr, *nxt = nullptr;
    int r = GPR_OK;
    gpr_run_info inf = {};
    do {
        if ((r = dev_alloc((void **)&lev, 4ull * k * n, "msbfs levels"))) break;
        if ((r = dev_alloc((void **)&seen, 8ull * n, "msbfs seen"))) break;
        if ((r = dev_alloc((void **)&cur, 8ull * n, "msbfs frontier"))) break;
        if ((r = dev_alloc((void **)&nxt, 8ull * n, "msbfs next"))) break;
        if ((r = dev_alloc((void **)&list[0], 4ull * n, "msbfs list"))) break;
        if ((r = dev_alloc((void **)&list[1], 4ull * n, "msbfs list"))) break;
        if ((r = dev_alloc((void **)&src, 4ull * k, "msbfs sources"))) break;
        std::vector<uint32_t> s(sources, sources + k);
        std::vector<uint32_t> uniq(s);
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        if (hipMemcpyAsync(src, s.data(), 4ull * k, hipMemcpyHostToDevice, g->stream) != hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = map_ids(g, src, k, true))) break;
        // the first active list: the distinct sources
        if (hipMemcpyAsync(list[0], uniq.data(), 4ull * uniq.size(), hipMemcpyHostToDevice, g->stream) != hipSuccess) {
            r = GPR_E_HIP;
            break;
        }
        if ((r = map_ids(g, list[0], (uint32_t)uniq.size(), true))) break;
        if (!g->rev0) { (void)hipEventCreate(&g->rev0); (void)hipEventCreate(&g->rev1); }
        (void)hipEventRecord(g->rev0, g->stream);
        hipLaunchKernelGGL(k_ms_init, dim3(grid_for((uint64_t)k * n, 256, 16384)), dim3(256), 0, g->stream, n, (uint32_t)k,
                           lev, seen, cur, nxt);
        hipLaunchKernelGGL(k_ms_sources, dim3(1), dim3(64), 0, g->stream, src, (uint32_t)k, n, lev, seen, cur);
        if ((r = reset_counters(g))) break;
        // out-degree sum of the sources
        uint64_t nacte = 0;
        {
            std::vector<uint64_t> ix(2);
            for (uint32_t v : uniq) {
                uint32_t vi = v;
                if ((r = to_internal_id(g, v, &vi))) break;
                if (hipMemcpy(ix.data(), g->index + vi, 16, hipMemcpyDeviceToHost) != hipSuccess) { r = GPR_E_HIP; break; }
                nacte += ix[1] - ix[0];
            }
            if (r) break;
        }
        uint32_t nactv = (uint32_t)uniq.size();
        const uint64_t full = k == 64 ? ~0ull : ((1ull << k) - 1);
        int cl = 0;
        uint32_t level = 0;
        while (nactv > 0) {
            const bool sparse = dir_opt == GPR_DIR_PUSH ||
                                (dir_opt != GPR_DIR_PULL && threshold_sparse(pct, nactv, nacte, g->m));
            if (sparse) {
                if ((r = scan_frontier_degrees(g, list[cl], nactv))) break;
                if (nacte)
                    hipLaunchKernelGGL(k_ms_push, dim3(grid_for(nacte, MS_TILE, 8192)), dim3(MS_THREADS), 0, g->stream,
                                       list[cl], nactv, g->scan, g->scan_blk, nacte, g->index, g->nbrs, cur, seen, nxt);
                inf.push_iters++;
            } else {
                hipLaunchKernelGGL(k_ms_pull, dim3(grid_for((n + 63) / 64, MS_THREADS / 64, 2048)), dim3(MS_THREADS), 0,
                                   g->stream, n, g->cindex, g->cnbrs, cur, seen, full, nxt);
            }
            if ((r = reset_counters(g))) break;
            hipLaunchKernelGGL(k_ms_finalize, dim3(grid_for((n + 63) / 64, MS_THREADS / 64, 2048)), dim3(MS_THREADS), 0,
                               g->stream, n, level, lev, seen, cur, nxt, g->deg, list[cl ^ 1], g->ctr);
            if (hipGetLastError() != hipSuccess) { r = GPR_E_HIP; break; }
            if ((r = read_counters(g))) break;
            nactv = g->hctr->listpos;
            nacte = g->hctr->nacte;
            cl ^= 1;
            ++level;
            inf.iters++;
        }
        if (r) break;
        (void)hipEventRecord(g->rev1, g->stream);
        (void)hipEventSynchronize(g->rev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, g->rev0, g->rev1);
        inf.ms = ms;
        if ((r = reset_counters(g))) break;
        hipLaunchKernelGGL(k_ms_edges, dim3(grid_for(n, 256, 4096)), dim3(256), 0, g->stream, n, seen, g->deg, g->ctr);
        if ((r = read_counters(g))) break;
        inf.edges = g->hctr->edges;
        if (levels) {
            const uint32_t *outp = lev;
            uint32_t *tmp = nullptr;
            if (g->order) {
                if ((r = dev_alloc((void **)&tmp, 4ull * k * n, "msbfs levels"))) break;
                hipLaunchKernelGGL(k_ms_scatter, dim3(grid_for((uint64_t)k * n, 256, 16384)), dim3(256), 0, g->stream, lev,
                                   g->order, n, (uint32_t)k, tmp);
                outp = tmp;
            }
            hipError_t e = hipMemcpyAsync(levels, outp, 4ull * k * n, hipMemcpyDeviceToHost, g->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
            dev_free(tmp);
            if (e != hipSuccess) { r = GPR_E_HIP; break; }
        }
    } while (0);
    if (r == GPR_E_HIP) set_error("msbfs: HIP error %s", hipGetErrorString(hipGetLastError()));
    dev_free(lev);
    dev_free(seen);
    dev_free(cur);
    dev_free(nxt);
    dev_free(list[0]);
    dev_free(list[1]);
    dev_free(src);
    if (r == GPR_OK && info) *info = inf;
    return r;
}
