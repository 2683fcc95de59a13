// kcore.hip -- k-core decomposition: the coreness of KC_bucket
// (bench/KC_bucket.C:104-330, buckets bench/buckets.h).
//
// KC_bucket peels a symmetric graph: the count-down degree of every vertex
// starts at its degree; the bucket of the smallest count K is taken as the
// frontier, and every out-edge of the frontier counts its destination down
// by one unless it is already at K (count_down_value, dsl/ast/redop.h:
// 1131-1200).  A destination that reaches K joins the current bucket; when
// the bucket is empty K moves to the next non-empty one.  The final counts
// are the core numbers, which are unique: any peeling order gives the same
// output.
//
// Here the whole peeling is ONE cooperative launch (the levels are many and
// mostly small: hundreds of distinct core values and several rounds each
// on RMAT graphs), with grid barriers between its phases:
//   level K, pass A  minimum count over the candidate list (vertices not yet
//                    peeled; level 0 takes every vertex without a list);
//   level K, pass B  candidates with count K form the frontier, those above
//                    K the next level's candidate list (peeled ones drop);
//   round            the frontier's out-edges count their destinations down;
//                    the one decrement that takes a destination from K+1 to
//                    K appends it to the next round's frontier, so every
//                    vertex is listed once and no dedup bitmap is needed.
// The count-down is a plain peek (counts only fall, so a stale value that is
// at most K already proves there is nothing to do), then atomicSub; a
// decrement that finds the count at or below K (a race with another
// decrement at the floor) is undone with atomicAdd before the round's
// barrier, so after every barrier each count is exact.
// Frontier lists as in the fused push: rows up to KC_LIGHT_MAX edges one
// lane per vertex, longer rows one wave per KC_SLICE-edge slice; appends are
// wave-aggregated.  List counts rotate over three slots per round, candidate
// counts and level minima over two slots per level, each cleared one phase
// after its last read and at least one barrier before its next write.
// Memory ordering and the barrier: grid_sync.hpp.  Termination: every block
// reads the same minimum after a barrier and leaves when no candidate
// remains; a barrier that makes no progress for the spin limit sets the
// error flag, which every wait also watches, so the grid always drains.
#include "internal.hpp"
#include "device.hpp"
#include "grid_sync.hpp"
#include "kcore.hpp"

#include <algorithm>
#include <cstddef>

namespace gpr {

constexpr int KC_THREADS = 1024;
constexpr int KC_WAVES = KC_THREADS / 64;
constexpr uint32_t KC_LIGHT_MAX = 32;
constexpr uint32_t KC_SLICE = 256;
constexpr int KC_U = 4;
static_assert(KC_SLICE == 64 * KC_U, "a slice is one pass of the wave");

struct KcCtl {
    unsigned long long arrive;        // groups arrived at the barrier (monotonic)
    unsigned long long nlight[3];     // frontier list lengths, rotating per round
    unsigned long long nheavy[3];
    unsigned long long ncand[2];      // candidate list lengths, alternating per level
    unsigned long long edges;         // out-edges of peeled vertices examined
    unsigned int error;               // 1: a barrier wait made no progress
    unsigned int levels;              // levels with a non-empty frontier
    unsigned int rounds;              // rounds run
    unsigned int kmax;                // largest core number
    unsigned int kmin[2];             // level minima, alternating
    unsigned long long pad[GS_LINE - 13];
    unsigned long long group[GS_MAX_GROUPS * GS_LINE];
};
static_assert(offsetof(KcCtl, group) == GS_LINE * sizeof(unsigned long long), "one line of counters");

struct KcLists {
    uint32_t *cand[2];
    uint32_t *light[2];
    uint64_t *heavy[2];               // (vertex << 32) | slice
};

__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t kc_slices(uint32_t dg) {
    return dg <= KC_LIGHT_MAX ? 0u : (dg + KC_SLICE - 1) / KC_SLICE;
}

// Append the vertices v[j] with put[j] to the frontier lists of slot c3 /
// buffer cb (light or heavy by degree).  Wave-uniform: every lane calls it.
__device__ __forceinline__ void kc_append(const bool (&put)[KC_U], const uint32_t (&v)[KC_U],
                                          const uint32_t *__restrict__ deg, const KcLists &L, int cb,
                                          unsigned long long *nlight, unsigned long long *nheavy) {
    uint32_t dg[KC_U];
    uint32_t cl = 0, ch = 0;
#pragma unroll
    for (int j = 0; j < KC_U; ++j) {
        dg[j] = put[j] ? deg[v[j]] : 0u;
        if (put[j]) {
            const uint32_t s = kc_slices(dg[j]);
            cl += s == 0;
            ch += s;
        }
    }
    if (!__ballot(cl | ch)) return;
    const uint32_t il = wave_incl_scan(cl), ih = wave_incl_scan(ch);
    const uint32_t tl = __shfl(il, 63, 64), th = __shfl(ih, 63, 64);
    unsigned long long bl = 0, bh = 0;
    if (lane() == 0) {
        if (tl) bl = __hip_atomic_fetch_add(nlight, (unsigned long long)tl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (th) bh = __hip_atomic_fetch_add(nheavy, (unsigned long long)th, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    bl = bcast64(bl) + (il - cl);
    bh = bcast64(bh) + (ih - ch);
#pragma unroll
    for (int j = 0; j < KC_U; ++j) {
        if (put[j]) {
            const uint32_t s = kc_slices(dg[j]);
            if (s == 0) {
                __hip_atomic_store(&L.light[cb][bl++], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                for (uint32_t k = 0; k < s; ++k)
                    __hip_atomic_store(&L.heavy[cb][bh++], ((uint64_t)v[j] << 32) | k, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Count down up to KC_U destinations per lane (where act[j]) with floor K;
// those that reach K go to the next frontier.  Wave-uniform.
__device__ __forceinline__ void kc_count_down(const bool (&act)[KC_U], const uint32_t (&d)[KC_U], uint32_t K,
                                              uint32_t *core, const uint32_t *__restrict__ deg, const KcLists &L,
                                              int nb, unsigned long long *nlight, unsigned long long *nheavy) {
    bool need[KC_U];
#pragma unroll
    for (int j = 0; j < KC_U; ++j) need[j] = act[j] && core[d[j]] > K;
    uint32_t old[KC_U];
#pragma unroll
    for (int j = 0; j < KC_U; ++j) old[j] = need[j] ? atomicSub(&core[d[j]], 1u) : 0u;
    bool put[KC_U];
#pragma unroll
    for (int j = 0; j < KC_U; ++j) {
        if (need[j] && old[j] <= K) atomicAdd(&core[d[j]], 1u);   // raced at the floor: undo
        put[j] = need[j] && old[j] == K + 1;
    }
    kc_append(put, d, deg, L, nb, nlight, nheavy);
}

__global__ void __launch_bounds__(KC_THREADS)
k_kcore(uint32_t n, const uint64_t *__restrict__ index, const uint32_t *__restrict__ nbrs,
        const uint32_t *__restrict__ deg, uint32_t *core, KcLists L, KcCtl *ctl, unsigned spin_limit) {
    __shared__ int s_ok;
    const unsigned ln = lane();
    const unsigned long long W = (unsigned long long)gridDim.x * KC_WAVES;
    // wave rank interleaved over the blocks: small phases land on many CUs
    const unsigned long long wid = (unsigned long long)(threadIdx.x / 64) * gridDim.x + blockIdx.x;
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    unsigned ph = 0;          // barriers passed
    unsigned r = 0;           // rounds run
    uint32_t lo = 0;          // vertices not yet peeled have a count >= lo
    unsigned long long my_edges = 0;
    bool ok = true;
    for (unsigned lvl = 0; ok; ++lvl) {
        const int lb = lvl & 1, nlb = (lvl + 1) & 1;
        const unsigned long long nc = lvl == 0 ? n : ld_agent(&ctl->ncand[lb]);
        const uint32_t *cand = L.cand[lb];
        if (leader) {
            // last read in level lvl-1 (before two barriers ago); written
            // again after this level's first barrier
            __hip_atomic_store(&ctl->ncand[nlb], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->kmin[nlb], ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // pass A: the smallest count among the candidates not yet peeled
        uint32_t mn = ~0u;
        for (unsigned long long i = wid * 64 + ln; i < nc; i += W * 64) {
            const uint32_t v = lvl == 0 ? (uint32_t)i : ld32(&cand[i]);
            const uint32_t c = ld32(&core[v]);
            if (c >= lo && c < mn) mn = c;
        }
        mn = wave_min(mn);
        if (ln == 0 && mn != ~0u)
            __hip_atomic_fetch_min(&ctl->kmin[lb], mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!grid_barrier(&ctl->arrive, &ctl->error, ctl->group, ph++, spin_limit, &s_ok)) break;
        const uint32_t K = __hip_atomic_load(&ctl->kmin[lb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (K == ~0u) break;                       // every vertex peeled
        // pass B: count K -> this level's frontier; above K -> candidates
        {
            const int c3 = r % 3, cb = r & 1;
            for (unsigned long long base = wid * 64; base < nc; base += W * 64) {
                const unsigned long long i = base + ln;
                uint32_t v = 0, c = 0;
                if (i < nc) {
                    v = lvl == 0 ? (uint32_t)i : ld32(&cand[i]);
                    c = ld32(&core[v]);
                }
                const bool live = i < nc && c >= lo;
                const bool front = live && c == K, keep = live && c > K;
                bool put[KC_U] = {front, false, false, false};
                uint32_t vv[KC_U] = {v, v, v, v};
                kc_append(put, vv, deg, L, cb, &ctl->nlight[c3], &ctl->nheavy[c3]);
                const uint64_t km = __ballot(keep);
                if (km) {
                    unsigned long long b = 0;
                    if (ln == 0)
                        b = __hip_atomic_fetch_add(&ctl->ncand[nlb], (unsigned long long)__popcll(km),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    b = bcast64(b);
                    if (keep)
                        __hip_atomic_store(&L.cand[nlb][b + mbcnt(km)], v, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (!grid_barrier(&ctl->arrive, &ctl->error, ctl->group, ph++, spin_limit, &s_ok)) break;
        // rounds of level K
        for (;;) {
            const int c3 = r % 3, cb = r & 1;
            const int n3 = (r + 1) % 3, nb = (r + 1) & 1;
            const unsigned long long nL = ld_agent(&ctl->nlight[c3]);
            const unsigned long long nH = ld_agent(&ctl->nheavy[c3]);
            if (nL == 0 && nH == 0) break;
            if (leader) {
                // slot (r+2)%3 was last read in round r-1
                const int o3 = (r + 2) % 3;
                __hip_atomic_store(&ctl->nlight[o3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&ctl->nheavy[o3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // light rows: one lane per vertex, KC_U edges in flight
            for (unsigned long long base = wid * 64; base < nL; base += W * 64) {
                const unsigned long long i = base + ln;
                uint64_t e = 0, e1 = 0;
                if (i < nL) {
                    const uint32_t v = ld32(&L.light[cb][i]);
                    e = index[v];
                    e1 = index[v + 1];
                    my_edges += e1 - e;
                }
                while (__ballot(e < e1)) {
                    bool act[KC_U];
                    uint32_t d[KC_U];
#pragma unroll
                    for (int j = 0; j < KC_U; ++j) {
                        act[j] = e + j < e1;
                        d[j] = act[j] ? nbrs[e + j] : 0u;
                    }
                    e = e + KC_U < e1 ? e + KC_U : e1;
                    kc_count_down(act, d, K, core, deg, L, nb, &ctl->nlight[n3], &ctl->nheavy[n3]);
                }
            }
            // heavy rows: one wave per slice
            for (unsigned long long j = wid; j < nH; j += W) {
                const uint64_t h = ld64(&L.heavy[cb][j]);
                const uint32_t v = (uint32_t)(h >> 32), sl = (uint32_t)h;
                const uint64_t rb = index[v], re = index[v + 1];
                const uint64_t eb = rb + (uint64_t)sl * KC_SLICE;
                const uint64_t ee = eb + KC_SLICE < re ? eb + KC_SLICE : re;
                if (ln == 0) my_edges += ee - eb;
                bool act[KC_U];
                uint32_t d[KC_U];
#pragma unroll
                for (int u = 0; u < KC_U; ++u) {
                    const uint64_t e = eb + (uint64_t)u * 64 + ln;
                    act[u] = e < ee;
                    d[u] = act[u] ? nbrs[e] : 0u;
                }
                kc_count_down(act, d, K, core, deg, L, nb, &ctl->nlight[n3], &ctl->nheavy[n3]);
            }
            ++r;
            if (!grid_barrier(&ctl->arrive, &ctl->error, ctl->group, ph++, spin_limit, &s_ok)) {
                ok = false;
                break;
            }
        }
        lo = K + 1;
        if (leader) {
            __hip_atomic_store(&ctl->kmax, K, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->levels, lvl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->rounds, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    my_edges = wave_sum(my_edges);
    if (ln == 0 && my_edges)
        __hip_atomic_fetch_add(&ctl->edges, my_edges, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(256) k_kc_init(uint32_t n, const uint32_t *__restrict__ deg, uint32_t *core,
                                                 KcCtl *ctl) {
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) core[v] = deg[v];
    if (blockIdx.x == 0) {
        unsigned long long *w = (unsigned long long *)ctl;
        for (unsigned i = threadIdx.x; i < sizeof(KcCtl) / 8; i += 256) w[i] = 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            ctl->kmin[0] = ~0u;
            ctl->kmin[1] = ~0u;
        }
    }
}

// ---- host side ------------------------------------------------------------------
struct KcBuf {
    size_t cand, light, heavy, total;   // byte offsets / size
    size_t nheavy;                      // heavy entries per buffer
};

static KcBuf kc_layout(const gpr_graph *g) {
    KcBuf b;
    const size_t n = std::max<size_t>(g->n, 1);
    b.nheavy = n + g->m / KC_SLICE + 64;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    b.cand = 0;
    b.light = al(2 * n * sizeof(uint32_t));
    b.heavy = b.light + al(2 * n * sizeof(uint32_t));
    b.total = b.heavy + 2 * b.nheavy * sizeof(uint64_t);
    return b;
}

static int kc_blocks(unsigned *nblocks) {
    static int per_cu = 0;
    if (!per_cu) GPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_kcore, KC_THREADS, 0));
    if (per_cu < 1) {
        set_error("kcore: the kernel cannot be resident (occupancy 0)");
        return GPR_E_HIP;
    }
    const double want = tun().fusion_blocks;
    unsigned cap = (unsigned)(per_cu * num_cus());
    if (cap > GS_GROUP * GS_MAX_GROUPS) cap = GS_GROUP * GS_MAX_GROUPS;
    const unsigned nb = want >= 1 ? (unsigned)want : (unsigned)num_cus();
    *nblocks = nb < cap ? nb : cap;
    return GPR_OK;
}

int kcore_peel(gpr_graph *g, gpr_prop *core, KcStats *st) {
    GPR_REQUIRE(g->symmetric, "kcore: the graph must be symmetric (KC_bucket peels an undirected graph)");
    GPR_REQUIRE(core && core->g == g && core->dtype == GPR_U32, "kcore: coreness must be a u32 property");
    const KcBuf b = kc_layout(g);
    if (g->kbuf_bytes < b.total) {
        dev_free(g->kbuf);
        g->kbuf = nullptr;
        g->kbuf_bytes = 0;
        GPR_TRY(dev_alloc(&g->kbuf, b.total, "kcore lists"));
        g->kbuf_bytes = b.total;
    }
    if (!g->kctl) GPR_TRY(dev_alloc(&g->kctl, sizeof(KcCtl), "kcore control"));
    KcCtl *ctl = (KcCtl *)g->kctl;
    char *base = (char *)g->kbuf;
    const size_t n = std::max<size_t>(g->n, 1);
    KcLists L;
    L.cand[0] = (uint32_t *)(base + b.cand);
    L.cand[1] = L.cand[0] + n;
    L.light[0] = (uint32_t *)(base + b.light);
    L.light[1] = L.light[0] + n;
    L.heavy[0] = (uint64_t *)(base + b.heavy);
    L.heavy[1] = L.heavy[0] + b.nheavy;
    uint32_t *cp = (uint32_t *)core->d;
    hipLaunchKernelGGL(k_kc_init, dim3(grid_for(g->n, 256, 4096)), dim3(256), 0, g->stream, g->n, g->deg, cp, ctl);
    GPR_HIP(hipGetLastError());
    const double sl = tun().fusion_spin_limit;
    unsigned spin = sl < 0 ? 0u : sl > 2147483647.0 ? 2147483647u : (unsigned)sl;
    unsigned nblocks = 0;
    GPR_TRY(kc_blocks(&nblocks));
    {
        ProfScope ps(g, "kcore");
        uint32_t nn = g->n;
        const uint64_t *index = g->index;
        const uint32_t *nbrs = g->nbrs, *deg = g->deg;
        void *args[] = {(void *)&nn, (void *)&index, (void *)&nbrs, (void *)&deg, (void *)&cp, (void *)&L,
                        (void *)&ctl, (void *)&spin};
        GPR_HIP(hipLaunchCooperativeKernel((const void *)k_kcore, dim3(nblocks), dim3(KC_THREADS), args, 0,
                                           g->stream));
    }
    KcCtl h;
    GPR_HIP(hipMemcpyAsync(&h, ctl, GS_LINE * sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    if (h.error) {
        set_error("kcore: a grid barrier made no progress (level %u, round %u)", h.levels, h.rounds);
        return GPR_E_HIP;
    }
    st->kmax = g->n ? h.kmax : 0;
    st->levels = h.levels;
    st->rounds = h.rounds;
    st->edges = h.edges;
    return GPR_OK;
}

} // namespace gpr
