// fusion.hip -- fused sparse push: Graptor's api::fusion (api/fusion.h:17-140,
// dsl/emap/fusion.h:17-375), where a push step processes every vertex it
// activates at once instead of returning it in the output frontier.  For the
// min relaxes of BFSLVLv and CCv (fusion op "changed -> process now",
// BFSLVLv.C:200-206, CCv.C:562-566) the result is the fixpoint of the relax
// reachable from the input frontier, and the output frontier is empty.
//
// Graptor runs one work-stealing queue per CPU thread.  A shared GPU queue
// serialises every claim on one word, so here the fused step is instead
// ONE cooperative launch that runs the level-synchronous push inside the
// kernel: every level relaxes the out-edges of the current list and builds
// the next one, and a grid barrier separates the levels.  What it saves is
// the host round trip of every level (tens of microseconds: the launches,
// the frontier counts read back), which dominates on graphs of high
// diameter (road networks, meshes), where the levels are many and small.
//
// Per level, two lists:
//   light  vertex IDs of rows with at most FU_LIGHT_MAX edges: one lane per
//          vertex, FU_U edges in flight per lane;
//   heavy  (vertex, slice) pairs of longer rows, one per FU_SLICE edges:
//          one wave per slice, so a hub's row is spread over many waves.
// Appends are wave-aggregated (one atomicAdd per wave and list), and a
// per-level bitmap keeps each vertex at most once in the list being built.
// Two bitmaps alternate: level L appends under bitmap (L+1)&1 and clears,
// for every vertex it processes, its bit in bitmap L&1, so that bitmap is
// clear before level L+1 appends under it.  Vertices without out-edges are
// never listed.  List counts rotate over three slots: level L reads slot
// L%3, appends to (L+1)%3 and clears (L+2)%3 (last read in level L-1).
// Level 0 reads the caller's list in place (only its long rows are listed
// again, as slices), so a large input frontier costs no copy.
//
// A vertex lowered while it is in the current list is appended to the next
// one, so it is processed again with its new value: the result equals the
// unfused steps' whatever the interleaving (BFS levels and CC labels
// bit-exact; BFS visits each vertex exactly once as in the unfused push).
//
// Memory ordering: every word that one block writes and another reads
// within the launch (prop, lists, counts, bitmaps, the barrier counter) is
// accessed with agent-scope atomics or agent-scope (coherent) loads and
// stores.  Before arriving at the barrier each wave waits until its memory
// operations are acknowledged (s_waitcnt vmcnt(0)); after the barrier a
// block reads only through coherent accesses.  No acquire/release fence is
// used: on this part each would write back or invalidate an XCD's L2.
//
// Termination: all blocks read the same list counts after a barrier and
// leave together when both are zero.  A barrier wait that sees no group
// arrive for `spin_limit` polls in a row sets the error flag; every wait also
// leaves when the flag is set, so the grid always drains.
#include "internal.hpp"
#include "device.hpp"
#include "grid_sync.hpp"

namespace gpr {

constexpr int FU_THREADS = 1024;
constexpr int FU_WAVES = FU_THREADS / 64;
constexpr uint32_t FU_LIGHT_MAX = 32;   // rows up to this length: lane per vertex
constexpr uint32_t FU_SLICE = 256;      // longer rows: one wave per slice of this many edges
constexpr int FU_U = 4;                 // edges per lane in flight (FU_SLICE = 64 * FU_U)
static_assert(FU_SLICE == 64 * FU_U, "a slice is one pass of the wave");
constexpr int FU_GROUP = GS_GROUP;
constexpr int FU_MAX_GROUPS = GS_MAX_GROUPS;
constexpr int FU_LINE = GS_LINE;

struct FusedCtl {
    unsigned long long arrive;       // groups arrived at the barrier (monotonic)
    unsigned long long nlight[3];    // list lengths, rotating per level
    unsigned long long nheavy[3];
    unsigned long long nedge[3];     // listed vertices plus their out-edges
    unsigned long long edges;        // edges relaxed
    unsigned long long pops;         // vertices processed
    unsigned int levels;             // levels run
    unsigned int error;              // 1: a barrier wait made no progress
    unsigned int dense;              // 1: stopped at a dense level, handed back
    unsigned int pad32;
    unsigned long long nout;         // vertices of the handed-back frontier
    unsigned long long pad[FU_LINE - 15];
    unsigned long long group[FU_MAX_GROUPS * FU_LINE];   // blocks arrived, per group (monotonic)
};

struct FusedLists {
    uint32_t *light[2];
    uint64_t *heavy[2];              // (vertex << 32) | slice
    uint64_t *bits[2];
};


__device__ __forceinline__ uint32_t slices_of(uint32_t dg) {
    return dg <= FU_LIGHT_MAX ? 0u : (dg + FU_SLICE - 1) / FU_SLICE;
}

// Append the destinations d[j] with push[j] set to the next lists (light or
// heavy by degree dg[j]).  Wave-uniform: every lane calls it.  With
// dense_at set, a wave that has long rows to list first adds its edges to
// the level's count; once that count is above dense_at the level will be
// handed back unprocessed (the check at the next barrier reads the same,
// only growing, count), so its long rows are listed whole, as light
// entries, instead of one entry per slice: a hub-heavy level that ends the
// step costs one entry per vertex, not one per FU_SLICE edges.
__device__ __forceinline__ void append_next(const bool (&push)[FU_U], const uint32_t (&d)[FU_U],
                                            const uint32_t (&dg)[FU_U], const FusedLists &L, int nb,
                                            unsigned long long *nlight, unsigned long long *nheavy,
                                            unsigned long long *nedge, unsigned long long dense_at) {
    uint32_t cl = 0, ch = 0, cv = 0;
    unsigned long long ce = 0;
#pragma unroll
    for (int j = 0; j < FU_U; ++j) {
        if (push[j]) {
            const uint32_t s = slices_of(dg[j]);
            cl += s == 0;
            ch += s;
            cv += 1;
            ce += dg[j] + 1ull;
        }
    }
    if (!__ballot(cl | ch)) return;
    ce = wave_sum(ce);
    const uint32_t th0 = wave_sum(ch);
    int whole = 0;   // list long rows whole: the level is already dense
    if (lane() == 0) {
        if (th0 && dense_at) {
            const unsigned long long before =
                __hip_atomic_fetch_add(nedge, ce, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            whole = before + ce > dense_at;
        } else {
            __hip_atomic_fetch_add(nedge, ce, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    whole = __shfl(whole, 0, 64);
    if (whole) {
        cl = cv;
        ch = 0;
    }
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
    for (int j = 0; j < FU_U; ++j) {
        if (push[j]) {
            const uint32_t s = whole ? 0u : slices_of(dg[j]);
            if (s == 0) {
                __hip_atomic_store(&L.light[nb][bl++], d[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                for (uint32_t k = 0; k < s; ++k)
                    __hip_atomic_store(&L.heavy[nb][bh++], ((uint64_t)d[j] << 32) | k, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Relax up to FU_U edges per lane (destinations d[j] where act[j]; act[j]
// implies act[0], and an idle slot's d[j] repeats a valid destination of the
// lane) with the value `val`; a lowered destination with out-edges whose bit
// in the next bitmap was clear is appended.  Wave-uniform.
//
// Values only decrease, so a plain (possibly stale) load that already holds
// at most `val` proves the atomicMin would change nothing: only the other
// slots go to the atomics.  A lane's atomics are issued together, idle slots
// with the identity of the operation (min with ~0, or with 0), so no slot
// waits for the one before.
template <int RELAX>
__device__ __forceinline__ void relax_edges(const bool (&act)[FU_U], const uint32_t (&d)[FU_U], uint32_t val,
                                            uint32_t *prop, const uint32_t *__restrict__ deg, const FusedLists &L,
                                            int nb, unsigned long long *nlight, unsigned long long *nheavy,
                                            unsigned long long *nedge, unsigned long long dense_at) {
    bool need[FU_U];
    bool any_need = false;
#pragma unroll
    for (int j = 0; j < FU_U; ++j) {
        const uint32_t peek = prop[d[j]];
        need[j] = act[j] && val < peek;
        any_need |= need[j];
    }
    uint32_t old[FU_U], dg[FU_U];
#pragma unroll
    for (int j = 0; j < FU_U; ++j) {
        old[j] = ~0u;
        dg[j] = 0;
    }
    if (any_need) {
#pragma unroll
        for (int j = 0; j < FU_U; ++j) {
            dg[j] = deg[d[j]];
            old[j] = atomicMin(&prop[d[j]], need[j] ? val : ~0u);
        }
    }
    bool low[FU_U];
    bool any = false;
#pragma unroll
    for (int j = 0; j < FU_U; ++j) {
        low[j] = need[j] && val < old[j] && dg[j] != 0;
        any |= low[j];
    }
    uint64_t ow[FU_U];
#pragma unroll
    for (int j = 0; j < FU_U; ++j) ow[j] = 0;
    if (any) {
#pragma unroll
        for (int j = 0; j < FU_U; ++j)
            ow[j] = __hip_atomic_fetch_or(&L.bits[nb][d[j] >> 6], low[j] ? 1ull << (d[j] & 63) : 0ull,
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    bool push[FU_U];
#pragma unroll
    for (int j = 0; j < FU_U; ++j) push[j] = low[j] && !((ow[j] >> (d[j] & 63)) & 1ull);
    if (__ballot(any)) append_next(push, d, dg, L, nb, nlight, nheavy, nedge, dense_at);
}

template <int RELAX>
__global__ void __launch_bounds__(FU_THREADS)
k_fused(const uint64_t *__restrict__ index, const uint32_t *__restrict__ nbrs, const uint32_t *__restrict__ deg,
        uint32_t *prop, FusedLists L, const uint32_t *__restrict__ in_list, uint32_t nin, FusedCtl *ctl,
        unsigned spin_limit, unsigned long long dense_at) {
    __shared__ int s_ok;
    const unsigned ln = lane();
    const unsigned long long W = (unsigned long long)gridDim.x * FU_WAVES;
    // wave rank interleaved over the blocks: a small level lands on many
    // compute units, not on the waves of the first blocks
    const unsigned long long w = (unsigned long long)(threadIdx.x / 64) * gridDim.x + blockIdx.x;
    unsigned long long edges = 0, pops = 0;
    unsigned lvl = 0;
    for (;; ++lvl) {
        const int c = lvl % 3, nx = (lvl + 1) % 3, stale = (lvl + 2) % 3;
        const int cb = lvl & 1, nb = cb ^ 1;
        // level 0 is the caller's list itself (its long rows also listed as
        // slices by the seed); a repeated ID there is processed twice, which
        // min-relaxes make harmless
        const unsigned long long nL = lvl == 0 ? nin : ld_agent(&ctl->nlight[c]);
        const unsigned long long nH = ld_agent(&ctl->nheavy[c]);
        const uint32_t *light = lvl == 0 ? in_list : L.light[cb];
        if (nL + nH == 0) break;   // the same counts in every block: all leave together
        // a dense level goes back to the caller (vertices + edges above the
        // push/pull threshold)
        if (lvl > 0 && dense_at && ld_agent(&ctl->nedge[c]) > dense_at) {
            if (blockIdx.x == 0 && threadIdx.x == 0) ctl->dense = 1;
            break;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            __hip_atomic_store(&ctl->nlight[stale], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->nheavy[stale], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->nedge[stale], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned long long *nlight = &ctl->nlight[nx], *nheavy = &ctl->nheavy[nx], *nedge = &ctl->nedge[nx];
        // ---- light rows: lane per vertex, waves from the front
        for (unsigned long long base = w * 64; base < nL; base += W * 64) {
            const unsigned long long i = base + ln;
            const bool on = i < nL;
            const uint32_t v = __hip_atomic_load(&light[on ? i : base], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t val = 0;
            uint64_t e = 0, e1 = 0;
            if (on) {
                if (lvl)
                    __hip_atomic_fetch_and(&L.bits[cb][v >> 6], ~(1ull << (v & 63)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t pv = __hip_atomic_load(&prop[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                val = RELAX == GPR_RELAX_MIN_PLUS1 ? pv + 1 : pv;
                e = index[v];
                e1 = index[v + 1];
                if (e1 - e > FU_LIGHT_MAX) e = e1;   // level 0 only: a long row, done by its slices
                else ++pops;
            }
            while (__ballot(e < e1)) {
                const uint64_t any = e < e1 ? e : 0;   // a valid edge for the clamped loads
                bool act[FU_U];
                uint32_t d[FU_U];
#pragma unroll
                for (int j = 0; j < FU_U; ++j) {
                    act[j] = e + j < e1;
                    d[j] = nbrs[act[j] ? e + j : any];
                    edges += act[j];
                }
                e = e + FU_U < e1 ? e + FU_U : e1;
                relax_edges<RELAX>(act, d, val, prop, deg, L, nb, nlight, nheavy, nedge, dense_at);
            }
        }
        // ---- heavy rows: wave per slice, waves from the back
        for (unsigned long long j = W - 1 - w; j < nH; j += W) {
            const uint64_t ent = __hip_atomic_load(&L.heavy[cb][j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t v = (uint32_t)(ent >> 32), s = (uint32_t)ent;
            if (s == 0 && ln == 0) {
                if (lvl)
                    __hip_atomic_fetch_and(&L.bits[cb][v >> 6], ~(1ull << (v & 63)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                ++pops;
            }
            const uint32_t pv = __hip_atomic_load(&prop[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t val = RELAX == GPR_RELAX_MIN_PLUS1 ? pv + 1 : pv;
            const uint64_t r1 = index[v + 1];
            const uint64_t e0 = index[v] + (uint64_t)s * FU_SLICE;
            bool act[FU_U];
            uint32_t d[FU_U];
#pragma unroll
            for (int k = 0; k < FU_U; ++k) {
                const uint64_t e = e0 + ln + 64 * k;
                act[k] = e < r1;
                d[k] = nbrs[act[k] ? e : act[0] ? e0 + ln : e0];   // clamped: the load is unconditional
                edges += act[k];
            }
            relax_edges<RELAX>(act, d, val, prop, deg, L, nb, nlight, nheavy, nedge, dense_at);
        }
        if (!grid_barrier(&ctl->arrive, &ctl->error, ctl->group, lvl, spin_limit, &s_ok)) break;
    }
    // one pair of atomics per block, none for a block without work: a
    // small step leaves most of the 4096 waves idle, and an atomic from
    // each of them on one line cost more than the step's work
    __shared__ unsigned long long red[FU_WAVES];
    edges = block_sum(edges, red);
    pops = block_sum(pops, red);
    if (threadIdx.x == 0) {
        if (edges) atomicAdd(&ctl->edges, edges);
        if (pops) atomicAdd(&ctl->pops, pops);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->levels = lvl;
}

__global__ void k_fused_init(FusedCtl *ctl) {
    ctl->arrive = 0;
    for (int i = 0; i < 3; ++i) ctl->nlight[i] = ctl->nheavy[i] = ctl->nedge[i] = 0;
    ctl->edges = 0;
    ctl->pops = 0;
    ctl->levels = 0;
    ctl->error = 0;
    ctl->dense = 0;
    ctl->nout = 0;
    for (int i = threadIdx.x; i < FU_MAX_GROUPS; i += blockDim.x) ctl->group[i * FU_LINE] = 0;
}

// Level 0, long rows: the slices of every input ID with more than
// FU_LIGHT_MAX out-edges (the short rows are read from the input list).
__global__ void k_fused_seed(const uint32_t *__restrict__ list, uint32_t k, const uint32_t *__restrict__ deg,
                             FusedLists L, FusedCtl *ctl) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < k; i0 += stride) {
        const uint32_t i = i0 + threadIdx.x;
        bool push[FU_U] = {};
        uint32_t d[FU_U] = {}, dg[FU_U] = {};
        if (i < k) {
            d[0] = list[i];
            dg[0] = deg[d[0]];
            push[0] = dg[0] > FU_LIGHT_MAX;
        }
        if (__ballot(push[0]))
            append_next(push, d, dg, L, 0, &ctl->nlight[0], &ctl->nheavy[0], &ctl->nedge[0], 0ull);
    }
}

// The handed-back level as a frontier list: the light IDs, then the vertex
// of every heavy row (its slice 0).
__global__ void k_fused_out(const uint32_t *__restrict__ light, unsigned long long nL,
                            const uint64_t *__restrict__ heavy, unsigned long long nH, uint32_t *out,
                            FusedCtl *ctl) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < nL + nH; i += stride) {
        if (i < nL) {
            out[i] = light[i];
        } else {
            const uint64_t ent = heavy[i - nL];
            if ((uint32_t)ent == 0) out[nL + atomicAdd(&ctl->nout, 1ull)] = (uint32_t)(ent >> 32);
        }
    }
}

// Buffers: two light lists of n IDs, two heavy lists, the second bitmap (the
// first is the graph's dedup bitmap, clear between steps).  A level's heavy
// list holds at most sum over distinct rows longer than FU_LIGHT_MAX of
// ceil(len/FU_SLICE) <= m/FU_SLICE + min(n, m/(FU_LIGHT_MAX+1)) entries.
struct FusedBuf {
    uint64_t heavy_cap;
    size_t off_light1, off_heavy0, off_heavy1, off_bits, bytes;
};

static FusedBuf fused_layout(const gpr_graph *g) {
    FusedBuf b;
    const uint64_t long_rows = g->m / (FU_LIGHT_MAX + 1);
    b.heavy_cap = g->m / FU_SLICE + (long_rows < g->n ? long_rows : g->n) + 64;
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    b.off_light1 = up(4ull * g->n + 4);
    b.off_heavy0 = b.off_light1 + up(4ull * g->n + 4);
    b.off_heavy1 = b.off_heavy0 + up(8ull * b.heavy_cap);
    b.off_bits = b.off_heavy1 + up(8ull * b.heavy_cap);
    b.bytes = b.off_bits + up(8ull * nwords(g->n));
    return b;
}

static FusedLists fused_lists(gpr_graph *g, const FusedBuf &b) {
    char *p = (char *)g->fbuf;
    FusedLists L;
    L.light[0] = (uint32_t *)p;
    L.light[1] = (uint32_t *)(p + b.off_light1);
    L.heavy[0] = (uint64_t *)(p + b.off_heavy0);
    L.heavy[1] = (uint64_t *)(p + b.off_heavy1);
    L.bits[0] = g->dedup;
    L.bits[1] = (uint64_t *)(p + b.off_bits);
    return L;
}

static int fused_ensure(gpr_graph *g, const FusedBuf &b) {
    if (!g->fctl) GPR_TRY(dev_alloc((void **)&g->fctl, sizeof(FusedCtl), "fusion control"));
    if (g->fbuf && g->fbuf_bytes >= b.bytes) return GPR_OK;
    dev_free(g->fbuf);
    g->fbuf = nullptr;
    g->fbuf_bytes = 0;
    GPR_TRY(dev_alloc(&g->fbuf, b.bytes, "fusion lists"));
    g->fbuf_bytes = b.bytes;
    GPR_TRY(fill_words(g, (uint64_t *)((char *)g->fbuf + b.off_bits), nwords(g->n), 0ull));
    return GPR_OK;
}

// Blocks of the cooperative grid: tunable fusion_blocks (0: one per compute
// unit), at most what can be resident at once (the launch fails rather than
// waits if they cannot all be).
static int fused_blocks(const void *kernel, unsigned *nblocks) {
    static const void *known[2] = {};
    static int known_per_cu[2] = {};
    int per_cu = 0;
    for (int i = 0; i < 2; ++i)
        if (known[i] == kernel) per_cu = known_per_cu[i];
    if (!per_cu) {
        GPR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, FU_THREADS, 0));
        for (int i = 0; i < 2; ++i)
            if (!known[i]) {
                known[i] = kernel;
                known_per_cu[i] = per_cu;
                break;
            }
    }
    if (per_cu < 1) {
        set_error("fused push: the kernel cannot be resident (occupancy 0)");
        return GPR_E_HIP;
    }
    const double want = tun().fusion_blocks;
    unsigned cap = (unsigned)(per_cu * num_cus());
    if (cap > FU_GROUP * FU_MAX_GROUPS) cap = FU_GROUP * FU_MAX_GROUPS;
    unsigned nb = want >= 1 ? (unsigned)want : (unsigned)num_cus();
    *nblocks = nb < cap ? nb : cap;
    return GPR_OK;
}

// dense_at: a level whose vertices + out-edges exceed it ends the step and
// becomes its output frontier (0: run to the fixpoint; also when out is null).
int fused_push(gpr_graph *g, int relax, uint32_t *prop, gpr_frontier *in, gpr_frontier *out, uint64_t dense_at,
               gpr_iter_stats *s) {
    const FusedBuf b = fused_layout(g);
    GPR_TRY(fused_ensure(g, b));
    FusedCtl *ctl = (FusedCtl *)g->fctl;
    FusedLists L = fused_lists(g, b);
    const uint32_t k = in->nactv;
    const double sl = tun().fusion_spin_limit;
    unsigned spin = sl < 0 ? 0u : sl > 2147483647.0 ? 2147483647u : (unsigned)sl;
    unsigned long long dense = out ? dense_at : 0;
    hipLaunchKernelGGL(k_fused_init, dim3(1), dim3(64), 0, g->stream, ctl);
    if (k)
        hipLaunchKernelGGL(k_fused_seed, dim3(grid_for(k, 256, 4096)), dim3(256), 0, g->stream, in->list, k, g->deg,
                           L, ctl);
    GPR_HIP(hipGetLastError());
    const void *kern = relax == GPR_RELAX_MIN_PLUS1 ? (const void *)k_fused<GPR_RELAX_MIN_PLUS1>
                                                     : (const void *)k_fused<GPR_RELAX_MIN>;
    unsigned nblocks = 0;
    GPR_TRY(fused_blocks(kern, &nblocks));
    {
        ProfScope ps(g, "fused");
        const uint64_t *index = g->index;
        const uint32_t *nbrs = g->nbrs, *deg = g->deg;
        const uint32_t *in_list = in->list;
        void *args[] = {(void *)&index, (void *)&nbrs, (void *)&deg, (void *)&prop, (void *)&L, (void *)&in_list,
                        (void *)&k, (void *)&ctl, (void *)&spin, (void *)&dense};
        if (tun().fusion_cooperative >= 0.5)
            GPR_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(FU_THREADS), args, 0, g->stream));
        else
            GPR_HIP(hipLaunchKernel(kern, dim3(nblocks), dim3(FU_THREADS), args, 0, g->stream));
    }
    // the first line of the control block: counts and flags
    FusedCtl h;
    GPR_HIP(hipMemcpyAsync(&h, ctl, FU_LINE * sizeof(unsigned long long), hipMemcpyDeviceToHost, g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    if (h.error) {
        // leave both bitmaps clear for the next step
        GPR_TRY(fill_words(g, L.bits[0], nwords(g->n), 0ull));
        GPR_TRY(fill_words(g, L.bits[1], nwords(g->n), 0ull));
        GPR_HIP(hipStreamSynchronize(g->stream));
        set_error("fused push: a grid barrier made no progress (level %u)", h.levels);
        return GPR_E_HIP;
    }
    s->edges_scanned = h.edges;
    s->levels = (int)h.levels;
    s->nactv_out = 0;
    s->nacte_out = 0;
    if (out) {
        out->kind = GPR_FRONTIER_EMPTY;
        out->nactv = 0;
        out->nacte = 0;
    }
    if (h.dense) {
        // level h.levels was not run: it is the output frontier; its IDs
        // carry bits in bitmap h.levels & 1, cleared here
        const int c = h.levels % 3, cb = h.levels & 1;
        const unsigned long long nL = h.nlight[c], nH = h.nheavy[c];
        GPR_TRY(frontier_ensure_list(out));
        hipLaunchKernelGGL(k_fused_out, dim3(grid_for(nL + nH, 256, 8192)), dim3(256), 0, g->stream, L.light[cb], nL,
                           L.heavy[cb], nH, out->list, ctl);
        GPR_HIP(hipGetLastError());
        GPR_TRY(fill_words(g, L.bits[cb], nwords(g->n), 0ull));
        unsigned long long nhv = 0;
        GPR_HIP(hipMemcpyAsync(&nhv, &ctl->nout, sizeof(nhv), hipMemcpyDeviceToHost, g->stream));
        GPR_HIP(hipStreamSynchronize(g->stream));
        out->kind = GPR_FRONTIER_LIST;
        out->nactv = (uint32_t)(nL + nhv);
        out->nacte = h.nedge[c] - out->nactv;
        s->nactv_out = out->nactv;
        s->nacte_out = out->nacte;
    }
    return GPR_OK;
}

} // namespace gpr
