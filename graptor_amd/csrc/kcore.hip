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
//   pass     over the candidates (vertices not yet peeled; each wave owns
//            one segment of them for the whole launch, taken in the first
//            pass as 64-vertex chunks dealt round-robin over the waves, so
//            that the hubs, which survive longest, are spread over all of
//            them): count K -> the level's frontier, above
//            K -> kept, compacted in place (the count lives in the wave's
//            registers: no atomics), with the smallest kept count;
//   round    the frontier's out-edges count their destinations down; the
//            one decrement that takes a destination from K+1 to K appends
//            it to the next round's frontier, so every vertex is listed
//            once and no dedup bitmap is needed.  Counts decremented to
//            above K are folded into a minimum as well.
// The next level's K is the smaller of the two minima: never above the
// smallest remaining count (every vertex not peeled either kept its count
// since the pass or was last seen at a decremented one), and below it only
// when a vertex counted down was then peeled; the pass then finds no vertex
// at K, no round runs, and its own kept minimum is exact.  So no separate
// scan for the minimum is needed.
// The count-down is a plain peek (counts only fall, so a stale value that is
// at most K already proves there is nothing to do), then atomicSub; a
// decrement that finds the count at or below K (a race with another
// decrement at the floor) is undone with atomicAdd before the round's
// barrier, so after every barrier each count is exact.
// Frontier lists as in the fused push: rows up to KC_LIGHT_MAX edges one
// lane per vertex, longer rows one wave per KC_SLICE-edge slice.  Light
// appends are staged per wave in LDS and reserved in the shared list once
// per KC_STAGE entries, heavy slices are reserved per wave step.  List
// counts rotate over three slots per round and level minima over two slots
// per level, each cleared one phase after its last read and at least one
// barrier before its next write.
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
constexpr int KC_STAGE = 512;           // light appends staged per wave in LDS
constexpr int KC_PU = 4;                // candidate chunks per wave step in passes A and B
constexpr uint32_t KC_HUBS = 24576;      // count-downs of vertices below this ID add up in LDS
static_assert(KC_HUBS % KC_THREADS == 0, "whole passes over the hub counters");
constexpr unsigned KC_MAX_WAVES = GS_GROUP * GS_MAX_GROUPS * KC_WAVES;
static_assert(KC_SLICE == 64 * KC_U, "a slice is one pass of the wave");
// phase kinds timed by block 0: the candidate pass, its barrier, round
// work, round barriers
constexpr int KC_PROF = 4;

struct KcCtl {
    unsigned long long arrive;        // groups arrived at the barrier (monotonic)
    unsigned long long nlight[3];     // frontier list lengths, rotating per round
    unsigned long long nheavy[3];
    unsigned long long edges;         // out-edges of peeled vertices examined
    unsigned int error;               // 1: a barrier wait made no progress
    unsigned int levels;              // levels with a non-empty frontier
    unsigned int rounds;              // rounds run
    unsigned int kmax;                // largest core number
    unsigned int kmin_s[3];           // minimum count kept by a pass, rotating per iteration
    unsigned int kmin_d[3];           // minimum count a round decremented to above K, rotating
    unsigned long long pad[GS_LINE - 13];
    unsigned long long group[GS_MAX_GROUPS * GS_LINE];
    unsigned long long prof[KC_PROF];   // block 0's wall-clock ticks per phase kind
};
static_assert(offsetof(KcCtl, group) == GS_LINE * sizeof(unsigned long long), "one line of counters");

struct KcLists {
    uint32_t *cand;                   // per-wave segments of the candidates
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

// Light appends of one wave are staged in its own LDS buffer and written
// out with one reservation per KC_STAGE entries at most, so the shared list
// counter sees one atomic per buffer instead of one per wave step.  The
// stage's count is wave-uniform; each phase flushes before its barrier.
struct KcStage {
    uint32_t *buf;   // this wave's KC_STAGE entries of LDS
    uint32_t n;      // entries staged
};

__device__ __forceinline__ void kc_flush(KcStage &st, uint32_t *list, unsigned long long *nlight) {
    if (st.n == 0) return;
    __builtin_amdgcn_wave_barrier();
    unsigned long long b = 0;
    if (lane() == 0)
        b = __hip_atomic_fetch_add(nlight, (unsigned long long)st.n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b = bcast64(b);
    for (uint32_t i = lane(); i < st.n; i += 64)
        __hip_atomic_store(&list[b + i], st.buf[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_wave_barrier();
    st.n = 0;
}

// The end-of-phase flush of every wave's stage with one reservation per
// block: the waves' counts are prefixed in LDS and thread 0 reserves their
// total.  Block-uniform: every thread calls it.
__device__ __forceinline__ void kc_flush_block(KcStage &st, uint32_t *list, unsigned long long *nlight,
                                               uint32_t *s_cnt, unsigned long long *s_base) {
    const unsigned w = threadIdx.x / 64;
    if (lane() == 0) s_cnt[w] = st.n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < KC_WAVES; ++i) tot += s_cnt[i];
        unsigned long long b = 0;
        if (tot) b = __hip_atomic_fetch_add(nlight, tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int i = 0; i < KC_WAVES; ++i) {
            s_base[i] = b;
            b += s_cnt[i];
        }
    }
    __syncthreads();
    const unsigned long long b = s_base[w];
    for (uint32_t i = lane(); i < st.n; i += 64)
        __hip_atomic_store(&list[b + i], st.buf[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st.n = 0;
}

// Append the vertices v[j] with put[j] to the frontier lists of buffer cb:
// light rows through the stage, heavy rows' slices directly.  Wave-uniform:
// every lane calls it.
__device__ __forceinline__ void kc_append(const bool (&put)[KC_U], const uint32_t (&v)[KC_U],
                                          const uint32_t *__restrict__ deg, const KcLists &L, int cb,
                                          unsigned long long *nlight, unsigned long long *nheavy, KcStage &st) {
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
    if (tl) {
        if (st.n + tl > (uint32_t)KC_STAGE) kc_flush(st, L.light[cb], nlight);
        uint32_t p = st.n + (il - cl);
#pragma unroll
        for (int j = 0; j < KC_U; ++j)
            if (put[j] && kc_slices(dg[j]) == 0) st.buf[p++] = v[j];
        st.n += tl;
        __builtin_amdgcn_wave_barrier();
    }
    if (th) {
        unsigned long long bh = 0;
        if (lane() == 0)
            bh = __hip_atomic_fetch_add(nheavy, (unsigned long long)th, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bh = bcast64(bh) + (ih - ch);
#pragma unroll
        for (int j = 0; j < KC_U; ++j) {
            if (put[j]) {
                const uint32_t s = kc_slices(dg[j]);
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
                                              int nb, unsigned long long *nlight, unsigned long long *nheavy,
                                              KcStage &st, uint32_t *hub, uint32_t &dmin) {
    bool need[KC_U];
#pragma unroll
    for (int j = 0; j < KC_U; ++j) need[j] = act[j] && core[d[j]] > K;
    // hubs: counted in this block's LDS, applied once per round (kc_hub_flush)
#pragma unroll
    for (int j = 0; j < KC_U; ++j)
        if (need[j] && d[j] < KC_HUBS) {
            atomicAdd(&hub[d[j]], 1u);
            need[j] = false;
        }
    uint32_t old[KC_U];
#pragma unroll
    for (int j = 0; j < KC_U; ++j) old[j] = need[j] ? atomicSub(&core[d[j]], 1u) : 0u;
    bool put[KC_U];
#pragma unroll
    for (int j = 0; j < KC_U; ++j) {
        if (need[j] && old[j] <= K) atomicAdd(&core[d[j]], 1u);   // raced at the floor: undo
        put[j] = need[j] && old[j] == K + 1;
        if (need[j] && old[j] > K + 1 && old[j] - 1 < dmin) dmin = old[j] - 1;
    }
    kc_append(put, d, deg, L, nb, nlight, nheavy, st);
}

// Apply this block's hub count-downs of the round: subtract each sum delta
// at once.  The count falls to max(K, c - delta) as delta single steps would:
// a subtraction that starts at or below K is undone, one that crosses K is
// topped back up to K, and that crossing one (there is exactly one per hub
// and level) appends the hub.  Block-uniform: every thread calls it.
__device__ __forceinline__ void kc_hub_flush(uint32_t *hub, uint32_t n, uint32_t K, uint32_t *core,
                                             const uint32_t *__restrict__ deg, const KcLists &L, int nb,
                                             unsigned long long *nlight, unsigned long long *nheavy, KcStage &st,
                                             uint32_t &dmin) {
    __syncthreads();
    for (uint32_t base = 0; base < KC_HUBS; base += KC_THREADS) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t dl = hub[j];
        bool put[KC_U] = {false, false, false, false};
        uint32_t v[KC_U] = {j, j, j, j};
        if (dl) {
            hub[j] = 0;
            const uint32_t old = atomicSub(&core[j], dl);
            if (old <= K) {
                atomicAdd(&core[j], dl);
            } else if (old - dl <= K) {
                if (old - dl < K) atomicAdd(&core[j], K - (old - dl));
                put[0] = true;
            } else if (old - dl < dmin) {
                dmin = old - dl;
            }
        }
        kc_append(put, v, deg, L, nb, nlight, nheavy, st);
    }
}

__global__ void __launch_bounds__(KC_THREADS)
k_kcore(uint32_t n, const uint64_t *__restrict__ index, const uint32_t *__restrict__ nbrs,
        const uint32_t *__restrict__ deg, uint32_t *core, KcLists L, KcCtl *ctl, unsigned spin_limit) {
    __shared__ int s_ok;
    __shared__ uint32_t s_stage[KC_WAVES * KC_STAGE];
    __shared__ uint32_t s_hub[KC_HUBS];
    for (uint32_t j = threadIdx.x; j < KC_HUBS; j += KC_THREADS) s_hub[j] = 0;
    __syncthreads();
    const unsigned ln = lane();
    const unsigned long long W = (unsigned long long)gridDim.x * KC_WAVES;
    // wave rank interleaved over the blocks: small phases land on many CUs
    const unsigned long long wid = (unsigned long long)(threadIdx.x / 64) * gridDim.x + blockIdx.x;
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    KcStage st = {&s_stage[(threadIdx.x / 64) * KC_STAGE], 0u};
    // this wave's segment of the candidates.  The first pass takes, without
    // a list, the 64-vertex chunks wid, wid+W, wid+2W, ... (coalesced, and
    // every wave gets a share of every degree range: after relabelling by
    // degree the low IDs are the hubs, which survive to the highest levels);
    // each pass compacts the survivors in place, so the count stays in a
    // register and needs no atomic
    const unsigned long long nchunks = ((unsigned long long)n + 63) / 64;
    const unsigned long long S = (nchunks + W - 1) / W * 64;
    uint32_t *seg = L.cand + wid * S;
    uint32_t cnt = (uint32_t)(wid < nchunks ? (nchunks - wid + W - 1) / W * 64 : 0ull);
    unsigned ph = 0;          // barriers passed
    unsigned r = 0;           // rounds run
    unsigned levels = 0;      // iterations with a non-empty frontier
    // K, the count of this iteration's frontier, never exceeds the smallest
    // count of a vertex not yet peeled: it is the minimum of the counts the
    // last pass kept and of those the last level's rounds decremented to
    // above its floor.  A vertex counted there and then peeled in the same
    // level can make it too small; the pass then finds no vertex at K, no
    // round runs, and the minimum it kept is exact.
    uint32_t K = 0;
    unsigned long long my_edges = 0;
    __shared__ uint32_t s_dmin, s_dflushed, s_smin;
    __shared__ uint32_t s_wcnt[KC_WAVES];
    __shared__ unsigned long long s_wbase[KC_WAVES];
    // block 0's phase times live in LDS (registers are the scarce resource)
    __shared__ unsigned long long s_prof[KC_PROF + 1];
    if (leader) {
        for (int k = 0; k < KC_PROF; ++k) s_prof[k] = 0;
        s_prof[KC_PROF] = wall_clock64();
    }
    auto mark = [&](int k) {
        if (leader) {
            const unsigned long long t = wall_clock64();
            s_prof[k] += t - s_prof[KC_PROF];
            s_prof[KC_PROF] = t;
        }
    };
    bool ok = true;
    for (unsigned it = 0; ok; ++it) {
        const int s3 = it % 3;
        if (leader) {
            // slot (it+1)%3 was last read after iteration it-2's last
            // barrier, so before iteration it-1's first; it is written next
            // in iteration it+1, after this iteration's barriers
            __hip_atomic_store(&ctl->kmin_s[(it + 1) % 3], ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->kmin_d[(it + 1) % 3], ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (threadIdx.x == 0) {
            s_dmin = ~0u;
            s_dflushed = ~0u;
            s_smin = ~0u;
        }
        // the pass: count K -> this level's frontier; above K -> kept in the
        // segment (written at or below the position read, so in place), and
        // their smallest count; below K -> peeled before, dropped
        uint32_t smin = ~0u;
        {
            const int c3 = r % 3, cb = r & 1;
            uint32_t p = 0;
            for (uint32_t base = 0; base < cnt; base += 64 * KC_PU) {
                uint32_t v[KC_U] = {0u, 0u, 0u, 0u}, c[KC_U];
                bool in[KC_U] = {false, false, false, false};
#pragma unroll
                for (int u = 0; u < KC_PU; ++u) {
                    const uint32_t i = base + u * 64 + ln;
                    if (it == 0) {
                        // chunk wid + (i/64)*W, vertex i%64 of it (the last chunk may be short)
                        const unsigned long long vv = 64 * (wid + (unsigned long long)(i / 64) * W) + ln;
                        in[u] = i < cnt && vv < n;
                        v[u] = in[u] ? (uint32_t)vv : 0u;
                    } else {
                        in[u] = i < cnt;
                        v[u] = in[u] ? ld32(&seg[i]) : 0u;
                    }
                }
#pragma unroll
                for (int u = 0; u < KC_PU; ++u) c[u] = in[u] ? ld32(&core[v[u]]) : 0u;
                bool put[KC_U] = {false, false, false, false}, keep[KC_U];
#pragma unroll
                for (int u = 0; u < KC_PU; ++u) {
                    put[u] = in[u] && c[u] == K;
                    keep[u] = in[u] && c[u] > K;
                    if (keep[u] && c[u] < smin) smin = c[u];
                }
                kc_append(put, v, deg, L, cb, &ctl->nlight[c3], &ctl->nheavy[c3], st);
                // every read of these KC_PU chunks precedes the writes, which
                // land at or below the chunks' own positions
#pragma unroll
                for (int u = 0; u < KC_PU; ++u) {
                    const uint64_t km = __ballot(keep[u]);
                    if (keep[u])
                        __hip_atomic_store(&seg[p + mbcnt(km)], v[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    p += (uint32_t)__popcll(km);
                }
            }
            cnt = p;
            // the block's minimum through LDS, then one global atomic; the
            // flush's barriers order the LDS minimum before thread 0 reads it
            smin = wave_min(smin);
            if (ln == 0 && smin != ~0u) atomicMin(&s_smin, smin);
            kc_flush_block(st, L.light[cb], &ctl->nlight[c3], s_wcnt, s_wbase);
        }
        if (threadIdx.x == 0 && s_smin != ~0u)
            __hip_atomic_fetch_min(&ctl->kmin_s[s3], s_smin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mark(0);
        if (!grid_barrier(&ctl->arrive, &ctl->error, ctl->group, ph++, spin_limit, &s_ok)) break;
        mark(1);
        {
            const int c3 = r % 3;
            if (ld_agent(&ctl->nlight[c3]) == 0 && ld_agent(&ctl->nheavy[c3]) == 0) {
                // no vertex at K: no round runs, the kept counts are exact
                const uint32_t ks = __hip_atomic_load(&ctl->kmin_s[s3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (ks == ~0u) break;              // every vertex peeled
                K = ks;
                continue;
            }
        }
        // rounds of level K
        uint32_t dmin = ~0u;   // per lane: counts decremented to above K
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
                    kc_count_down(act, d, K, core, deg, L, nb, &ctl->nlight[n3], &ctl->nheavy[n3], st, s_hub, dmin);
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
                kc_count_down(act, d, K, core, deg, L, nb, &ctl->nlight[n3], &ctl->nheavy[n3], st, s_hub, dmin);
            }
            kc_hub_flush(s_hub, n, K, core, deg, L, nb, &ctl->nlight[n3], &ctl->nheavy[n3], st, dmin);
            kc_flush_block(st, L.light[nb], &ctl->nlight[n3], s_wcnt, s_wbase);
            // the block's smallest count decremented to above K, sent on when
            // it improved (one global atomic per block and round at most)
            {
                const uint32_t wm = wave_min(dmin);
                if (ln == 0 && wm != ~0u) atomicMin(&s_dmin, wm);
                __syncthreads();
                if (threadIdx.x == 0 && s_dmin < s_dflushed) {
                    __hip_atomic_fetch_min(&ctl->kmin_d[s3], s_dmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_dflushed = s_dmin;
                }
            }
            ++r;
            mark(2);
            if (!grid_barrier(&ctl->arrive, &ctl->error, ctl->group, ph++, spin_limit, &s_ok)) {
                ok = false;
                break;
            }
            mark(3);
        }
        if (!ok) break;
        ++levels;
        if (leader) {
            __hip_atomic_store(&ctl->kmax, K, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->levels, levels, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctl->rounds, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint32_t ks = __hip_atomic_load(&ctl->kmin_s[s3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t kd = __hip_atomic_load(&ctl->kmin_d[s3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t kn = ks < kd ? ks : kd;
        if (kn == ~0u) break;                      // nothing left above K
        K = kn;
    }
    if (leader)
        for (int k = 0; k < KC_PROF; ++k)
            __hip_atomic_store(&ctl->prof[k], s_prof[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        if (threadIdx.x < 3) {
            ctl->kmin_s[threadIdx.x] = ~0u;
            ctl->kmin_d[threadIdx.x] = ~0u;
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
    b.light = al((n + 64 * (size_t)KC_MAX_WAVES) * sizeof(uint32_t));
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
    L.cand = (uint32_t *)(base + b.cand);
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
    if (g->prof.on) {
        unsigned long long t[KC_PROF];
        GPR_HIP(hipMemcpy(t, ctl->prof, sizeof(t), hipMemcpyDeviceToHost));
        int dev = 0, khz = 0;
        GPR_HIP(hipGetDevice(&dev));
        GPR_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        static const char *names[KC_PROF] = {"kcore_pass", "kcore_wait_pass", "kcore_rounds", "kcore_wait_rounds"};
        for (int k = 0; k < KC_PROF; ++k) {
            auto &e = g->prof.acc[names[k]];
            e.first += khz > 0 ? (double)t[k] / (double)khz : 0.0;
            e.second += k < 2 ? h.levels : h.rounds;
        }
    }
    st->kmax = g->n ? h.kmax : 0;
    st->levels = h.levels;
    st->rounds = h.rounds;
    st->edges = h.edges;
    return GPR_OK;
}

} // namespace gpr
