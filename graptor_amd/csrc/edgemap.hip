// edgemap.hip -- push and pull traversal kernels and the edgemap dispatcher.
//
// Reference semantics:
//   dispatcher      api::edgemap, include/graptor/api.h:2143-2415, with the
//                   thresholds frac_threshold / default_threshold :130-160
//   sparse push     csr_sparse_with_f, dsl/edgemap.h:1397-1740, edge-balanced
//                   blocks of EMAP_BLOCK_SIZE edges (:29-31, :1434-1436) and
//                   edge_partition::process_push, dsl/emap/edgechunk.h:347-410
//   dense pull      GraphCSx_csc scalar pull, dsl/emap/GraphCSx_csc.h:50-140
//                   (destination filter re-checked per edge, :110-111)
//   relax           redop_min / atomic_min CAS returning "changed",
//                   dsl/ast/redop.h:628-805, simd/vector_ref.h:205-225
//
// GPU design:
//   push: the frontier's out-degrees are prefix-summed (two small kernels),
//         then each workgroup takes a tile of PUSH_THREADS*PUSH_ITEMS
//         consecutive frontier edges (load-balanced search: the tile's
//         frontier slice is staged in LDS and each lane binary-searches it),
//         so hubs and leaves cost the same per edge and neighbour reads are
//         coalesced.  Relax is a u32 atomicMin; first activation is decided
//         by atomicOr on a bitmap word (output bitmap or a dedup bitmap).
//   pull: one wave per 64 consecutive destinations, lane per destination,
//         so a bitmap output word is one ballot and one store.  Lanes with
//         at most 32 in-edges loop alone with early exit; the wave walks the
//         in-edges of every heavier destination together (64 edges per
//         step, wave-min, early exit once the destination filter fails).
#include "internal.hpp"
#include "device.hpp"
#include "policy_dev.hpp"
#include "grid_sync.hpp"

#include <algorithm>

namespace gpr {

// ---- degree prefix of a frontier list -------------------------------------
constexpr int SCAN_THREADS = SCAN_BLOCK / 4;

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_local(const uint32_t *__restrict__ list, uint32_t k, const uint32_t *__restrict__ deg,
             uint64_t *__restrict__ pos, uint64_t *__restrict__ btot) {
    __shared__ uint64_t wsum[SCAN_THREADS / 64];
    uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * 4;
    uint64_t d[4];
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t i = base + j;
        d[j] = i < k ? deg[list[i]] : 0;
        s += d[j];
    }
    uint64_t inc = wave_incl_scan(s);
    unsigned w = threadIdx.x / 64;
    if (lane() == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t off = 0;
    for (unsigned j = 0; j < w; ++j) off += wsum[j];
    uint64_t run = off + inc - s;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t i = base + j;
        if (i < k) pos[i] = run;
        run += d[j];
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        if (gridDim.x == 1) {
            // one block: its exclusive offset 0 and the total, as the top
            // scan would write them (which then need not run)
            btot[0] = 0;
            btot[1] = off + inc;
        } else {
            btot[blockIdx.x] = off + inc;
        }
    }
}

// Exclusive scan of the block totals in place; total written to a[count].
__global__ void __launch_bounds__(1024) k_scan_top(uint64_t *a, uint32_t count) {
    __shared__ uint64_t carry;
    __shared__ uint64_t wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < count; b += blockDim.x) {
        uint32_t i = b + threadIdx.x;
        uint64_t v = i < count ? a[i] : 0;
        uint64_t inc = wave_incl_scan(v);
        unsigned w = threadIdx.x / 64;
        if (lane() == 63) wsum[w] = inc;
        __syncthreads();
        uint64_t off = carry;
        for (unsigned j = 0; j < w; ++j) off += wsum[j];
        if (i < count) a[i] = off + inc - v;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = off + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) a[count] = carry;
}

template <int RELAX> __device__ __forceinline__ uint32_t relax_val(uint32_t ps) {
    return RELAX == GPR_RELAX_MIN_PLUS1 ? ps + 1u : ps;
}

// setif (BFSv.C:146): a destination takes a value only while its top bit
// is set (unset: ~0); the value is the source's (internal) ID
constexpr uint32_t SETIF_OPEN = 0x80000000u;


// ---- push ------------------------------------------------------------------
constexpr int PUSH_TILE = PUSH_THREADS * PUSH_ITEMS;
constexpr int PUSH_LDS = PUSH_TILE + 1;

// Push activation of a destination whose relax just changed it
// (csr_sparse_with_f, dsl/edgemap.h:1497-1507: relax && update): the
// reduction activates it; the method form (and reduction_or_method, which
// pushes with the method) only when p[d], read right after the relax,
// differs from prev[d].
__device__ __forceinline__ bool push_records(int record, const uint32_t *prop, const uint32_t *prev, uint32_t d) {
    if (record == GPR_RECORD_REDUCTION) return true;
    const uint32_t now = __hip_atomic_load(&prop[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return now != prev[d];
}

// First activation of d on the bitmap `bits` (the output or the dedup
// bitmap): every lane that wants to record its destination sets d's bit,
// and the lane whose set finds the bit clear records d.  Wave-uniform.
// The edges of one source are consecutive lanes and its row is sorted, so
// a hub's destinations come as runs of lanes on the same bitmap word; one
// atomic per lane would queue on those words.  When the wave has such runs
// (8 or more lanes joining the run of the lane before), each run ORs its
// bits together (a segmented scan), its last lane sets them with one
// atomic, and every lane of the run reads that atomic's old word; a lane
// whose bit a lower lane of its run also holds leaves d to that lane.
// Otherwise each lane sets its own bit.
__device__ __forceinline__ bool bitmap_first(bool want, uint32_t d, uint64_t *bits) {
    const uint32_t ln = lane();
    const uint32_t key = want ? d >> 6 : 0xffffffffu - ln;   // idle lanes: a word of their own
    const uint64_t bit = want ? 1ull << (d & 63) : 0ull;
    const uint32_t pk = __shfl_up(key, 1, 64);
    const bool head = ln == 0 || pk != key;
    if (__popcll(__ballot(!head)) < 8) {
        if (!want) return false;
        const uint64_t ow = atomicOr((unsigned long long *)&bits[key], bit);
        return !(ow & bit);
    }
    // inclusive OR within each run; flag: a run head lies in the lanes
    // already combined
    uint64_t inc = bit;
    bool flag = head;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t ob = __shfl_up(inc, off, 64);
        const int of = __shfl_up((int)flag, off, 64);
        if (ln >= (uint32_t)off && !flag) {
            inc |= ob;
            flag = of != 0;
        }
    }
    const uint64_t before = __shfl_up(inc, 1, 64);
    const uint64_t excl = head ? 0ull : before;
    const uint32_t nk = __shfl_down(key, 1, 64);
    const bool tail = ln == 63 || nk != key;
    uint64_t old = 0;
    if (tail && want) old = atomicOr((unsigned long long *)&bits[key], inc);
    const uint64_t tails = __ballot(tail);
    const int my_tail = __ffsll((long long)(tails & (~0ull << ln))) - 1;
    const uint64_t ow = __shfl(old, my_tail, 64);
    return want && !(ow & bit) && !(excl & bit);
}

// bits == nullptr: the step records nothing (no record argument, or no
// output frontier); the relaxes still run.
template <int RELAX>
__global__ void __launch_bounds__(PUSH_THREADS)
k_push(const uint32_t *__restrict__ flist, uint32_t k, const uint64_t *__restrict__ pos,
       const uint64_t *__restrict__ btot, uint64_t E, const uint64_t *__restrict__ index,
       const uint32_t *__restrict__ nbrs, const uint32_t *__restrict__ deg, uint32_t *prop, int dkind,
       uint32_t darg, int record, const uint32_t *__restrict__ prev, uint64_t *bits, uint32_t *out_list,
       DevCounters *ctr) {
    __shared__ uint64_t s_pos[PUSH_LDS];
    __shared__ uint32_t s_v[PUSH_LDS];
    __shared__ uint32_t q[PUSH_TILE];
    __shared__ uint32_t s_i0, s_cnt, qn, qbase;
    __shared__ uint64_t red[PUSH_THREADS / 64];
    __shared__ uint32_t red32[PUSH_THREADS / 64];
    const uint64_t ntiles = (E + PUSH_TILE - 1) / PUSH_TILE;
    uint32_t my_v = 0;
    uint64_t my_e = 0;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t e0 = t * PUSH_TILE;
        const uint64_t e1 = min(E, e0 + PUSH_TILE);
        if (threadIdx.x < 64) {
            // the tile's first and last sources: the last frontier entry
            // starting at or before e0, and at or before e1 - 1
            const uint32_t i0 = wave_ub_pos(pos, btot, k, e0, 0u) - 1;
            const uint32_t i1 = wave_ub_pos(pos, btot, k, e1 - 1, i0 + 1) - 1;
            if (threadIdx.x == 0) {
                s_i0 = i0;
                s_cnt = i1 - i0 + 1;
                qn = 0;
            }
        }
        __syncthreads();
        const uint32_t i0 = s_i0, cnt = s_cnt;
        const bool in_lds = cnt <= (uint32_t)PUSH_LDS;
        if (in_lds)
            for (uint32_t j = threadIdx.x; j < cnt; j += PUSH_THREADS) {
                s_pos[j] = frontier_pos(pos, btot, i0 + j);
                s_v[j] = flist[i0 + j];
            }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < PUSH_ITEMS; ++it) {
            const uint64_t e = e0 + (uint64_t)it * PUSH_THREADS + threadIdx.x;
            bool want = false;   // the relax changed d and the step records it
            uint32_t d = 0;
            if (e < e1) {
                uint64_t pj;
                uint32_t s;
                if (in_lds) {
                    uint32_t lo = 0, hi = cnt;
                    while (lo < hi) {
                        uint32_t mid = (lo + hi) >> 1;
                        if (s_pos[mid] <= e) lo = mid + 1; else hi = mid;
                    }
                    pj = s_pos[lo - 1];
                    s = s_v[lo - 1];
                } else {
                    uint32_t j = ub_pos(pos, btot, k, e) - 1;
                    pj = frontier_pos(pos, btot, j);
                    s = flist[j];
                }
                d = nbrs[index[s] + (e - pj)];
                const uint32_t pd = prop[d];
                if constexpr (RELAX == GPR_RELAX_SETIF) {
                    // single trigger: the CAS that replaces the open value
                    // wins, so each destination is set and activated once
                    want = (pd & SETIF_OPEN) && dst_ok(pd, dkind, darg) && atomicCAS(&prop[d], pd, s) == pd &&
                           bits && push_records(record, prop, prev, d);
                } else if (dst_ok(pd, dkind, darg)) {
                    const uint32_t nv = relax_val<RELAX>(prop[s]);
                    if (nv < pd) {
                        const uint32_t old = atomicMin(&prop[d], nv);
                        want = old > nv && bits && push_records(record, prop, prev, d);
                    }
                }
            }
            const bool act = bits ? bitmap_first(want, d, bits) : false;
            if (act) {
                ++my_v;
                my_e += deg[d];
            }
            if (out_list) {
                // one LDS atomic per wave for its activations
                const uint64_t bal = __ballot(act);
                if (bal) {
                    uint32_t wb = 0;
                    if (lane() == 0) wb = atomicAdd(&qn, (uint32_t)__popcll(bal));
                    wb = __shfl(wb, 0);
                    if (act) q[wb + mbcnt(bal)] = d;
                }
            }
        }
        __syncthreads();
        if (out_list) {
            const uint32_t c = qn;
            if (threadIdx.x == 0 && c) qbase = atomicAdd(&ctr->listpos, c);
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < c; j += PUSH_THREADS) out_list[qbase + j] = q[j];
        }
        __syncthreads();
    }
    const uint32_t tv = block_sum(my_v, red32);
    const uint64_t te = block_sum(my_e, red);
    if (threadIdx.x == 0) {
        if (tv) atomicAdd(&ctr_stripe(ctr)->nactv, tv);
        if (te) atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)te);
    }
}

// ---- pull ------------------------------------------------------------------
constexpr int PULL_THREADS = 256;
constexpr uint32_t PULL_QCAP = 1024;   // list output: LDS queue entries per wave
constexpr size_t PULL_QBYTES = sizeof(uint32_t) * PULL_QCAP * (PULL_THREADS / 64);
constexpr uint32_t PULL_SERIAL_MAX = 32;
// The destination filter usually stops a row after a few edges (early
// exit), so the lane-serial path loads one neighbour at a time: loading
// ahead costs more frontier-bitmap lines than it hides latency.
constexpr int PULL_UNROLL = 1;
constexpr int PULL_HEAVY_UNROLL = 2;

// IX: the CSC offsets' type (a 32-bit copy when m < 2^32: the pull reads
// two offsets per destination it tests, a tenth of its first-step traffic
// on RMAT-26 with 64-bit offsets)
template <int RELAX, class IX>
__global__ void __launch_bounds__(PULL_THREADS)
k_pull(uint32_t n, const IX *__restrict__ cindex, const uint32_t *__restrict__ cnbrs,
       const uint32_t *__restrict__ deg, uint32_t *prop, const uint32_t *__restrict__ prev, int dkind,
       uint32_t darg, int record, const uint64_t *__restrict__ in_bits, uint64_t *out_bits,
       uint32_t *out_list, DevCounters *ctr, uint32_t serial_max, uint32_t list_cap) {
    __shared__ uint64_t red[PULL_THREADS / 64];
    __shared__ uint32_t red32[PULL_THREADS / 64];
    extern __shared__ uint32_t pull_q[];   // list output: PULL_QCAP entries per wave
    const uint32_t ngroups = (n + 63) / 64;
    const uint32_t wid = (blockIdx.x * PULL_THREADS + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * (PULL_THREADS / 64);
    const unsigned ln = lane();
    WaveQueue wq{pull_q + (threadIdx.x / 64) * PULL_QCAP, PULL_QCAP, 0, list_cap};
    uint32_t my_v = 0;
    uint64_t my_e = 0, my_scan = 0;
    for (uint32_t grp = wid; grp < ngroups; grp += nwaves) {
        const uint32_t d = grp * 64 + ln;
        const bool valid = d < n;
        const uint32_t orig = valid ? prop[d] : 0u;
        uint32_t cur = orig;
        const bool act = valid && dst_ok(cur, dkind, darg) && (RELAX != GPR_RELAX_SETIF || (cur & SETIF_OPEN));
        uint64_t beg = 0, end = 0;
        if (act) {
            beg = cindex[d];
            end = cindex[d + 1];
        }
        const bool heavy = act && (end - beg) > serial_max;
        if (act && !heavy) {
            // PULL_UNROLL independent neighbour loads (and frontier tests)
            // in flight per lane; edges past an early exit are loaded but
            // not applied
            bool live = true;
            for (uint64_t e = beg; live && e < end; e += PULL_UNROLL) {
                uint32_t s[PULL_UNROLL];
                bool hit[PULL_UNROLL];
#pragma unroll
                for (int u = 0; u < PULL_UNROLL; ++u) s[u] = e + u < end ? cnbrs[e + u] : 0xffffffffu;
#pragma unroll
                for (int u = 0; u < PULL_UNROLL; ++u)
                    hit[u] = s[u] != 0xffffffffu && (!in_bits || bit_test(in_bits, s[u]));
#pragma unroll
                for (int u = 0; u < PULL_UNROLL; ++u) {
                    if (live && hit[u]) {
                        if constexpr (RELAX == GPR_RELAX_SETIF) {
                            cur = s[u];   // the first frontier in-neighbour
                            live = false;
                        } else {
                            const uint32_t nv = relax_val<RELAX>(prop[s[u]]);
                            if (nv < cur) {
                                cur = nv;
                                if (!dst_ok(cur, dkind, darg)) live = false;
                            }
                        }
                    }
                    if (live && s[u] != 0xffffffffu) ++my_scan;
                }
            }
        }
        uint64_t hm = __ballot(heavy);
        while (hm) {
            const int L = __ffsll((long long)hm) - 1;
            hm &= hm - 1;
            const uint64_t b = __shfl(beg, L, 64), en = __shfl(end, L, 64);
            uint32_t best = __shfl(cur, L, 64);
            for (uint64_t base = b; base < en; base += 64 * PULL_HEAVY_UNROLL) {
                uint32_t s[PULL_HEAVY_UNROLL];
#pragma unroll
                for (int u = 0; u < PULL_HEAVY_UNROLL; ++u) {
                    const uint64_t e = base + (uint64_t)u * 64 + ln;
                    s[u] = e < en ? cnbrs[e] : 0xffffffffu;
                }
                uint32_t v = 0xffffffffu;
#pragma unroll
                for (int u = 0; u < PULL_HEAVY_UNROLL; ++u) {
                    if (s[u] != 0xffffffffu) {
                        ++my_scan;
                        if (!in_bits || bit_test(in_bits, s[u])) {
                            // setif: the smallest frontier ID of the first
                            // step with one (the open value is ~0)
                            const uint32_t nv = RELAX == GPR_RELAX_SETIF ? s[u] : relax_val<RELAX>(prop[s[u]]);
                            v = nv < v ? nv : v;
                        }
                    }
                }
                v = wave_min(v);
                best = v < best ? v : best;
                if (RELAX == GPR_RELAX_SETIF ? !(best & SETIF_OPEN) : !dst_ok(best, dkind, darg)) break;
            }
            if ((int)ln == L) cur = best;
        }
        const bool changed = valid && cur < orig;
        if (changed) prop[d] = cur;
        // a pull privatises its destinations, so reduction_or_method
        // records the reduction (api.h:1401-1416)
        const bool rec = record == GPR_RECORD_REDUCTION || record == GPR_RECORD_REDUCTION_OR_METHOD ? changed
                         : record == GPR_RECORD_METHOD_NE ? (valid && cur != prev[d])
                                                          : false;
        const uint64_t word = __ballot(rec);
        if (out_bits && ln == 0) out_bits[grp] = word;
        if (out_list) wq.push(rec, word, d, out_list, &ctr->listpos);
        if (ln == 0) my_v += (uint32_t)__popcll(word);
        if (rec) my_e += deg[d];
    }
    if (out_list) {
        // the waves' remaining queues go out together: one append per
        // workgroup (out_list is the same for the whole grid)
        __shared__ uint32_t qc[PULL_THREADS / 64], qbase;
        const unsigned w = threadIdx.x / 64;
        if (ln == 0) qc[w] = wq.full ? 0u : wq.cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int j = 0; j < PULL_THREADS / 64; ++j) t += qc[j];
            qbase = t ? atomicAdd(&ctr->listpos, t) : 0u;
        }
        __syncthreads();
        uint32_t off = qbase;
        for (unsigned j = 0; j < w; ++j) off += qc[j];
        if (!wq.full)
            for (uint32_t j = ln; j < wq.cnt; j += 64) out_list[off + j] = wq.q[j];
    }
    const uint32_t tv = block_sum(my_v, red32);
    const uint64_t te = block_sum(my_e, red);
    const uint64_t ts = block_sum(my_scan, red);
    if (threadIdx.x == 0) {
        if (tv) atomicAdd(&ctr_stripe(ctr)->nactv, tv);
        if (te) atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)te);
        if (ts) atomicAdd(&ctr_stripe(ctr)->edges, (unsigned long long)ts);
    }
}

// ---- vertex maps -------------------------------------------------------------
__global__ void k_vmap_fill32(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_vmap_fill64(uint64_t *p, uint32_t n, uint64_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_vmap_iota(uint32_t *p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}
__global__ void k_vmap_fill_list32(uint32_t *p, const uint32_t *__restrict__ list, uint32_t k, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) p[list[i]] = v;
}
__global__ void k_vmap_fill_bits32(uint32_t *p, const uint64_t *__restrict__ bits, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (bit_test(bits, i)) p[i] = v;
}
__global__ void k_vmap_copy_list32(uint32_t *p, const uint32_t *__restrict__ q, const uint32_t *__restrict__ list,
                                   uint32_t k) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const uint32_t v = list[i];
        p[v] = q[v];
    }
}
__global__ void k_vmap_copy_bits32(uint32_t *p, const uint32_t *__restrict__ q, const uint64_t *__restrict__ bits,
                                   uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (bit_test(bits, i)) p[i] = q[i];
}

// vertex_filter over a bitmap or all-true frontier: out bit v = in(v) &&
// ((p[v] == value) != neg).  One wave per output word.
__global__ void __launch_bounds__(256)
k_vfilter_bits(uint32_t n, const uint32_t *__restrict__ p, uint32_t value, int neg,
               const uint64_t *__restrict__ in_bits, uint64_t *out_bits, const uint32_t *__restrict__ deg,
               DevCounters *ctr) {
    __shared__ uint64_t red[4];
    __shared__ uint32_t red32[4];
    const uint32_t ngroups = (n + 63) / 64;
    const uint32_t wid = (blockIdx.x * 256 + threadIdx.x) / 64, nw = gridDim.x * 4;
    uint32_t my_v = 0;
    uint64_t my_e = 0;
    for (uint32_t grp = wid; grp < ngroups; grp += nw) {
        const uint32_t v = grp * 64 + lane();
        bool keep = v < n && (!in_bits || bit_test(in_bits, v));
        if (keep) keep = ((p[v] == value) != (neg != 0));
        const uint64_t w = __ballot(keep);
        if (lane() == 0) {
            out_bits[grp] = w;
            my_v += (uint32_t)__popcll(w);
        }
        if (keep) my_e += deg[v];
    }
    const uint32_t tv = block_sum(my_v, red32);
    const uint64_t te = block_sum(my_e, red);
    if (threadIdx.x == 0) {
        if (tv) atomicAdd(&ctr_stripe(ctr)->nactv, tv);
        if (te) atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)te);
    }
}

// Vertex filter over a list, order-preserving, in two phases (as the
// bitmap-to-list conversion): phase 0 counts the entries each workgroup
// keeps among its VF_TILE consecutive list entries (wave w takes the w-th
// run of 64 * VF_ITEMS) and adds their number and out-degree sum to the
// counters; after an exclusive scan of those counts, phase 1 writes the
// kept entries at their offsets (one workgroup: both in one launch).  No
// shared append position is touched.
constexpr int VF_THREADS = 256, VF_ITEMS = 8, VF_TILE = VF_THREADS * VF_ITEMS;
__global__ void __launch_bounds__(VF_THREADS)
k_vfilter_list(const uint32_t *__restrict__ in, uint32_t k, const uint32_t *__restrict__ p, uint32_t value,
               int neg, uint32_t *out, const uint32_t *__restrict__ deg, unsigned long long *blk, int phase,
               DevCounters *ctr) {
    __shared__ uint64_t red[VF_THREADS / 64];
    __shared__ uint32_t wtot[VF_THREADS / 64];
    const unsigned w = threadIdx.x / 64, ln = lane();
    const uint64_t base = (uint64_t)blockIdx.x * VF_TILE + (uint64_t)w * 64 * VF_ITEMS;
    uint32_t v[VF_ITEMS];
#pragma unroll
    for (int it = 0; it < VF_ITEMS; ++it) {
        const uint64_t i = base + it * 64 + ln;
        v[it] = i < k ? in[i] : 0u;
    }
    uint64_t m[VF_ITEMS];
    uint32_t cnt = 0;
#pragma unroll
    for (int it = 0; it < VF_ITEMS; ++it) {
        const uint64_t i = base + it * 64 + ln;
        const bool keep = i < k && (p[v[it]] == value) != (neg != 0);
        m[it] = __ballot(keep);
        cnt += (uint32_t)__popcll(m[it]);
    }
    if (ln == 0) wtot[w] = cnt;
    if (phase != 1) {
        // counting (phase 0; phase 2 = a single workgroup, which counts and
        // writes at offset 0 in one launch)
        uint64_t es = 0;
#pragma unroll
        for (int it = 0; it < VF_ITEMS; ++it)
            if ((m[it] >> ln) & 1ull) es += deg[v[it]];
        const uint64_t te = block_sum(es, red);   // its barriers also publish wtot
        if (threadIdx.x == 0) {
            uint32_t c = 0;
            for (int j = 0; j < VF_THREADS / 64; ++j) c += wtot[j];
            if (phase == 0) blk[blockIdx.x] = c;
            if (c) atomicAdd(&ctr_stripe(ctr)->nactv, c);
            if (te) atomicAdd(&ctr_stripe(ctr)->nacte, (unsigned long long)te);
        }
        if (phase == 0) return;
    }
    __syncthreads();
    uint64_t off = phase == 2 ? 0 : blk[blockIdx.x];
    for (unsigned j = 0; j < w; ++j) off += wtot[j];
#pragma unroll
    for (int it = 0; it < VF_ITEMS; ++it) {
        if ((m[it] >> ln) & 1ull) out[off + mbcnt(m[it])] = v[it];
        off += (uint32_t)__popcll(m[it]);
    }
}

// ---- host side ----------------------------------------------------------------
// Degree prefix of a vertex list into g->scan / g->scan_blk (pos(i) =
// scan[i] + scan_blk[i / SCAN_BLOCK]).
int scan_frontier_degrees(gpr_graph *g, const uint32_t *list, uint32_t k) {
    const uint32_t nb = (k + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb) hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(SCAN_THREADS), 0, g->stream, list, k, g->deg, g->scan, g->scan_blk);
    if (nb != 1) hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, g->stream, g->scan_blk, nb);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

static unsigned pull_grid(uint32_t n) {
    uint32_t groups = (n + 63) / 64;
    return grid_for(groups, PULL_THREADS / 64, (unsigned)std::max(1.0, tun().pull_blocks_per_cu * num_cus()));
}

int pr_edgemap(gpr_graph *g, const gpr_emap_desc *desc, gpr_frontier *in, gpr_iter_stats *st);
int policy_action(gpr_graph *g, uint32_t nactv, uint64_t nacte);
void policy_log_features(gpr_graph *g, uint32_t nactv, uint64_t nacte);
double host_clock_ms();
PolicyCtx policy_ctx(gpr_graph *g, uint64_t nacte_in);
void policy_after_step(gpr_graph *g, uint64_t nacte_in, int decision);

// The push/pull switch.  pct <= 0: default_threshold::is_sparse, the
// integer rule nactv + nacte <= m/20 (api.h:150-153); pct > 0:
// frac_threshold(pct)::is_sparse with its float and double roundings
// (api.h:135-139).
bool threshold_sparse(float pct, uint64_t nactv, uint64_t nacte, uint64_t m) {
    if (pct <= 0.f) return nactv + nacte <= m / 20;
    return (double)(float)(nactv + nacte) <= (double)(float)m / (100.0 / (double)pct);
}
// The largest nactv + nacte that threshold_sparse calls sparse.
uint64_t threshold_limit(float pct, uint64_t m) {
    if (pct <= 0.f) return m / 20;
    return (uint64_t)((double)(float)m / (100.0 / (double)pct));
}

// The CSC offsets as 32-bit values, when every offset fits (m < 2^32):
// built once per CSR (graph_replace_csr drops the copy), on the graph's
// stream; *out stays null otherwise.
__global__ void k_narrow_offsets(const uint64_t *__restrict__ ix, uint32_t count, uint32_t *__restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        out[i] = (uint32_t)ix[i];
}
static int csc_offsets32(gpr_graph *g, const uint32_t **out) {
    *out = nullptr;
    if (g->m >= (1ull << 32) || (uint64_t)g->n + 1 >= (1ull << 32)) return GPR_OK;
    if (!g->cindex32) {
        GPR_TRY(dev_alloc((void **)&g->cindex32, sizeof(uint32_t) * ((size_t)g->n + 1), "32-bit CSC index"));
        hipLaunchKernelGGL(k_narrow_offsets, dim3(grid_for((uint64_t)g->n + 1, 256, 8192)), dim3(256), 0, g->stream,
                           g->cindex, g->n + 1, g->cindex32);
        GPR_HIP(hipGetLastError());
    }
    *out = g->cindex32;
    return GPR_OK;
}

int edgemap_minrelax(gpr_graph *g, const gpr_emap_desc *desc, gpr_frontier *in, gpr_frontier *out,
                     gpr_iter_stats *st) {
    gpr_frontier *const out_arg = out;
    GPR_REQUIRE(desc->prop && desc->prop->dtype == GPR_U32 && desc->prop->g == g,
                "edgemap: relax needs a u32 property of this graph");
    GPR_REQUIRE(desc->record == GPR_RECORD_NONE || out, "edgemap: record needs an output frontier");
    GPR_REQUIRE(!out || out != in, "edgemap: output frontier must differ from the input");
    GPR_REQUIRE(desc->record >= GPR_RECORD_NONE && desc->record <= GPR_RECORD_REDUCTION_OR_METHOD,
                "edgemap: unknown record %d", desc->record);
    const bool method = desc->record == GPR_RECORD_METHOD_NE || desc->record == GPR_RECORD_REDUCTION_OR_METHOD;
    GPR_REQUIRE(!method || (desc->prev && desc->prev->dtype == GPR_U32 && desc->prev->g == g),
                "edgemap: record method needs a u32 prev property of this graph");
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const int in_kind = in ? in->kind : GPR_FRONTIER_ALL;
    const uint32_t nactv = in ? in->nactv : n;
    const uint64_t nacte = in ? in->nacte : m;
    gpr_iter_stats s = {};
    s.nactv_in = nactv;
    s.nacte_in = nacte;
    s.levels = 1;
    if (in_kind == GPR_FRONTIER_EMPTY || nactv == 0) {
        if (out) { out->kind = GPR_FRONTIER_EMPTY; out->nactv = 0; out->nacte = 0; }
        if (st) *st = s;
        return GPR_OK;
    }
    int dir = desc->direction;
    int okind = desc->out_kind;
    int fusion = desc->fusion;
    // trace runs force one step's action (no inference)
    bool list_or_bits = false;   // pull: a list while it stays short, else the bitmap
    const bool forced = dir == GPR_DIR_POLICY && g->pol_force_step >= 0;
    const bool use_pol = dir == GPR_DIR_POLICY && (g->pol_params || forced);
    if (dir == GPR_DIR_POLICY) {
        // the learned policy's action; without a network, the rule
        int action = 0;
        if (forced) {
            if (g->pol_log) policy_log_features(g, nactv, nacte);
            if (g->pol_iter == g->pol_force_step) {
                action = g->pol_force_action;
                g->pol_force_t0 = host_clock_ms();
            } else if (g->pol_persist && g->pol_iter > g->pol_force_step) {
                action = g->pol_force_action;   // the forced action holds to the end
            } else if (g->pol_follow) {
                action = policy_action(g, nactv, nacte);   // the network's step (host forward pass)
            }
        } else if (use_pol) {
            action = policy_action(g, nactv, nacte);
        }
        switch (action) {
        case 1:   // unfused sparse push; where the rule pulls, only while the
                  // frontier's edges stay within twice its threshold (a push
                  // of a larger frontier can cost many pulls)
            if (threshold_sparse(desc->threshold_pct, nactv, nacte, m) ||
                nactv + nacte <= 2 * threshold_limit(desc->threshold_pct, m)) {
                dir = GPR_DIR_PUSH;
                fusion = GPR_FUSION_OFF;
            } else {
                dir = GPR_DIR_AUTO;
            }
            break;
        case 2:   // push fused while its levels stay small
            dir = GPR_DIR_PUSH;
            fusion = GPR_FUSION_SMALL;
            break;
        case 3:   // a pull the rule takes anyway, writing a list (a pull's
                  // cost grows with n, which the features do not carry, so
                  // the network may not turn a sparse step into one); the
                  // bitmap is written too, and taken when the list would
                  // be long (list_limit)
            if (threshold_sparse(desc->threshold_pct, nactv, nacte, m)) {
                dir = GPR_DIR_AUTO;
            } else {
                dir = GPR_DIR_PULL;
                okind = GPR_FRONTIER_LIST;
                list_or_bits = true;
            }
            break;
        default:  // the threshold rule's own step
            dir = GPR_DIR_AUTO;
        }
    }
    PolicyCtx pol = {};
    if (use_pol && !forced) pol = policy_ctx(g, nacte);
    if (dir == GPR_DIR_AUTO) dir = threshold_sparse(desc->threshold_pct, nactv, nacte, m) ? GPR_DIR_PUSH : GPR_DIR_PULL;
    // the automatic output representation: a bitmap for a pull, and for a
    // push whose input has more than push_bitmap_frac * n out-edges (a
    // large output).  The next step reads a bitmap back as a list sorted
    // by ID, whose neighbouring sources share the cache lines of their
    // rows and labels; a push's list comes out in arrival order.  On a
    // 16.8M-vertex grid this made pushes of frontiers above 500K vertices
    // 15-20% faster; below n/8 edges the bitmap's clearing and scan (both
    // over n bits) cost more than that gains.
    if (okind == 0)
        okind = dir == GPR_DIR_PUSH && (double)nacte <= tun().push_bitmap_frac * (double)n ? GPR_FRONTIER_LIST
                                                                                      : GPR_FRONTIER_BITMAP;
    // no record argument: nothing is activated and the output stays empty
    // (csr_sparse_no_f / csc_sparse_aset_no_f, api.h:2249-2252, 2321-2322)
    if (desc->record == GPR_RECORD_NONE) out = nullptr;
    if (!out) okind = 0;
    if (!out) list_or_bits = false;
    if (okind == GPR_FRONTIER_LIST) GPR_TRY(frontier_ensure_list(out));
    if (okind == GPR_FRONTIER_BITMAP || list_or_bits) GPR_TRY(frontier_ensure_bits(out));
    const uint32_t list_limit = list_or_bits ? std::max<uint32_t>(65536u, n / 32) : 0u;
    uint32_t *prop = (uint32_t *)desc->prop->d;
    const uint32_t *prev = desc->prev ? (const uint32_t *)desc->prev->d : nullptr;
    GPR_REQUIRE(in || dir != GPR_DIR_PUSH,
                "edgemap: push over the implicit all-true frontier needs an explicit frontier");
    // api::fusion: a sparse step that processes what it activates at once
    // (fusion_select(true), or default_fusion_select: >= 4096 edges,
    // api.h:227-248); the record and dst filter are implied by the relax.
    // The fused kernel runs one workgroup per CU, so a large level costs it
    // more than an unfused step: under default_fusion_select a frontier
    // above fusion_small_limit runs unfused, and a level above it ends the
    // fused step, as under GPR_FUSION_SMALL
    const uint64_t small = (uint64_t)tun().fusion_small_limit;
    const bool small_in = (uint64_t)nactv + nacte <= small;
    const bool fuse_small = (fusion == GPR_FUSION_SMALL || (fusion == GPR_FUSION_DEFAULT && nacte >= 4096)) && small_in;
    const bool fuse = dir == GPR_DIR_PUSH && in && desc->relax != GPR_RELAX_SETIF &&
                      (fusion == GPR_FUSION_ON || fuse_small);
    if (fuse) {
        GPR_HIP(hipEventRecord(g->ev0, g->stream));
        if (in_kind != GPR_FRONTIER_LIST) GPR_TRY(frontier_to_list(in));
        // with the direction left to the engine, a level that turns dense
        // ends the fused step and is returned, so the next step can pull;
        // small-level fusion also hands back the first large level
        uint64_t dense_at = 0;
        if (desc->direction == GPR_DIR_AUTO || desc->direction == GPR_DIR_POLICY)
            dense_at = threshold_limit(desc->threshold_pct, m);
        if (fuse_small) dense_at = dense_at ? std::min(dense_at, small) : small;
        GPR_TRY(fused_push(g, desc->relax, (uint32_t *)desc->prop->d, in, out, dense_at, &s));
        GPR_HIP(hipEventRecord(g->ev1, g->stream));
        GPR_HIP(hipEventSynchronize(g->ev1));
        float ms = 0.f;
        GPR_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
        if (use_pol) policy_after_step(g, nacte, -1);   // no inference in the fused kernel: the host's next
        s.direction = GPR_DIR_PUSH;
        s.out_kind = out ? out->kind : GPR_FRONTIER_EMPTY;
        s.fused = 1;
        s.ms = ms;
        if (st) *st = s;
        return GPR_OK;
    }
    GPR_HIP(hipEventRecord(g->ev0, g->stream));
    // representation changes first: they use the counters themselves
    const uint64_t *in_bits = nullptr;
    if (dir == GPR_DIR_PUSH && in_kind != GPR_FRONTIER_LIST) GPR_TRY(frontier_to_list(in));
    if (dir == GPR_DIR_PULL && in_kind != GPR_FRONTIER_ALL) {
        GPR_TRY(frontier_to_bits(in));
        in_bits = in->bits;
    }
    GPR_TRY(reset_counters(g));
    if (dir == GPR_DIR_PUSH) {
        const uint32_t k = in->nactv;
        const uint64_t E = in->nacte;
        const uint32_t nb = (k + SCAN_BLOCK - 1) / SCAN_BLOCK;
        ProfScope ps(g, "push");
        if (nb)
            hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(SCAN_THREADS), 0, g->stream, in->list, k, g->deg,
                               g->scan, g->scan_blk);
        if (nb != 1) hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, g->stream, g->scan_blk, nb);
        // activations are deduplicated on the output bitmap, or for a list
        // output on the graph's dedup bitmap, which is cleared again from
        // the list afterwards; without an output nothing is activated, so
        // the dedup bitmap stays clear
        uint64_t *bits = okind == GPR_FRONTIER_LIST ? g->dedup : nullptr;
        if (okind == GPR_FRONTIER_BITMAP) {
            GPR_TRY(fill_words(g, out->bits, nwords(n), 0ull));
            bits = out->bits;
        }
        uint32_t *olist = okind == GPR_FRONTIER_LIST ? out->list : nullptr;
        if (E) {
            const unsigned grid = grid_for(E, PUSH_TILE, (unsigned)std::max(1.0, tun().push_blocks_per_cu * num_cus()));
#define PUSH(R)                                                                                                \
    hipLaunchKernelGGL(k_push<R>, dim3(grid), dim3(PUSH_THREADS), 0, g->stream, in->list, k, g->scan, g->scan_blk, E, \
                       g->index, g->nbrs, g->deg, prop, desc->dst_filter, desc->dst_arg, desc->record, prev, bits, \
                       olist, g->ctr)
            if (desc->relax == GPR_RELAX_MIN_PLUS1) PUSH(GPR_RELAX_MIN_PLUS1);
            else if (desc->relax == GPR_RELAX_MIN) PUSH(GPR_RELAX_MIN);
            else PUSH(GPR_RELAX_SETIF);
#undef PUSH
        }
        if (olist) GPR_TRY(clear_dedup_of_list(g, olist, E < n ? E : n));
        s.edges_scanned = E;
    } else {
        uint64_t *ob = okind == GPR_FRONTIER_BITMAP || list_or_bits ? out->bits : nullptr;
        uint32_t *ol = okind == GPR_FRONTIER_LIST ? out->list : nullptr;
        int record = out ? desc->record : GPR_RECORD_NONE;
        ProfScope ps(g, "pull");
        const uint32_t *cix32 = nullptr;
        GPR_TRY(csc_offsets32(g, &cix32));
#define PULL(R, IX, CIX)                                                                                        \
    hipLaunchKernelGGL((k_pull<R, IX>), dim3(pull_grid(n)), dim3(PULL_THREADS), ol ? PULL_QBYTES : 0, g->stream, n,     \
                       CIX, g->cnbrs, g->deg,                                                                    \
                       prop, prev, desc->dst_filter, desc->dst_arg, record, in_bits, ob, ol, g->ctr,                 \
                       (uint32_t)tun().pull_serial_max, list_limit)
        if (cix32) {
            if (desc->relax == GPR_RELAX_MIN_PLUS1) PULL(GPR_RELAX_MIN_PLUS1, uint32_t, cix32);
            else if (desc->relax == GPR_RELAX_MIN) PULL(GPR_RELAX_MIN, uint32_t, cix32);
            else PULL(GPR_RELAX_SETIF, uint32_t, cix32);
        } else {
            if (desc->relax == GPR_RELAX_MIN_PLUS1) PULL(GPR_RELAX_MIN_PLUS1, uint64_t, g->cindex);
            else if (desc->relax == GPR_RELAX_MIN) PULL(GPR_RELAX_MIN, uint64_t, g->cindex);
            else PULL(GPR_RELAX_SETIF, uint64_t, g->cindex);
        }
#undef PULL
    }
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipEventRecord(g->ev1, g->stream));
    // the counts' reduction also runs the network on them (the next step's
    // choice), in the same one-wave kernel
    GPR_TRY(read_counters(g, pol.params ? &pol : nullptr));
    float ms = 0.f;
    GPR_HIP(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    if (desc->record == GPR_RECORD_NONE && out_arg) {
        out_arg->kind = GPR_FRONTIER_EMPTY;
        out_arg->nactv = 0;
        out_arg->nacte = 0;
    }
    if (list_or_bits && g->hctr->listpos > list_limit) okind = GPR_FRONTIER_BITMAP;   // the list is incomplete
    if (out) {
        out->kind = okind;
        out->nactv = okind == GPR_FRONTIER_LIST ? g->hctr->listpos : g->hctr->nactv;
        out->nacte = g->hctr->nacte;
        if (out->nactv == 0) out->kind = GPR_FRONTIER_EMPTY;
        s.nactv_out = out->nactv;
        s.nacte_out = out->nacte;
    }
    if (dir == GPR_DIR_PULL) s.edges_scanned = g->hctr->edges;
    if (use_pol) policy_after_step(g, nacte, forced ? -1 : (int)g->hctr->decision);
    s.direction = dir;
    s.out_kind = okind;
    s.ms = ms;
    if (st) *st = s;
    return GPR_OK;
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_edgemap(gpr_graph *g, const gpr_emap_desc *desc, gpr_frontier *in, gpr_frontier *out,
                           gpr_iter_stats *st) {
    if (!g || !desc) return GPR_E_INVAL;
    GPR_REQUIRE(!in || in->g == g, "edgemap: input frontier belongs to another graph");
    GPR_REQUIRE(!out || out->g == g, "edgemap: output frontier belongs to another graph");
    switch (desc->relax) {
    case GPR_RELAX_MIN_PLUS1:
    case GPR_RELAX_MIN:
    case GPR_RELAX_SETIF:
        return edgemap_minrelax(g, desc, in, out, st);
    case GPR_RELAX_PR_SUM:
        return pr_edgemap(g, desc, in, st);
    default:
        set_error("edgemap: unknown relax %d", desc->relax);
        return GPR_E_INVAL;
    }
}

extern "C" int gpr_vertex_map(gpr_graph *g, int op, gpr_prop *p, const gpr_prop *q, gpr_frontier *f,
                              uint64_t value_bits) {
    if (!g || !p || p->g != g) return GPR_E_INVAL;
    const uint32_t n = g->n;
    const unsigned grid = grid_for(n, 256, 8192);
    switch (op) {
    case GPR_VMAP_FILL:
        if (p->elem == 4)
            hipLaunchKernelGGL(k_vmap_fill32, dim3(grid), dim3(256), 0, g->stream, (uint32_t *)p->d, n, (uint32_t)value_bits);
        else if (p->elem == 8)
            hipLaunchKernelGGL(k_vmap_fill64, dim3(grid), dim3(256), 0, g->stream, (uint64_t *)p->d, n, value_bits);
        else {
            set_error("vertex_map fill: unsupported element size %zu", p->elem);
            return GPR_E_INVAL;
        }
        break;
    case GPR_VMAP_COPY:
        GPR_REQUIRE(q && q->elem == p->elem && q->g == g, "vertex_map copy: mismatched properties");
        GPR_HIP(hipMemcpyAsync(p->d, q->d, p->elem * (size_t)n, hipMemcpyDeviceToDevice, g->stream));
        break;
    case GPR_VMAP_IOTA:
        GPR_REQUIRE(p->dtype == GPR_U32, "vertex_map iota needs u32");
        hipLaunchKernelGGL(k_vmap_iota, dim3(grid), dim3(256), 0, g->stream, (uint32_t *)p->d, n);
        break;
    case GPR_VMAP_FILL_FRONTIER:
        GPR_REQUIRE(p->elem == 4 && f && f->g == g, "vertex_map over a frontier needs a u32 property");
        if (f->kind == GPR_FRONTIER_ALL)
            hipLaunchKernelGGL(k_vmap_fill32, dim3(grid), dim3(256), 0, g->stream, (uint32_t *)p->d, n, (uint32_t)value_bits);
        else if (f->kind == GPR_FRONTIER_LIST && f->nactv)
            hipLaunchKernelGGL(k_vmap_fill_list32, dim3(grid_for(f->nactv, 256, 8192)), dim3(256), 0, g->stream,
                               (uint32_t *)p->d, f->list, f->nactv, (uint32_t)value_bits);
        else if (f->kind == GPR_FRONTIER_BITMAP)
            hipLaunchKernelGGL(k_vmap_fill_bits32, dim3(grid), dim3(256), 0, g->stream, (uint32_t *)p->d, f->bits, n,
                               (uint32_t)value_bits);
        break;
    case GPR_VMAP_COPY_FRONTIER:
        GPR_REQUIRE(p->elem == 4 && q && q->elem == 4 && q->g == g && f && f->g == g,
                    "vertex_map copy over a frontier needs two u32 properties and a frontier");
        if (f->kind == GPR_FRONTIER_ALL)
            GPR_HIP(hipMemcpyAsync(p->d, q->d, 4 * (size_t)n, hipMemcpyDeviceToDevice, g->stream));
        else if (f->kind == GPR_FRONTIER_LIST && f->nactv)
            hipLaunchKernelGGL(k_vmap_copy_list32, dim3(grid_for(f->nactv, 256, 8192)), dim3(256), 0, g->stream,
                               (uint32_t *)p->d, (const uint32_t *)q->d, f->list, f->nactv);
        else if (f->kind == GPR_FRONTIER_BITMAP)
            hipLaunchKernelGGL(k_vmap_copy_bits32, dim3(grid), dim3(256), 0, g->stream, (uint32_t *)p->d,
                               (const uint32_t *)q->d, f->bits, n);
        break;
    default:
        set_error("vertex_map: unknown op %d", op);
        return GPR_E_INVAL;
    }
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

extern "C" int gpr_vertex_filter(gpr_graph *g, gpr_frontier *in, gpr_prop *p, uint32_t value, int neg,
                                 gpr_frontier *out) {
    if (!g || !p || !out || p->dtype != GPR_U32 || out == in) return GPR_E_INVAL;
    const int kind = in ? in->kind : GPR_FRONTIER_ALL;
    GPR_TRY(reset_counters(g));
    if (kind == GPR_FRONTIER_EMPTY) {
        out->kind = GPR_FRONTIER_EMPTY;
        out->nactv = 0;
        out->nacte = 0;
        return GPR_OK;
    }
    if (kind == GPR_FRONTIER_LIST) {
        GPR_TRY(frontier_ensure_list(out));
        const uint32_t k = in->nactv;
        const unsigned blocks = (unsigned)(((uint64_t)k + VF_TILE - 1) / VF_TILE);
        GPR_TRY(ensure_partials(g, blocks + 1));
        unsigned long long *cnt = (unsigned long long *)g->partial;
        if (blocks == 1) {
            hipLaunchKernelGGL(k_vfilter_list, dim3(1), dim3(VF_THREADS), 0, g->stream, in->list, k,
                               (const uint32_t *)p->d, value, neg, out->list, g->deg, cnt, 2, g->ctr);
        } else if (blocks) {
            hipLaunchKernelGGL(k_vfilter_list, dim3(blocks), dim3(VF_THREADS), 0, g->stream, in->list, k,
                               (const uint32_t *)p->d, value, neg, out->list, g->deg, cnt, 0, g->ctr);
            excl_scan_counts(g, cnt, blocks);
            hipLaunchKernelGGL(k_vfilter_list, dim3(blocks), dim3(VF_THREADS), 0, g->stream, in->list, k,
                               (const uint32_t *)p->d, value, neg, out->list, g->deg, cnt, 1, g->ctr);
        }
        GPR_HIP(hipGetLastError());
        GPR_TRY(read_counters(g));
        out->kind = GPR_FRONTIER_LIST;
        out->nactv = g->hctr->nactv;
        out->nacte = g->hctr->nacte;
    } else {
        GPR_TRY(frontier_ensure_bits(out));
        hipLaunchKernelGGL(k_vfilter_bits, dim3(pull_grid(g->n)), dim3(256), 0, g->stream, g->n, (const uint32_t *)p->d,
                           value, neg, kind == GPR_FRONTIER_BITMAP ? in->bits : nullptr, out->bits, g->deg, g->ctr);
        GPR_HIP(hipGetLastError());
        GPR_TRY(read_counters(g));
        out->kind = GPR_FRONTIER_BITMAP;
        out->nactv = g->hctr->nactv;
        out->nacte = g->hctr->nacte;
    }
    if (out->nactv == 0) out->kind = GPR_FRONTIER_EMPTY;
    return GPR_OK;
}
