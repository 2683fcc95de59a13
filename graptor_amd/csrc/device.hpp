// device.hpp -- wave64 helpers for gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace gpr {

__device__ __forceinline__ unsigned lane() { return __lane_id(); }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ unsigned mbcnt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T> __device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T t = __shfl_xor(v, o, 64);
        v = t < v ? t : v;
    }
    return v;
}
template <class T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T t = __shfl_xor(v, o, 64);
        v = t > v ? t : v;
    }
    return v;
}
// Inclusive prefix sum across the wave.
template <class T> __device__ __forceinline__ T wave_incl_scan(T v) {
    unsigned l = lane();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T t = __shfl_up(v, o, 64);
        if (l >= (unsigned)o) v += t;
    }
    return v;
}

// Block-wide sum of a per-thread value; result valid in every thread.
// `red` must hold blockDim.x/64 elements.
template <class T> __device__ __forceinline__ T block_sum(T v, T *red) {
    v = wave_sum(v);
    unsigned w = threadIdx.x / 64, nw = blockDim.x / 64;
    __syncthreads();
    if (lane() == 0) red[w] = v;
    __syncthreads();
    T s = 0;
    for (unsigned i = 0; i < nw; ++i) s += red[i];
    return s;
}

__device__ __forceinline__ bool dst_ok(uint32_t v, int kind, uint32_t arg) {
    return kind == GPR_DST_GT ? v > arg : kind == GPR_DST_NE ? v != arg : true;
}

__device__ __forceinline__ bool bit_test(const uint64_t *bits, uint32_t v) {
    return (bits[v >> 6] >> (v & 63)) & 1ull;
}

// Edge position of frontier entry i from the blocked degree prefix
// (scan[i] + scan_blk[i / SCAN_BLOCK]).
__device__ __forceinline__ uint64_t frontier_pos(const uint64_t *pos, const uint64_t *btot, uint32_t i) {
    return pos[i] + btot[i / SCAN_BLOCK];
}

// First index i in [0,k) with pos(i) > e (k if none).
__device__ __forceinline__ uint32_t ub_pos(const uint64_t *pos, const uint64_t *btot, uint32_t k, uint64_t e) {
    uint32_t lo = 0, hi = k;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (frontier_pos(pos, btot, mid) <= e) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// First index i in [lo, k) with pos(i) > e (k if none), by one wave: 64
// probes per round, evenly spaced over the remaining range, and a ballot of
// pos <= e (a prefix of the lanes, pos being non-decreasing) narrows the
// range 64-fold; about log64(k) rounds of one load each, where a one-lane
// binary search waits on log2(k) loads in a row.  All 64 lanes take part
// and get the result.
__device__ __forceinline__ uint32_t wave_ub_pos(const uint64_t *pos, const uint64_t *btot, uint32_t k, uint64_t e,
                                                uint32_t lo) {
    uint32_t hi = k;   // the answer lies in [lo, hi]
    const uint32_t l = lane();
    while (lo < hi) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t p = lo + l * step;
        const bool le = p < hi && frontier_pos(pos, btot, p) <= e;
        const uint32_t c = (uint32_t)__popcll(__ballot(le));
        if (c == 0) break;   // pos(lo) > e
        // the answer lies in (lo + (c-1) step, lo + c step], clipped to hi
        const uint32_t nhi = min(lo + c * step, hi);
        lo = lo + (c - 1) * step + 1;
        hi = nhi;
    }
    return lo;
}

// A wave's private queue of list entries in LDS (cap entries, cap a
// multiple of 64 and >= 128).  Entries go in 64 lanes at a time, as a
// ballot word; the queue is appended to the output list with ONE atomic on
// the list position when the next word might not fit, and at the end, so
// that a list of k entries takes about k / (cap - 64) atomics on that one
// word instead of one per wave word.  The order of the entries in the
// list is not specified.  All 64 lanes call push and flush together.
//
// With a limit (list_cap > 0), a wave that sees the list position pass it
// stops queueing (full): the caller then takes another representation of
// the same set (the list is incomplete), and the list costs at most about
// list_cap / (cap - 64) + one atomic per wave.
struct WaveQueue {
    uint32_t *q;     // this wave's cap entries of LDS
    uint32_t cap;
    uint32_t cnt;    // wave-uniform
    uint32_t list_cap = 0;
    bool full = false;
    __device__ __forceinline__ void flush(uint32_t *out, unsigned int *listpos) {
        if (!cnt) return;
        __builtin_amdgcn_wave_barrier();
        uint32_t base = 0;
        if (lane() == 0) base = atomicAdd(listpos, cnt);
        base = __shfl(base, 0, 64);
        for (uint32_t j = lane(); j < cnt; j += 64) out[base + j] = q[j];
        __builtin_amdgcn_wave_barrier();
        if (list_cap && base + cnt > list_cap) full = true;
        cnt = 0;
    }
    __device__ __forceinline__ void push(bool keep, uint64_t word, uint32_t v, uint32_t *out,
                                         unsigned int *listpos) {
        if (!word || full) return;
        if (cnt + 64 > cap) {
            flush(out, listpos);
            if (full) return;
        }
        if (keep) q[cnt + mbcnt(word)] = v;
        cnt += (uint32_t)__popcll(word);
    }
};

} // namespace gpr
