// grid_sync.hpp -- the grid-wide barrier of the persistent cooperative
// launches (the fused push, the k-core peeling).
//
// Every word one block writes and another reads inside such a launch is
// accessed with agent-scope atomics or coherent loads and stores; before
// arriving at the barrier each wave waits until its memory operations are
// acknowledged, so after the barrier every block sees them.  No
// acquire/release fence is used: on this part each would write back or
// invalidate an XCD's L2.
#pragma once
#include "device.hpp"

namespace gpr {

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): waits until
// this wave's earlier memory operations, stores and atomics included, are
// acknowledged
constexpr int WAIT_VM0 = 0x0f70;

constexpr int GS_GROUP = 16;            // blocks per first-level barrier counter
constexpr int GS_MAX_GROUPS = 64;
constexpr int GS_LINE = 16;             // u64 per 128-byte line: one counter per line

__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long bcast64(unsigned long long v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Grid barrier in two levels.  Block b arrives on the counter of group
// b % ngroups (at most GS_GROUP blocks each); the last block of a group to
// arrive at level `lvl` counts the group on `arrive`, and level `lvl` ends
// when (lvl+1)*ngroups groups are counted.  All counters are monotonic.
// Returns false (in every thread) when the error flag is set or the wait saw
// no new arrival for `spin_limit` polls in a row.
__device__ __forceinline__ bool grid_barrier(unsigned long long *arrive, unsigned int *error, unsigned long long *group,
                                             unsigned lvl, unsigned spin_limit, int *s_ok) {
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned ngroups = (gridDim.x + GS_GROUP - 1) / GS_GROUP;
        const unsigned grp = blockIdx.x % ngroups;
        const unsigned gsize = gridDim.x / ngroups + (grp < gridDim.x % ngroups ? 1u : 0u);
        const unsigned long long got = __hip_atomic_fetch_add(&group[grp * GS_LINE], 1ull, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
        if (got + 1 == (unsigned long long)(lvl + 1) * gsize)
            __hip_atomic_fetch_add(arrive, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long target = (unsigned long long)(lvl + 1) * ngroups;
        unsigned long long seen = ~0ull;
        unsigned spin = 0;
        int ok = 1;
        for (;;) {
            const unsigned long long a = ld_agent(arrive);
            if (a >= target) break;
            if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (a != seen) {
                seen = a;
                spin = 0;
            } else if (++spin > spin_limit) {
                __hip_atomic_fetch_or(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

} // namespace gpr
