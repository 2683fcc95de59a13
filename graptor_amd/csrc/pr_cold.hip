// pr_cold.hip -- PageRank edges from "cold" sources, by propagation blocking
// with one value per (source, destination bin) pair.
//
// A pull gathers contrib[s] for every in-edge (s, d).  After degree-ordered
// relabelling the contributions of the first sources stay in the L2, but a
// gather from any other source is a random read that costs a whole cache
// line (64 B for a 4 B value); those gathers bound the plain pull.  The
// edges whose source is at or above a split point `from` (all edges when
// from = 0) are therefore summed by streaming instead:
//
//   k_cold_write    per chunk of a tile of consecutive sources: the tile's
//                   contributions are staged in LDS; its pairs are read in
//                   (bin, source) order (a coalesced 2 B word: the source
//                   offset, bit 15 on the first pair of its (tile, bin)
//                   cell) and each pair's value is written once to its slot
//                   in the bin-major value array, where the tile's pairs of
//                   one bin form a contiguous run.  A pair's cell is the
//                   cell of its 64-pair block's first pair plus a ballot
//                   prefix count of the cell starts, and the cell gives the
//                   run's offset.  The first chunk of a tile also sums
//                   |contribution| x out-degree over the tile's sources (the
//                   mass of its edges).
//   k_cold_accum    per chunk of a bin of COLD_BIN destinations: the bin's
//                   entries (2 B destinations, one per edge, grouped by
//                   pair) are streamed; bit 15 marks the first entry of a
//                   pair, so an entry's pair is the pair of its 64-entry
//                   block's first entry plus a ballot prefix count; a wave
//                   loads the block's next 64 pair values beside its
//                   destinations and hands each lane its value by shuffle.
//                   The sums go into LDS as 64-bit fixed-point integers, so
//                   the result does not depend on the order of the atomics.  A bin in one
//                   chunk updates y directly; the chunks of a heavy bin (the
//                   hub destinations) write integer partials that
//   k_cold_combine_groups and k_cold_combine add up per destination in two
//                   levels (groups of COLD_GROUP partials, then the group
//                   sums), so a hub bin of hundreds of chunks is summed by
//                   many workgroups instead of one dependent chain each.
//
// After degree relabelling a source's edges fall into few bins relative to
// its degree (RMAT: 0.10 / 0.16 / 0.22 pairs per edge at scales 20 / 22 /
// 24), so a value crosses HBM once per pair instead of once per edge; the
// per-edge traffic is the 2 B destination (and per pair a 2 B write word,
// the 4 B value written and the 4 B value read back).  This is the
// "partition-centric" form of propagation blocking (Lakhotia, Kannan and
// Prasanna, USENIX ATC 2018).
//
// The fixed-point unit 2^E is chosen per iteration from the total mass M (a
// deterministic FP64 sum of the tiles' masses): every partial sum is at most
// M <= 2^61 units, and a contribution loses less than one unit.  Write
// chunks go to the XCDs in contiguous ranges of tiles, so that the runs of
// neighbouring tiles, which share cache lines at their ends, are written
// through the same L2.
//
// The layout is static for a graph and split point: built once on the
// device (edge counts per (tile, bin), a stable radix sort of the edges by
// (tile, bin) that keeps each source's edges together, pair flags and their
// prefix sum, pair counts per cell), planned on the host (run offsets,
// the ordinals of the non-empty cells, chunks) and kept until the graph's
// CSR changes.  Entries are the CSR edges s -> d with s >= from, so directed
// graphs work the same way.
#include "internal.hpp"
#include "device.hpp"
#include "pr_codec.hpp"

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <vector>

namespace gpr {

// ---- layout construction ---------------------------------------------------
__device__ __forceinline__ uint32_t tile_of(const uint32_t *tile_lo, uint32_t ntiles, uint32_t u) {
    uint32_t lo = 0, hi = ntiles;   // last t with tile_lo[t] <= u
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tile_lo[mid] <= u) lo = mid; else hi = mid;
    }
    return lo;
}

// Row of edge e: the last r in [lo, hi) with index[r] <= e.
__device__ __forceinline__ uint32_t cold_row_of(const uint64_t *index, uint32_t lo, uint32_t hi, uint64_t e) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (index[mid] <= e) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

// Over the cold edges in tiles of CK_EDGES consecutive edges (balanced
// whatever the degrees; the rows of a tile found through LDS-staged offsets
// when they fit): the sort key tile*nbins + bin and the word (source offset
// in the tile) << 14 | (destination in bin) at the edge's position.
constexpr int CK_EDGES = 4096;
__global__ void __launch_bounds__(256)
k_cold_keys(const uint64_t *__restrict__ index, const uint32_t *__restrict__ nbrs,
            const uint32_t *__restrict__ tile_lo, uint32_t ntiles, uint32_t nbins, uint32_t *keys, uint32_t *vals) {
    __shared__ uint64_t sidx[CK_EDGES + 1];
    __shared__ uint32_t s_r0, s_cnt, s_t0;
    const uint32_t first = tile_lo[0], last = tile_lo[ntiles];
    const uint64_t E0 = index[first], E1 = index[last];
    const uint64_t nblk = (E1 - E0 + CK_EDGES - 1) / CK_EDGES;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint64_t e0 = E0 + k * CK_EDGES, e1 = min(E1, e0 + CK_EDGES);
        if (threadIdx.x == 0) {
            const uint32_t r0 = cold_row_of(index, first, last + 1, e0);
            const uint32_t r1 = cold_row_of(index, r0, last + 1, e1 - 1);
            s_r0 = r0;
            s_cnt = r1 - r0 + 2;
            s_t0 = tile_of(tile_lo, ntiles, r0);
        }
        __syncthreads();
        const uint32_t r0 = s_r0, rc = s_cnt, t0 = s_t0;
        const bool in_lds = rc <= CK_EDGES + 1;
        if (in_lds)
            for (uint32_t j = threadIdx.x; j < rc; j += 256) sidx[j] = index[r0 + j];
        __syncthreads();
        for (uint64_t e = e0 + threadIdx.x; e < e1; e += 256) {
            uint32_t u;
            if (in_lds) {
                uint32_t lo = 0, hi = rc;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (sidx[mid] <= e) lo = mid + 1; else hi = mid;
                }
                u = r0 + lo - 1;
            } else {
                u = cold_row_of(index, r0, r0 + rc, e);
            }
            uint32_t t = t0;   // the rows of one edge block span few tiles
            while (t + 1 < ntiles && tile_lo[t + 1] <= u) ++t;
            const uint32_t v = nbrs[e];
            const uint32_t b = v / COLD_BIN;
            const uint32_t cell = t * nbins + b;   // < 2^32: checked by the host
            keys[e - E0] = cell;
            vals[e - E0] = ((u - tile_lo[t]) << 14) | (v - b * COLD_BIN);
        }
        __syncthreads();
    }
}

// Sorted entry p (tile-major, a source's edges of one bin together): 1 when
// it starts a pair (the first entry of its cell, or a new source); the
// first entry of a cell also records its position (no atomics: each cell
// has one first entry).
__global__ void k_cold_flags(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ words, uint64_t count,
                             uint32_t *flag, uint32_t *cstart) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < count; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = keys[p];
        const bool first = p == 0 || k != keys[p - 1];
        const bool f = first || (words[p] >> 14) != (words[p - 1] >> 14);
        flag[p] = f ? 1u : 0u;
        if (first) cstart[k] = (uint32_t)p;
    }
}

// Per cell with entries: the pairs before it (its first entry starts a
// pair, so incl there counts it too).
__global__ void k_cold_cellpairs(const uint32_t *__restrict__ cstart, uint64_t cells,
                                 const uint32_t *__restrict__ incl, uint32_t *pstart) {
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < cells; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = cstart[c];
        pstart[c] = s == 0xffffffffu ? 0xffffffffu : incl[s] - 1u;
    }
}

// Sorted entry p, with incl = inclusive prefix sum of the flags: its
// destination (and pair flag) at its bin-major slot p + edelta[cell]; a
// pair's write word (its source offset, bit 15 on the first pair of its
// cell) at the pair's tile-major position incl[p] - 1, and for a pair on a
// 64-pair boundary its cell's ordinal; and for an entry on a 64-entry
// boundary of the bin-major order, the block's pair base (the bin-major
// pair of that entry).
__global__ void k_cold_finish(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ words,
                              const uint32_t *__restrict__ incl, uint64_t count, const uint32_t *__restrict__ edelta,
                              const uint32_t *__restrict__ pdelta, const uint32_t *__restrict__ ord, uint16_t *wsrc,
                              uint32_t *wbase, uint16_t *cdst, uint32_t *bbase) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < count; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t cell = keys[p], w = words[p];
        const uint32_t q = incl[p] - 1;   // tile-major pair
        const uint32_t f = incl[p] - (p ? incl[p - 1] : 0u);
        const uint32_t E = (uint32_t)p + edelta[cell];   // modulo 2^32, as the offsets
        cdst[E] = (uint16_t)((w & (COLD_BIN - 1)) | (f << 15));
        if (f) {
            const bool first = p == 0 || keys[p - 1] != cell;   // a cell's first entry starts a pair
            wsrc[q] = (uint16_t)((w >> 14) | (first ? 0x8000u : 0u));
            if ((q & 63u) == 0) wbase[q >> 6] = ord[cell];
        }
        if ((E & 63u) == 0) bbase[E >> 6] = q + pdelta[cell];
    }
}

void cold_free(gpr_graph *g) {
    ColdLayout *L = g->cold;
    if (!L) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    dev_free(L->tile_lo);
    dev_free(L->roff);
    dev_free(L->tcell);
    dev_free(L->wsrc);
    dev_free(L->wbase);
    dev_free(L->cdst);
    dev_free(L->bbase);
    dev_free(L->vals);
    dev_free(L->mass);
    dev_free(L->wchunk);
    dev_free(L->achunk);
    dev_free(L->multi);
    dev_free(L->partials);
    dev_free(L->groups);
    dev_free(L->partials2);
    dev_free(L->gaps);
    delete L;
    g->cold = nullptr;
}

// Tiles over [from, n): each starts at a source with edges and grows while
// it spans < COLD_TILE_V sources and its edges stay <= COLD_TILE_E (a
// single source with more edges gets a tile of its own).  lo[t+1] is the
// start of the next tile (or the end); sources without edges between two
// tiles belong to neither.
static void make_tiles(const std::vector<uint64_t> &idx, uint32_t from, uint32_t n, std::vector<uint32_t> &lo) {
    lo.clear();
    uint32_t u = from;
    while (u < n) {
        while (u < n && idx[u + 1 - from] == idx[u - from]) ++u;   // skip sources without edges
        if (u >= n) break;
        const uint32_t t0 = u;
        uint64_t e = 0;
        while (u < n && u - t0 < COLD_TILE_V) {
            const uint64_t d = idx[u + 1 - from] - idx[u - from];
            if (e > 0 && e + d > COLD_TILE_E) break;
            e += d;
            ++u;
        }
        lo.push_back(t0);
    }
    lo.push_back(u);   // end of the last tile
}

static uint32_t bits_for(uint64_t v) {
    uint32_t b = 1;
    while (b < 64 && (1ull << b) <= v) ++b;
    return b;
}

// Run offsets of a cell-major count table (cells t*nbins + b): the
// bin-major position of cell (t,b) minus its tile-major position, modulo
// 2^32 (both positions are < 2^32); bstart gets each bin's first position.
static void plan_runs(const std::vector<uint32_t> &cnt, uint32_t ntiles, uint32_t nbins, std::vector<uint32_t> &delta,
                      std::vector<uint64_t> &bstart) {
    std::vector<uint64_t> bsize(nbins, 0);
    for (uint32_t t = 0; t < ntiles; ++t)
        for (uint32_t b = 0; b < nbins; ++b) bsize[b] += cnt[(size_t)t * nbins + b];
    bstart.assign(nbins + 1, 0);
    for (uint32_t b = 0; b < nbins; ++b) bstart[b + 1] = bstart[b] + bsize[b];
    std::vector<uint64_t> bpos(bstart.begin(), bstart.end() - 1);   // next slot of each bin
    delta.resize((size_t)ntiles * nbins);
    uint64_t pos = 0;
    for (uint32_t t = 0; t < ntiles; ++t)
        for (uint32_t b = 0; b < nbins; ++b) {
            const size_t cell = (size_t)t * nbins + b;
            delta[cell] = (uint32_t)(bpos[b] - pos);
            bpos[b] += cnt[cell];
            pos += cnt[cell];
        }
}

int cold_prepare(gpr_graph *g, uint32_t from) {
    if (g->cold && g->cold->from == from) return GPR_OK;
    cold_free(g);
    ColdLayout *L = new ColdLayout();
    L->from = from;
    g->cold = L;
    const uint32_t n = g->n;
    if (from >= n) return GPR_OK;   // nothing cold: ntiles = 0
    const uint32_t nbins = (uint32_t)(((uint64_t)n + COLD_BIN - 1) / COLD_BIN);
    std::vector<uint64_t> idx((size_t)(n - from) + 1);
    GPR_HIP(hipMemcpyAsync(idx.data(), g->index + from, sizeof(uint64_t) * idx.size(), hipMemcpyDeviceToHost,
                           g->stream));
    GPR_HIP(hipStreamSynchronize(g->stream));
    const uint64_t ncold = idx.back() - idx.front();
    // no cold edges, or more than 32-bit entry positions address: the pull
    // handles every edge
    if (ncold == 0 || ncold >= 0xffffffffull) return GPR_OK;
    std::vector<uint32_t> lo;
    make_tiles(idx, from, n, lo);
    // the sources outside every tile's span [lo[t], lo[t] + w_t): their
    // ranks are updated by a kernel of their own when the PageRank tail is
    // fused into the write stage
    std::vector<uint32_t> gap_lo, gap_len;
    {
        uint32_t at = from;
        for (size_t t = 0; t + 1 < lo.size(); ++t) {
            if (lo[t] > at) { gap_lo.push_back(at); gap_len.push_back(lo[t] - at); }
            at = lo[t] + std::min<uint32_t>(lo[t + 1] - lo[t], COLD_TILE_V);
        }
        if (n > at) { gap_lo.push_back(at); gap_len.push_back(n - at); }
    }
    idx.clear();
    idx.shrink_to_fit();
    const uint32_t ntiles = (uint32_t)lo.size() - 1;
    const size_t cells = (size_t)ntiles * nbins;
    if (cells >= 0xffffffffull) return GPR_OK;   // cell keys are 32-bit: the pull handles every edge
    uint32_t *cnt = nullptr, *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr, *edelta = nullptr;
    uint32_t *pdelta = nullptr, *dord = nullptr;
    void *tmp = nullptr;
    int r = GPR_OK;
    std::vector<uint32_t> ecnt, pcnt, edel, pdel, ord, roff, tcell;
    std::vector<uint32_t> wch, ach, multi, groups;   // plans, 4 words per item
    uint32_t nslots = 0, nslots2 = 0;
    uint64_t npairs = 0;
    const unsigned pg = grid_for(ncold, CK_EDGES, 65536);
    const unsigned eg = grid_for(ncold, 256, 65536);
    do {
        if ((r = dev_alloc((void **)&L->tile_lo, sizeof(uint32_t) * lo.size(), "cold tiles"))) break;
        if (hipMemcpyAsync(L->tile_lo, lo.data(), sizeof(uint32_t) * lo.size(), hipMemcpyHostToDevice, g->stream) !=
            hipSuccess) { r = GPR_E_HIP; break; }
        // ---- edges sorted by (tile, bin), stable: a source's edges of one
        // bin stay together
        if ((r = dev_alloc((void **)&k0, 4 * ncold, "cold sort keys"))) break;
        if ((r = dev_alloc((void **)&k1, 4 * ncold, "cold sort keys"))) break;
        if ((r = dev_alloc((void **)&v0, 4 * ncold, "cold sort words"))) break;
        if ((r = dev_alloc((void **)&v1, 4 * ncold, "cold sort words"))) break;
        hipLaunchKernelGGL(k_cold_keys, dim3(pg), dim3(256), 0, g->stream, g->index, g->nbrs, L->tile_lo, ntiles,
                           nbins, k0, v0);
        rocprim::double_buffer<uint32_t> kb(k0, k1);
        rocprim::double_buffer<uint32_t> vb(v0, v1);
        const unsigned end_bit = bits_for(cells);
        size_t tsz = 0;
        if (rocprim::radix_sort_pairs(nullptr, tsz, kb, vb, (size_t)ncold, 0, end_bit, g->stream) != hipSuccess) {
            r = GPR_E_HIP;
            break;
        }
        if ((r = dev_alloc(&tmp, tsz, "cold sort workspace"))) break;
        if (rocprim::radix_sort_pairs(tmp, tsz, kb, vb, (size_t)ncold, 0, end_bit, g->stream) != hipSuccess) {
            r = GPR_E_HIP;
            break;
        }
        dev_free(tmp);
        tmp = nullptr;
        // ---- pair flags (into the sort's spare key buffer), cell starts,
        // the flags' prefix sum (into the spare word buffer), pairs before
        // each cell
        uint32_t *flag = kb.alternate(), *incl = vb.alternate();
        if ((r = dev_alloc((void **)&cnt, sizeof(uint32_t) * 2 * cells, "cold cell starts"))) break;
        uint32_t *cstart = cnt, *pstart = cnt + cells;
        if (hipMemsetAsync(cstart, 0xff, sizeof(uint32_t) * cells, g->stream) != hipSuccess) { r = GPR_E_HIP; break; }
        hipLaunchKernelGGL(k_cold_flags, dim3(eg), dim3(256), 0, g->stream, kb.current(), vb.current(), ncold, flag,
                           cstart);
        tsz = 0;
        if (rocprim::inclusive_scan(nullptr, tsz, flag, incl, (size_t)ncold, rocprim::plus<uint32_t>(), g->stream) !=
            hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc(&tmp, tsz, "cold scan workspace"))) break;
        if (rocprim::inclusive_scan(tmp, tsz, flag, incl, (size_t)ncold, rocprim::plus<uint32_t>(), g->stream) !=
            hipSuccess) { r = GPR_E_HIP; break; }
        dev_free(tmp);
        tmp = nullptr;
        hipLaunchKernelGGL(k_cold_cellpairs, dim3(grid_for(cells, 256, 65536)), dim3(256), 0, g->stream, cstart,
                           (uint64_t)cells, incl, pstart);
        std::vector<uint32_t> hstart(2 * cells);
        uint32_t ptotal = 0;
        if (hipMemcpyAsync(hstart.data(), cnt, sizeof(uint32_t) * 2 * cells, hipMemcpyDeviceToHost, g->stream) !=
                hipSuccess ||
            hipMemcpyAsync(&ptotal, incl + (ncold - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream) !=
                hipSuccess ||
            hipStreamSynchronize(g->stream) != hipSuccess) { r = GPR_E_HIP; break; }
        dev_free(cnt);
        cnt = nullptr;
        // ---- counts per cell from the starts of consecutive non-empty cells
        // (cells are numbered in the sorted order)
        ecnt.assign(cells, 0);
        pcnt.assign(cells, 0);
        {
            size_t prev = cells;
            for (size_t c = 0; c <= cells; ++c) {
                if (c < cells && hstart[c] == 0xffffffffu) continue;
                const uint64_t es = c < cells ? hstart[c] : ncold;
                const uint64_t ps = c < cells ? hstart[cells + c] : ptotal;
                if (prev < cells) {
                    ecnt[prev] = (uint32_t)(es - hstart[prev]);
                    pcnt[prev] = (uint32_t)(ps - hstart[cells + prev]);
                }
                prev = c;
            }
        }
        hstart.clear();
        hstart.shrink_to_fit();
        // ---- entry plan: bin-major entry positions, accumulation chunks
        std::vector<uint64_t> ebstart;
        plan_runs(ecnt, ntiles, nbins, edel, ebstart);
        ecnt.clear();
        ecnt.shrink_to_fit();
        for (uint32_t b = 0; b < nbins; ++b) {
            const uint64_t e0 = ebstart[b], e1 = ebstart[b + 1];
            const uint32_t k = e1 > e0 ? (uint32_t)((e1 - e0 + COLD_CHUNK_E - 1) / COLD_CHUNK_E) : 1u;
            for (uint32_t i = 0; i < k; ++i) {
                ach.push_back(b);
                ach.push_back((uint32_t)std::min<uint64_t>(e1, e0 + (uint64_t)i * COLD_CHUNK_E));
                ach.push_back((uint32_t)std::min<uint64_t>(e1, e0 + (uint64_t)(i + 1) * COLD_CHUNK_E));
                ach.push_back(k == 1 ? COLD_DIRECT : nslots + i);
            }
            if (k > 1) {
                // the combine adds a hub bin's k chunk partials in two levels:
                // groups of <= COLD_GROUP slots, then the groups' sums
                multi.push_back(b);
                multi.push_back(nslots);
                multi.push_back(k);
                multi.push_back(nslots2);
                for (uint32_t g0 = 0; g0 < k; g0 += COLD_GROUP) {
                    groups.push_back(nslots + g0);
                    groups.push_back(std::min<uint32_t>(COLD_GROUP, k - g0));
                    groups.push_back(nslots2++);
                    groups.push_back(0);
                }
                nslots += k;
            }
        }
        // ---- pair plan: bin-major pair slots, write chunks per tile
        {
            std::vector<uint64_t> pbstart;
            plan_runs(pcnt, ntiles, nbins, pdel, pbstart);
            npairs = pbstart[nbins];
            uint64_t pos = 0;
            // the non-empty cells in tile-major order: each cell's ordinal,
            // the run offsets by ordinal, each tile's first ordinal (every
            // tile has pairs)
            ord.assign(cells, 0u);
            tcell.assign(ntiles + 1, 0u);
            for (uint32_t t = 0; t < ntiles; ++t) {
                tcell[t] = (uint32_t)roff.size();
                for (uint32_t b = 0; b < nbins; ++b) {
                    const size_t cell = (size_t)t * nbins + b;
                    if (!pcnt[cell]) continue;
                    ord[cell] = (uint32_t)roff.size();
                    roff.push_back(pdel[cell]);
                }
            }
            tcell[ntiles] = (uint32_t)roff.size();
            for (uint32_t t = 0; t < ntiles; ++t) {
                const uint64_t t0 = pos;
                for (uint32_t b = 0; b < nbins; ++b) pos += pcnt[(size_t)t * nbins + b];
                for (uint64_t q = t0; q < pos; q += COLD_WCHUNK) {
                    wch.push_back(t);
                    wch.push_back((uint32_t)q);
                    wch.push_back((uint32_t)std::min<uint64_t>(pos, q + COLD_WCHUNK));
                    wch.push_back(q == t0 ? 1u : 0u);
                }
            }
        }
        pcnt.clear();
        pcnt.shrink_to_fit();
        if ((r = dev_alloc((void **)&L->roff, sizeof(uint32_t) * roff.size(), "cold pair run offsets"))) break;
        if ((r = dev_alloc((void **)&L->tcell, sizeof(uint32_t) * tcell.size(), "cold tile cells"))) break;
        if ((r = dev_alloc((void **)&pdelta, sizeof(uint32_t) * cells, "cold cell pair offsets"))) break;
        if ((r = dev_alloc((void **)&dord, sizeof(uint32_t) * cells, "cold cell ordinals"))) break;
        if ((r = dev_alloc((void **)&edelta, sizeof(uint32_t) * cells, "cold entry run offsets"))) break;
        if (hipMemcpyAsync(L->roff, roff.data(), sizeof(uint32_t) * roff.size(), hipMemcpyHostToDevice, g->stream) !=
                hipSuccess ||
            hipMemcpyAsync(L->tcell, tcell.data(), sizeof(uint32_t) * tcell.size(), hipMemcpyHostToDevice,
                           g->stream) != hipSuccess ||
            hipMemcpyAsync(pdelta, pdel.data(), sizeof(uint32_t) * cells, hipMemcpyHostToDevice, g->stream) !=
                hipSuccess ||
            hipMemcpyAsync(dord, ord.data(), sizeof(uint32_t) * cells, hipMemcpyHostToDevice, g->stream) !=
                hipSuccess ||
            hipMemcpyAsync(edelta, edel.data(), sizeof(uint32_t) * cells, hipMemcpyHostToDevice, g->stream) !=
                hipSuccess) { r = GPR_E_HIP; break; }
        if ((r = dev_alloc((void **)&L->wsrc, 2 * npairs, "cold write words"))) break;
        if ((r = dev_alloc((void **)&L->wbase, 4 * (npairs / 64 + 1), "cold write bases"))) break;
        // zero tails: the accumulation reads up to two loop steps past a
        // chunk's end without bounds checks
        if ((r = dev_alloc((void **)&L->cdst, 2 * (ncold + COLD_PAD_E), "cold destinations"))) break;
        if ((r = dev_alloc((void **)&L->bbase, 4 * (ncold / 64 + 1 + COLD_PAD_E / 64), "cold pair bases"))) break;
        if (hipMemsetAsync(L->cdst, 0, 2 * (ncold + COLD_PAD_E), g->stream) != hipSuccess ||
            hipMemsetAsync(L->bbase, 0, 4 * (ncold / 64 + 1 + COLD_PAD_E / 64), g->stream) != hipSuccess) {
            r = GPR_E_HIP;
            break;
        }
        hipLaunchKernelGGL(k_cold_finish, dim3(eg), dim3(256), 0, g->stream, kb.current(), vb.current(), incl, ncold,
                           edelta, pdelta, dord, L->wsrc, L->wbase, L->cdst, L->bbase);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess) { r = GPR_E_HIP; break; }
        dev_free(k0);
        dev_free(k1);
        dev_free(v0);
        dev_free(v1);
        dev_free(edelta);
        dev_free(pdelta);
        dev_free(dord);
        k0 = k1 = v0 = v1 = edelta = pdelta = dord = nullptr;
        if ((r = dev_alloc((void **)&L->vals, sizeof(uint32_t) * (npairs + 64), "cold values"))) break;
        L->nwchunks = (uint32_t)(wch.size() / 4);
        L->nachunks = (uint32_t)(ach.size() / 4);
        L->nmulti = (uint32_t)(multi.size() / 4);
        L->ngroups = (uint32_t)(groups.size() / 4);
        if ((r = dev_alloc((void **)&L->wchunk, sizeof(uint32_t) * wch.size(), "cold write chunks"))) break;
        if ((r = dev_alloc((void **)&L->achunk, sizeof(uint32_t) * ach.size(), "cold chunks"))) break;
        if ((r = dev_alloc((void **)&L->mass, sizeof(double) * L->nwchunks, "cold mass"))) break;
        if (L->nmulti) {
            if ((r = dev_alloc((void **)&L->multi, sizeof(uint32_t) * multi.size(), "cold multi-chunk bins"))) break;
            if ((r = dev_alloc((void **)&L->partials, sizeof(long long) * COLD_BIN * nslots, "cold partials"))) break;
            if ((r = dev_alloc((void **)&L->groups, sizeof(uint32_t) * groups.size(), "cold combine groups"))) break;
            if ((r = dev_alloc((void **)&L->partials2, sizeof(long long) * COLD_BIN * nslots2, "cold group partials")))
                break;
            if (hipMemcpy(L->multi, multi.data(), sizeof(uint32_t) * multi.size(), hipMemcpyHostToDevice) !=
                    hipSuccess ||
                hipMemcpy(L->groups, groups.data(), sizeof(uint32_t) * groups.size(), hipMemcpyHostToDevice) !=
                    hipSuccess) { r = GPR_E_HIP; break; }
        }
        if (hipMemcpy(L->wchunk, wch.data(), sizeof(uint32_t) * wch.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(L->achunk, ach.data(), sizeof(uint32_t) * ach.size(), hipMemcpyHostToDevice) != hipSuccess) {
            r = GPR_E_HIP;
            break;
        }
    } while (0);
    dev_free(cnt);
    dev_free(k0);
    dev_free(k1);
    dev_free(v0);
    dev_free(v1);
    dev_free(edelta);
    dev_free(pdelta);
    dev_free(dord);
    dev_free(tmp);
    if (r != GPR_OK) {
        if (r == GPR_E_HIP) set_error("PageRank cold-edge layout build failed");
        cold_free(g);
        return r;
    }
    L->ngaps = (uint32_t)gap_lo.size();
    if (L->ngaps) {
        std::vector<uint32_t> h(gap_lo);
        h.push_back(0);
        for (size_t i = 0; i < gap_len.size(); ++i) h.push_back(h.back() + gap_len[i]);
        if ((r = dev_alloc((void **)&L->gaps, sizeof(uint32_t) * h.size(), "cold gaps")) == GPR_OK &&
            hipMemcpyAsync(L->gaps, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice, g->stream) !=
                hipSuccess)
            r = GPR_E_HIP;
        if (r == GPR_OK && hipStreamSynchronize(g->stream) != hipSuccess) r = GPR_E_HIP;
        if (r != GPR_OK) {
            if (r == GPR_E_HIP) set_error("PageRank cold-edge layout build failed");
            cold_free(g);
            return r;
        }
    }
    L->ntiles = ntiles;
    L->nbins = nbins;
    L->ncold = ncold;
    L->npairs = npairs;
    return GPR_OK;
}

// ---- per-iteration kernels -------------------------------------------------
constexpr int CW_THREADS = 1024;
constexpr int CW_UNROLL = 16;
constexpr int CW_STAGE = 8;               // window loads per thread in flight
constexpr uint32_t CW_DELTA_LDS = 4096;   // a tile's run offsets staged in LDS up to this many cells

// Workgroup i -> chunk: the XCD that runs workgroup i (i mod 8 on this
// part) takes a contiguous range of chunks, in order.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t i, uint32_t count) {
    const uint32_t q = count / 8, r = count % 8, x = i % 8, k = i / 8;
    return x * q + min(x, r) + k;
}

// The copy loop of one write chunk: pairs [p0, p1) of the tile, values
// from the LDS window, run offsets by cell ordinal - o0 (< nc) from ro (LDS or
// global, fixed per instantiation so that every access has a known address
// space).  A wave handles 64 consecutive pairs: the ordinal of a lane's
// cell is the block's base (the ordinal of its first pair's cell) plus the
// cell starts among lanes 1..l.  Lanes before p0 read their real words (the
// previous chunk's pairs, whose starts count); lanes at or past p1 come
// after every lane in the chunk and read p0's word.
template <int FMT, class D>
__device__ __forceinline__ void cold_copy(uint32_t p0, uint32_t p1, const uint16_t *__restrict__ wsrc,
                                          const uint32_t *wbase, D ro, uint32_t o0, uint32_t nc,
                                          const typename Codec<FMT>::store_t *win,
                                          typename Codec<FMT>::store_t *__restrict__ vals) {
    // the block bases are uniform: scalar loads through the constant
    // address space
    const __attribute__((address_space(4))) uint32_t *wbs = (const __attribute__((address_space(4))) uint32_t *)wbase;
    const uint32_t wo = __builtin_amdgcn_readfirstlane(threadIdx.x / 64) * 64;
    // from p0 rounded down to 64: every wave reads whole cache lines
    for (uint32_t base = p0 & ~63u; base < p1; base += CW_THREADS * CW_UNROLL) {
        uint32_t wv[CW_UNROLL], bw[CW_UNROLL];
#pragma unroll
        for (int u = 0; u < CW_UNROLL; ++u) {
            const uint32_t p = base + u * CW_THREADS + threadIdx.x;
            wv[u] = __builtin_nontemporal_load(&wsrc[p < p1 ? p : p0]);
            const uint32_t bs = base + u * CW_THREADS + wo;   // the wave's block
            bw[u] = wbs[(bs < p1 ? bs : p0) >> 6];
        }
        uint32_t dl[CW_UNROLL];
#pragma unroll
        for (int u = 0; u < CW_UNROLL; ++u) {
            const uint32_t p = base + u * CW_THREADS + threadIdx.x;
            const uint64_t m = __ballot(wv[u] & 0x8000u);
            const uint32_t o = bw[u] + mbcnt(m >> 1);
            const uint32_t i = o - o0;
            dl[u] = ro[p >= p0 && p < p1 && i < nc ? i : 0u];
        }
#pragma unroll
        for (int u = 0; u < CW_UNROLL; ++u) {
            const uint32_t p = base + u * CW_THREADS + threadIdx.x;
            if (p >= p0 && p < p1) vals[p + dl[u]] = win[wv[u] & 0x7fffu];
        }
    }
}

// The previous iteration's vertex-map tail, fused into the staging (the
// update of gpr_pr_finish without its residual): x' = (y + w) / s and the
// contribution damping * x' / deg, from the scalars of the last reduction.
struct PrTail {
    const float *y;
    float *x;
    float damping;
};

// Stage a tile's w contributions in LDS, CW_STAGE loads per thread at a
// time (clamped, so that all are in flight together); with MASS (the
// tile's first chunk), also this thread's share of sum |contribution| x
// out-degree.  With UPD the contributions are computed from y (PrTail)
// exactly as the tail computes them, and the first chunk writes x'.
template <int FMT, bool MASS, bool UPD>
__device__ __forceinline__ double cold_stage(const typename Codec<FMT>::store_t *__restrict__ c,
                                             const uint32_t *__restrict__ deg, uint32_t w,
                                             typename Codec<FMT>::store_t *win, const PrTail &tl,
                                             const PrScalars *__restrict__ prs) {
    using S = typename Codec<FMT>::store_t;
    double ms = 0.0;
    double pw = 0.0, pinv = 0.0;
    float sc = 1.0f;
    if constexpr (UPD) {
        pw = prs->w;
        pinv = prs->inv_s;
        sc = FMT == GPR_CONTRIB_F16 ? prs->enc_scale : 1.0f;
    }
    for (uint32_t i0 = 0; i0 < w; i0 += CW_THREADS * CW_STAGE) {
        S v[CW_STAGE];
        uint32_t dg[CW_STAGE];
        float yv[CW_STAGE];
#pragma unroll
        for (int k = 0; k < CW_STAGE; ++k) {
            const uint32_t i = i0 + k * CW_THREADS + threadIdx.x, ic = i < w ? i : 0u;
            if constexpr (UPD) {
                yv[k] = tl.y[ic];
                dg[k] = deg[ic];
            } else {
                v[k] = c[ic];
                if constexpr (MASS) dg[k] = deg[ic];
            }
        }
#pragma unroll
        for (int k = 0; k < CW_STAGE; ++k) {
            const uint32_t i = i0 + k * CW_THREADS + threadIdx.x;
            if (i < w) {
                if constexpr (UPD) {
                    const float xn = (float)(((double)yv[k] + pw) * pinv);
                    if constexpr (MASS) {
                        if (tl.x) tl.x[i] = xn;
                    }
                    const float cv = dg[k] ? tl.damping * xn / (float)dg[k] : 0.0f;
                    v[k] = Codec<FMT>::enc(cv * sc);
                }
                win[i] = v[k];
                if constexpr (MASS) ms += (double)fabsf(Codec<FMT>::dec(v[k])) * (double)dg[k];
            }
        }
    }
    return ms;
}

template <int FMT, bool UPD>
__global__ void __launch_bounds__(CW_THREADS)
k_cold_write(const uint32_t *__restrict__ wchunk, uint32_t nwchunks, const uint32_t *__restrict__ tile_lo,
             const uint32_t *__restrict__ tcell, const uint16_t *__restrict__ wsrc, const uint32_t *wbase,
             const uint32_t *__restrict__ roff, uint32_t stage_max,
             const typename Codec<FMT>::store_t *__restrict__ contrib, typename Codec<FMT>::store_t *__restrict__ vals,
             const uint32_t *__restrict__ deg, double *__restrict__ mass, const PrScalars *__restrict__ prs,
             PrTail tail) {
    using S = typename Codec<FMT>::store_t;
    __shared__ S win[COLD_TILE_V];
    __shared__ uint32_t dls[CW_DELTA_LDS];
    __shared__ double red[CW_THREADS / 64];
    const uint32_t c = xcd_chunk(blockIdx.x, nwchunks);
    const uint32_t t = wchunk[4 * c], p0 = wchunk[4 * c + 1], p1 = wchunk[4 * c + 2], first = wchunk[4 * c + 3];
    // a tile's sources span < COLD_TILE_V IDs; tile_lo[t+1] may lie further
    // on when sources without edges follow the tile
    const uint32_t lo = tile_lo[t], w = min(tile_lo[t + 1] - lo, COLD_TILE_V);
    // the tile's first chunk: the mass of its edges, sum of |contribution|
    // x out-degree over its sources (a value is added once per edge)
    PrTail tl = tail;
    if constexpr (UPD) {
        tl.y += lo;
        if (tl.x) tl.x += lo;
    }
    const double ms0 = first ? cold_stage<FMT, true, UPD>(contrib + lo, deg + lo, w, win, tl, prs)
                             : cold_stage<FMT, false, UPD>(contrib + lo, deg + lo, w, win, tl, prs);
    const uint32_t o0 = tcell[t], nc = tcell[t + 1] - o0;   // the tile's non-empty cells
    const bool staged = nc <= stage_max;                     // their run offsets fit in LDS (stage_max <= CW_DELTA_LDS)
    if (staged)
        for (uint32_t i = threadIdx.x; i < nc; i += CW_THREADS) dls[i] = roff[o0 + i];
    __syncthreads();
    if (staged) cold_copy<FMT>(p0, p1, wsrc, wbase, (const uint32_t *)dls, o0, nc, win, vals);
    else cold_copy<FMT>(p0, p1, wsrc, wbase, roff + o0, o0, nc, win, vals);
    const double ms = block_sum(ms0, red);
    if (threadIdx.x == 0) mass[c] = FMT == GPR_CONTRIB_F16 ? ms / (double)prs->enc_scale : ms;
}

constexpr int CA_THREADS = 1024;
constexpr int CA_U = 8;                          // 64-entry blocks of one wave in flight
constexpr uint32_t nw_max = CA_THREADS / 64;

// v * 2^-E truncated toward zero, as an integer (sh = -E - 32).  The unit
// is chosen so that |v| * 2^-E < 2^61 for every contribution, so
// t = |v| * 2^-E / 2^32 < 2^29; scaling by a power of two, floor and the
// subtraction of the floor are exact in FP32 (t has 24 significant bits),
// so hi = floor(t) and lo = trunc((t - hi) * 2^32) are the two 32-bit halves
// of the integer.  (A result below 2^-126 may lose bits to a subnormal, but
// its integer part is 0 either way.)
__device__ __forceinline__ uint64_t to_fixed_mag(float v, int sh) {
    const float t = ldexpf(fabsf(v), sh);
    const float h = floorf(t);
    const uint32_t hi = (uint32_t)h;
    const uint32_t lo = (uint32_t)ldexpf(t - h, 32);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ long long to_fixed(float v, int sh) {
    const uint64_t mag = to_fixed_mag(v, sh);
    return v < 0.0f ? -(long long)mag : (long long)mag;
}

// y[d] (+)= s * unit, and the running sum of y and maximum of y/deg.
__device__ __forceinline__ void cold_store(float *y, uint32_t d, long long s, double unit, int add, double &my_sum,
                                           float &my_max, const uint32_t *deg, bool f16) {
    const double c = (double)s * unit;
    const float yv = (float)(add ? (double)y[d] + c : c);
    y[d] = yv;
    my_sum += (double)yv;
    if (f16) {
        const uint32_t dg = deg[d];
        if (dg) my_max = fmaxf(my_max, yv / (float)dg);
    }
}

template <int NT>
__device__ __forceinline__ void block_emit(double my_sum, float my_max, double *part, float *part_max, uint32_t slot,
                                           double *dred, float *fred) {
    const double ts = block_sum(my_sum, dred);
    my_max = wave_max(my_max);
    __syncthreads();
    if (lane() == 0) fred[threadIdx.x / 64] = my_max;
    __syncthreads();
    if (threadIdx.x == 0) {
        float mx = fred[0];
        for (int i = 1; i < NT / 64; ++i) mx = fmaxf(mx, fred[i]);
        part[slot] = ts;
        part_max[slot] = mx;
    }
}

// Raw buffer descriptor over `bytes` bytes from p (gfx9 word 3: 32-bit
// untyped data); a load past the range returns 0.
constexpr int BUF_WORD3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, BUF_WORD3);
}
constexpr int BUF_NT = 2;   // cache policy of a streamed read (slc: non-temporal)

__device__ __forceinline__ uint32_t buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ uint32_t buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, uint16_t) {
    return (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}

// A stored value's bits in a 32-bit register and back (shuffles move bits).
__device__ __forceinline__ uint32_t store_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ uint32_t store_bits(uint16_t v) { return v; }
template <class S> __device__ __forceinline__ S store_from_bits(uint32_t b);
template <> __device__ __forceinline__ float store_from_bits<float>(uint32_t b) { return __uint_as_float(b); }
template <> __device__ __forceinline__ uint16_t store_from_bits<uint16_t>(uint32_t b) { return (uint16_t)b; }

template <int FMT>
__global__ void __launch_bounds__(CA_THREADS)
k_cold_accum(const uint32_t *__restrict__ achunk, uint32_t n, const uint16_t *cdst, const uint32_t *bbase,
             const typename Codec<FMT>::store_t *vals, uint64_t npairs,
             const uint32_t *__restrict__ deg, float *__restrict__ y, long long *__restrict__ partials, int add,
             double *__restrict__ part, float *__restrict__ part_max, const PrScalars *__restrict__ prs) {
    using C = Codec<FMT>;
    using S = typename C::store_t;
    __shared__ unsigned long long acc[COLD_BIN];
    __shared__ double dred[CA_THREADS / 64];
    __shared__ float fred[CA_THREADS / 64];
    const uint32_t b = achunk[4 * blockIdx.x], p0 = achunk[4 * blockIdx.x + 1], p1 = achunk[4 * blockIdx.x + 2];
    const uint32_t slot = achunk[4 * blockIdx.x + 3];
    for (uint32_t i = threadIdx.x; i < COLD_BIN; i += CA_THREADS) acc[i] = 0ull;
    __syncthreads();
    const double unit = prs->cold_unit;
    const float dscale = FMT == GPR_CONTRIB_F16 ? 1.0f / prs->enc_scale : 1.0f;
    // the wave index through readfirstlane: known uniform, so the offsets
    // and pair bases derived from it live in scalar registers
    const unsigned w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), ln = lane(), nw = CA_THREADS / 64;
    const int shift = -prs->cold_exp - 32;
    // 32-bit offsets from the chunk's first entry rounded down to 64, so
    // that every wave reads whole cache lines no other wave reads, and a
    // wave's 64 entries are one block of the pair bases.  Reads run past the
    // chunk (into the next one, or the arrays' zero tails) without bounds
    // checks; only the atomics are limited to the chunk.
    //
    // A block's entries use pairs pb .. pb+63 at most (pb = the pair of its
    // first entry), so lane l loads vals[pb + l] beside its destination and
    // takes the value of pair pb + r from lane r, r = the pair starts among
    // lanes 1..l (entries past the chunk sit in higher lanes and never
    // count): no load waits for another.
    const uint32_t a0 = p0 & ~63u, lo = p0 - a0, len = p1 - a0;
    const uint32_t *bb = bbase + (a0 >> 6);
    const uint32_t step = nw * 64 * CA_U;
    static_assert(3 * nw_max * 64 * CA_U <= COLD_PAD_E, "zero tail shorter than three loop steps");
    // Buffer descriptors from the chunk's first entry and first pair, each
    // covering what the chunk's blocks use: the destinations up to the
    // chunk's end rounded up to a block, the pairs up to the window of the
    // first block past that end (all pairs from the first on, with their
    // tail, when that base is the zero tail).  The whole offset of a load is
    // its vector offset, which the range check covers, so the loads of the
    // step prefetched past the chunk's end return 0 without reading memory.
    // A block past the end of the pair-base array reads base 0 (its zero
    // tail); bases are clamped to the chunk's first pair, so no offset is
    // negative.
    const uint32_t lenr = (len + 63u) & ~63u;
    const __amdgpu_buffer_rsrc_t krs = buf_rsrc(cdst + a0, 2u * lenr);
    const uint32_t pfirst = __builtin_amdgcn_readfirstlane(bb[0]);
    const uint32_t pend = __builtin_amdgcn_readfirstlane(bb[lenr >> 6]);
    const uint64_t vall = npairs + 64 - pfirst;
    const uint64_t vcnt = pend > pfirst && (uint64_t)(pend - pfirst) + 64 < vall ? (uint64_t)(pend - pfirst) + 64 : vall;
    const uint64_t vbytes = vcnt * (uint64_t)sizeof(S);
    const __amdgpu_buffer_rsrc_t vrs = buf_rsrc(vals + pfirst, vbytes < 0xffffffffull ? (uint32_t)vbytes : 0xffffffffu);
    // Two steps in flight: the loads of the next step are issued before
    // the shuffles and atomics of this one, so that a wave's LDS work
    // overlaps its own memory latency (one workgroup per CU, four waves per
    // SIMD, leave the registers for it).  The bases run two steps ahead.
    uint32_t pbn[CA_U], key[CA_U], vv[CA_U];
    auto issue = [&](uint32_t at, uint32_t *kk, uint32_t *vl) {
#pragma unroll
        for (int u = 0; u < CA_U; ++u) {
            kk[u] = (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(krs, (int)(2 * (ln + at + u * nw * 64)), 0, BUF_NT);
            // the base was loaded a step ago and is uniform by construction;
            // readfirstlane puts it, and the offset derived from it, in
            // scalar registers (an offset left in a vector register, as a
            // mixed signed/unsigned max() or a fused clamp-subtract produce,
            // makes the compiler treat it as non-uniform)
            const uint32_t pbr = (uint32_t)__builtin_amdgcn_readfirstlane(pbn[u]);
            const uint32_t pb = pbr > pfirst ? pbr : pfirst;
            const uint32_t soff = (uint32_t)__builtin_amdgcn_readfirstlane((int)(sizeof(S) * (pb - pfirst)));
            vl[u] = buf_load(vrs, sizeof(S) * ln + soff, 0u, S());
        }
    };
    // the bases are uniform: read through the constant address space they
    // are scalar loads, counted apart from the vector loads, so waiting for
    // them does not wait for the step in flight
    const __attribute__((address_space(4))) uint32_t *bbs = (const __attribute__((address_space(4))) uint32_t *)bb;
    auto bases = [&](uint32_t at) {
#pragma unroll
        for (int u = 0; u < CA_U; ++u) pbn[u] = bbs[(at + u * nw * 64) >> 6];
    };
    // LDS atomics are exact whatever the order and whatever the number of
    // lanes on one address (the hub destinations).  A block wholly inside
    // the chunk needs no per-lane range test, and the sign is applied only
    // in a wave with a negative value (PageRank's are not).
    auto process = [&](uint32_t base, const uint32_t *kk, const uint32_t *vl) {
        // every lane's value first (the shuffles are in flight together),
        // then the atomics
        uint32_t vb[CA_U];
#pragma unroll
        for (int u = 0; u < CA_U; ++u) {
            const uint64_t m = __ballot(kk[u] & 0x8000u);
            const uint32_t r = mbcnt(m >> 1);
            vb[u] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(r << 2), (int)vl[u]);
        }
#pragma unroll
        for (int u = 0; u < CA_U; ++u) {
            const uint32_t ob = base + u * nw * 64, o = ob + ln;
            const float v = C::dec(store_from_bits<S>(vb[u])) * dscale;
            uint64_t fx = to_fixed_mag(v, shift);
            if (__builtin_expect(__ballot(v < 0.0f) != 0, 0)) fx = v < 0.0f ? (uint64_t)(-(long long)fx) : fx;
            bool in = true;
            if (!(ob >= 64 && ob + 64 <= len)) in = o >= lo && o < len;
            if (in) atomicAdd(&acc[kk[u] & (COLD_BIN - 1)], (unsigned long long)fx);
        }
    };
    // two register sets that take turns (unrolled by two, so that no set is
    // copied into the other: a copy would wait for its loads); the asm
    // keeps each step's loads issued before the other step's uses (it stops
    // loads from being sunk to their uses; the arrays are not __restrict__
    // for this reason)
    uint32_t key2[CA_U], vv2[CA_U];
    bases(w * 64);
    issue(w * 64, key, vv);
    bases(w * 64 + step);
    for (uint32_t base = w * 64; base < len; base += 2 * step) {
        issue(base + step, key2, vv2);
        bases(base + 2 * step);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        process(base, key, vv);
        if (base + step >= len) break;
        issue(base + 2 * step, key, vv);
        bases(base + 3 * step);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        process(base + step, key2, vv2);
    }
    __syncthreads();
    const uint32_t d0 = b * COLD_BIN, nd = min(COLD_BIN, n - d0);
    if (slot != COLD_DIRECT) {
        // one of several chunks of this bin: integer partials for the combine
        long long *pp = partials + (size_t)slot * COLD_BIN;
        for (uint32_t i = threadIdx.x; i < COLD_BIN; i += CA_THREADS) pp[i] = (long long)acc[i];
        if (threadIdx.x == 0) {
            part[b] = 0.0;
            part_max[b] = 0.0f;
        }
        return;
    }
    double my_sum = 0.0;
    float my_max = 0.0f;
    for (uint32_t i = threadIdx.x; i < nd; i += CA_THREADS)
        cold_store(y, d0 + i, (long long)acc[i], unit, add, my_sum, my_max, deg, FMT == GPR_CONTRIB_F16);
    block_emit<CA_THREADS>(my_sum, my_max, part, part_max, b, dred, fred);
}

constexpr int CC_THREADS = 1024;
constexpr uint32_t CC_PER_BIN = COLD_BIN / CC_THREADS;

// Bins in several chunks, first level: one workgroup per (group of <=
// COLD_GROUP chunk partials, CC_THREADS destinations) adds the group's
// partials, all loads in flight together, into the group's slot.
__global__ void __launch_bounds__(CC_THREADS)
k_cold_combine_groups(const uint32_t *__restrict__ groups, const long long *__restrict__ partials,
                      long long *__restrict__ partials2) {
    const uint32_t gi = blockIdx.x / CC_PER_BIN, sub = blockIdx.x % CC_PER_BIN;
    const uint32_t first = groups[4 * gi], cnt = groups[4 * gi + 1], slot2 = groups[4 * gi + 2];
    const uint32_t i = sub * CC_THREADS + threadIdx.x;
    long long t[COLD_GROUP];
#pragma unroll
    for (uint32_t c = 0; c < COLD_GROUP; ++c) t[c] = c < cnt ? partials[(size_t)(first + c) * COLD_BIN + i] : 0ll;
    long long s = 0;
#pragma unroll
    for (uint32_t c = 0; c < COLD_GROUP; ++c) s += t[c];
    partials2[(size_t)slot2 * COLD_BIN + i] = s;
}

// Second level: one workgroup per CC_THREADS destinations of a bin adds its
// groups' sums and stores y; its y sum and y/deg maximum go to slot
// nbins + blockIdx.x.  Integer sums: exact in any grouping.
__global__ void __launch_bounds__(CC_THREADS)
k_cold_combine(const uint32_t *__restrict__ multi, uint32_t nbins, uint32_t n, const long long *__restrict__ partials2,
               const uint32_t *__restrict__ deg, float *__restrict__ y, int add, double *__restrict__ part,
               float *__restrict__ part_max, const PrScalars *__restrict__ prs, int f16) {
    __shared__ double dred[CC_THREADS / 64];
    __shared__ float fred[CC_THREADS / 64];
    const uint32_t mb = blockIdx.x / CC_PER_BIN, sub = blockIdx.x % CC_PER_BIN;
    const uint32_t b = multi[4 * mb], first = multi[4 * mb + 3];
    const uint32_t k = (multi[4 * mb + 2] + COLD_GROUP - 1) / COLD_GROUP;
    const uint32_t i = sub * CC_THREADS + threadIdx.x;
    const uint32_t d0 = b * COLD_BIN, nd = min(COLD_BIN, n - d0);
    double my_sum = 0.0;
    float my_max = 0.0f;
    if (i < nd) {
        long long s = 0;
        uint32_t c = 0;
        for (; c + 8 <= k; c += 8) {
            long long t[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = partials2[(size_t)(first + c + j) * COLD_BIN + i];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += t[j];
        }
        for (; c < k; ++c) s += partials2[(size_t)(first + c) * COLD_BIN + i];
        cold_store(y, d0 + i, s, prs->cold_unit, add, my_sum, my_max, deg, f16 != 0);
    }
    block_emit<CC_THREADS>(my_sum, my_max, part, part_max, nbins + blockIdx.x, dred, fred);
}

// x' = (y + w) / s for the sources in no tile (no out-edges), whose
// contributions no edge reads: the rest of the fused tail.  zero_y: the
// graph is symmetric, so these vertices have no in-edges either and y is 0
// for them (y is never read: x' = (0 + w) / s, the same value).
__global__ void __launch_bounds__(256)
k_pr_update_gaps(const uint32_t *__restrict__ gaps, uint32_t ngaps, const float *__restrict__ y, float *x,
                 const PrScalars *__restrict__ prs, int zero_y) {
    const uint32_t *pre = gaps + ngaps;
    const uint32_t total = pre[ngaps];
    const double w = prs->w, inv_s = prs->inv_s;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
        uint32_t lo = 0, hi = ngaps;   // last r with pre[r] <= i
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= i) lo = mid; else hi = mid;
        }
        const uint32_t v = gaps[lo] + (i - pre[lo]);
        x[v] = (float)(((zero_y ? 0.0 : (double)y[v]) + w) * inv_s);
    }
}

bool cold_tail_fusable(const gpr_graph *g) { return g->cold && g->cold->from == 0 && g->cold->ntiles > 0; }

int cold_launch_write(gpr_graph *g, int fmt, const void *contrib, const float *y, float *x, float damping) {
    ColdLayout *L = g->cold;
    const uint32_t stage_max = (uint32_t)std::min<double>(CW_DELTA_LDS, tun().pr_write_stage_cells);
    // with GPR_PR_SKIP_RANKS the tail computes the contributions but writes
    // no rank (x stays null in the kernel)
    const PrTail tail = {y, g->pr_skip_ranks ? nullptr : x, damping};
    const bool upd = x != nullptr;
    if (upd && L->ngaps && !g->pr_skip_unlinked)
        hipLaunchKernelGGL(k_pr_update_gaps, dim3(grid_for(g->n, 256, 2048)), dim3(256), 0, g->stream, L->gaps,
                           L->ngaps, y, x, g->prs, g->symmetric ? 1 : 0);
#define CW_(F, U)                                                                                               \
    hipLaunchKernelGGL((k_cold_write<F, U>), dim3(L->nwchunks), dim3(CW_THREADS), 0, g->stream, L->wchunk,      \
                       L->nwchunks, L->tile_lo, L->tcell, L->wsrc, L->wbase, L->roff, stage_max,                \
                       (const typename Codec<F>::store_t *)contrib, (typename Codec<F>::store_t *)L->vals, g->deg,  \
                       L->mass, g->prs, tail)
#define CW(F)                                                                                                   \
    if (upd) CW_(F, true); else CW_(F, false)
    switch (fmt) {
    case GPR_CONTRIB_F32: CW(GPR_CONTRIB_F32); break;
    case GPR_CONTRIB_BF16: CW(GPR_CONTRIB_BF16); break;
    case GPR_CONTRIB_F16: CW(GPR_CONTRIB_F16); break;
    case GPR_CONTRIB_CFP16: CW(GPR_CONTRIB_CFP16); break;
    case GPR_CONTRIB_BF16T: CW(GPR_CONTRIB_BF16T); break;
    case GPR_CONTRIB_SCFP16: CW(GPR_CONTRIB_SCFP16); break;
    }
#undef CW
#undef CW_
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

int cold_launch_accum(gpr_graph *g, int fmt, float *y, int add, double *part, float *part_max, uint32_t *nparts) {
    ColdLayout *L = g->cold;
#define CA(F)                                                                                                    \
    hipLaunchKernelGGL(k_cold_accum<F>, dim3(L->nachunks), dim3(CA_THREADS), 0, g->stream, L->achunk, g->n,       \
                       L->cdst, L->bbase, (const typename Codec<F>::store_t *)L->vals, L->npairs, g->deg, y,      \
                       L->partials, add, part, part_max, g->prs)
    switch (fmt) {
    case GPR_CONTRIB_F32: CA(GPR_CONTRIB_F32); break;
    case GPR_CONTRIB_BF16: CA(GPR_CONTRIB_BF16); break;
    case GPR_CONTRIB_F16: CA(GPR_CONTRIB_F16); break;
    case GPR_CONTRIB_CFP16: CA(GPR_CONTRIB_CFP16); break;
    case GPR_CONTRIB_BF16T: CA(GPR_CONTRIB_BF16T); break;
    case GPR_CONTRIB_SCFP16: CA(GPR_CONTRIB_SCFP16); break;
    }
#undef CA
    if (L->nmulti) {
        hipLaunchKernelGGL(k_cold_combine_groups, dim3(L->ngroups * CC_PER_BIN), dim3(CC_THREADS), 0, g->stream,
                           L->groups, (const long long *)L->partials, L->partials2);
        hipLaunchKernelGGL(k_cold_combine, dim3(L->nmulti * CC_PER_BIN), dim3(CC_THREADS), 0, g->stream, L->multi,
                           L->nbins, g->n, (const long long *)L->partials2, g->deg, y, add, part, part_max, g->prs,
                           fmt == GPR_CONTRIB_F16 ? 1 : 0);
    }
    GPR_HIP(hipGetLastError());
    *nparts = L->nbins + L->nmulti * CC_PER_BIN;
    return GPR_OK;
}

} // namespace gpr
