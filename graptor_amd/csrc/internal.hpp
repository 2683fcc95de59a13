// internal.hpp -- shared definitions of the gfx950 traversal engine.
//
// Data layout in HBM (per graph handle):
//   index  u64[n+1]   CSR offsets (Graptor EID), nbrs u32[m] (Graptor VID)
//   cindex/cnbrs      CSC; aliases the CSR for symmetric graphs
//   deg    u32[n]     out-degree
//   bucket u32[nb+1]  row starts of the pull "stream" buckets over the CSC:
//                     rows r with floor((cindex[r] + r) / STREAM_ITEMS) == b
//   long rows         CSC rows with in-degree > LONG_DEG, cut into
//                     LONG_CHUNK-edge chunks processed by whole workgroups
// Frontiers are either a u32 vertex list or a bitmap of u64 words (bit v of
// word v/64), one wave64 ballot per word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>

#include "../../include/gpr.h"

namespace gpr {

// ---- tunables (the autotuner may override the launch-time ones) ---------
constexpr int WAVE = 64;
constexpr int STREAM_ITEMS = 2048;  // rows + edges per pull bucket
constexpr int STREAM_THREADS = 256;
constexpr uint32_t LONG_DEG = 2048;   // rows above this are "long"
constexpr uint32_t LONG_CHUNK = 8192; // edges per long-row chunk
constexpr int PUSH_THREADS = 256;
constexpr int PUSH_ITEMS = 4;         // edges per thread per push tile
constexpr int SCAN_BLOCK = 1024;      // frontier entries per scan block

// ---- error handling ------------------------------------------------------
void set_error(const char *fmt, ...);
#define GPR_HIP(expr)                                                       \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            ::gpr::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #expr,    \
                             hipGetErrorString(e_));                        \
            return GPR_E_HIP;                                               \
        }                                                                   \
    } while (0)
#define GPR_TRY(expr)                                                       \
    do {                                                                    \
        int r_ = (expr);                                                    \
        if (r_ != GPR_OK) return r_;                                        \
    } while (0)
#define GPR_REQUIRE(cond, ...)                                              \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::gpr::set_error(__VA_ARGS__);                                  \
            return GPR_E_INVAL;                                             \
        }                                                                   \
    } while (0)

int dev_alloc(void **p, size_t bytes, const char *what);
void dev_free(void *p);

// ---- runtime tunables (set by the autotuner through gpr_tune_set) --------
struct Tunables {
    double pull_serial_max = 32;              // lane-serial rows up to this degree
    double pull_blocks_per_cu = 8;            // pull grid = CUs * this
    double push_blocks_per_cu = 32;           // push grid cap = CUs * this
    double pr_split_bytes = 0;                // FP32 bytes of pulled sources (0: all binned; >= 4n: no split)
    double fusion_spin_limit = 1 << 22;       // fused push: polls without progress before it gives up
    double fusion_blocks = 0;                 // fused push: cooperative grid size (0: one per CU)
    double pr_write_stage_cells = 4096;       // PR write: a tile's run offsets go to LDS up to this many cells
    double fusion_small_limit = 1 << 17;      // GPR_FUSION_SMALL / _DEFAULT: frontier size that fuses / level size that ends it
    double fusion_cooperative = 1;            // fused push: 1 = hipLaunchCooperativeKernel, 0 = a plain launch
    double push_bitmap_frac = 0.125;          // auto output of a push: a bitmap when the input's edges exceed this * n
};
Tunables &tun();
int num_cus();

// ---- device-resident counters of one handle ------------------------------
// The frontier counts are striped: workgroup b adds its totals to stripe
// b % CTR_STRIPES (one 128-byte line each, right after the DevCounters
// line), so no single line takes an atomic from every workgroup of a large
// grid; a one-wave kernel sums the stripes when the host reads the counters
// back (k_read_counters).  The list append position stays one word (it
// allocates positions).  The DevCounters line of the host mirror receives
// the totals and the policy's choice.
constexpr int CTR_STRIPES = 64;
struct alignas(128) DevCounters {
    unsigned long long nacte;   // host mirror: edges of the frontier produced
    unsigned long long edges;   // host mirror: edges examined
    unsigned int nactv;         // host mirror: vertices of the frontier produced
    unsigned int listpos;       // append position of the list being produced
    unsigned int pad0;
    unsigned int ticket2;       // last-workgroup ticket of the FP64 reductions (reset by that workgroup)
    unsigned int decision;      // host mirror: the policy's choice for the next step
    unsigned int pad;
};
struct alignas(128) CtrStripe {
    unsigned long long nacte;
    unsigned long long edges;
    unsigned int nactv;
    unsigned int pad;
};
constexpr size_t COUNTER_BYTES = sizeof(DevCounters) + CTR_STRIPES * sizeof(CtrStripe);
__device__ __forceinline__ CtrStripe *ctr_stripe(DevCounters *c) {
    return reinterpret_cast<CtrStripe *>(c + 1) + (blockIdx.x % CTR_STRIPES);
}

// Scalars of the PageRank iteration, kept on the device so that an
// iteration needs no host round trip.
struct PrScalars {
    double sum_y;      // sum of pulled ranks
    double w;          // (1 - sum_y) / n
    double inv_s;      // 1 / sum(y + w)
    double delta;      // L1 residual of the last update
    double sum_x;      // sum of the normalised ranks
    float max_ratio;   // max over v of y[v]/deg[v] (for the FP16 scale)
    float enc_scale;   // contribution encode scale 2^k (FP16 only)
    float dec_scale;   // decode scale 2^-k of the stored contributions
    int cold_exp;      // E of the fixed-point unit 2^E of the binned sums
    double cold_unit;  // that unit, 2^E
};

// Static propagation-blocking layout of the PageRank edges whose source is
// at or above `from`: the sources outside the cache-resident prefix of a
// degree-ordered graph, whose gathers would each cost a cache line (all
// sources when from = 0).  Cold edges are grouped into tiles of consecutive
// sources (<= COLD_TILE_V sources, <= COLD_TILE_E edges) and bins of
// COLD_BIN destinations; the edges of one source into one bin form a pair,
// which carries one value.  The write kernel reads each tile's pairs in
// (bin, source) order and writes each pair's value to its slot in the
// bin-major value array, where a bin's pairs are contiguous (its runs, one
// per tile); the accumulation kernel streams each bin's entries (one per
// edge, grouped by pair) into LDS.  Bins with more than COLD_CHUNK_E entries
// are split into chunks whose integer partials a third kernel adds up.
constexpr uint32_t COLD_BIN = 16384;
constexpr uint32_t COLD_TILE_V = 32768;
constexpr uint64_t COLD_TILE_E = 1ull << 24;
constexpr uint64_t COLD_WCHUNK = 131072;   // pairs per write chunk
constexpr uint64_t COLD_CHUNK_E = 1u << 20;
constexpr uint64_t COLD_PAD_E = 65536;          // zero tail of the destination array (and of the pair bases / 64)
constexpr uint32_t COLD_DIRECT = 0xffffffffu;   // chunk slot of a bin in one chunk
constexpr uint32_t COLD_GROUP = 16;             // chunk partials summed per first-level combine group
struct ColdLayout {
    uint32_t from = 0;              // first cold source
    uint32_t ntiles = 0, nbins = 0;
    uint64_t ncold = 0;             // cold edges (entries)
    uint64_t npairs = 0;            // distinct (source, bin) pairs over the cold edges
    uint32_t *tile_lo = nullptr;    // ntiles+1 source boundaries
    // cells (t,b) with pairs are numbered in tile-major order (ordinals)
    uint32_t *roff = nullptr;       // per non-empty cell: bin-major pair slot - tile-major pair position
    uint32_t *tcell = nullptr;      // ntiles+1: ordinal of each tile's first non-empty cell
    uint16_t *wsrc = nullptr;       // per pair, tile-major: source - tile_lo[t], bit 15 = first pair of its cell
    uint32_t *wbase = nullptr;      // per 64 tile-major pairs: ordinal of the first pair's cell
    uint16_t *cdst = nullptr;       // per entry, bin-major: destination - b*COLD_BIN, bit 15 = first entry of a pair;
                                    // COLD_PAD_E zero entries follow
    uint32_t *bbase = nullptr;      // per 64 bin-major entries: the pair of the first entry (zero tail)
    uint32_t *vals = nullptr;       // per pair, bin-major: the staged contribution (4 or 2 bytes)
    uint32_t nwchunks = 0, nachunks = 0, nmulti = 0;
    uint32_t *wchunk = nullptr;     // per write chunk: tile, first pair, end pair (tile-major), first chunk of tile
    uint32_t *achunk = nullptr;     // per accumulation chunk: bin, first entry, end entry (bin-major), slot
    uint32_t *multi = nullptr;      // per bin in several chunks: bin, first slot, chunks, first group slot
    long long *partials = nullptr;  // COLD_BIN fixed-point sums per slot
    uint32_t ngroups = 0;
    uint32_t *groups = nullptr;     // per group of <= COLD_GROUP slots of one bin: first slot, slots, group slot, 0
    long long *partials2 = nullptr; // COLD_BIN fixed-point sums per group
    double *mass = nullptr;         // per write chunk: sum over its tile's edges of |contribution| (first chunk)
    // sources in no tile's span (no out-edges): ngaps ranges, gaps[0..ngaps)
    // their first sources, gaps[ngaps..2*ngaps] the prefix of their lengths
    uint32_t ngaps = 0;
    uint32_t *gaps = nullptr;
};

struct Profile {
    bool on = false;
    std::map<std::string, std::pair<double, uint64_t>> acc;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

} // namespace gpr

struct gpr_graph {
    uint32_t n = 0;
    uint64_t m = 0;
    int symmetric = 1;
    uint64_t *index = nullptr;
    uint32_t *nbrs = nullptr;
    uint64_t *cindex = nullptr;   // == index when symmetric
    uint32_t *cnbrs = nullptr;
    uint32_t *cindex32 = nullptr; // 32-bit copy of cindex for the pull when m < 2^32 (built on first use)
    uint32_t *deg = nullptr;
    // pull partitions over the CSC
    uint32_t nbuckets = 0;
    uint32_t *bucket = nullptr;     // nbuckets+1 row starts
    uint32_t nlong = 0;
    uint32_t *long_rows = nullptr;  // nlong ids, ascending
    uint32_t *long_base = nullptr;  // nlong+1 chunk prefix
    uint32_t nlong_chunks = 0;
    uint32_t hdv = 0;               // highest-degree vertex (original ID)
    // vertex relabelling: internal i <-> original order[i]; perm = inverse.
    // All IDs and per-vertex arrays crossing the C ABI are original.
    uint32_t *order = nullptr;
    uint32_t *perm = nullptr;
    hipStream_t stream = nullptr;
    // scratch
    gpr::DevCounters *ctr = nullptr;     // device
    gpr::DevCounters *hctr = nullptr;    // host mirror: fine-grained pinned memory the device writes
    gpr::DevCounters *hctr_dev = nullptr;  // its device address
    uint64_t *dedup = nullptr;           // n/64 words, all zero between steps
    uint64_t *scan = nullptr;            // frontier degree prefix (n+1)
    uint64_t *scan_blk = nullptr;        // per-scan-block totals
    double *partial = nullptr;           // per-workgroup partial sums
    size_t partial_cap = 0;
    float *partial_f = nullptr;          // per-workgroup partial maxima
    double *long_partial = nullptr;      // per long chunk
    double *dsum_slices = nullptr;       // 64 slice sums of deterministic_sum
    gpr::PrScalars *prs = nullptr;       // device
    // trace of the last run
    std::vector<gpr_iter_stats> trace;
    gpr::Profile prof;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;    // per edgemap step
    hipEvent_t rev0 = nullptr, rev1 = nullptr;  // per application run
    // recycled frontier buffers (avoid hipMalloc per iteration)
    std::vector<void *> pool_list, pool_bits;
    // learned policy state of the current run
    float *pol_params = nullptr;    // device copy of the network (null: none)
    float pol_margin = 0.f;         // gate margin loaded with the network
    std::vector<float> pol_host;    // host copy (first-step decision)
    bool pol_have = false;          // a device decision for the next step exists
    int pol_next = 0;
    int pol_iter = 0;
    double pol_unvisited = 0;
    // trace runs: step pol_force_step takes action pol_force_action (with
    // pol_persist, so does every later step) and every other step the
    // rule's (the loaded network's with pol_follow); with pol_log set, each
    // step's features are logged
    bool ctr_clean = true;           // counters zero (set by their read-back, cleared by a reset's kernels)
    bool pr_skip_unlinked = false;   // gpr_pr_iterate_fused with GPR_PR_SKIP_UNLINKED, for its call
    bool pr_skip_ranks = false;      // gpr_pr_iterate_fused with GPR_PR_SKIP_RANKS, for its call
    int pol_force_step = -1;
    int pol_force_action = 0;
    double pol_force_t0 = -1.0;   // host clock (ms) when the forced step began
    bool pol_log = false;
    bool pol_follow = false;
    bool pol_persist = false;
    std::vector<float> pol_feat_log;
    uint32_t max_deg = 0;
    gpr::ColdLayout *cold = nullptr;   // PageRank cold-edge layout (built on first use)
    void *fbuf = nullptr;              // fused push: level lists and second bitmap (built on first use)
    size_t fbuf_bytes = 0;
    void *fctl = nullptr;              // fused push: control block
    void *kbuf = nullptr;              // k-core peeling: candidate and frontier lists (built on first use)
    size_t kbuf_bytes = 0;
    void *kctl = nullptr;              // k-core peeling: control block
};

struct gpr_prop {
    gpr_graph *g = nullptr;
    int dtype = GPR_U32;
    size_t elem = 4;
    void *d = nullptr;
};

struct gpr_frontier {
    gpr_graph *g = nullptr;
    int kind = GPR_FRONTIER_EMPTY;
    uint32_t nactv = 0;
    uint64_t nacte = 0;
    uint32_t *list = nullptr;   // capacity n
    uint64_t *bits = nullptr;   // (n+63)/64 words
};

namespace gpr {

inline size_t nwords(uint32_t n) { return ((size_t)n + 63) / 64; }
inline unsigned grid_for(uint64_t work, unsigned per_block, unsigned cap = 0x7fffffffu) {
    uint64_t g = (work + per_block - 1) / per_block;
    if (g == 0) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

// Profiling helpers: record events around a launch on the handle's stream.
struct ProfScope {
    gpr_graph *g;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(gpr_graph *g_, const char *n_) : g(g_), name(n_) {
        if (g->prof.on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, g->stream);
        }
    }
    ~ProfScope() {
        if (g->prof.on) {
            (void)hipEventRecord(b, g->stream);
            g->prof.pending.push_back({name, {a, b}});
        }
    }
};
void prof_flush(gpr_graph *g);

// Fused sparse push (fusion.hip): process the frontier and everything it
// activates in one launch; `in` must be a list.
// In-place exclusive scan of per-workgroup counts (one workgroup, on g's stream).
void excl_scan_counts(gpr_graph *g, unsigned long long *counts, uint32_t count);
int fused_push(gpr_graph *g, int relax, uint32_t *prop, gpr_frontier *in, gpr_frontier *out, uint64_t dense_at,
               gpr_iter_stats *s);

// PageRank cold-edge layout (pr_cold.hip)
int cold_prepare(gpr_graph *g, uint32_t from);   // (re)build for this split point
void cold_free(gpr_graph *g);
int scan_frontier_degrees(gpr_graph *g, const uint32_t *list, uint32_t k);
// x == nullptr: stage the stored contributions; else compute them from y
// with the previous iteration's tail fused in (x' written, cold_tail_fusable)
int cold_launch_write(gpr_graph *g, int fmt, const void *contrib, const float *y, float *x, float damping);
bool cold_tail_fusable(const gpr_graph *g);
// y (+)= the binned sums; *nparts receives the number of y partial slots
int cold_launch_accum(gpr_graph *g, int fmt, float *y, int add, double *part, float *part_max, uint32_t *nparts);

// Host entry points shared between translation units.
int frontier_ensure_list(gpr_frontier *f);
int frontier_ensure_bits(gpr_frontier *f);
int frontier_to_bits(gpr_frontier *f);            // convert LIST/ALL -> BITMAP
int frontier_to_list(gpr_frontier *f);            // convert BITMAP/ALL -> LIST
int fill_words(gpr_graph *g, uint64_t *p, size_t words, uint64_t v);
int clear_dedup_of_list(gpr_graph *g, const uint32_t *list, uint64_t max_count);
int ids_download(gpr_graph *g, const gpr_prop *p, uint32_t *host);   // ID-valued property, original IDs
struct PolicyCtx;
// totals of the counters into g->hctr (and with pol, the policy's next
// choice into hctr->decision); synchronises
int read_counters(gpr_graph *g, const PolicyCtx *pol = nullptr);
int reset_counters(gpr_graph *g);
int ensure_partials(gpr_graph *g, size_t count);
int graph_setup(gpr_graph *g);                    // deg, buckets, long rows
// ID mapping between original (ABI) and internal (relabelled) vertex IDs.
int map_ids(gpr_graph *g, uint32_t *d_ids, uint32_t k, bool to_internal);
int to_internal_id(gpr_graph *g, uint32_t orig, uint32_t *internal);
int deterministic_sum(gpr_graph *g, const double *parts, uint32_t count,
                      double *dst);               // device-side, async
int batch_noalloc_setup(int dev);                 // scratch address of gpr_batch_cc_noalloc (batch.hip)
// push/pull switch (edgemap.hip): pct <= 0 is default_threshold (m/20),
// pct > 0 frac_threshold(pct)
bool threshold_sparse(float pct, uint64_t nactv, uint64_t nacte, uint64_t m);
uint64_t threshold_limit(float pct, uint64_t m);

} // namespace gpr
