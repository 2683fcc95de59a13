// runtime.hip -- errors, device selection, allocation, counters, profiling.
#include "internal.hpp"
#include "device.hpp"
#include "grid_sync.hpp"
#include "policy_dev.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

namespace gpr {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int dev_alloc(void **p, size_t bytes, const char *what) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) for %s failed: %s", bytes, what, hipGetErrorString(e));
        *p = nullptr;
        return GPR_E_OOM;
    }
    return GPR_OK;
}

void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

__global__ void k_reset_counters(DevCounters *c) {
    if (threadIdx.x == 0) {
        c->nacte = 0;
        c->nactv = 0;
        c->edges = 0;
        c->listpos = 0;
    }
    if (threadIdx.x < CTR_STRIPES) {
        CtrStripe *st = reinterpret_cast<CtrStripe *>(c + 1) + threadIdx.x;
        st->nacte = 0;
        st->edges = 0;
        st->nactv = 0;
    }
}

// The read-back below leaves the counters zero, so a reset right after it
// (the usual order: read at the end of one step, reset at the start of the
// next) launches nothing.  g->ctr_clean: no kernel has touched the
// counters since they were last zeroed.
int reset_counters(gpr_graph *g) {
    if (g->ctr_clean) {
        g->ctr_clean = false;   // the caller's kernels follow
        return GPR_OK;
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, g->stream, g->ctr);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

// The frontier statistics' reduction: one wave sums the stripes (lane i
// reads stripe i) and writes the totals and the list position to the host
// mirror, and zeroes what it read for the next step.  With a network (pol.params), the same wave then runs the
// learned policy on the totals (policy_dev.hpp), its weights loaded beside
// the stripes so that both arrive together, and writes the next step's
// choice: the inference costs the step no launch, no host work and nothing
// in the step's own kernels.
__global__ void __launch_bounds__(64) k_read_counters(DevCounters *c, DevCounters *h, PolicyCtx pol) {
    MlpWave net;
    if (pol.params) net.load(pol.params);
    CtrStripe *st = reinterpret_cast<CtrStripe *>(c + 1) + threadIdx.x;
    const unsigned long long e0 = st->nacte, e1 = st->edges;
    const unsigned int v0 = st->nactv;
    const unsigned int lp = c->listpos;
    st->nacte = 0;
    st->edges = 0;
    st->nactv = 0;
    const unsigned long long ne = wave_sum(e0), ed = wave_sum(e1);
    const unsigned int nv = wave_sum(v0);
    unsigned int decision = 0;
    if (pol.params) {
        float f[6];
        policy_features(pol, nv, ne, f);
        decision = (unsigned)net.run(f, pol.margin);
    }
    if (threadIdx.x == 0) {
        h->nacte = ne;
        h->edges = ed;
        h->nactv = nv;
        h->listpos = lp;
        h->decision = decision;
        c->listpos = 0;
    }
}

int read_counters(gpr_graph *g, const PolicyCtx *pol) {
    static_assert(CTR_STRIPES == 64, "one stripe per lane of the read-back wave");
    const PolicyCtx none = {};
    hipLaunchKernelGGL(k_read_counters, dim3(1), dim3(64), 0, g->stream, g->ctr, g->hctr_dev, pol ? *pol : none);
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipStreamSynchronize(g->stream));
    g->ctr_clean = true;
    return GPR_OK;
}

int ensure_partials(gpr_graph *g, size_t count) {
    if (count <= g->partial_cap) return GPR_OK;
    dev_free(g->partial);
    dev_free(g->partial_f);
    g->partial = nullptr;
    g->partial_f = nullptr;
    size_t cap = count + count / 4 + 1024;
    GPR_TRY(dev_alloc((void **)&g->partial, cap * sizeof(double), "partials"));
    GPR_TRY(dev_alloc((void **)&g->partial_f, cap * sizeof(float), "partials"));
    g->partial_cap = cap;
    return GPR_OK;
}

// Deterministic FP64 sum of `count` partials: every workgroup folds a fixed
// slice in a fixed order, the last workgroup (by ticket) folds the slice
// sums in index order.  The result does not depend on scheduling.
constexpr int DSUM_THREADS = 256;
constexpr int DSUM_BLOCKS = 64;
__global__ void __launch_bounds__(DSUM_THREADS)
k_deterministic_sum(const double *__restrict__ parts, uint32_t count, double *slices,
                    unsigned int *ticket, double *dst) {
    __shared__ double red[DSUM_THREADS];
    __shared__ bool last;
    uint32_t per = (count + gridDim.x - 1) / gridDim.x;
    uint32_t lo = blockIdx.x * per, hi = min(count, lo + per);
    double s = 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) s += parts[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    // the slice as an agent-scope store, acknowledged before the ticket; the
    // last workgroup reads the slices with agent-scope loads (no fence)
    if (threadIdx.x == 0) {
        __hip_atomic_store(&slices[blockIdx.x], red[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(WAIT_VM0);
        unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (unsigned i = 0; i < gridDim.x; ++i)
            tot += __hip_atomic_load(&slices[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *dst = tot;
        *ticket = 0;
    }
}

int deterministic_sum(gpr_graph *g, const double *parts, uint32_t count, double *dst) {
    // slices live after the partials: reserve DSUM_BLOCKS doubles at the end
    double *slices = g->dsum_slices;
    hipLaunchKernelGGL(k_deterministic_sum, dim3(DSUM_BLOCKS), dim3(DSUM_THREADS), 0, g->stream,
                       parts, count, slices, &g->ctr->ticket2, dst);
    GPR_HIP(hipGetLastError());
    return GPR_OK;
}

Tunables &tun() {
    static Tunables t;
    return t;
}

int num_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

std::vector<gpr_iter_stats> &trace_of(gpr_graph *g) { return g->trace; }
void trace_clear(gpr_graph *g) { g->trace.clear(); }

// Device time of an application run: events on the handle's stream at the
// start and end of run() (the reference times run() only, driver.C:245-255).
double stream_time_ms(gpr_graph *g, bool start) {
    if (!g->rev0) {
        hipEventCreate(&g->rev0);
        hipEventCreate(&g->rev1);
    }
    if (start) {
        hipEventRecord(g->rev0, g->stream);
        return 0.0;
    }
    hipEventRecord(g->rev1, g->stream);
    hipEventSynchronize(g->rev1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, g->rev0, g->rev1);
    return ms;
}

void prof_flush(gpr_graph *g) {
    if (g->prof.pending.empty()) return;
    hipStreamSynchronize(g->stream);
    for (auto &p : g->prof.pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.second.first, p.second.second);
        auto &a = g->prof.acc[p.first];
        a.first += ms;
        a.second += 1;
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    g->prof.pending.clear();
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return GPR_E_INVAL;
    snprintf(buf, len, "%s", g_err);
    return GPR_OK;
}

extern "C" const char *gpr_version(void) { return "gpr 0.1 gfx950"; }

extern "C" int gpr_init(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= device) {
        set_error("no HIP device %d (count %d): %s", device, count, hipGetErrorString(e));
        return GPR_E_NODEV;
    }
    hipDeviceProp_t p;
    GPR_HIP(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s, this library is built for gfx950", device, p.gcnArchName);
        return GPR_E_NODEV;
    }
    GPR_HIP(hipSetDevice(device));
    return batch_noalloc_setup(device);
}

static double *tunable(const char *name) {
    Tunables &t = tun();
    if (!strcmp(name, "pull_serial_max")) return &t.pull_serial_max;
    if (!strcmp(name, "pull_blocks_per_cu")) return &t.pull_blocks_per_cu;
    if (!strcmp(name, "push_blocks_per_cu")) return &t.push_blocks_per_cu;
    if (!strcmp(name, "pr_split_bytes")) return &t.pr_split_bytes;
    if (!strcmp(name, "fusion_spin_limit")) return &t.fusion_spin_limit;
    if (!strcmp(name, "fusion_blocks")) return &t.fusion_blocks;
    if (!strcmp(name, "pr_write_stage_cells")) return &t.pr_write_stage_cells;
    if (!strcmp(name, "fusion_small_limit")) return &t.fusion_small_limit;
    if (!strcmp(name, "fusion_cooperative")) return &t.fusion_cooperative;
    if (!strcmp(name, "push_bitmap_frac")) return &t.push_bitmap_frac;
    return nullptr;
}

extern "C" int gpr_tune_set(const char *name, double value) {
    double *p = name ? tunable(name) : nullptr;
    if (!p) {
        set_error("unknown tunable '%s'", name ? name : "(null)");
        return GPR_E_INVAL;
    }
    if (!(value >= 0) || value > 1e12) {
        set_error("tunable %s: value %g out of range", name, value);
        return GPR_E_INVAL;
    }
    *p = value;
    return GPR_OK;
}

extern "C" int gpr_tune_get(const char *name, double *value) {
    double *p = name ? tunable(name) : nullptr;
    if (!p || !value) {
        set_error("unknown tunable '%s'", name ? name : "(null)");
        return GPR_E_INVAL;
    }
    *value = *p;
    return GPR_OK;
}

extern "C" int gpr_profile_enable(gpr_graph *g, int on) {
    if (!g) return GPR_E_INVAL;
    prof_flush(g);
    g->prof.on = on != 0;
    return GPR_OK;
}

extern "C" int gpr_profile_get(const gpr_graph *gc, const char *name, double *ms, uint64_t *launches) {
    if (!gc || !name) return GPR_E_INVAL;
    gpr_graph *g = const_cast<gpr_graph *>(gc);
    prof_flush(g);
    auto it = g->prof.acc.find(name);
    if (ms) *ms = it == g->prof.acc.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == g->prof.acc.end() ? 0 : it->second.second;
    return GPR_OK;
}

extern "C" int gpr_profile_reset(gpr_graph *g) {
    if (!g) return GPR_E_INVAL;
    prof_flush(g);
    g->prof.acc.clear();
    return GPR_OK;
}

extern "C" int gpr_trace_get(const gpr_graph *g, gpr_iter_stats *out, int cap, int *k) {
    if (!g || !k) return GPR_E_INVAL;
    int cnt = (int)g->trace.size();
    *k = cnt;
    if (out)
        for (int i = 0; i < cnt && i < cap; ++i) out[i] = g->trace[i];
    return GPR_OK;
}
