// policy.hip -- the per-iteration choice of traversal direction, output
// frontier representation and fusion.
//
// Fallback and baseline: the reference's threshold rule (sparse push iff
// nactv + nacte <= m/20, api.h:150-153; a list output for push and a dense
// output for pull, dsl/edgemap.h:1571-1601) with the benchmark's own
// fusion setting.  Learned policy: a 6-32-32-4 MLP (policy_dev.hpp) whose
// forward pass runs in the one-wave reduction of each push/pull step's
// frontier counts; the host reads its choice together with the counts it
// already reads (after a fused step, which has no such reduction, the host
// evaluates it).  Its four actions are the rule's own step, an unfused
// push, a push fused while its levels stay small, and a pull with a list
// output; another action replaces the rule's only with a logit lead of at
// least the loaded margin.  Outputs are identical whichever action is
// taken (the relax is a min, the frontier a set); only the time differs.
//
// Training data: gpr_policy_trace runs the BFSLVLv or CCv benchmark once
// under the rule (logging each step's features), then again with one step
// forced to each action, at evenly spread steps; the time from the start
// of that step to the end of the run says what each action costs.  Training:
// full-batch Adam on the GPU with this file's forward/backward kernels
// (per-sample soft targets and weights from the timings).
#include "internal.hpp"
#include "device.hpp"
#include "policy_dev.hpp"

#include <math.h>
#include <string.h>

#include <cmath>

#include <algorithm>
#include <chrono>
#include <vector>

namespace gpr {

// ---- host mirror of the feature map and the forward pass --------------------
static void host_features(const gpr_graph *g, uint32_t nactv, uint64_t nacte, double unvisited, int iter, float f[6]) {
    const double mean = g->n ? (double)g->m / (double)g->n : 0.0;
    const float lm = log2f(1.0f + (float)g->m);
    f[0] = log2f(1.0f + (float)nactv) / 32.0f;
    f[1] = 1.0f + (log2f(1.0f + (float)nacte) - lm) / 32.0f;
    f[2] = 1.0f + ((float)log2(1.0 + (unvisited > 0 ? unvisited : 0.0)) - lm) / 32.0f;
    f[3] = (float)iter / 64.0f;
    f[4] = (float)log2(1.0 + mean) / 8.0f;
    f[5] = (float)log2(1.0 + (mean > 0 ? g->max_deg / mean : 0.0)) / 24.0f;
}

static int host_forward(const float *p, float margin, const float f[6]) {
    const MlpView P(p);
    float h1[MLP_H], h2[MLP_H], z[MLP_OUT];
    for (int j = 0; j < MLP_H; ++j) {
        float a = P.b1[j];
        for (int i = 0; i < MLP_IN; ++i) a += f[i] * P.w1[i * MLP_H + j];
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < MLP_H; ++j) {
        float a = P.b2[j];
        for (int i = 0; i < MLP_H; ++i) a += h1[i] * P.w2[i * MLP_H + j];
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int k = 0; k < MLP_OUT; ++k) {
        float a = P.b3[k];
        for (int i = 0; i < MLP_H; ++i) a += h2[i] * P.w3[i * MLP_OUT + k];
        z[k] = a;
    }
    return policy_gate(z, margin);
}

PolicyCtx policy_ctx(gpr_graph *g, uint64_t nacte_in) {
    PolicyCtx c = {};
    c.params = g->pol_params;
    c.margin = g->pol_margin;
    const double mean = g->n ? (double)g->m / (double)g->n : 0.0;
    c.mean_deg_feat = (float)log2(1.0 + mean) / 8.0f;
    c.skew_feat = (float)log2(1.0 + (mean > 0 ? g->max_deg / mean : 0.0)) / 24.0f;
    c.iter_feat = (float)(g->pol_iter + 1) / 64.0f;
    c.log_m = log2f(1.0f + (float)g->m);
    c.unvisited_before = g->pol_unvisited - (double)nacte_in;
    return c;
}

// The action of the current step (a network is loaded): the choice the
// device made at the end of the previous step, or the host's own for the
// first step and after a fused step.
int policy_action(gpr_graph *g, uint32_t nactv, uint64_t nacte) {
    if (g->pol_have) return g->pol_next;
    float f[6];
    host_features(g, nactv, nacte, g->pol_unvisited - (double)nacte, g->pol_iter, f);
    return host_forward(g->pol_host.data(), g->pol_margin, f);
}

// Monotonic host clock in milliseconds.
double host_clock_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Trace runs: the features of the step about to run.
void policy_log_features(gpr_graph *g, uint32_t nactv, uint64_t nacte) {
    float f[6];
    host_features(g, nactv, nacte, g->pol_unvisited - (double)nacte, g->pol_iter, f);
    g->pol_feat_log.insert(g->pol_feat_log.end(), f, f + 6);
}

// Book-keeping after a step run under the policy; decision < 0: none was
// made on the device.
void policy_after_step(gpr_graph *g, uint64_t nacte_in, int decision) {
    g->pol_unvisited -= (double)nacte_in;
    if (g->pol_unvisited < 0) g->pol_unvisited = 0;
    g->pol_iter++;
    g->pol_next = decision < 0 ? 0 : decision;
    g->pol_have = decision >= 0;
}

// ---- MLP training kernels ------------------------------------------------------------
constexpr int MB = 64;   // samples per workgroup

// Forward + backward of MB samples per workgroup.  Loss per sample:
// w * sum_k -t_k log softmax(z)_k.  Writes the workgroup's parameter
// gradient (sum over its samples) and loss to gpart/lpart.
__global__ void __launch_bounds__(MB)
k_mlp_fwd_bwd(const float *__restrict__ params, const float *__restrict__ X, const float *__restrict__ T,
              const float *__restrict__ Wt, int count, float *gpart, float *lpart, float *logits) {
    __shared__ float sx[MB][MLP_IN];
    __shared__ float sh1[MB][MLP_H], sh2[MB][MLP_H], sd1[MB][MLP_H], sd2[MB][MLP_H];
    __shared__ float sdz[MB][MLP_OUT];
    __shared__ float sloss[MB];
    const MlpView P(params);
    const int t = threadIdx.x;
    const int s = blockIdx.x * MB + t;
    const bool valid = s < count;
    float x[MLP_IN], h1[MLP_H], h2[MLP_H], z[MLP_OUT];
    for (int i = 0; i < MLP_IN; ++i) x[i] = valid ? X[(size_t)s * MLP_IN + i] : 0.f;
    for (int j = 0; j < MLP_H; ++j) {
        float a = P.b1[j];
        for (int i = 0; i < MLP_IN; ++i) a += x[i] * P.w1[i * MLP_H + j];
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < MLP_H; ++j) {
        float a = P.b2[j];
        for (int i = 0; i < MLP_H; ++i) a += h1[i] * P.w2[i * MLP_H + j];
        h2[j] = a > 0.f ? a : 0.f;
    }
    float zmax = -3.0e38f;
    for (int k = 0; k < MLP_OUT; ++k) {
        float a = P.b3[k];
        for (int i = 0; i < MLP_H; ++i) a += h2[i] * P.w3[i * MLP_OUT + k];
        z[k] = a;
        zmax = fmaxf(zmax, a);
        if (logits && valid) logits[(size_t)s * MLP_OUT + k] = a;
    }
    float se = 0.f;
    for (int k = 0; k < MLP_OUT; ++k) se += expf(z[k] - zmax);
    const float lse = zmax + logf(se);
    const float w = valid && Wt ? Wt[s] : (valid ? 1.f : 0.f);
    float loss = 0.f, dz[MLP_OUT];
    for (int k = 0; k < MLP_OUT; ++k) {
        const float tk = valid && T ? T[(size_t)s * MLP_OUT + k] : 0.f;
        loss += -tk * (z[k] - lse);
        dz[k] = w * (expf(z[k] - lse) - tk);   // targets sum to 1
    }
    loss *= w;
    float d2[MLP_H], d1[MLP_H];
    for (int j = 0; j < MLP_H; ++j) {
        float a = 0.f;
        for (int k = 0; k < MLP_OUT; ++k) a += dz[k] * P.w3[j * MLP_OUT + k];
        d2[j] = h2[j] > 0.f ? a : 0.f;
    }
    for (int i = 0; i < MLP_H; ++i) {
        float a = 0.f;
        for (int j = 0; j < MLP_H; ++j) a += d2[j] * P.w2[i * MLP_H + j];
        d1[i] = h1[i] > 0.f ? a : 0.f;
    }
    for (int i = 0; i < MLP_IN; ++i) sx[t][i] = x[i];
    for (int j = 0; j < MLP_H; ++j) {
        sh1[t][j] = h1[j];
        sh2[t][j] = h2[j];
        sd1[t][j] = d1[j];
        sd2[t][j] = d2[j];
    }
    for (int k = 0; k < MLP_OUT; ++k) sdz[t][k] = dz[k];
    sloss[t] = loss;
    __syncthreads();
    // parameter gradients: each entry a sum over the workgroup's samples
    float *gout = gpart + (size_t)blockIdx.x * MLP_PARAMS;
    for (int p = t; p < MLP_PARAMS; p += MB) {
        float acc = 0.f;
        int q = p;
        if (q < MLP_IN * MLP_H) {                         // W1[i][j]
            const int i = q / MLP_H, j = q % MLP_H;
            for (int r = 0; r < MB; ++r) acc += sx[r][i] * sd1[r][j];
        } else if ((q -= MLP_IN * MLP_H) < MLP_H) {       // b1[j]
            for (int r = 0; r < MB; ++r) acc += sd1[r][q];
        } else if ((q -= MLP_H) < MLP_H * MLP_H) {        // W2[i][j]
            const int i = q / MLP_H, j = q % MLP_H;
            for (int r = 0; r < MB; ++r) acc += sh1[r][i] * sd2[r][j];
        } else if ((q -= MLP_H * MLP_H) < MLP_H) {        // b2[j]
            for (int r = 0; r < MB; ++r) acc += sd2[r][q];
        } else if ((q -= MLP_H) < MLP_H * MLP_OUT) {      // W3[i][k]
            const int i = q / MLP_OUT, k = q % MLP_OUT;
            for (int r = 0; r < MB; ++r) acc += sh2[r][i] * sdz[r][k];
        } else {                                          // b3[k]
            q -= MLP_H * MLP_OUT;
            for (int r = 0; r < MB; ++r) acc += sdz[r][q];
        }
        gout[p] = acc;
    }
    if (t == 0) {
        float l = 0.f;
        for (int r = 0; r < MB; ++r) l += sloss[r];
        lpart[blockIdx.x] = l;
    }
}

// Sum the workgroup gradients in order (deterministic), divide by the total
// weight; mode 0: write the gradient; mode 1: Adam step.
__global__ void __launch_bounds__(256)
k_mlp_reduce(const float *__restrict__ gpart, const float *__restrict__ lpart, int nblocks, float wsum,
             float *params, float *grad, float *mom, float *vel, float lr, int step, int mode, float *loss) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int p = threadIdx.x; p < MLP_PARAMS; p += blockDim.x) {
        float g = 0.f;
        for (int b = 0; b < nblocks; ++b) g += gpart[(size_t)b * MLP_PARAMS + p];
        g /= wsum;
        if (mode == 0) {
            grad[p] = g;
        } else {
            const float m = b1 * mom[p] + (1.f - b1) * g;
            const float v = b2 * vel[p] + (1.f - b2) * g * g;
            mom[p] = m;
            vel[p] = v;
            const float mh = m / (1.f - powf(b1, (float)step));
            const float vh = v / (1.f - powf(b2, (float)step));
            params[p] -= lr * mh / (sqrtf(vh) + eps);
        }
    }
    if (threadIdx.x == 0) {
        float l = 0.f;
        for (int b = 0; b < nblocks; ++b) l += lpart[b];
        *loss = l / wsum;
    }
}

static hipStream_t mlp_stream() {
    static hipStream_t s = nullptr;
    if (!s) (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    return s;
}

struct MlpBuffers {
    float *p = nullptr, *x = nullptr, *t = nullptr, *w = nullptr, *gpart = nullptr, *lpart = nullptr;
    float *grad = nullptr, *mom = nullptr, *vel = nullptr, *loss = nullptr, *logits = nullptr;
    ~MlpBuffers() {
        for (float *q : {p, x, t, w, gpart, lpart, grad, mom, vel, loss, logits}) dev_free(q);
    }
};

static int mlp_upload(MlpBuffers &B, const float *params, const float *X, const float *T, const float *W, int count) {
    const int nb = (count + MB - 1) / MB;
    GPR_TRY(dev_alloc((void **)&B.p, 4ull * MLP_PARAMS, "mlp params"));
    GPR_TRY(dev_alloc((void **)&B.x, 4ull * MLP_IN * std::max(count, 1), "mlp inputs"));
    GPR_TRY(dev_alloc((void **)&B.gpart, 4ull * MLP_PARAMS * std::max(nb, 1), "mlp gradients"));
    GPR_TRY(dev_alloc((void **)&B.lpart, 4ull * std::max(nb, 1), "mlp losses"));
    GPR_TRY(dev_alloc((void **)&B.grad, 4ull * MLP_PARAMS, "mlp gradient"));
    GPR_TRY(dev_alloc((void **)&B.loss, 4, "mlp loss"));
    hipStream_t st = mlp_stream();
    GPR_HIP(hipMemcpyAsync(B.p, params, 4ull * MLP_PARAMS, hipMemcpyHostToDevice, st));
    GPR_HIP(hipMemcpyAsync(B.x, X, 4ull * MLP_IN * count, hipMemcpyHostToDevice, st));
    if (T) {
        GPR_TRY(dev_alloc((void **)&B.t, 4ull * MLP_OUT * std::max(count, 1), "mlp targets"));
        GPR_HIP(hipMemcpyAsync(B.t, T, 4ull * MLP_OUT * count, hipMemcpyHostToDevice, st));
    }
    if (W) {
        GPR_TRY(dev_alloc((void **)&B.w, 4ull * std::max(count, 1), "mlp weights"));
        GPR_HIP(hipMemcpyAsync(B.w, W, 4ull * count, hipMemcpyHostToDevice, st));
    }
    return GPR_OK;
}

} // namespace gpr

using namespace gpr;

extern "C" int gpr_policy_set(gpr_graph *g, const float *params, int count) {
    if (!g) return GPR_E_INVAL;
    if (!params) {
        dev_free(g->pol_params);
        g->pol_params = nullptr;
        g->pol_host.clear();
        g->pol_margin = 0.f;
        return GPR_OK;
    }
    GPR_REQUIRE(count == MLP_PARAMS || count == MLP_PARAMS + 1,
                "policy: expected %d parameters (or one more, the gate margin), got %d", MLP_PARAMS, count);
    for (int i = 0; i < count; ++i) GPR_REQUIRE(std::isfinite(params[i]), "policy: parameter %d is not finite", i);
    const float margin = count > MLP_PARAMS ? params[MLP_PARAMS] : 0.f;
    GPR_REQUIRE(margin >= 0.f, "policy: the gate margin must not be negative");
    if (!g->pol_params) GPR_TRY(dev_alloc((void **)&g->pol_params, 4ull * MLP_PARAMS, "policy"));
    GPR_HIP(hipMemcpy(g->pol_params, params, 4ull * MLP_PARAMS, hipMemcpyHostToDevice));
    g->pol_host.assign(params, params + MLP_PARAMS);
    g->pol_margin = margin;
    return GPR_OK;
}

extern "C" int gpr_policy_reset(gpr_graph *g) {
    if (!g) return GPR_E_INVAL;
    g->pol_have = false;
    g->pol_iter = 0;
    g->pol_unvisited = (double)g->m;
    return GPR_OK;
}

extern "C" int gpr_mlp_param_count(void) { return MLP_PARAMS; }

extern "C" int gpr_mlp_forward(const float *params, const float *x, int count, float *logits) {
    if (!params || !x || !logits || count <= 0) return GPR_E_INVAL;
    MlpBuffers B;
    GPR_TRY(mlp_upload(B, params, x, nullptr, nullptr, count));
    GPR_TRY(dev_alloc((void **)&B.logits, 4ull * MLP_OUT * count, "mlp logits"));
    const int nb = (count + MB - 1) / MB;
    hipStream_t st = mlp_stream();
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(MB), 0, st, B.p, B.x, (const float *)nullptr,
                       (const float *)nullptr, count, B.gpart, B.lpart, B.logits);
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(logits, B.logits, 4ull * MLP_OUT * count, hipMemcpyDeviceToHost, st));
    GPR_HIP(hipStreamSynchronize(st));
    return GPR_OK;
}

extern "C" int gpr_mlp_grad(const float *params, const float *x, const float *targets, const float *weights,
                            int count, float *grad, float *loss) {
    if (!params || !x || !targets || !grad || count <= 0) return GPR_E_INVAL;
    MlpBuffers B;
    GPR_TRY(mlp_upload(B, params, x, targets, weights, count));
    const int nb = (count + MB - 1) / MB;
    float wsum = 0.f;
    for (int i = 0; i < count; ++i) wsum += weights ? weights[i] : 1.f;
    GPR_REQUIRE(wsum > 0.f, "mlp: weights sum to zero");
    hipStream_t st = mlp_stream();
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(MB), 0, st, B.p, B.x, B.t, B.w, count, B.gpart, B.lpart,
                       (float *)nullptr);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(1), dim3(256), 0, st, B.gpart, B.lpart, nb, wsum, B.p, B.grad,
                       (float *)nullptr, (float *)nullptr, 0.f, 0, 0, B.loss);
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(grad, B.grad, 4ull * MLP_PARAMS, hipMemcpyDeviceToHost, st));
    float l = 0.f;
    GPR_HIP(hipMemcpyAsync(&l, B.loss, 4, hipMemcpyDeviceToHost, st));
    GPR_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l;
    return GPR_OK;
}

extern "C" int gpr_mlp_train(float *params, const float *x, const float *targets, const float *weights, int count,
                             int epochs, float lr, float *loss) {
    if (!params || !x || !targets || count <= 0 || epochs <= 0) return GPR_E_INVAL;
    MlpBuffers B;
    GPR_TRY(mlp_upload(B, params, x, targets, weights, count));
    GPR_TRY(dev_alloc((void **)&B.mom, 4ull * MLP_PARAMS, "adam"));
    GPR_TRY(dev_alloc((void **)&B.vel, 4ull * MLP_PARAMS, "adam"));
    hipStream_t st = mlp_stream();
    GPR_HIP(hipMemsetAsync(B.mom, 0, 4ull * MLP_PARAMS, st));
    GPR_HIP(hipMemsetAsync(B.vel, 0, 4ull * MLP_PARAMS, st));
    const int nb = (count + MB - 1) / MB;
    float wsum = 0.f;
    for (int i = 0; i < count; ++i) wsum += weights ? weights[i] : 1.f;
    GPR_REQUIRE(wsum > 0.f, "mlp: weights sum to zero");
    for (int e = 1; e <= epochs; ++e) {
        hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(MB), 0, st, B.p, B.x, B.t, B.w, count, B.gpart, B.lpart,
                           (float *)nullptr);
        hipLaunchKernelGGL(k_mlp_reduce, dim3(1), dim3(256), 0, st, B.gpart, B.lpart, nb, wsum, B.p, B.grad, B.mom,
                           B.vel, lr, e, 1, B.loss);
    }
    GPR_HIP(hipGetLastError());
    GPR_HIP(hipMemcpyAsync(params, B.p, 4ull * MLP_PARAMS, hipMemcpyDeviceToHost, st));
    float l = 0.f;
    GPR_HIP(hipMemcpyAsync(&l, B.loss, 4, hipMemcpyDeviceToHost, st));
    GPR_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l;
    return GPR_OK;
}

// Trace generation: the benchmarks themselves (algo 0 = BFSLVLv from
// `start`, 1 = CCv), each with its default settings.  A first run under
// the rule (with GPR_TRACE_FOLLOW, under the loaded network) logs every
// step's features and counts the steps; then, at up to cap evenly spread
// steps i, the run is repeated with step i forced to each action (with
// GPR_TRACE_PERSIST, step i and every later step) and every other step the
// rule's (the network's).  The runs share the steps before i, so
// times[k][a] (the best of three runs' host-clock time from the start of
// step i to the end) differ by what action a from step i on costs.
extern "C" int gpr_bfslvl(gpr_graph *g, uint32_t start, const gpr_run_opts *o, uint32_t *levels, gpr_run_info *info);
extern "C" int gpr_ccv(gpr_graph *g, const gpr_run_opts *o, uint32_t *labels, gpr_run_info *info);

extern "C" int gpr_policy_trace(gpr_graph *g, int algo, uint32_t start, int cap, float *features, float *times,
                                int *k) {
    const bool follow = (algo & GPR_TRACE_FOLLOW) != 0;
    const bool persist = (algo & GPR_TRACE_PERSIST) != 0;
    algo &= ~(GPR_TRACE_FOLLOW | GPR_TRACE_PERSIST);
    if (!g || !k || cap <= 0 || (algo != 0 && algo != 1)) return GPR_E_INVAL;
    GPR_REQUIRE(algo == 1 || start < g->n, "trace: start out of range");
    GPR_REQUIRE(!follow || !g->pol_host.empty(), "trace: GPR_TRACE_FOLLOW without a loaded network");
    *k = 0;
    float *saved = g->pol_params;
    g->pol_params = nullptr;
    g->pol_follow = follow;
    g->pol_persist = persist;
    gpr_run_opts o = {};
    o.direction = GPR_DIR_POLICY;
    o.max_iter = 100;
    o.tol = 1e-7;
    // one run with step `step` forced to `action`; the best of `reps` times
    // from the start of that step to the end of the run (host clock: what
    // the step and the steps after it cost, launches and waits included,
    // without the noise of the steps before it), or of the whole run when
    // the run ends before that step
    auto run = [&](int step, int action, int reps, double *ms) -> int {
        g->pol_force_step = step;
        g->pol_force_action = action;
        *ms = 1e30;
        for (int rep = 0; rep < reps; ++rep) {
            gpr_run_info info = {};
            g->pol_force_t0 = -1.0;
            const double t_start = host_clock_ms();
            const int rc = algo == 0 ? gpr_bfslvl(g, start, &o, nullptr, &info) : gpr_ccv(g, &o, nullptr, &info);
            if (rc != GPR_OK) return rc;
            const double t_end = host_clock_ms();
            *ms = std::min(*ms, t_end - (g->pol_force_t0 >= 0 ? g->pol_force_t0 : t_start));
        }
        return GPR_OK;
    };
    int r = GPR_OK;
    do {
        // the unforced run: features of every step
        double t0 = 0;
        g->pol_feat_log.clear();
        g->pol_log = true;
        r = run(0x7fffffff, 0, 1, &t0);
        g->pol_log = false;
        if (r) break;
        const int steps = (int)(g->pol_feat_log.size() / 6);
        const std::vector<float> feats = g->pol_feat_log;
        const int take = std::min(steps, cap);
        for (int j = 0; j < take && !r; ++j) {
            // evenly spread, first and last step included
            const int i = take == 1 ? 0 : (int)((int64_t)j * (steps - 1) / (take - 1));
            for (int a = 0; a < 4 && !r; ++a) {
                double ms = 0;
                if ((r = run(i, a, 3, &ms))) break;
                if (times) times[j * 4 + a] = (float)ms;
            }
            if (features)
                for (int q = 0; q < 6; ++q) features[j * 6 + q] = feats[(size_t)i * 6 + q];
            *k = j + 1;
        }
    } while (0);
    g->pol_force_step = -1;
    g->pol_force_action = 0;
    g->pol_log = false;
    g->pol_follow = false;
    g->pol_persist = false;
    g->pol_feat_log.clear();
    g->pol_params = saved;
    return r;
}
