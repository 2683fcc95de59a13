// policy_dev.hpp -- the direction/representation policy network on the
// device: a 6-32-32-4 MLP (ReLU, FP32) whose forward pass runs in the
// one-wave kernel that reduces a step's frontier statistics (the striped
// counters, runtime.hip k_read_counters), right after their totals, so the
// next step's choice costs no launch of its own.
//
// Features (all of the frontier just produced; the edge counts relative
// to m, as the threshold rule sees them, so that the features of a state
// do not move with the size of the graph):
//   f0 log2(1 + nactv) / 32                        frontier size
//   f1 1 + log2((1 + nacte) / (1 + m)) / 32        frontier out-edge sum
//   f2 1 + log2((1 + unvisited) / (1 + m)) / 32    edges of vertices not yet settled
//   f3 iteration / 64
//   f4 log2(1 + m/n) / 8                           mean degree
//   f5 log2(1 + maxdeg/(m/n)) / 24                 degree skew
// Outputs, the next step's action (argmax, gated):
//   0 the threshold rule's own step (the baseline: direction by the rule,
//     the caller's fusion setting, the default output representation)
//   1 an unfused sparse push (where the rule pulls, only while the
//     frontier's edges stay within twice the rule's threshold)
//   2 a push that fuses while its levels stay small (GPR_FUSION_SMALL)
//   3 a pull that writes a list output, where the rule pulls (elsewhere
//     the rule's step: a pull costs in proportion to n, which the
//     features do not carry)
// Gate: another action replaces the rule's only when its logit exceeds the
// rule's by at least the margin loaded with the network (calibrated on
// held-back traces), so an uncertain network takes the rule's step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpr {

constexpr int MLP_IN = 6, MLP_H = 32, MLP_OUT = 4;
constexpr int MLP_PARAMS = MLP_IN * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_OUT + MLP_OUT;  // 1412

// parameter layout: W1[in][h], b1[h], W2[h][h], b2[h], W3[h][out], b3[out]
struct MlpView {
    const float *w1, *b1, *w2, *b2, *w3, *b3;
    __host__ __device__ explicit MlpView(const float *p)
        : w1(p), b1(p + MLP_IN * MLP_H), w2(b1 + MLP_H), b2(w2 + MLP_H * MLP_H), w3(b2 + MLP_H),
          b3(w3 + MLP_H * MLP_OUT) {}
};

// Constants of the graph the features need, filled on the host.
struct PolicyCtx {
    const float *params;      // device, MLP_PARAMS floats (null: no policy)
    float margin;             // gate: logit lead an action needs over action 0
    float mean_deg_feat;      // f4
    float skew_feat;          // f5
    float iter_feat;          // f3 for the NEXT step
    float log_m;              // log2(1 + m)
    double unvisited_before;  // unvisited edges before this step
};

__device__ __forceinline__ void policy_features(const PolicyCtx &c, uint32_t nactv, uint64_t nacte, float f[6]) {
    double unv = c.unvisited_before - (double)nacte;
    if (unv < 0) unv = 0;
    f[0] = log2f(1.0f + (float)nactv) / 32.0f;
    f[1] = 1.0f + (log2f(1.0f + (float)nacte) - c.log_m) / 32.0f;
    f[2] = 1.0f + ((float)log2(1.0 + unv) - c.log_m) / 32.0f;
    f[3] = c.iter_feat;
    f[4] = c.mean_deg_feat;
    f[5] = c.skew_feat;
}

// The gated choice from the output logits: the argmax, unless it is not
// action 0 and leads action 0 by less than the margin.
__host__ __device__ __forceinline__ int policy_gate(const float z[MLP_OUT], float margin) {
    int best = 0;
    for (int k = 1; k < MLP_OUT; ++k)
        if (z[k] > z[best]) best = k;
    return best != 0 && z[best] - z[0] < margin ? 0 : best;
}

// One wave (64 lanes, all active) evaluates the network: lane j holds
// hidden unit j (j < 32; the upper half repeats the lower), activations
// move between lanes by shuffles.  load() issues every weight a lane needs
// (73 of them) in one batch, so that a caller can issue its own loads
// (the frontier counts) behind them and wait once for all; run() returns
// the gated choice in every lane.
struct MlpWave {
    float w1[MLP_IN], w2[MLP_H], w3[MLP_H], b1, b2, b3;
    __device__ __forceinline__ void load(const float *params) {
        const MlpView P(params);
        const unsigned ln = __lane_id();
        const unsigned j = ln & (MLP_H - 1), k = ln & (MLP_OUT - 1);
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) w1[i] = P.w1[i * MLP_H + j];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) w2[i] = P.w2[i * MLP_H + j];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) w3[i] = P.w3[i * MLP_OUT + k];
        b1 = P.b1[j];
        b2 = P.b2[j];
        b3 = P.b3[k];
    }
    __device__ __forceinline__ int run(const float f[6], float margin) const {
        float a1 = b1;
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a1 += f[i] * w1[i];
        const float h1 = a1 > 0.f ? a1 : 0.f;
        float a2 = b2;
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) a2 += __shfl(h1, i, 64) * w2[i];
        const float h2 = a2 > 0.f ? a2 : 0.f;
        float a3 = b3;
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) a3 += __shfl(h2, i, 64) * w3[i];
        float z[MLP_OUT];
#pragma unroll
        for (int q = 0; q < MLP_OUT; ++q) z[q] = __shfl(a3, q, 64);
        return policy_gate(z, margin);
    }
};

} // namespace gpr
