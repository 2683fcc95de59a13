// policy_dev.hpp -- the direction/representation policy network on the
// device: a 6-32-32-4 MLP (ReLU, FP32) whose forward pass runs inside the
// last workgroup of a traversal kernel, right after the frontier statistics
// of that step are complete, so the next step's choice costs no launch.
//
// Features (all of the frontier just produced):
//   f0 log2(1 + nactv) / 32          frontier size
//   f1 log2(1 + nacte) / 40          frontier out-edge sum
//   f2 log2(1 + unvisited) / 40      edges of vertices not yet settled
//   f3 iteration / 64
//   f4 log2(1 + m/n) / 8             mean degree
//   f5 log2(1 + maxdeg/(m/n)) / 24   degree skew
// Outputs (argmax): 0 push + list output, 1 push + bitmap output,
//                   2 pull + list output, 3 pull + bitmap output.
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
    float mean_deg_feat;      // f4
    float skew_feat;          // f5
    float iter_feat;          // f3 for the NEXT step
    double unvisited_before;  // unvisited edges before this step
};

__device__ __forceinline__ void policy_features(const PolicyCtx &c, uint32_t nactv, uint64_t nacte, float f[6]) {
    double unv = c.unvisited_before - (double)nacte;
    if (unv < 0) unv = 0;
    f[0] = log2f(1.0f + (float)nactv) / 32.0f;
    f[1] = log2f(1.0f + (float)nacte) / 40.0f;
    f[2] = (float)log2(1.0 + unv) / 40.0f;
    f[3] = c.iter_feat;
    f[4] = c.mean_deg_feat;
    f[5] = c.skew_feat;
}

// The whole (last) workgroup evaluates the network: the parameters are
// staged in LDS with one load round, thread j computes hidden unit j.
// Every thread of the workgroup must call it; returns the argmax.
struct PolicyLds {
    float p[MLP_PARAMS];
    float h[2 * MLP_H];
    float z[MLP_OUT];
};
__device__ __forceinline__ int policy_forward_block(const float *params, const float f[6], PolicyLds &L) {
    for (int i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) L.p[i] = params[i];
    __syncthreads();
    const MlpView P(L.p);
    const int j = threadIdx.x;
    if (j < MLP_H) {
        float a = P.b1[j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a += f[i] * P.w1[i * MLP_H + j];
        L.h[j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (j < MLP_H) {
        float a = P.b2[j];
#pragma unroll 8
        for (int i = 0; i < MLP_H; ++i) a += L.h[i] * P.w2[i * MLP_H + j];
        L.h[MLP_H + j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (j < MLP_OUT) {
        float a = P.b3[j];
#pragma unroll 8
        for (int i = 0; i < MLP_H; ++i) a += L.h[MLP_H + i] * P.w3[i * MLP_OUT + j];
        L.z[j] = a;
    }
    __syncthreads();
    int best = 0;
    for (int k = 1; k < MLP_OUT; ++k)
        if (L.z[k] > L.z[best]) best = k;
    return best;
}

// One wave (64 lanes, all active) evaluates the network with lane j holding
// hidden unit j (activations move between lanes by shuffles); returns the
// argmax in every lane.
__device__ __forceinline__ int policy_forward_wave(const float *params, const float f[6]) {
    const MlpView P(params);
    const unsigned ln = __lane_id();
    float h1 = 0.f;
    if (ln < MLP_H) {
        float a = P.b1[ln];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a += f[i] * P.w1[i * MLP_H + ln];
        h1 = a > 0.f ? a : 0.f;
    }
    float a2 = ln < MLP_H ? P.b2[ln] : 0.f;
    for (int i = 0; i < MLP_H; ++i) {
        const float hi = __shfl(h1, i, 64);
        if (ln < MLP_H) a2 += hi * P.w2[i * MLP_H + ln];
    }
    const float h2 = a2 > 0.f ? a2 : 0.f;
    float z = ln < MLP_OUT ? P.b3[ln] : -3.0e38f;
    for (int i = 0; i < MLP_H; ++i) {
        const float hi = __shfl(h2, i, 64);
        if (ln < MLP_OUT) z += hi * P.w3[i * MLP_OUT + ln];
    }
    // argmax over lanes 0..3 (ties to the lower index)
    int best = 0;
    float bz = __shfl(z, 0, 64);
    for (int k = 1; k < MLP_OUT; ++k) {
        const float zk = __shfl(z, k, 64);
        if (zk > bz) {
            bz = zk;
            best = k;
        }
    }
    return best;
}

} // namespace gpr
