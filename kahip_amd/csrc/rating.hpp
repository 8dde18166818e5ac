// Edge ratings that replace edge weights in the LP clustering: the plain expansion*^2 rating
// (reference: lib/partition/coarsening/edge_rating/edge_ratings.cpp:171-183) and a learned rating
// (an 8 -> 64 -> 64 -> 1 MLP over integer-derived edge features, BF16 on MFMA, FP32 accumulate),
// both quantised to int32 scores in [1, kScoreMax].
#pragma once

#include "common.hpp"
#include "devbuf.hpp"
#include "lp.hpp"

namespace kahip {

constexpr int kFeatures = 8;
constexpr int kHidden = 64;
constexpr int kScoreMax = 4096;  // max score: degree 4e5 x 4096 stays below 2^31
constexpr int kMlpParams = kHidden * kFeatures + kHidden + kHidden * kHidden + kHidden + kHidden + 1;

// MLP parameters in FP32, row-major: W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3.
struct MlpWeights {
    float p[kMlpParams];
};
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kFeatures;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden;

enum class RatingMode : int { None = 0, ExpansionStar2 = 1, Learned = 2 };

class RatingWorkspace {
public:
    // features[e * 8 + f] (FP32) of every directed edge of g; `seed` drives the 3 one-round LP
    // sweeps (cluster bound `bound`) whose agreement is feature 6.
    void features(const DGraph& g, int bound, uint32_t seed, float* features, hipStream_t s,
                  LpWorkspace& lp, Profiler* prof);
    // logits[e] of the MLP on the features (MFMA, BF16 weights and activations).
    void mlp_forward(const float* features, long long m, const MlpWeights& w, float* logits,
                     hipStream_t s, Profiler* prof);
    // Integer scores of the chosen rating for every directed edge of g, shifted right by
    // score_shift(g) (at least 1) so that a node's score sum stays below 2^31.
    void scores(RatingMode mode, const DGraph& g, int bound, uint32_t seed, const MlpWeights* w,
                int* score, hipStream_t s, LpWorkspace& lp, Profiler* prof);
    // One optimisation step (Adam, step >= 1) of the mean binary cross-entropy of the MLP on
    // (features, labels); params / m / v / grads are FP32 device arrays of kMlpParams, grads
    // receives the mean gradient and *loss (device, may be null) the mean loss; lr = 0 only
    // computes them. Deterministic: fixed-order reductions only.
    void train_step(const float* features, const float* labels, int batch, float* params,
                    float* m, float* v, float* grads, int step, float lr, float* loss,
                    hipStream_t s);

    // Smallest shift with max_degree(g) * (kScoreMax >> shift) < 2^31 (0 below degree 524288).
    int score_shift(const DGraph& g, hipStream_t s);

private:
    void sweeps_and_sketch(const DGraph& g, int bound, uint32_t seed, hipStream_t s,
                           LpWorkspace& lp, Profiler* prof);
    void upload_weights(const MlpWeights& w, hipStream_t s);
    DevBuf<unsigned long long> sig_;
    DevBuf<int> lab_, size_, map_, heavy_, rec_;
    DevBuf<float> feat_, logit_, partial_, loss_;
    DevBuf<uint16_t> wpack_, featb_;
};

// score = clamp(round(logit * 128) + 2048, 1, kScoreMax); the same formula on the host.
__host__ __device__ inline int quantize_logit(float logit) {
    float x = logit * 128.0f + 2048.0f;
    if (!(x >= 1.0f)) return 1;  // also maps NaN to 1
    if (x >= (float)kScoreMax) return kScoreMax;
    return (int)__builtin_rintf(x);
}

}  // namespace kahip
