// Band predictor (MLP 24 -> 64 -> 64 -> 5): parameter layout and launchers.
#pragma once
#include "bsx_internal.h"

namespace bsx {

constexpr int kBandFeatures = 24;  // features per task
constexpr int kFeatPad = 32;       // stored features per task (BF16, zero padded)
constexpr int kFeatScan = 16;      // query bases covered by the ungapped feature scans
constexpr int kHidden = 64;
constexpr int kOffW1 = 0;                                   // W1[64][24]
constexpr int kOffB1 = kOffW1 + kHidden * kBandFeatures;    // b1[64]
constexpr int kOffW2 = kOffB1 + kHidden;                    // W2[64][64]
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;          // b2[64]
constexpr int kOffW3 = kOffB2 + kHidden;                    // W3[5][64]
constexpr int kOffB3 = kOffW3 + kBandClasses * kHidden;     // b3[5]
constexpr int kParamCount = kOffB3 + kBandClasses;          // 6085
static_assert(kParamCount == 6085, "parameter count");
constexpr int kMaxGradWaves = 1024;  // partial-gradient rows of one gradient pass

struct MlpDev {
  float* params;   // [kParamCount] FP32 master
  float* adam_m;   // [kParamCount]
  float* adam_v;   // [kParamCount]
  uint16_t* w1p;   // [64][32]  BF16 operand copies
  uint16_t* w2p;   // [64][64]
  uint16_t* w3p;   // [16][64]
  uint16_t* w2t;   // [64][64]  transposes for the backward pass
  uint16_t* w3t;   // [64][16]
};

hipError_t mlp_launch_features(const ExtArgs& a, uint16_t* out, int num_cu, hipStream_t st);
hipError_t mlp_launch_predict_tasks(const ExtArgs& a, const MlpDev& m, uint8_t* cls, int num_cu, hipStream_t st);
hipError_t mlp_launch_forward(const MlpDev& m, const uint16_t* feats, int64_t n, float* logits, uint8_t* cls,
                              int num_cu, hipStream_t st);
int mlp_grad_waves(int64_t n);
hipError_t mlp_launch_grad(const MlpDev& m, const uint16_t* feats, const uint8_t* labels, int64_t n, float* part,
                           float* grad, hipStream_t st);
hipError_t mlp_launch_adam(const MlpDev& m, const float* grad, float lr, int step, hipStream_t st);
hipError_t mlp_launch_pack(const MlpDev& m, hipStream_t st);

hipError_t launch_band_label_merge(const ExtArgs& a, const ExtWork& wk, int c, uint8_t* labels, int first,
                                   int num_cu, hipStream_t st);
hipError_t launch_band_stats(int64_t n, const ExtWork& wk, unsigned long long* stats, int num_cu, hipStream_t st);

}  // namespace bsx
