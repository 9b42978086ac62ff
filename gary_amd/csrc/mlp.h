// mlp.h — the route-cost MLP shared by the triangle selector (selector.hip) and the 4-cycle
// vertex routing (hom4_kernels.hip): parameter layout, forward pass, device-side training, and the
// reader of the saved model files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gary.h"

namespace gary {

constexpr int kIn = GARY_MLP_IN, kHid = GARY_MLP_HIDDEN, kOut = GARY_MLP_OUT;
constexpr int kParams = GARY_MLP_PARAMS;
// parameter layout: W1[kHid][kIn] b1[kHid] W2[kHid][kHid] b2[kHid] W3[kOut][kHid] b3[kOut]
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid,
              oW3 = oB2 + kHid, oB3 = oW3 + kOut * kHid;
static_assert(oB3 + kOut == kParams, "parameter layout");

// y = W3 relu(W2 relu(W1 x + b1) + b2) + b3; h1, h2 receive the hidden activations.
__device__ __forceinline__ void mlp_forward(const float *__restrict__ P, const float *x, float *h1,
                                            float *h2, float *y) {
  for (int k = 0; k < kHid; ++k) {
    float a = P[oB1 + k];
    for (int i = 0; i < kIn; ++i) a += P[oW1 + k * kIn + i] * x[i];
    h1[k] = a > 0.f ? a : 0.f;
  }
  for (int k = 0; k < kHid; ++k) {
    float a = P[oB2 + k];
    for (int i = 0; i < kHid; ++i) a += P[oW2 + k * kHid + i] * h1[i];
    h2[k] = a > 0.f ? a : 0.f;
  }
  for (int o = 0; o < kOut; ++o) {
    float a = P[oB3 + o];
    for (int i = 0; i < kHid; ++i) a += P[oW3 + o * kHid + i] * h2[i];
    y[o] = a;
  }
}

// Adam state of one parameter vector on the device (kParams floats each).
struct MlpTrainState {
  float *P = nullptr, *M1 = nullptr, *M2 = nullptr, *grad = nullptr, *loss = nullptr;
  int64_t steps = 0;
};

// Allocates the parameters and Adam state on the current device (all zero) unless allocated
// already; they live until the process ends.
int mlp_alloc(MlpTrainState &s);

// He-uniform initialisation from a splitmix64 stream (zero biases), written to the device
// parameters; resets the Adam moments.
int mlp_init(MlpTrainState &s, uint64_t seed);

// `epochs` full-batch Adam steps on the masked squared error over n rows (host arrays), then
// the final loss to *h_loss. Uses the device workspace of the current device.
int mlp_train(MlpTrainState &s, const float *h_x, const float *h_t, const float *h_mask, int32_t n,
              int32_t epochs, float lr, float *h_loss);

// The numbers of a top-level key of a model file (a number or an array of numbers, possibly
// nested): false when the key is absent or holds anything else.
bool selector_file_numbers(const std::string &doc, const char *key, std::vector<double> &out);

}  // namespace gary
