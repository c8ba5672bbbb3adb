// Learned space adapter network (internal): 8 -> 64 -> 64 -> 1 MLP on BF16 MFMA.
#pragma once

#include "nc_internal.h"

namespace nc {
constexpr int kNP = 8 * 64 + 64 + 64 * 64 + 64 + 64 + 1;  // parameters
}  // namespace nc

struct nc_mlp_s {
  nc::DevBuf<float> params, m, v, grad, loss;  // FP32 master weights, Adam moments, gradient
  nc::DevBuf<float> part, loss_part;           // per-workgroup gradient partials
  nc::Stream stream;
  int64_t step = 0;
};

namespace nc {
// logits of n samples (n x 8 features, FP32, device) on stream s
void mlp_forward_device(nc_mlp_s* M, int64_t n, const float* d_x, float* d_logits, hipStream_t s);
// mean binary cross-entropy gradient into M->grad; returns the mean loss
double mlp_grad_device(nc_mlp_s* M, int64_t n, const float* d_x, const float* d_y, hipStream_t s);
}  // namespace nc
