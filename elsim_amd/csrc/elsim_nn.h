// elsim_nn.h -- interface of the learned warm start shared by the SBM step
// driver: the network object and its inference call.
#pragma once

#include <cstdint>

#include "elsim_common.h"

namespace elsim {

constexpr int NN_NLAYERS = 7;  // e0 e1 e2 e3 d2 d1 d0 (3x3x3 convolutions)
constexpr int NN_IN_CH = 16;   // stored input channels (features zero-padded)

// bfloat16 <-> float, round to nearest even.
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct NnNet;
int64_t nn_param_count();
// pred[v] (FP32) for the input features x (bf16 [voxel][in_used], in_used
// 16 or 8; with 8, the network's input channels 8..15 are zero); grid dims
// multiples of 8.
int nn_forward(NnNet* net, int n1, int n2, int n3, const uint16_t* x, float* pred, int in_used);
// The same prediction written as FP64 into the interior of a ghosted
// (n1 + 2) x (n2 + 2) x (n3 + 2) box (ghosts untouched).
int nn_forward_g(NnNet* net, int n1, int n2, int n3, const uint16_t* x, double* pred_g, int in_used);
// The network of a handle issued by elsim_nn_create (nullptr if unknown or
// destroyed).
NnNet* nn_lookup(int64_t handle);
// Detach `net` from every SBM state that holds it (called when the network is
// destroyed).
void sbm_forget_net(NnNet* net);

}  // namespace elsim
