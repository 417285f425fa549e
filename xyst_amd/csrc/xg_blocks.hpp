// Device helpers shared by the node-block kernels (xg_riemann.hip,
// xg_kozak.hip): the XCD-aware block order and the halo fix-up pass.
#pragma once
#include <hip/hip_runtime.h>

#include "xg_ctx.hpp"

namespace {

// Workgroup i runs on XCD i % 8; give each XCD a contiguous range of node
// blocks so that neighbouring blocks share that XCD's L2.
__device__ inline int xcd_block(int i, int n) {
  int x = i & 7, j = i >> 3;
  int q = n >> 3, r = n & 7;
  return x * q + min(x, r) + j;
}

// Adds halo partials (NC components per node) to their owners, block order.
template <int NC>
__global__ __launch_bounds__(xgd::kBlock) void k_fixup(
    int nfx, const int* __restrict__ fx_node, const int* __restrict__ fx_off,
    const int* __restrict__ fx_slot, const double* __restrict__ S,
    double* __restrict__ R) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfx) return;
  double* r = R + (size_t)fx_node[i] * NC;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = r[c];
  for (int k = fx_off[i]; k < fx_off[i + 1]; ++k) {
    const double* s = S + (size_t)fx_slot[k] * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += s[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) r[c] = acc[c];
}

}  // namespace
