// Device helpers shared by the node-block kernels (xg_riemann.hip,
// xg_kozak.hip): the XCD-aware block order, the halo fix-up pass and the
// CFL time step reduction.
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

__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// per-node CFL limit of the explicit solvers' dt, unsteady branch
// (RieCG.cpp:827-839, KozCG.cpp:648-660), block minimum; nc components per
// node, the flow variables first
[[maybe_unused]] __global__ __launch_bounds__(xgd::kBlock) void k_dt_partial(
    int npoin, int nc, const double* __restrict__ U, const double* __restrict__ vol,
    double g, double* __restrict__ part) {
  __shared__ double sm[xgd::kBlock / 64];
  double m = 1.7976931348623157e308;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npoin;
       p += (int64_t)gridDim.x * blockDim.x) {
    const double* s = U + p * nc;
    double r = s[0];
    double u = s[1] / r, v = s[2] / r, w = s[3] / r;
    double pr = (s[4] - 0.5 * r * (u * u + v * v + w * w)) * (g - 1.0);
    double c = sqrt(g * fmax(pr, 0.0) / r);
    double L = cbrt(vol[p]);
    double vel = sqrt(u * u + v * v + w * w);
    m = fmin(m, L / fmax(vel + c, 1.0e-8));
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int k = 1; k < xgd::kBlock / 64; ++k) b = fmin(b, sm[k]);
    part[blockIdx.x] = b;
  }
}

[[maybe_unused]] __global__ __launch_bounds__(xgd::kRedBlocks) void k_min_final(
    int n, const double* __restrict__ part, double scale,
    double* __restrict__ out) {
  __shared__ double sm[xgd::kRedBlocks / 64];
  double m = 1.7976931348623157e308;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, part[i]);
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int k = 1; k < xgd::kRedBlocks / 64; ++k) b = fmin(b, sm[k]);
    out[0] = b * scale;
  }
}

// tp += dtp, the local times of steady runs (RieCG.cpp:1047-1050,
// KozCG.cpp:1189-1193)
[[maybe_unused]] __global__ void k_axpy1(int64_t n, const double* __restrict__ d, double* __restrict__ t) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] += d[i];
}

// steady dt of the explicit solvers (RieCG.cpp:810-824, KozCG.cpp:636-649):
// dtp per node and a block minimum; nc components per node
[[maybe_unused]] __global__ __launch_bounds__(xgd::kBlock) void k_dt_steady(
    int npoin, int nc, const double* __restrict__ U, const double* __restrict__ vol,
    double g, double cfl, double* __restrict__ dtp, double* __restrict__ part) {
  __shared__ double sm[xgd::kBlock / 64];
  double m = 1.7976931348623157e308;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npoin;
       p += (int64_t)gridDim.x * blockDim.x) {
    const double* s = U + p * nc;
    double r = s[0];
    double u = s[1] / r, v = s[2] / r, w = s[3] / r;
    double pr = (s[4] - 0.5 * r * (u * u + v * v + w * w)) * (g - 1.0);
    double c = sqrt(g * fmax(pr, 0.0) / r);
    double L = cbrt(vol[p]);
    double vel = sqrt(u * u + v * v + w * w);
    double d = L / fmax(vel + c, 1.0e-8) * cfl;
    dtp[p] = d;
    m = fmin(m, d);
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int k = 1; k < xgd::kBlock / 64; ++k) b = fmin(b, sm[k]);
    part[blockIdx.x] = b;
  }
}

// Symmetry BCs (physics::symbc, BC.cpp:110-136): removes the normal part of the
// 3-vector at V[node * stride + off], one (node, normal) pair after the other
// in list order; the pairs are grouped by node (off[i]..off[i+1]).
[[maybe_unused]] __global__ void k_symbc(int ns, const int* __restrict__ node,
                                         const int* __restrict__ off,
                                         const double* __restrict__ nrm,
                                         double* __restrict__ V, int stride, int voff) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  double* u = V + (size_t)node[i] * stride + voff;
  double a = u[0], b = u[1], c = u[2];
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const double* n = nrm + (size_t)k * 3;
    double vn = a * n[0] + b * n[1] + c * n[2];
    a -= vn * n[0];
    b -= vn * n[1];
    c -= vn * n[2];
  }
  u[0] = a; u[1] = b; u[2] = c;
}

}  // namespace
