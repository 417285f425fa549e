// Learned graph preconditioner for the pressure-Poisson CG (no counterpart
// in the reference): a 3-layer message-passing network over the matrix graph,
// 32 channels, applied inside the flexible CG of xg_cg.hip, trained on the
// device by the forward and backward kernels below.
//
//   s = ||D^-1/2 r|| / sqrt(N),  x_i = r_i / (sqrt|d_i| s)
//   h0 = relu(we x + be)                                   (1 -> 32)
//   l = 1..3:  m = Ahat h,  Ahat = |D|^-1/2 A |D|^-1/2      (SpMM, 32 channels)
//              h += relu(W_l [h; m] + b_l)                 (64 -> 32, BF16 MFMA,
//                                                           FP32 accumulation)
//   y = wo . h3 + bo,  z_i = r_i / d_i + s y_i / sqrt|d_i|   (Jacobi plus a
// learned correction; |d| because the ChoCG matrix -K is negative definite.)
// Decoupled rows
// (the diagonal their only nonzero: the Dirichlet rows) take z_i = r_i / d_i,
// which is exact there; the network acts on the coupled rows only.
// Activations are stored in BF16, sums are FP32.  Training minimises the
// energy-norm error of z relative to that of the true error e (A e = r):
//   L = (z^T A z - 2 z^T r) / (e^T r) + 1,  dL/dz = 2 (A z - r) / (e^T r).
// The backward pass is FP32; weight gradients are per-workgroup partials
// summed in a fixed order (deterministic).  Adam on FP32 master weights.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "xg_ctx.hpp"
#include "xg_gnn.hpp"

using namespace xgd;

namespace {

constexpr int kC = 32;                 // channels
constexpr int kL = 3;                  // message-passing layers
constexpr int kWsz = kC * 2 * kC;      // 2048 weights of a layer
constexpr int kOffWe = 0, kOffBe = 32;
constexpr int kOffW0 = 64;             // layer l: W at kOffW0 + l*2080, b after
constexpr int kLayerSz = kWsz + kC;
constexpr int kOffWo = kOffW0 + kL * kLayerSz;   // 6304
constexpr int kOffBo = kOffWo + kC;              // 6336
constexpr int kThreads = 256;
constexpr int kParts = 2048;           // max workgroups of reduction kernels
constexpr int kChunk = 64;             // nodes per backward chunk

static_assert(kOffBo + 1 == XG_GNN_NPARAM, "parameter layout");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline uint16_t f2bf(float f) {   // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline double block_sum(double v, double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) s += sm[k];
  return s;
}

__device__ inline double sum_parts(const double* __restrict__ p, int n, double* sm) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) v += p[i];
  return block_sum(v, sm);
}

// Ahat_ik = a_ik / sqrt|d_i d_k| (FP32), sqrt|d|, and the decoupled-row flag
__global__ __launch_bounds__(kThreads) void k_gnn_prep(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const double* __restrict__ a, const double* __restrict__ d,
    float* __restrict__ ah, float* __restrict__ sqd, uint8_t* __restrict__ dec) {
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const double di = fabs(d[i]);
  sqd[i] = (float)sqrt(di);
  int nz = 0;
  for (int k = ia[i]; k < ia[i + 1]; ++k) {
    ah[k] = (float)(a[k] / sqrt(di * fabs(d[ja[k]])));
    nz += a[k] != 0.0 ? 1 : 0;
  }
  dec[i] = (nz == 1 && di != 0.0) ? 1 : 0;
}

// partials of sum r_i^2 / |d_i|
__global__ __launch_bounds__(kThreads) void k_gnn_scale_part(
    int64_t n, const double* __restrict__ r, const double* __restrict__ d,
    double* __restrict__ part) {
  __shared__ double sm[kThreads / 64];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    v += r[i] * r[i] / fabs(d[i]);
  v = block_sum(v, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// s from the partials; x = r/(sqrt(d) s); h0 = relu(we x + be); thread per
// (node, channel)
__global__ __launch_bounds__(kThreads) void k_gnn_encode(
    int64_t n, const double* __restrict__ r, const float* __restrict__ sqd,
    const double* __restrict__ part, int nparts, const float* __restrict__ P,
    float* __restrict__ xin, uint16_t* __restrict__ H0, double* __restrict__ sc) {
  __shared__ double sm[kThreads / 64];
  const double s2 = sum_parts(part, nparts, sm) / (double)n;
  const double s = s2 > 0.0 ? sqrt(s2) : 1.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[0] = s;
  const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t i = g / kC;
  const int c = (int)(g % kC);
  if (i >= n) return;
  const float x = (float)(r[i] / ((double)sqd[i] * s));
  if (c == 0) xin[i] = x;
  H0[i * kC + c] = f2bf(fmaxf(P[kOffWe + c] * x + P[kOffBe + c], 0.0f));
}

// m = Ahat h (BF16 in, FP32 sum, BF16 out); thread per (node, channel)
__global__ __launch_bounds__(kThreads) void k_gnn_agg(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const float* __restrict__ ah, const uint16_t* __restrict__ H,
    uint16_t* __restrict__ M) {
  const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t i = g / kC;
  const int c = (int)(g % kC);
  if (i >= nrow) return;
  float acc = 0.0f;
  for (int k = ia[i]; k < ia[i + 1]; ++k) acc += ah[k] * bf2f(H[(int64_t)ja[k] * kC + c]);
  M[i * kC + c] = f2bf(acc);
}

// h_out = h + relu(W [h; m] + b) for 32 nodes per wave on gfx950's
// 32x32x16 BF16 MFMA (FP32 accumulation), four steps over the 64 inputs:
// lane l supplies A[l%32][k0..k0+7] (node row) and B[k0..k0+7][l%32]
// (= W[l%32][k0..]) with k0 = 16 kb + 8 (l/32); it receives D[i][l%32] for
// i = 8 (v/4) + 4 (l/32) + v%4, v = 0..15.
__global__ __launch_bounds__(kThreads) void k_gnn_mlp(
    int64_t n, const uint16_t* __restrict__ Hin, const uint16_t* __restrict__ Min,
    const uint16_t* __restrict__ Wb, const float* __restrict__ b,
    uint16_t* __restrict__ Hout, uint16_t* __restrict__ Aout) {
  const int lane = threadIdx.x & 63;
  const int64_t i0 = ((int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)) * 32;
  if (i0 >= n) return;
  const int ri = lane & 31, kq = lane >> 5;
  const int64_t node = i0 + ri;
  const bool valid = node < n;
  floatx16 acc = {};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const int k0 = kb * 16 + kq * 8;     // eight inputs, all from h or all from m
    u16x8 av = {0, 0, 0, 0, 0, 0, 0, 0};
    if (valid) {
      const uint16_t* src = k0 < kC ? Hin + node * kC + k0 : Min + node * kC + (k0 - kC);
      av = *(const u16x8*)src;
    }
    const u16x8 bv = *(const u16x8*)(Wb + ri * (2 * kC) + k0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                  __builtin_bit_cast(bf16x8, bv), acc, 0, 0,
                                                  0);
  }
  const float bj = b[ri];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int i = 8 * (v >> 2) + 4 * kq + (v & 3);
    const int64_t nd = i0 + i;
    if (nd < n) {
      const float a = acc[v] + bj;
      Aout[nd * kC + ri] = f2bf(a);
      Hout[nd * kC + ri] = f2bf(bf2f(Hin[nd * kC + ri]) + fmaxf(a, 0.0f));
    }
  }
}

// y = wo . h3 + bo; z = r/d + s y / sqrt|d| (FP64 out), r/d on decoupled rows
__global__ __launch_bounds__(kThreads) void k_gnn_decode(
    int64_t n, const uint16_t* __restrict__ H3, const float* __restrict__ P,
    const float* __restrict__ sqd, const uint8_t* __restrict__ dec,
    const double* __restrict__ r, const double* __restrict__ d,
    const double* __restrict__ sc, double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  if (dec[i]) {
    z[i] = r[i] / d[i];
    return;
  }
  float y = P[kOffBo];
#pragma unroll
  for (int c = 0; c < kC; ++c) y += P[kOffWo + c] * bf2f(H3[i * kC + c]);
  z[i] = r[i] / d[i] + sc[0] * (double)y / (double)sqd[i];
}

// dz = 2 (A z - r) / etr;  dy = dz s / sqrt(d);  partials of z^T A z, z^T r
__global__ __launch_bounds__(kThreads) void k_gnn_dz(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const double* __restrict__ a, const double* __restrict__ z,
    const double* __restrict__ r, double etr, const float* __restrict__ sqd,
    const uint8_t* __restrict__ dec, const double* __restrict__ sc,
    float* __restrict__ dy, double* __restrict__ part) {
  __shared__ double sm[kThreads / 64];
  double zaz = 0.0, zr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < nrow;
       i += (int64_t)gridDim.x * kThreads) {
    double az = 0.0;
    for (int k = ia[i]; k < ia[i + 1]; ++k) az += a[k] * z[ja[k]];
    const double dz = 2.0 * (az - r[i]) / etr;
    dy[i] = dec[i] ? 0.0f : (float)(dz * sc[0] / (double)sqd[i]);
    zaz += z[i] * az;
    zr += z[i] * r[i];
  }
  const double A = block_sum(zaz, sm);
  const double B = block_sum(zr, sm);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = A;
    part[kParts + blockIdx.x] = B;
  }
}

// loss of the pair into sc[1] (one workgroup)
__global__ __launch_bounds__(kThreads) void k_gnn_loss(
    const double* __restrict__ part, int nparts, double etr, double* __restrict__ sc) {
  __shared__ double sm[kThreads / 64];
  const double zaz = sum_parts(part, nparts, sm);
  const double zr = sum_parts(part + kParts, nparts, sm);
  if (threadIdx.x == 0) sc[1] = (zaz - 2.0 * zr) / etr + 1.0;
}

// decoder backward: dH3 = dy wo; partials of dwo (32) and dbo (1) per workgroup
__global__ __launch_bounds__(kThreads) void k_gnn_dec_back(
    int64_t n, const uint16_t* __restrict__ H3, const float* __restrict__ dy,
    const float* __restrict__ P, float* __restrict__ dH, float* __restrict__ wpart) {
  __shared__ float sm[kThreads];
  const int c = threadIdx.x % kC;
  float aw = 0.0f, ab = 0.0f;
  for (int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x; g < n * kC;
       g += (int64_t)gridDim.x * kThreads) {
    const int64_t i = g / kC;
    const float d = dy[i];
    dH[g] = d * P[kOffWo + c];
    aw += d * bf2f(H3[g]);
    if (c == 0) ab += d;
  }
  // sum the 8 thread groups of each channel in a fixed order
  sm[threadIdx.x] = aw;
  __syncthreads();
  if (threadIdx.x < kC) {
    float s = 0.0f;
    for (int k = 0; k < kThreads / kC; ++k) s += sm[k * kC + threadIdx.x];
    wpart[(int64_t)blockIdx.x * (kC + 1) + threadIdx.x] = s;
  }
  __syncthreads();
  sm[threadIdx.x] = ab;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int k = 0; k < kThreads / kC; ++k) s += sm[k * kC];
    wpart[(int64_t)blockIdx.x * (kC + 1) + kC] = s;
  }
}

// layer backward (FP32): da = dH * (a > 0); dW += da^T [h; m], db += sum da;
// d[h;m] = W^T da; dHprev = dH + d[h;m][:32]; dM = d[h;m][32:].
// Chunks of 64 nodes staged in LDS; per workgroup partials of dW and db.
__global__ __launch_bounds__(kThreads) void k_gnn_layer_back(
    int64_t n, const uint16_t* __restrict__ Hprev, const uint16_t* __restrict__ Mv,
    const uint16_t* __restrict__ Apre, const float* __restrict__ W,
    const float* __restrict__ dH, float* __restrict__ dHprev, float* __restrict__ dM,
    float* __restrict__ wpart) {
  __shared__ float sW[kWsz];
  __shared__ float sDa[kChunk * kC];
  __shared__ float sHM[kChunk * 2 * kC];
  const int t = threadIdx.x;
  for (int k = t; k < kWsz; k += kThreads) sW[k] = W[k];
  float accw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float accb = 0.0f;
  const int wj = t / 8, wk = (t % 8) * 8;      // dW outputs of this thread
  const int64_t nchunk = (n + kChunk - 1) / kChunk;
  for (int64_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const int64_t base = ch * kChunk;
    __syncthreads();
    for (int k = t; k < kChunk * kC; k += kThreads) {
      const int64_t nd = base + k / kC;
      float v = 0.0f;
      if (nd < n) {
        const int64_t g = nd * kC + k % kC;
        v = bf2f(Apre[g]) > 0.0f ? dH[g] : 0.0f;
      }
      sDa[k] = v;
    }
    for (int k = t; k < kChunk * 2 * kC; k += kThreads) {
      const int64_t nd = base + k / (2 * kC);
      const int kk = k % (2 * kC);
      float v = 0.0f;
      if (nd < n) v = kk < kC ? bf2f(Hprev[nd * kC + kk]) : bf2f(Mv[nd * kC + kk - kC]);
      sHM[k] = v;
    }
    __syncthreads();
    for (int i = 0; i < kChunk; ++i) {
      const float da = sDa[i * kC + wj];
#pragma unroll
      for (int q = 0; q < 8; ++q) accw[q] += da * sHM[i * 2 * kC + wk + q];
    }
    if (t < kC)
      for (int i = 0; i < kChunk; ++i) accb += sDa[i * kC + t];
    // d[h;m] for node i = t/4, outputs k = (t%4)*16 .. +15
    const int ni = t / 4, k0 = (t % 4) * 16;
    const int64_t nd = base + ni;
    if (nd < n) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = k0 + q;
        float s = 0.0f;
        for (int j = 0; j < kC; ++j) s += sW[j * 2 * kC + k] * sDa[ni * kC + j];
        if (k < kC)
          dHprev[nd * kC + k] = dH[nd * kC + k] + s;
        else
          dM[nd * kC + k - kC] = s;
      }
    }
  }
  float* out = wpart + (int64_t)blockIdx.x * kLayerSz;
#pragma unroll
  for (int q = 0; q < 8; ++q) out[wj * 2 * kC + wk + q] = accw[q];
  if (t < kC) out[kWsz + t] = accb;
}

// dHprev += Ahat dM (Ahat symmetric); thread per (node, channel)
__global__ __launch_bounds__(kThreads) void k_gnn_agg_back(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const float* __restrict__ ah, const float* __restrict__ dM,
    float* __restrict__ dHprev) {
  const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t i = g / kC;
  const int c = (int)(g % kC);
  if (i >= nrow) return;
  float acc = 0.0f;
  for (int k = ia[i]; k < ia[i + 1]; ++k) acc += ah[k] * dM[(int64_t)ja[k] * kC + c];
  dHprev[g] += acc;
}

// encoder backward: dpre = dH0 * (we x + be > 0); partials of dwe, dbe
__global__ __launch_bounds__(kThreads) void k_gnn_enc_back(
    int64_t n, const float* __restrict__ dH0, const float* __restrict__ xin,
    const float* __restrict__ P, float* __restrict__ wpart) {
  __shared__ float sm[kThreads];
  const int c = threadIdx.x % kC;
  float aw = 0.0f, ab = 0.0f;
  for (int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x; g < n * kC;
       g += (int64_t)gridDim.x * kThreads) {
    const float x = xin[g / kC];
    const float pre = P[kOffWe + c] * x + P[kOffBe + c];
    const float dp = pre > 0.0f ? dH0[g] : 0.0f;
    aw += dp * x;
    ab += dp;
  }
  sm[threadIdx.x] = aw;
  __syncthreads();
  if (threadIdx.x < kC) {
    float s = 0.0f;
    for (int k = 0; k < kThreads / kC; ++k) s += sm[k * kC + threadIdx.x];
    wpart[(int64_t)blockIdx.x * 2 * kC + threadIdx.x] = s;
  }
  __syncthreads();
  sm[threadIdx.x] = ab;
  __syncthreads();
  if (threadIdx.x < kC) {
    float s = 0.0f;
    for (int k = 0; k < kThreads / kC; ++k) s += sm[k * kC + threadIdx.x];
    wpart[(int64_t)blockIdx.x * 2 * kC + kC + threadIdx.x] = s;
  }
}

// dst[o] += sum_b part[b*stride + o], b in order
__global__ void k_gnn_sum_parts(const float* __restrict__ part, int nblk, int stride,
                                int count, float* __restrict__ dst) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= count) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * stride + o];
  dst[o] += s;
}

// Adam on the master parameters with the accumulated gradient scaled by
// `gscale`; then the BF16 copies of the layer weights
__global__ void k_gnn_adam(float* __restrict__ P, float* __restrict__ G,
                           float* __restrict__ m1, float* __restrict__ m2, int np,
                           float lr, float b1, float b2, float eps, float c1, float c2,
                           float gscale) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= np) return;
  const float g = G[o] * gscale;
  const float a = b1 * m1[o] + (1.0f - b1) * g;
  const float v = b2 * m2[o] + (1.0f - b2) * g * g;
  m1[o] = a;
  m2[o] = v;
  P[o] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
  G[o] = 0.0f;
}

__global__ void k_gnn_weights_bf16(const float* __restrict__ P, uint16_t* __restrict__ Wb) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= kL * kWsz) return;
  const int l = o / kWsz, k = o % kWsz;
  Wb[o] = f2bf(P[kOffW0 + l * kLayerSz + k]);
}

// z = r / d where the learned z is no descent direction: (z, r) must have
// the sign of the Jacobi direction's (r/d, r) (negative for the ChoCG matrix
// -K) and be finite; every workgroup decides alike from the partials
__global__ __launch_bounds__(kThreads) void k_gnn_zr_part(
    int64_t n, const double* __restrict__ z, const double* __restrict__ r,
    const double* __restrict__ d, double* __restrict__ part) {
  __shared__ double sm[kThreads / 64];
  double v = 0.0, w = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    v += z[i] * r[i];
    w += r[i] * r[i] / d[i];
  }
  v = block_sum(v, sm);
  w = block_sum(w, sm);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v;
    part[kParts + blockIdx.x] = w;
  }
}

__global__ __launch_bounds__(kThreads) void k_gnn_fallback(
    int64_t n, double* __restrict__ z, const double* __restrict__ r,
    const double* __restrict__ d, const double* __restrict__ part, int nparts,
    double* __restrict__ sc) {
  __shared__ double sm[kThreads / 64];
  const double zr = sum_parts(part, nparts, sm);
  const double jr = sum_parts(part + kParts, nparts, sm);
  const bool bad = !(zr * jr > 0.0) || !isfinite(zr);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[2] = bad ? 1.0 : 0.0;
  if (!bad) return;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads)
    z[i] = r[i] / d[i];
}

int grid_for(int64_t work) {
  int64_t g = (work + kThreads - 1) / kThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kParts));
}

}  // namespace

// ----------------------------------------------------------------- GnnNet --

struct GnnNet {
  int64_t n = 0, nnz = 0;
  std::vector<DevArr> allocs;
  float *P = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t* Wb = nullptr;
  float *ah = nullptr, *sqd = nullptr, *xin = nullptr, *dy = nullptr;
  uint8_t* dec = nullptr;     // decoupled (Dirichlet) rows
  uint16_t* H[kL + 1] = {};
  uint16_t* M[kL] = {};
  uint16_t* Apre[kL] = {};
  float *dH = nullptr, *dH2 = nullptr, *dM = nullptr;
  double *part = nullptr, *sc = nullptr, *zbuf = nullptr;
  float* wpart = nullptr;
  double* hsc = nullptr;      // pinned host mirror of sc
  int64_t adam_t = 0;
  int64_t pending = 0;        // pairs accumulated since the last Adam step
  bool prepared = false;
};

namespace {

int gnn_alloc(xg_ctx* c, GnnNet* g, int64_t n, int64_t nnz) {
  auto& L = g->allocs;
  int rc;
  const size_t nc = (size_t)n * kC;
  if ((rc = dev_alloc(c, L, (void**)&g->P, XG_GNN_NPARAM * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->G, XG_GNN_NPARAM * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->m1, XG_GNN_NPARAM * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->m2, XG_GNN_NPARAM * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->Wb, kL * kWsz * 2)) ||
      (rc = dev_alloc(c, L, (void**)&g->ah, (size_t)nnz * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->sqd, (size_t)n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->dec, (size_t)n)) ||
      (rc = dev_alloc(c, L, (void**)&g->xin, (size_t)n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->dy, (size_t)n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->dH, nc * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->dH2, nc * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->dM, nc * 4)) ||
      (rc = dev_alloc(c, L, (void**)&g->part, sizeof(double) * 2 * kParts)) ||
      (rc = dev_alloc(c, L, (void**)&g->sc, sizeof(double) * 8)) ||
      (rc = dev_alloc(c, L, (void**)&g->zbuf, (size_t)n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&g->wpart, sizeof(float) * kParts * kLayerSz)))
    return rc;
  for (int l = 0; l <= kL; ++l)
    if ((rc = dev_alloc(c, L, (void**)&g->H[l], nc * 2))) return rc;
  for (int l = 0; l < kL; ++l)
    if ((rc = dev_alloc(c, L, (void**)&g->M[l], nc * 2)) ||
        (rc = dev_alloc(c, L, (void**)&g->Apre[l], nc * 2)))
      return rc;
  if (hipHostMalloc((void**)&g->hsc, sizeof(double) * 8) != hipSuccess)
    return fail(c, XG_EHIP, "hipHostMalloc");
  XG_HIP(c, hipMemset(g->G, 0, XG_GNN_NPARAM * 4));
  XG_HIP(c, hipMemset(g->m1, 0, XG_GNN_NPARAM * 4));
  XG_HIP(c, hipMemset(g->m2, 0, XG_GNN_NPARAM * 4));
  XG_HIP(c, hipMemset(g->sc, 0, sizeof(double) * 8));
  return XG_OK;
}

int gnn_forward(xg_ctx* c, GnnNet* g, const int* ia, const int* ja, const double* r,
                const double* d, double* z) {
  const int64_t n = g->n;
  const int gs = grid_for(n);
  hipLaunchKernelGGL(k_gnn_scale_part, dim3(gs), dim3(kThreads), 0, c->stream, n, r, d,
                     g->part);
  hipLaunchKernelGGL(k_gnn_encode, dim3((unsigned)((n * kC + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, c->stream, n, r, g->sqd, g->part, gs, g->P, g->xin,
                     g->H[0], g->sc);
  const unsigned gnc = (unsigned)((n * kC + kThreads - 1) / kThreads);
  const unsigned gmlp = (unsigned)((n + 32 * (kThreads / 64) - 1) / (32 * (kThreads / 64)));
  for (int l = 0; l < kL; ++l) {
    hipLaunchKernelGGL(k_gnn_agg, dim3(gnc), dim3(kThreads), 0, c->stream, n, ia, ja,
                       g->ah, g->H[l], g->M[l]);
    hipLaunchKernelGGL(k_gnn_mlp, dim3(gmlp), dim3(kThreads), 0, c->stream, n, g->H[l],
                       g->M[l], g->Wb + l * kWsz, g->P + kOffW0 + l * kLayerSz + kWsz,
                       g->H[l + 1], g->Apre[l]);
  }
  hipLaunchKernelGGL(k_gnn_decode, dim3(grid_for(n) > 0 ? (unsigned)((n + kThreads - 1) / kThreads) : 1),
                     dim3(kThreads), 0, c->stream, n, g->H[kL], g->P, g->sqd, g->dec, r, d,
                     g->sc, z);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int gnn_backward(xg_ctx* c, GnnNet* g, const int* ia, const int* ja, const double* a,
                 const double* r, double etr, const double* z) {
  const int64_t n = g->n;
  const int gs = grid_for(n);
  const unsigned gnc = (unsigned)((n * kC + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_gnn_dz, dim3(gs), dim3(kThreads), 0, c->stream, n, ia, ja, a, z, r,
                     etr, g->sqd, g->dec, g->sc, g->dy, g->part);
  hipLaunchKernelGGL(k_gnn_loss, dim3(1), dim3(kThreads), 0, c->stream, g->part, gs, etr,
                     g->sc);
  const int gd = grid_for(n * kC);
  hipLaunchKernelGGL(k_gnn_dec_back, dim3(gd), dim3(kThreads), 0, c->stream, n, g->H[kL],
                     g->dy, g->P, g->dH, g->wpart);
  hipLaunchKernelGGL(k_gnn_sum_parts, dim3(1), dim3(64), 0, c->stream, g->wpart, gd,
                     kC + 1, kC + 1, g->G + kOffWo);
  float* dh = g->dH;
  float* dhp = g->dH2;
  const int64_t nchunk = (n + kChunk - 1) / kChunk;
  const int gl = (int)std::min<int64_t>(nchunk, kParts);
  for (int l = kL - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_gnn_layer_back, dim3(gl), dim3(kThreads), 0, c->stream, n,
                       g->H[l], g->M[l], g->Apre[l], g->P + kOffW0 + l * kLayerSz, dh,
                       dhp, g->dM, g->wpart);
    hipLaunchKernelGGL(k_gnn_sum_parts, dim3((kLayerSz + 255) / 256), dim3(256), 0,
                       c->stream, g->wpart, gl, kLayerSz, kLayerSz,
                       g->G + kOffW0 + l * kLayerSz);
    hipLaunchKernelGGL(k_gnn_agg_back, dim3(gnc), dim3(kThreads), 0, c->stream, n, ia, ja,
                       g->ah, g->dM, dhp);
    std::swap(dh, dhp);
  }
  hipLaunchKernelGGL(k_gnn_enc_back, dim3(gd), dim3(kThreads), 0, c->stream, n, dh,
                     g->xin, g->P, g->wpart);
  hipLaunchKernelGGL(k_gnn_sum_parts, dim3(1), dim3(64), 0, c->stream, g->wpart, gd,
                     2 * kC, 2 * kC, g->G + kOffWe);
  XG_HIP(c, hipGetLastError());
  ++g->pending;
  return XG_OK;
}

int read_sc(xg_ctx* c, GnnNet* g) {
  XG_HIP(c, hipMemcpyAsync(g->hsc, g->sc, sizeof(double) * 8, hipMemcpyDeviceToHost,
                           c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

}  // namespace

// ------------------------------------------------- interface for xg_cg.hip --

int xg_gnn_prepare_internal(xg_ctx* c, const int* ia, const int* ja, const double* a,
                            const double* d, int64_t n) {
  GnnNet* g = c->gnn;
  if (!g || g->n != n) return fail(c, XG_ENOMESH, "no learned preconditioner for this system");
  hipLaunchKernelGGL(k_gnn_prep, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, c->stream, n, ia, ja, a, d, g->ah, g->sqd, g->dec);
  XG_HIP(c, hipGetLastError());
  g->prepared = true;
  return XG_OK;
}

int xg_gnn_apply_internal(xg_ctx* c, const int* ia, const int* ja, const double* r,
                          const double* d, double* z) {
  GnnNet* g = c->gnn;
  int rc = gnn_forward(c, g, ia, ja, r, d, z);
  if (rc) return rc;
  const int gs = grid_for(g->n);
  hipLaunchKernelGGL(k_gnn_zr_part, dim3(gs), dim3(kThreads), 0, c->stream, g->n, z, r,
                     d, g->part);
  hipLaunchKernelGGL(k_gnn_fallback, dim3(gs), dim3(kThreads), 0, c->stream, g->n, z, r, d,
                     g->part, gs, g->sc);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int xg_gnn_fallback_flag(xg_ctx* c, int* used) {
  GnnNet* g = c->gnn;
  int rc = read_sc(c, g);
  if (rc) return rc;
  *used = g->hsc[2] != 0.0;
  return XG_OK;
}

void xg_gnn_free(xg_ctx* c) {
  GnnNet* g = c->gnn;
  if (!g) return;
  dev_free_all(g->allocs);
  if (g->hsc) (void)hipHostFree(g->hsc);
  delete g;
  c->gnn = nullptr;
}

int xg_gnn_forward_internal(xg_ctx* c, const int* ia, const int* ja, const double* r,
                            const double* d, double* z) {
  return gnn_forward(c, c->gnn, ia, ja, r, d, z);
}

int xg_gnn_backward_internal(xg_ctx* c, const int* ia, const int* ja, const double* a,
                             const double* r, double etr, const double* z, double* loss) {
  GnnNet* g = c->gnn;
  int rc = gnn_backward(c, g, ia, ja, a, r, etr, z);
  if (rc) return rc;
  if (loss) {
    if ((rc = read_sc(c, g))) return rc;
    *loss = g->hsc[1];
  }
  return XG_OK;
}

double* xg_gnn_zbuf(xg_ctx* c) { return c->gnn ? c->gnn->zbuf : nullptr; }

extern "C" {

int xg_gnn_create(xg_ctx* c, int64_t n, int64_t nnz) {
  if (!c) return XG_EINVAL;
  if (n <= 0 || nnz <= 0) return fail(c, XG_EINVAL, "xg_gnn_create: bad sizes");
  XG_HIP(c, hipSetDevice(c->device));
  auto* g = new GnnNet;
  g->n = n;
  g->nnz = nnz;
  int rc = gnn_alloc(c, g, n, nnz);
  if (rc) {
    dev_free_all(g->allocs);
    if (g->hsc) (void)hipHostFree(g->hsc);
    delete g;
    return rc;
  }
  xg_gnn_free(c);
  c->gnn = g;
  return XG_OK;
}

int xg_gnn_set_params(xg_ctx* c, const double* p, int64_t np) {
  if (!c || !c->gnn) return c ? fail(c, XG_ENOMESH, "no learned preconditioner") : XG_EINVAL;
  if (np != XG_GNN_NPARAM || !p) return fail(c, XG_EINVAL, "xg_gnn_set_params: size");
  std::vector<float> f(p, p + np);
  GnnNet* g = c->gnn;
  XG_HIP(c, hipMemcpy(g->P, f.data(), np * 4, hipMemcpyHostToDevice));
  XG_HIP(c, hipMemset(g->G, 0, np * 4));
  XG_HIP(c, hipMemset(g->m1, 0, np * 4));
  XG_HIP(c, hipMemset(g->m2, 0, np * 4));
  g->adam_t = 0;
  g->pending = 0;
  hipLaunchKernelGGL(k_gnn_weights_bf16, dim3((kL * kWsz + 255) / 256), dim3(256), 0,
                     c->stream, g->P, g->Wb);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

int xg_gnn_get_params(xg_ctx* c, double* p, double* grad, int64_t np) {
  if (!c || !c->gnn) return c ? fail(c, XG_ENOMESH, "no learned preconditioner") : XG_EINVAL;
  if (np != XG_GNN_NPARAM) return fail(c, XG_EINVAL, "xg_gnn_get_params: size");
  GnnNet* g = c->gnn;
  std::vector<float> f(np);
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (p) {
    XG_HIP(c, hipMemcpy(f.data(), g->P, np * 4, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < np; ++k) p[k] = f[k];
  }
  if (grad) {
    XG_HIP(c, hipMemcpy(f.data(), g->G, np * 4, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < np; ++k) grad[k] = f[k];
  }
  return XG_OK;
}

int xg_gnn_adam(xg_ctx* c, double lr, double beta1, double beta2, double eps) {
  if (!c || !c->gnn) return c ? fail(c, XG_ENOMESH, "no learned preconditioner") : XG_EINVAL;
  GnnNet* g = c->gnn;
  if (g->pending == 0) return XG_OK;
  ++g->adam_t;
  const float c1 = 1.0f - (float)std::pow(beta1, (double)g->adam_t);
  const float c2 = 1.0f - (float)std::pow(beta2, (double)g->adam_t);
  hipLaunchKernelGGL(k_gnn_adam, dim3((XG_GNN_NPARAM + 255) / 256), dim3(256), 0, c->stream,
                     g->P, g->G, g->m1, g->m2, XG_GNN_NPARAM, (float)lr, (float)beta1,
                     (float)beta2, (float)eps, c1, c2, 1.0f / (float)g->pending);
  hipLaunchKernelGGL(k_gnn_weights_bf16, dim3((kL * kWsz + 255) / 256), dim3(256), 0,
                     c->stream, g->P, g->Wb);
  XG_HIP(c, hipGetLastError());
  g->pending = 0;
  return XG_OK;
}

}  // extern "C"
