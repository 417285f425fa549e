// Device side of the smoothed-aggregation AMG preconditioner (no
// counterpart in the reference; setup in xg_amg.cpp).
//
// The hierarchy lives in HBM in FP32 (values) with int32 CSR indices: per
// level the operator B_l, 1/diag(B_l), the prolongator P_l and R_l = P_l^T;
// the coarsest operator's dense inverse.  One application is a symmetric
// V-cycle on B = S A (S = sign(diag A)) for the right-hand side S r:
// damped-Jacobi pre-smoothing from zero, residual, restriction, coarse
// correction, prolongation, the same number of post-smoothing sweeps.  Every
// kernel has a fixed summation order (8 lanes per CSR row, fixed butterfly),
// so applications are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "xg_amg.hpp"
#include "xg_ctx.hpp"

using namespace xgd;

struct AmgDevLevel {
  int n = 0;
  int *ia = nullptr, *ja = nullptr;
  float *a = nullptr, *dinv = nullptr;
  float omega = 0.f;
  int *pia = nullptr, *pja = nullptr, *ria = nullptr, *rja = nullptr;
  float *pa = nullptr, *ra = nullptr;
  float *x = nullptr, *x2 = nullptr, *b = nullptr, *r = nullptr;
};

struct AmgDev {
  int64_t version = -1;
  std::vector<AmgDevLevel> lev;
  float* cinv = nullptr;
  float* sign = nullptr;        // S of the fine level
  std::vector<DevArr> allocs;
  double op_complexity = 0.0, setup_ms = 0.0;
  int nu = 1;                   // smoothing sweeps before and after
};

namespace {

constexpr int kT = 256;
constexpr int kLanes = 8;

__device__ inline float row8(float v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

__device__ inline float row_dot(const int* __restrict__ ia, const int* __restrict__ ja,
                                const float* __restrict__ a, const float* __restrict__ v,
                                int row, int lane) {
  float s = 0.f;
  for (int k = ia[row] + lane; k < ia[row + 1]; k += kLanes) s += a[k] * v[ja[k]];
  return row8(s);
}

int grid_rows(int n) {
  int64_t g = ((int64_t)n * kLanes + kT - 1) / kT;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
}

// x = w D^-1 b (first sweep from zero)
__global__ __launch_bounds__(kT) void k_amg_jacobi0(int n, float w,
                                                   const float* __restrict__ dinv,
                                                   const float* __restrict__ b,
                                                   float* __restrict__ x) {
  int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) x[i] = w * dinv[i] * b[i];
}

// xo = xi + w D^-1 (b - B xi)
__global__ __launch_bounds__(kT) void k_amg_jacobi(int n, const int* __restrict__ ia,
                                                  const int* __restrict__ ja,
                                                  const float* __restrict__ a,
                                                  const float* __restrict__ dinv, float w,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ xi,
                                                  float* __restrict__ xo) {
  const int lane = threadIdx.x & (kLanes - 1);
  for (int64_t row = ((int64_t)blockIdx.x * kT + threadIdx.x) / kLanes; row < n;
       row += (int64_t)gridDim.x * kT / kLanes) {
    const float s = row_dot(ia, ja, a, xi, (int)row, lane);
    if (lane == 0) xo[row] = xi[row] + w * dinv[row] * (b[row] - s);
  }
}

// r = b - B x
__global__ __launch_bounds__(kT) void k_amg_residual(int n, const int* __restrict__ ia,
                                                    const int* __restrict__ ja,
                                                    const float* __restrict__ a,
                                                    const float* __restrict__ b,
                                                    const float* __restrict__ x,
                                                    float* __restrict__ r) {
  const int lane = threadIdx.x & (kLanes - 1);
  for (int64_t row = ((int64_t)blockIdx.x * kT + threadIdx.x) / kLanes; row < n;
       row += (int64_t)gridDim.x * kT / kLanes) {
    const float s = row_dot(ia, ja, a, x, (int)row, lane);
    if (lane == 0) r[row] = b[row] - s;
  }
}

// y = M v (restriction), or y += M v (prolongation, add = 1)
__global__ __launch_bounds__(kT) void k_amg_spmv(int n, const int* __restrict__ ia,
                                                const int* __restrict__ ja,
                                                const float* __restrict__ a,
                                                const float* __restrict__ v,
                                                float* __restrict__ y, int add) {
  const int lane = threadIdx.x & (kLanes - 1);
  for (int64_t row = ((int64_t)blockIdx.x * kT + threadIdx.x) / kLanes; row < n;
       row += (int64_t)gridDim.x * kT / kLanes) {
    const float s = row_dot(ia, ja, a, v, (int)row, lane);
    if (lane == 0) y[row] = add ? y[row] + s : s;
  }
}

// x = C b with C the dense coarsest inverse (one workgroup per row)
__global__ __launch_bounds__(kT) void k_amg_dense(int n, const float* __restrict__ C,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ x) {
  __shared__ float sm[kT / 64];
  const int row = blockIdx.x;
  float s = 0.f;
  for (int j = threadIdx.x; j < n; j += kT) s += C[(size_t)row * n + j] * b[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < kT / 64; ++k) t += sm[k];
    x[row] = t;
  }
}

__global__ __launch_bounds__(kT) void k_amg_in(int n, const float* __restrict__ sgn,
                                              const double* __restrict__ r,
                                              float* __restrict__ b) {
  int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) b[i] = sgn[i] * (float)r[i];
}

__global__ __launch_bounds__(kT) void k_amg_out(int n, const float* __restrict__ x,
                                               double* __restrict__ z) {
  int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) z[i] = (double)x[i];
}

int nb(int n) { return std::max(1, (n + kT - 1) / kT); }

template <class T>
int up(xg_ctx* c, std::vector<DevArr>& L, T** d, const std::vector<T>& v) {
  return upload(c, L, d, v);
}

std::vector<int> to_i32(const std::vector<int64_t>& v) { return std::vector<int>(v.begin(), v.end()); }
std::vector<float> to_f32(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

void free_dev(AmgDev* m) {
  if (!m) return;
  dev_free_all(m->allocs);
  delete m;
}

void vcycle(xg_ctx* c, AmgDev* m, size_t l) {
  AmgDevLevel& L = m->lev[l];
  if (l + 1 == m->lev.size()) {
    hipLaunchKernelGGL(k_amg_dense, dim3(L.n), dim3(kT), 0, c->stream, L.n, m->cinv, L.b, L.x);
    return;
  }
  AmgDevLevel& C = m->lev[l + 1];
  const int gr = grid_rows(L.n);
  hipLaunchKernelGGL(k_amg_jacobi0, dim3(nb(L.n)), dim3(kT), 0, c->stream, L.n, L.omega,
                     L.dinv, L.b, L.x);
  for (int s = 1; s < m->nu; ++s) {
    hipLaunchKernelGGL(k_amg_jacobi, dim3(gr), dim3(kT), 0, c->stream, L.n, L.ia, L.ja, L.a,
                       L.dinv, L.omega, L.b, L.x, L.x2);
    std::swap(L.x, L.x2);
  }
  hipLaunchKernelGGL(k_amg_residual, dim3(gr), dim3(kT), 0, c->stream, L.n, L.ia, L.ja, L.a,
                     L.b, L.x, L.r);
  hipLaunchKernelGGL(k_amg_spmv, dim3(grid_rows(C.n)), dim3(kT), 0, c->stream, C.n, L.ria,
                     L.rja, L.ra, L.r, C.b, 0);
  vcycle(c, m, l + 1);
  hipLaunchKernelGGL(k_amg_spmv, dim3(gr), dim3(kT), 0, c->stream, L.n, L.pia, L.pja, L.pa,
                     C.x, L.x, 1);
  for (int s = 0; s < m->nu; ++s) {
    hipLaunchKernelGGL(k_amg_jacobi, dim3(gr), dim3(kT), 0, c->stream, L.n, L.ia, L.ja, L.a,
                       L.dinv, L.omega, L.b, L.x, L.x2);
    std::swap(L.x, L.x2);
  }
}

}  // namespace

void xg_amg_free(xg_ctx* c) {
  free_dev(c->amg);
  c->amg = nullptr;
}

int xg_amg_prepare_internal(xg_ctx* c, int64_t n, const int64_t* ia, const int32_t* ja,
                            const double* a, int64_t version) {
  if (c->amg && c->amg->version == version) return XG_OK;
  if (n >= (1ll << 31) || ia[n] >= (1ll << 31))
    return fail(c, XG_EUNSUPPORTED, "AMG: more than 2^31 rows or nonzeros");
  xg::AmgHierarchy h;
  xg::AmgOptions opt;
  std::string err;
  int rc = xg::amg_setup(n, ia, ja, a, opt, h, err);
  if (rc) return fail(c, rc, err);
  XG_HIP(c, hipStreamSynchronize(c->stream));
  xg_amg_free(c);
  AmgDev* m = new AmgDev;
  auto& L = m->allocs;
  std::vector<float> sg(n);
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t p = ia[i]; p < ia[i + 1]; ++p)
      if (ja[p] == i) d = a[p];
    sg[i] = d < 0.0 ? -1.f : 1.f;
  }
  if ((rc = up(c, L, &m->sign, sg))) { free_dev(m); return rc; }
  m->lev.resize(h.lev.size());
  for (size_t l = 0; l < h.lev.size(); ++l) {
    const xg::AmgLevel& H = h.lev[l];
    AmgDevLevel& D = m->lev[l];
    D.n = (int)H.B.nrow;
    D.omega = (float)H.omega;
    const size_t nn = (size_t)std::max(1, D.n);
    if ((rc = up(c, L, &D.ia, to_i32(H.B.ia))) || (rc = up(c, L, &D.ja, H.B.ja)) ||
        (rc = up(c, L, &D.a, to_f32(H.B.a))) || (rc = up(c, L, &D.dinv, to_f32(H.dinv))) ||
        (rc = dev_alloc(c, L, (void**)&D.x, nn * 4)) ||
        (rc = dev_alloc(c, L, (void**)&D.x2, nn * 4)) ||
        (rc = dev_alloc(c, L, (void**)&D.b, nn * 4)) ||
        (rc = dev_alloc(c, L, (void**)&D.r, nn * 4))) {
      free_dev(m);
      return rc;
    }
    if (l + 1 < h.lev.size() &&
        ((rc = up(c, L, &D.pia, to_i32(H.P.ia))) || (rc = up(c, L, &D.pja, H.P.ja)) ||
         (rc = up(c, L, &D.pa, to_f32(H.P.a))) || (rc = up(c, L, &D.ria, to_i32(H.R.ia))) ||
         (rc = up(c, L, &D.rja, H.R.ja)) || (rc = up(c, L, &D.ra, to_f32(H.R.a))))) {
      free_dev(m);
      return rc;
    }
  }
  if ((rc = up(c, L, &m->cinv, to_f32(h.cinv)))) { free_dev(m); return rc; }
  m->version = version;
  m->op_complexity = h.op_complexity;
  m->setup_ms = h.setup_ms;
  c->amg = m;
  return XG_OK;
}

int xg_amg_apply_internal(xg_ctx* c, const double* r, double* z) {
  AmgDev* m = c->amg;
  if (!m) return fail(c, XG_ENOMESH, "AMG hierarchy not built");
  AmgDevLevel& F = m->lev[0];
  hipLaunchKernelGGL(k_amg_in, dim3(nb(F.n)), dim3(kT), 0, c->stream, F.n, m->sign, r, F.b);
  vcycle(c, m, 0);
  hipLaunchKernelGGL(k_amg_out, dim3(nb(F.n)), dim3(kT), 0, c->stream, F.n, F.x, z);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

extern "C" int xg_amg_info(xg_ctx* c, int* levels, double* op_complexity, double* setup_ms) {
  if (!c) return XG_EINVAL;
  if (!c->amg) return fail(c, XG_ENOMESH, "AMG hierarchy not built");
  if (levels) *levels = (int)c->amg->lev.size();
  if (op_complexity) *op_complexity = c->amg->op_complexity;
  if (setup_ms) *setup_ms = c->amg->setup_ms;
  return XG_OK;
}
