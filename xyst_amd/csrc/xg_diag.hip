// MI355X device path of the node diagnostics (NodeDiagnostics.cpp): the
// volume-weighted sums the solvers hand to Transporter::diagnostics, which
// divides them by the total volume and takes square roots.
//
// Deterministic two-pass reduction: each of at most kRedBlocks workgroups
// sums a fixed strided subset of the nodes (wave shuffles + a fixed tree),
// then one workgroup adds the workgroup partials in index order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "xg_ctx.hpp"

using namespace xgd;

// Scratch memory of xg_diag_rho_dev_noalloc: the workgroup partials of at
// most kScratchSums sums over at most kRedBlocks workgroups (90,112 bytes),
// one copy per GPU, held until the process ends.  Host code never passes
// `scratch` itself to a kernel: its device address on each GPU comes from
// hipGetSymbolAddress (xg_diag_scratch_init).
constexpr int kScratchSums = 11;
constexpr size_t kScratchDoubles = (size_t)kScratchSums * kRedBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Called by xg_ctx_create with `device` current: records the device address
// of this GPU's `scratch` copy once per device.
int xg_diag_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return XG_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return XG_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return XG_EHIP;
  g_scratch_ptr[device] = static_cast<double*>(p);
  return XG_OK;
}

namespace {

double* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

__device__ inline double block_sum(double v, double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) s += sm[k];
  return s;
}

// rhocompute's sums (NodeDiagnostics.cpp:96-102) for ncomp = NC: U_c^2 V,
// (U_c - Un_c)^2 V and U_4 V (NC >= 5), into part[q * nb + block]
template <int NC>
__global__ __launch_bounds__(kBlock) void k_diag_rho(int64_t n, const double* __restrict__ U,
                                                     const double* __restrict__ Un,
                                                     const double* __restrict__ V,
                                                     double* __restrict__ part) {
  constexpr int NS = 2 * NC + 1;
  __shared__ double sm[kBlock / 64];
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double v = V[i];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double u = U[i * NC + c];
      const double d = u - Un[i * NC + c];
      acc[c] += u * u * v;
      acc[NC + c] += d * d * v;
    }
    if (NC >= 5) acc[2 * NC] += U[i * NC + 4] * v;
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const double s = block_sum(acc[q], sm);
    if (threadIdx.x == 0) part[(size_t)q * gridDim.x + blockIdx.x] = s;
  }
}

// precompute's sums (NodeDiagnostics.cpp:224-236): p^2 V, u_c^2 V, dp^2 V,
// (u_c - un_c)^2 V, NC components of U per node
template <int NC>
__global__ __launch_bounds__(kBlock) void k_diag_pre(int64_t n, const double* __restrict__ U,
                                                     const double* __restrict__ Un,
                                                     const double* __restrict__ P,
                                                     const double* __restrict__ dP,
                                                     const double* __restrict__ V,
                                                     double* __restrict__ part) {
  constexpr int NS = 2 * NC + 2;
  __shared__ double sm[kBlock / 64];
  double acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) acc[q] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double v = V[i];
    acc[0] += P[i] * P[i] * v;
    acc[NC + 1] += dP[i] * dP[i] * v;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double u = U[i * NC + c];
      const double d = u - Un[i * NC + c];
      acc[1 + c] += u * u * v;
      acc[NC + 2 + c] += d * d * v;
    }
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const double s = block_sum(acc[q], sm);
    if (threadIdx.x == 0) part[(size_t)q * gridDim.x + blockIdx.x] = s;
  }
}

// rhocompute's analytic-error sums (NodeDiagnostics.cpp:112-128): the
// numerical solution in primitive form (the scalars as they are) against
// An, (nu_c - An_c)^2 V and |nu_c - An_c| V, NC components per node; NC = 1,
// 3, 4: precompute's pressure, velocity and scalar errors
// (NodeDiagnostics.cpp:237-250), the values as they are
template <int NC>
__global__ __launch_bounds__(kBlock) void k_diag_err(int64_t n, const double* __restrict__ U,
                                                     const double* __restrict__ V,
                                                     const double* __restrict__ An,
                                                     double* __restrict__ part) {
  __shared__ double sm[kBlock / 64];
  double acc[2 * NC];
#pragma unroll
  for (int q = 0; q < 2 * NC; ++q) acc[q] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double v = V[i];
    double nu[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) nu[c] = U[i * NC + c];
    if constexpr (NC >= 5) {
      nu[1] /= nu[0];
      nu[2] /= nu[0];
      nu[3] /= nu[0];
      nu[4] = nu[4] / nu[0] - 0.5 * (nu[1] * nu[1] + nu[2] * nu[2] + nu[3] * nu[3]);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double du = nu[c] - An[i * NC + c];
      acc[c] += du * du * v;
      acc[NC + c] += fabs(du) * v;
    }
  }
#pragma unroll
  for (int q = 0; q < 2 * NC; ++q) {
    const double s = block_sum(acc[q], sm);
    if (threadIdx.x == 0) part[(size_t)q * gridDim.x + blockIdx.x] = s;
  }
}

// out[q] = sum of part[q * nb .. q * nb + nb) in index order
__global__ __launch_bounds__(kBlock) void k_diag_final(int ns, int nb,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ out) {
  __shared__ double sm[kBlock / 64];
  for (int q = 0; q < ns; ++q) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += kBlock) v += part[(size_t)q * nb + i];
    const double s = block_sum(v, sm);
    if (threadIdx.x == 0) out[q] = s;
  }
}

int diag_buffers(xg_ctx* c, int ns, int nb, double** part, double** out) {
  const size_t bytes = sizeof(double) * ((size_t)ns * nb + ns);
  if (c->diagbuf.bytes < bytes) {
    if (c->diagbuf.p) (void)hipFree(c->diagbuf.p);
    c->diagbuf.p = nullptr;
    c->diagbuf.bytes = 0;
    if (hipMalloc(&c->diagbuf.p, bytes) != hipSuccess)
      return fail(c, XG_EHIP, "hipMalloc (diagnostics)");
    c->diagbuf.bytes = bytes;
  }
  *part = (double*)c->diagbuf.p;
  *out = *part + (size_t)ns * nb;
  return XG_OK;
}

int finish(xg_ctx* c, int ns, int nb, double* part, double* dout, double* out) {
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, c->stream, ns, nb, part, dout);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipMemcpyAsync(out, dout, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

}  // namespace

extern "C" {

int xg_diag_rho_dev(xg_ctx* c, int64_t npoin, int ncomp, const double* U, const double* Un,
                    const double* V, double* out) {
  if (!c) return XG_EINVAL;
  if (npoin < 0 || !U || !Un || !V || !out)
    return fail(c, XG_EINVAL, "xg_diag_rho: bad arguments");
  if (ncomp != 1 && ncomp != 3 && (ncomp < 5 || ncomp > XG_MAX_NCOMP))
    return fail(c, XG_EUNSUPPORTED, "xg_diag_rho: ncomp must be 1, 3 or 5..XG_MAX_NCOMP");
  XG_HIP(c, hipSetDevice(c->device));
  const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(npoin)));
  const int ns = 2 * ncomp + 1;
  double *part, *dout;
  int rc = diag_buffers(c, ns, nb, &part, &dout);
  if (rc) return rc;
  static_assert(XG_MAX_NCOMP == 7, "one instantiation per component count");
  if (ncomp == 7)
    hipLaunchKernelGGL(k_diag_rho<7>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_diag_rho<6>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  else if (ncomp == 5)
    hipLaunchKernelGGL(k_diag_rho<5>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  else if (ncomp == 3)
    hipLaunchKernelGGL(k_diag_rho<3>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  else
    hipLaunchKernelGGL(k_diag_rho<1>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  return finish(c, ns, nb, part, dout, out);
}

int xg_diag_rho_dev_noalloc(xg_ctx* c, int64_t npoin, int ncomp, const double* U,
                            const double* Un, const double* V, double* out) {
  if (!c) return XG_EINVAL;
  if (npoin < 0 || !U || !Un || !V || !out)
    return fail(c, XG_EINVAL, "xg_diag_rho_noalloc: bad arguments");
  if (ncomp != 5 && ncomp != 3 && ncomp != 1)
    return fail(c, XG_EUNSUPPORTED, "xg_diag_rho_noalloc: ncomp must be 1, 3 or 5");
  const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(npoin)));
  const int ns = 2 * ncomp + 1;
  if ((size_t)ns * nb > kScratchDoubles)
    return fail(c, XG_EINVAL, "xg_diag_rho_noalloc: more partial sums than the "
                              "scratch array holds (11 x 1024)");
  double* part = scratch_ptr(c->device);
  if (!part)
    return fail(c, XG_EHIP, "xg_diag_rho_noalloc: no scratch address for this device");
  XG_HIP(c, hipSetDevice(c->device));
  if (ncomp == 5)
    hipLaunchKernelGGL(k_diag_rho<5>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  else if (ncomp == 3)
    hipLaunchKernelGGL(k_diag_rho<3>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  else
    hipLaunchKernelGGL(k_diag_rho<1>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, V, part);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, c->stream, ns, nb, part, out);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int xg_diag_err_dev(xg_ctx* c, int64_t npoin, int ncomp, const double* U, const double* V,
                    const double* An, double* out) {
  if (!c) return XG_EINVAL;
  if (npoin < 0 || !U || !V || !An || !out)
    return fail(c, XG_EINVAL, "xg_diag_err: bad arguments");
  if (ncomp != 1 && ncomp != 3 && ncomp != 4 && (ncomp < 5 || ncomp > XG_MAX_NCOMP))
    return fail(c, XG_EUNSUPPORTED, "xg_diag_err: ncomp must be 1, 3, 4 or 5..XG_MAX_NCOMP");
  XG_HIP(c, hipSetDevice(c->device));
  const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(npoin)));
  const int ns = 2 * ncomp;
  double *part, *dout;
  int rc = diag_buffers(c, ns, nb, &part, &dout);
  if (rc) return rc;
  if (ncomp == 1)
    hipLaunchKernelGGL(k_diag_err<1>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, V, An, part);
  else if (ncomp == 3)
    hipLaunchKernelGGL(k_diag_err<3>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, V, An, part);
  else if (ncomp == 4)
    hipLaunchKernelGGL(k_diag_err<4>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, V, An, part);
  else if (ncomp == 5)
    hipLaunchKernelGGL(k_diag_err<5>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, V, An, part);
  else if (ncomp == 6)
    hipLaunchKernelGGL(k_diag_err<6>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, V, An, part);
  else
    hipLaunchKernelGGL(k_diag_err<7>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, V, An, part);
  return finish(c, ns, nb, part, dout, out);
}

int xg_diag_pre_nc_dev(xg_ctx* c, int64_t npoin, int ncomp, const double* U,
                       const double* Un, const double* P, const double* dP, const double* V,
                       double* out) {
  if (!c) return XG_EINVAL;
  if (npoin < 0 || !U || !Un || !P || !dP || !V || !out)
    return fail(c, XG_EINVAL, "xg_diag_pre: bad arguments");
  if (ncomp != 3 && ncomp != 4)
    return fail(c, XG_EUNSUPPORTED, "xg_diag_pre: ncomp must be 3 or 4");
  XG_HIP(c, hipSetDevice(c->device));
  const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(npoin)));
  const int ns = 2 * ncomp + 2;
  double *part, *dout;
  int rc = diag_buffers(c, ns, nb, &part, &dout);
  if (rc) return rc;
  if (ncomp == 3)
    hipLaunchKernelGGL(k_diag_pre<3>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, P,
                       dP, V, part);
  else
    hipLaunchKernelGGL(k_diag_pre<4>, dim3(nb), dim3(kBlock), 0, c->stream, npoin, U, Un, P,
                       dP, V, part);
  return finish(c, ns, nb, part, dout, out);
}

int xg_diag_pre_dev(xg_ctx* c, int64_t npoin, const double* U, const double* Un,
                    const double* P, const double* dP, const double* V, double* out) {
  return xg_diag_pre_nc_dev(c, npoin, 3, U, Un, P, dP, V, out);
}

}  // extern "C"
