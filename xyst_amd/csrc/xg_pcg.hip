// Several partitions of one linear system on one device: the chares of a
// ConjugateGradients array that share a GPU (ConjugateGradients.cpp:
// 105-823).  Partition k holds the partial rows of its own elements (A_k),
// b_k with the full values at shared nodes, and x_k; its local-to-global
// node map gid_k takes the place of the node communication map.
//
// Layout: the partitions' rows are concatenated (partition-major), their
// matrices form one block-diagonal CSR (int32 indices, FP64 values).  The
// exchanges of the reference become one kernel over the shared rows: every
// copy of a shared (node, component) row receives the sum of all copies in
// partition order (comres, comd, comq: partial sums completed), or that sum
// divided by the number of copies (comx/x: the solution averaged).  Dot
// products weight each row by 1 or 0 (a node counts in the highest
// partition holding it, tk::slave, Reorder.cpp:393-412) and reduce with a
// fixed two-pass order.  The host drives the recurrence and reads the dot
// products back: this path serves several chares of one GPU, not speed (the
// single-partition solver is xg_cg_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/xg.h"

namespace {

constexpr int kT = 256;
constexpr int kParts = 256;   // workgroups of the dot-product pass

__global__ __launch_bounds__(kT) void k_pcg_spmv(int64_t n, const int* __restrict__ ia,
                                                 const int* __restrict__ ja,
                                                 const double* __restrict__ a,
                                                 const double* __restrict__ v,
                                                 double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = ia[i]; j < ia[i + 1]; ++j) s += a[j] * v[ja[j]];
  out[i] = s;
}

// every copy of shared row s gets the sum of its copies (partition order),
// divided by their count when avg
__global__ __launch_bounds__(kT) void k_pcg_complete(int64_t nsh, const int* __restrict__ off,
                                                     const int* __restrict__ row,
                                                     double* __restrict__ v, int avg) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= nsh) return;
  double t = 0.0;
  for (int k = off[s]; k < off[s + 1]; ++k) t += v[row[k]];
  if (avg) t /= (double)(off[s + 1] - off[s]);
  for (int k = off[s]; k < off[s + 1]; ++k) v[row[k]] = t;
}

__device__ inline double block_sum(double v, double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kT / 64; ++k) s += sm[k];
  return s;
}

// part[blk] = sum w a1 b1, part[kParts + blk] = sum w a2 b2 over a fixed
// strided subset of the rows
__global__ __launch_bounds__(kT) void k_pcg_dot2(int64_t n, const double* __restrict__ w,
                                                 const double* __restrict__ a1,
                                                 const double* __restrict__ b1,
                                                 const double* __restrict__ a2,
                                                 const double* __restrict__ b2,
                                                 double* __restrict__ part) {
  __shared__ double sm[kT / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)kParts * kT) {
    const double wi = w[i];
    s1 += wi * a1[i] * b1[i];
    s2 += wi * a2[i] * b2[i];
  }
  s1 = block_sum(s1, sm);
  s2 = block_sum(s2, sm);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[kParts + blockIdx.x] = s2;
  }
}

__global__ __launch_bounds__(kT) void k_pcg_final(const double* __restrict__ part,
                                                  double* __restrict__ out) {
  __shared__ double sm[kT / 64];
  const double a = block_sum(part[threadIdx.x], sm);
  const double b = block_sum(part[kParts + threadIdx.x], sm);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

// setup: r = b - r (r held A x), p = r; d = the row's diagonal (Jacobi) or
// 1/count of its node (no preconditioner), completed by the exchange after
__global__ __launch_bounds__(kT) void k_pcg_init(int64_t n, const int* __restrict__ ia,
                                                 const int* __restrict__ ja,
                                                 const double* __restrict__ a,
                                                 const double* __restrict__ b,
                                                 const double* __restrict__ cnt, int jacobi,
                                                 double* __restrict__ r, double* __restrict__ p,
                                                 double* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double ri = -r[i] + b[i];
  r[i] = ri;
  p[i] = ri;
  double di = 1.0 / cnt[i];
  if (jacobi) {
    di = 0.0;
    for (int j = ia[i]; j < ia[i + 1]; ++j)
      if (ja[j] == i) di = a[j];
  }
  d[i] = di;
}

__global__ __launch_bounds__(kT) void k_pcg_div(int64_t n, const double* __restrict__ r,
                                                const double* __restrict__ d,
                                                double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) z[i] = r[i] / d[i];
}

// p = z + alpha p
__global__ __launch_bounds__(kT) void k_pcg_dir(int64_t n, double alpha,
                                                const double* __restrict__ z,
                                                double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = z[i] + alpha * p[i];
}

// r -= alpha q, z = r / d
__global__ __launch_bounds__(kT) void k_pcg_res(int64_t n, double alpha,
                                                const double* __restrict__ q,
                                                const double* __restrict__ d,
                                                double* __restrict__ r, double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double ri = r[i] - alpha * q[i];
  r[i] = ri;
  z[i] = ri / d[i];
}

// x += alpha p
__global__ __launch_bounds__(kT) void k_pcg_sol(int64_t n, double alpha,
                                                const double* __restrict__ p,
                                                double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) x[i] += alpha * p[i];
}

unsigned grid(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kT - 1) / kT); }

}  // namespace

struct xg_pcg {
  int device = 0, nparts = 0, ncomp = 1;
  std::vector<int64_t> rows, roff;   // rows per partition, row offsets
  int64_t nrow = 0, nsh = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int *ia = nullptr, *ja = nullptr, *shoff = nullptr, *shrow = nullptr;
  double *a = nullptr, *w = nullptr, *cnt = nullptr;
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *p = nullptr,
         *q = nullptr, *d = nullptr, *part = nullptr, *out = nullptr;
  double normb = 0.0, rho = 0.0;
  bool ready = false;
  std::string err;
};

namespace {

int pfail(xg_pcg* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

template <class T>
int palloc(xg_pcg* s, T** dst, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess)
    return pfail(s, XG_EHIP, "xg_pcg: hipMalloc");
  s->allocs.push_back(p);
  *dst = static_cast<T*>(p);
  return XG_OK;
}

template <class T>
int pput(xg_pcg* s, T** dst, const std::vector<T>& v) {
  int rc = palloc(s, dst, v.size());
  if (rc) return rc;
  if (!v.empty() && hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) !=
                        hipSuccess)
    return pfail(s, XG_EHIP, "xg_pcg: hipMemcpy");
  return XG_OK;
}

// the two dot products, read back (the reference's reductions)
int dot2(xg_pcg* s, const double* a1, const double* b1, const double* a2, const double* b2,
         double out[2]) {
  hipLaunchKernelGGL(k_pcg_dot2, dim3(kParts), dim3(kT), 0, s->stream, s->nrow, s->w, a1, b1,
                     a2, b2, s->part);
  hipLaunchKernelGGL(k_pcg_final, dim3(1), dim3(kT), 0, s->stream, s->part, s->out);
  if (hipGetLastError() != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg: kernel launch");
  if (hipMemcpyAsync(out, s->out, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream) !=
          hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return pfail(s, XG_EHIP, "xg_pcg: reading a dot product");
  return XG_OK;
}

void complete(xg_pcg* s, double* v, int avg) {
  if (s->nsh)
    hipLaunchKernelGGL(k_pcg_complete, dim3(grid(s->nsh)), dim3(kT), 0, s->stream, s->nsh,
                       s->shoff, s->shrow, v, avg);
}

}  // namespace

extern "C" {

int xg_pcg_create(int device, int nparts, int ncomp, const int64_t* npoin,
                  const size_t* const* gid, const size_t* const* ia, const size_t* const* ja,
                  const double* const* a, int one_based, xg_pcg** out) {
  if (!out) return XG_EINVAL;
  xg_pcg* s = new xg_pcg;
  *out = s;
  if (nparts < 1 || ncomp < 1 || ncomp > 8 || !npoin || !gid || !ia || !ja || !a)
    return pfail(s, XG_EINVAL, "xg_pcg_create: bad arguments");
  s->device = device;
  s->nparts = nparts;
  s->ncomp = ncomp;
  const int base = one_based ? 1 : 0;
  std::vector<int> hia(1, 0), hja;
  std::vector<double> ha;
  std::map<size_t, std::vector<int64_t>> copies;   // global node -> first row of each copy
  s->roff.assign(1, 0);
  for (int k = 0; k < nparts; ++k) {
    const int64_t n = npoin[k] * ncomp;
    if (npoin[k] < 0 || (npoin[k] && (!gid[k] || !ia[k] || !ja[k] || !a[k])))
      return pfail(s, XG_EINVAL, "xg_pcg_create: bad partition arrays");
    const int64_t r0 = s->roff.back();
    if (r0 + n >= (int64_t)1 << 31) return pfail(s, XG_EINVAL, "xg_pcg_create: too many rows");
    for (int64_t i = 0; i < n; ++i) {
      if (ia[k][i + 1] < ia[k][i]) return pfail(s, XG_EINVAL, "xg_pcg_create: bad ia");
      for (size_t j = ia[k][i] - base; j < ia[k][i + 1] - base; ++j) {
        const size_t c = ja[k][j] - base;
        if (c >= (size_t)n) return pfail(s, XG_EINVAL, "xg_pcg_create: column out of range");
        hja.push_back((int)(r0 + c));
        ha.push_back(a[k][j]);
      }
      hia.push_back((int)hja.size());
    }
    for (int64_t i = 0; i < npoin[k]; ++i) {
      auto& v = copies[gid[k][i]];
      if (!v.empty() && v.back() >= r0)
        return pfail(s, XG_EINVAL, "xg_pcg_create: a global id repeats in a partition");
      v.push_back(r0 + i * ncomp);
    }
    s->rows.push_back(n);
    s->roff.push_back(r0 + n);
  }
  s->nrow = s->roff.back();
  // row weights (1 in the highest partition holding the node) and counts
  std::vector<double> w(s->nrow, 0.0), cnt(s->nrow, 1.0);
  std::vector<int> shoff(1, 0), shrow;
  for (const auto& [g, v] : copies) {
    for (int c = 0; c < ncomp; ++c) {
      w[v.back() + c] = 1.0;
      for (int64_t r : v) cnt[r + c] = (double)v.size();
      if (v.size() > 1) {
        for (int64_t r : v) shrow.push_back((int)(r + c));
        shoff.push_back((int)shrow.size());
      }
    }
  }
  s->nsh = (int64_t)shoff.size() - 1;
  if (hipSetDevice(device) != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg_create: hipSetDevice");
  int rc;
  if ((rc = pput(s, &s->ia, hia)) || (rc = pput(s, &s->ja, hja)) || (rc = pput(s, &s->a, ha)) ||
      (rc = pput(s, &s->w, w)) || (rc = pput(s, &s->cnt, cnt)) ||
      (rc = pput(s, &s->shoff, shoff)) || (rc = pput(s, &s->shrow, shrow)))
    return rc;
  const size_t n = (size_t)s->nrow;
  if ((rc = palloc(s, &s->x, n)) || (rc = palloc(s, &s->b, n)) || (rc = palloc(s, &s->r, n)) ||
      (rc = palloc(s, &s->z, n)) || (rc = palloc(s, &s->p, n)) || (rc = palloc(s, &s->q, n)) ||
      (rc = palloc(s, &s->d, n)) || (rc = palloc(s, &s->part, 2 * kParts)) ||
      (rc = palloc(s, &s->out, 2)))
    return rc;
  return XG_OK;
}

int xg_pcg_set_stream(xg_pcg* s, void* stream) {
  if (!s) return XG_EINVAL;
  s->stream = (hipStream_t)stream;
  return XG_OK;
}

int xg_pcg_setup(xg_pcg* s, const double* const* x, const double* const* b, int pc,
                 double* normb) {
  if (!s) return XG_EINVAL;
  if (!s->x) return pfail(s, XG_EINVAL, "xg_pcg_setup: not created");
  if (!b || (pc != 0 && pc != 1)) return pfail(s, XG_EINVAL, "xg_pcg_setup: bad arguments");
  if (hipSetDevice(s->device) != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg: hipSetDevice");
  for (int k = 0; k < s->nparts; ++k) {
    const size_t n = (size_t)s->rows[k], o = (size_t)s->roff[k];
    if (!n) continue;
    if (!b[k]) return pfail(s, XG_EINVAL, "xg_pcg_setup: null b of a partition");
    if (hipMemcpyAsync(s->b + o, b[k], n * 8, hipMemcpyHostToDevice, s->stream) != hipSuccess)
      return pfail(s, XG_EHIP, "xg_pcg_setup: hipMemcpy");
    hipError_t e = (x && x[k])
                       ? hipMemcpyAsync(s->x + o, x[k], n * 8, hipMemcpyHostToDevice, s->stream)
                       : hipMemsetAsync(s->x + o, 0, n * 8, s->stream);
    if (e != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg_setup: hipMemcpy");
  }
  // r = A x completed at shared rows, r = b - r, p = r, d completed, z = r/d
  const unsigned g = grid(s->nrow);
  hipLaunchKernelGGL(k_pcg_spmv, dim3(g), dim3(kT), 0, s->stream, s->nrow, s->ia, s->ja, s->a,
                     s->x, s->r);
  complete(s, s->r, 0);
  hipLaunchKernelGGL(k_pcg_init, dim3(g), dim3(kT), 0, s->stream, s->nrow, s->ia, s->ja, s->a,
                     s->b, s->cnt, pc, s->r, s->p, s->d);
  complete(s, s->d, 0);
  hipLaunchKernelGGL(k_pcg_div, dim3(g), dim3(kT), 0, s->stream, s->nrow, s->r, s->d, s->z);
  double v[2];
  int rc = dot2(s, s->r, s->z, s->b, s->b, v);
  if (rc) return rc;
  s->rho = v[0];
  s->normb = std::sqrt(v[1]);
  s->ready = true;
  if (normb) *normb = s->normb;
  return XG_OK;
}

int xg_pcg_solve(xg_pcg* s, int64_t maxit, double tol, double* normr, int64_t* it,
                 int* converged) {
  if (!s) return XG_EINVAL;
  if (!s->ready) return pfail(s, XG_ENOMESH, "xg_pcg_setup not called");
  if (maxit < 1 || !(tol >= 0.0)) return pfail(s, XG_EINVAL, "xg_pcg_solve: bad maxit/tol");
  if (hipSetDevice(s->device) != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg: hipSetDevice");
  const unsigned g = grid(s->nrow);
  const double eps = 2.220446049250313e-16;
  const double nb = s->normb > 1.0e-14 ? s->normb : 1.0;
  double rho0 = 0.0, nr = 0.0;
  int64_t k = 0;
  bool finished = false;
  for (;;) {
    // next(): p = z + alpha p (ConjugateGradients.cpp:583-595)
    const double beta = k == 0 ? 0.0 : s->rho / rho0;
    rho0 = s->rho;
    hipLaunchKernelGGL(k_pcg_dir, dim3(g), dim3(kT), 0, s->stream, s->nrow, beta, s->z, s->p);
    // qAp/comq/q(): q = A p completed at shared rows, then (p, q)
    hipLaunchKernelGGL(k_pcg_spmv, dim3(g), dim3(kT), 0, s->stream, s->nrow, s->ia, s->ja,
                       s->a, s->p, s->q);
    complete(s, s->q, 0);
    double v[2];
    int rc = dot2(s, s->p, s->q, s->p, s->q, v);
    if (rc) return rc;
    // pq(): breakdown test, r -= alpha q, z = r/d, (r,z), (r,r)
    double alpha;
    if (std::fabs(v[0]) < eps) {
      finished = true;
      alpha = 0.0;
    } else {
      alpha = s->rho / v[0];
    }
    hipLaunchKernelGGL(k_pcg_res, dim3(g), dim3(kT), 0, s->stream, s->nrow, alpha, s->q, s->d,
                       s->r, s->z);
    if ((rc = dot2(s, s->r, s->z, s->r, s->r, v))) return rc;
    s->rho = v[0];
    // rz()/comx/x(): x += alpha p, shared copies averaged
    hipLaunchKernelGGL(k_pcg_sol, dim3(g), dim3(kT), 0, s->stream, s->nrow, alpha, s->p, s->x);
    complete(s, s->x, 1);
    if (hipGetLastError() != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg: kernel launch");
    ++k;
    nr = std::sqrt(v[1]);
    if (finished || nr < tol * nb || k >= maxit) break;
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess)
    return pfail(s, XG_EHIP, "xg_pcg_solve: synchronize");
  if (normr) *normr = nr;
  if (it) *it = k;
  if (converged) *converged = nr > tol * nb ? 0 : 1;
  s->ready = false;
  return XG_OK;
}

int xg_pcg_solution(xg_pcg* s, int part, double* x) {
  if (!s) return XG_EINVAL;
  if (!s->x || part < 0 || part >= s->nparts || (!x && s->rows[part]))
    return pfail(s, XG_EINVAL, "xg_pcg_solution: bad arguments");
  if (hipSetDevice(s->device) != hipSuccess) return pfail(s, XG_EHIP, "xg_pcg: hipSetDevice");
  if (s->rows[part] &&
      (hipMemcpyAsync(x, s->x + s->roff[part], (size_t)s->rows[part] * 8,
                      hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
       hipStreamSynchronize(s->stream) != hipSuccess))
    return pfail(s, XG_EHIP, "xg_pcg_solution: hipMemcpy");
  return XG_OK;
}

const char* xg_pcg_error(const xg_pcg* s) { return s ? s->err.c_str() : "null solver"; }

int xg_pcg_destroy(xg_pcg* s) {
  if (!s) return XG_EINVAL;
  if (!s->allocs.empty()) {
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    for (void* p : s->allocs) (void)hipFree(p);
  }
  delete s;
  return XG_OK;
}

}  // extern "C"
