// MI355X device path of tk::ConjugateGradients (ConjugateGradients.cpp:
// 105-823) for one GPU-resident partition, plus a mixed-precision variant.
//
// One CG iteration of the reference is two kernels:
//   k_cg_spmv:   p = z + beta p (into the other p buffer), q = A p, partial
//                sums of (p,q); the last workgroup reduces them in index order
//                and sets alpha = rho/(p,q) (or flags the (p,q) ~ 0 breakdown).
//   k_cg_update: r -= alpha q, z = r/d, x += alpha p, partial sums of (r,z)
//                and (r,r); the last workgroup reduces them, advances the
//                iteration count, evaluates the stop test and beta.
// Scalars live on the device (CgState); the host reads them back only every
// few iterations; a finished solve turns the remaining launches into no-ops,
// so the iteration count and result are those of the reference recurrence.
// Reductions are deterministic: fixed per-workgroup partials summed in index
// order by whichever workgroup arrives last (an integer arrival counter; no
// floating-point atomics).
//
// SpMV: 8 lanes per CSR row (rows hold ~15 entries on tet meshes), lane
// partial sums combined by a fixed butterfly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "xg_csr.hpp"
#include "xg_ctx.hpp"

using namespace xgd;

namespace {

constexpr int kCgThreads = 256;
constexpr int kLanesPerRow = 8;
constexpr int kMaxParts = 2048;

struct CgState {
  double rho, beta, alpha, rr, normb;
  long long it;
  int finished, done, converged, pad;
  unsigned int count;            // arrival counter of the last-block reductions
  unsigned int pad2;
  double tol_nb;                 // tol * max(normb, 1 if normb <= 1e-14)
  long long maxit;
  double brk;                    // |(p,q)| below this ends the solve
};

}  // namespace

struct CgSolver {
  int ncomp = 1;
  int64_t nrow = 0, nnz = 0;
  // host copies: pattern and original values (restored semantics of apply)
  std::vector<int64_t> ia;
  std::vector<int32_t> ja;
  std::vector<double> a0;
  bool dev_bc = false;           // device values carry Dirichlet rows
  bool restore = false;          // restore the original values before reuse
  std::vector<int64_t> bc_key;   // Dirichlet (node, mask...) set of the device values
  std::vector<DevArr> allocs;
  int *dia = nullptr, *dja = nullptr;
  double* da = nullptr;          // values as used by the solve
  float* da32 = nullptr;         // FP32 copy for the mixed-precision inner solve
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *d = nullptr;
  double *p0 = nullptr, *p1 = nullptr, *q = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32a = nullptr, *p32b = nullptr,
        *q32 = nullptr, *d32 = nullptr, *e32 = nullptr;
  double* part = nullptr;        // 2 x kMaxParts partial sums
  CgState* st = nullptr;
  CgState* hst = nullptr;        // pinned host mirror
  int pbuf = 0;                  // which p buffer holds the current p
  int pc = 0;
  bool ready = false;            // setup done
};

namespace {

// ---------------------------------------------------------------- kernels --

template <class T>
__device__ inline T warp_row_sum(T v) {
  // butterfly over the 8 lanes of a row: fixed order, same result every run
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

__device__ inline double block_sum(double v, double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sm[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kCgThreads / 64; ++k) s += sm[k];
  __syncthreads();
  return s;
}

// Publishes this workgroup's partial sums; returns true in the workgroup that
// arrives last, which then sees every partial.
__device__ inline bool arrive(CgState* st, double* part, int nparts,
                              const double* vals, int nv) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    for (int k = 0; k < nv; ++k) part[k * kMaxParts + blockIdx.x] = vals[k];
    __threadfence();
    unsigned int prev = atomicAdd(&st->count, 1u);
    last = (prev == (unsigned int)nparts - 1);
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// Sum of partials [0, n) in index order, by the whole workgroup (fixed tree).
__device__ inline double sum_parts(const double* part, int n, double* sm) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kCgThreads)
    v += __builtin_nontemporal_load(part + i);
  return block_sum(v, sm);
}

// r = b - A x (setup, ConjugateGradients::residual/initres) and the diagonal
// preconditioner d (pc: 1 = jacobi diag(A), 0 = none -> 1).
__global__ __launch_bounds__(kCgThreads) void k_cg_residual(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const double* __restrict__ a, const double* __restrict__ x,
    const double* __restrict__ b, double* __restrict__ r,
    double* __restrict__ d, int pc) {
  const int lane = threadIdx.x & (kLanesPerRow - 1);
  for (int64_t row = ((int64_t)blockIdx.x * kCgThreads + threadIdx.x) / kLanesPerRow;
       row < nrow; row += (int64_t)gridDim.x * kCgThreads / kLanesPerRow) {
    double s = 0.0, dg = 0.0;
    for (int j = ia[row] + lane; j < ia[row + 1]; j += kLanesPerRow) {
      int c = ja[j];
      s += a[j] * x[c];
      if (c == row) dg = a[j];
    }
    s = warp_row_sum(s);
    dg = warp_row_sum(dg);
    if (lane == 0) {
      r[row] = b[row] - s;
      d[row] = pc ? dg : 1.0;
    }
  }
}

// z = r/d, p = r (initres); partial sums of (b,b) and (r,z)
__global__ __launch_bounds__(kCgThreads) void k_cg_init(
    int64_t n, const double* __restrict__ b, const double* __restrict__ r,
    const double* __restrict__ d, double* __restrict__ z,
    double* __restrict__ p, CgState* st, double* part, double tol) {
  __shared__ double sm[kCgThreads / 64];
  double bb = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    double zi = r[i] / d[i];
    z[i] = zi;
    p[i] = r[i];
    bb += b[i] * b[i];
    rz += r[i] * zi;
  }
  double v[2] = {block_sum(bb, sm), block_sum(rz, sm)};
  if (arrive(st, part, gridDim.x, v, 2)) {
    double BB = sum_parts(part, gridDim.x, sm);
    double RZ = sum_parts(part + kMaxParts, gridDim.x, sm);
    if (threadIdx.x == 0) {
      st->normb = sqrt(BB);
      st->rho = RZ;
      st->beta = 0.0;
      st->it = 0;
      st->finished = st->done = st->converged = 0;
      double nb = st->normb > 1.0e-14 ? st->normb : 1.0;
      st->tol_nb = tol * nb;
      st->brk = 2.220446049250313e-16;   // DBL_EPSILON, absolute
      st->count = 0;
    }
  }
}

// p' = z + beta p, q = A p', partial (p',q); last block: alpha
template <class T>
__global__ __launch_bounds__(kCgThreads) void k_cg_spmv(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const T* __restrict__ a, const T* __restrict__ z, const T* __restrict__ pold,
    T* __restrict__ pnew, T* __restrict__ q, CgState* st, double* part) {
  __shared__ double sm[kCgThreads / 64];
  if (st->done) return;
  const T beta = (T)st->beta;
  const int lane = threadIdx.x & (kLanesPerRow - 1);
  double pq = 0.0;
  for (int64_t row = ((int64_t)blockIdx.x * kCgThreads + threadIdx.x) / kLanesPerRow;
       row < nrow; row += (int64_t)gridDim.x * kCgThreads / kLanesPerRow) {
    T s = 0;
    for (int j = ia[row] + lane; j < ia[row + 1]; j += kLanesPerRow) {
      int c = ja[j];
      s += a[j] * (z[c] + beta * pold[c]);
    }
    s = warp_row_sum(s);
    if (lane == 0) {
      T pr = z[row] + beta * pold[row];
      pnew[row] = pr;
      q[row] = s;
      pq += (double)pr * (double)s;
    }
  }
  double v = block_sum(pq, sm);
  if (arrive(st, part, gridDim.x, &v, 1)) {
    double PQ = sum_parts(part, gridDim.x, sm);
    if (threadIdx.x == 0) {
      if (fabs(PQ) < st->brk) {      // ConjugateGradients.cpp:680-687
        st->finished = 1;
        st->alpha = 0.0;
      } else {
        st->alpha = st->rho / PQ;
      }
      st->count = 0;
    }
  }
}

// r -= alpha q, z = r/d, x += alpha p; partial (r,z), (r,r); last block: the
// stop test of ConjugateGradients::x (ConjugateGradients.cpp:772-823) and beta
template <class T>
__global__ __launch_bounds__(kCgThreads) void k_cg_update(
    int64_t n, const T* __restrict__ q, const T* __restrict__ d,
    const T* __restrict__ p, T* __restrict__ r, T* __restrict__ z,
    T* __restrict__ x, CgState* st, double* part) {
  __shared__ double sm[kCgThreads / 64];
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    T ri = r[i] - alpha * q[i];
    T zi = ri / d[i];
    r[i] = ri;
    z[i] = zi;
    x[i] += alpha * p[i];
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  double v[2] = {block_sum(rz, sm), block_sum(rr, sm)};
  if (arrive(st, part, gridDim.x, v, 2)) {
    double RZ = sum_parts(part, gridDim.x, sm);
    double RR = sum_parts(part + kMaxParts, gridDim.x, sm);
    if (threadIdx.x == 0) {
      double rho0 = st->rho;
      st->rho = RZ;
      st->rr = RR;
      st->beta = RZ / rho0;
      st->it += 1;
      double normr = sqrt(RR);
      if (st->finished || normr < st->tol_nb || st->it >= st->maxit) {
        st->done = 1;
        st->converged = normr > st->tol_nb ? 0 : 1;
      }
      st->count = 0;
    }
  }
}

// ---- mixed precision helpers ----

// x += (double)e
__global__ __launch_bounds__(kCgThreads) void k_add_f32(
    int64_t n, const float* __restrict__ e, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads)
    x[i] += (double)e[i];
}

// FP32 inner setup: r32 = rin, e = 0, z = r/d, p = r; (r,z), (r,r) -> state
__global__ __launch_bounds__(kCgThreads) void k_cg32_init(
    int64_t n, const double* __restrict__ rin, const float* __restrict__ d,
    float* __restrict__ r, float* __restrict__ z, float* __restrict__ p,
    float* __restrict__ e, CgState* st, double* part, double tol_abs,
    long long maxit) {
  __shared__ double sm[kCgThreads / 64];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    float ri = (float)rin[i];
    float zi = ri / d[i];
    r[i] = ri;
    z[i] = zi;
    p[i] = ri;
    e[i] = 0.0f;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  double v[2] = {block_sum(rz, sm), block_sum(rr, sm)};
  if (arrive(st, part, gridDim.x, v, 2)) {
    double RZ = sum_parts(part, gridDim.x, sm);
    double RR = sum_parts(part + kMaxParts, gridDim.x, sm);
    if (threadIdx.x == 0) {
      st->rho = RZ;
      st->rr = RR;
      st->beta = 0.0;
      st->it = 0;
      st->finished = st->converged = 0;
      st->done = sqrt(RR) < tol_abs ? 1 : 0;
      st->tol_nb = tol_abs;
      st->maxit = maxit;
      st->brk = 1.0e-300;           // inner solve: stop only at (p,q) ~ 0
      st->count = 0;
    }
  }
}

// r = b - A x in FP64 with partial (r,r); last block: rr, normb test
__global__ __launch_bounds__(kCgThreads) void k_residual_norm(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const double* __restrict__ a, const double* __restrict__ x,
    const double* __restrict__ b, double* __restrict__ r, CgState* st,
    double* part) {
  __shared__ double sm[kCgThreads / 64];
  const int lane = threadIdx.x & (kLanesPerRow - 1);
  double rr = 0.0;
  for (int64_t row = ((int64_t)blockIdx.x * kCgThreads + threadIdx.x) / kLanesPerRow;
       row < nrow; row += (int64_t)gridDim.x * kCgThreads / kLanesPerRow) {
    double s = 0.0;
    for (int j = ia[row] + lane; j < ia[row + 1]; j += kLanesPerRow) s += a[j] * x[ja[j]];
    s = warp_row_sum(s);
    if (lane == 0) {
      double ri = b[row] - s;
      r[row] = ri;
      rr += ri * ri;
    }
  }
  double v = block_sum(rr, sm);
  if (arrive(st, part, gridDim.x, &v, 1)) {
    double RR = sum_parts(part, gridDim.x, sm);
    if (threadIdx.x == 0) {
      st->rr = RR;
      st->count = 0;
    }
  }
}

__global__ void k_f32_from_f64(int64_t n, const double* __restrict__ s,
                               float* __restrict__ d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

// --------------------------------------------------------------- host side --

int grid_rows(int64_t nrow) {
  int64_t g = (nrow * kLanesPerRow + kCgThreads - 1) / kCgThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kMaxParts));
}

int grid_vec(int64_t n) {
  int64_t g = (n + kCgThreads - 1) / kCgThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1024));
}

void free_solver(CgSolver* s) {
  if (!s) return;
  dev_free_all(s->allocs);
  if (s->hst) (void)hipHostFree(s->hst);
  delete s;
}

// Device values back to the matrix as given to xg_cg_matrix.
int restore_values(xg_ctx* c) {
  CgSolver* s = c->cg;
  std::vector<float> a32(s->a0.begin(), s->a0.end());
  XG_HIP(c, hipMemcpy(s->da, s->a0.data(), s->a0.size() * 8, hipMemcpyHostToDevice));
  XG_HIP(c, hipMemcpy(s->da32, a32.data(), a32.size() * 4, hipMemcpyHostToDevice));
  s->dev_bc = false;
  s->restore = false;
  return XG_OK;
}

int check_cg(xg_ctx* c) {
  if (!c) return XG_EINVAL;
  if (!c->cg) return fail(c, XG_ENOMESH, "no CG matrix (xg_cg_matrix)");
  return XG_OK;
}

int read_state(xg_ctx* c) {
  CgSolver* s = c->cg;
  XG_HIP(c, hipMemcpyAsync(s->hst, s->st, sizeof(CgState), hipMemcpyDeviceToHost,
                           c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

// FP64 CG iterations on the current state until done (host checks every
// `batch` iterations).
int run_iterations(xg_ctx* c, int64_t maxit) {
  CgSolver* s = c->cg;
  const int gr = grid_rows(s->nrow), gv = grid_vec(s->nrow);
  const int batch = 8;
  for (int64_t k = 0; k < maxit; k += batch) {
    for (int i = 0; i < batch && k + i < maxit; ++i) {
      double* pold = s->pbuf ? s->p1 : s->p0;
      double* pnew = s->pbuf ? s->p0 : s->p1;
      hipLaunchKernelGGL(k_cg_spmv<double>, dim3(gr), dim3(kCgThreads), 0,
                         c->stream, s->nrow, s->dia, s->dja, s->da, s->z, pold,
                         pnew, s->q, s->st, s->part);
      hipLaunchKernelGGL(k_cg_update<double>, dim3(gv), dim3(kCgThreads), 0,
                         c->stream, s->nrow, s->q, s->d, pnew, s->r, s->z, s->x,
                         s->st, s->part);
      s->pbuf ^= 1;
    }
    XG_HIP(c, hipGetLastError());
    int rc = read_state(c);
    if (rc) return rc;
    if (s->hst->done) {
      // the no-op launches after the stop did not advance p
      int extra = (int)(std::min<int64_t>(batch, maxit - k) -
                        (s->hst->it - k));
      if (extra & 1) s->pbuf ^= 1;
      return XG_OK;
    }
  }
  return XG_OK;
}

}  // namespace

// =================================================================== C-ABI ==

extern "C" {

int xg_cg_matrix(xg_ctx* c, int ncomp, int64_t npoin, const size_t* ia,
                 const size_t* ja, const double* a, int base) {
  if (!c) return XG_EINVAL;
  if (ncomp < 1 || npoin <= 0 || !ia || !ja || !a || (base != 0 && base != 1))
    return fail(c, XG_EINVAL, "xg_cg_matrix: bad arguments");
  const int64_t nrow = npoin * ncomp;
  if (nrow >= (1ll << 31)) return fail(c, XG_EINVAL, "xg_cg_matrix: too many rows");
  if (ia[0] != (size_t)base) return fail(c, XG_EINVAL, "xg_cg_matrix: ia[0] != base");
  const int64_t nnz = (int64_t)ia[nrow] - base;
  if (nnz <= 0 || nnz >= (1ll << 31)) return fail(c, XG_EINVAL, "xg_cg_matrix: bad nnz");
  XG_HIP(c, hipSetDevice(c->device));
  auto* s = new CgSolver;
  s->ncomp = ncomp;
  s->nrow = nrow;
  s->nnz = nnz;
  s->ia.resize(nrow + 1);
  s->ja.resize(nnz);
  for (int64_t i = 0; i <= nrow; ++i) {
    s->ia[i] = (int64_t)ia[i] - base;
    if (i && s->ia[i] < s->ia[i - 1]) { free_solver(s); return fail(c, XG_EINVAL, "xg_cg_matrix: ia not monotone"); }
  }
  for (int64_t j = 0; j < nnz; ++j) {
    int64_t col = (int64_t)ja[j] - base;
    if (col < 0 || col >= nrow) { free_solver(s); return fail(c, XG_EINVAL, "xg_cg_matrix: column out of range"); }
    s->ja[j] = (int32_t)col;
  }
  s->a0.assign(a, a + nnz);
  std::vector<int32_t> ia32(s->ia.begin(), s->ia.end());
  std::vector<float> a32(nnz);
  for (int64_t j = 0; j < nnz; ++j) a32[j] = (float)a[j];
  int rc;
  auto& L = s->allocs;
  const size_t n = (size_t)nrow;
  if ((rc = upload(c, L, &s->dia, ia32)) || (rc = upload(c, L, &s->dja, s->ja)) ||
      (rc = upload(c, L, &s->da, s->a0)) || (rc = upload(c, L, &s->da32, a32)) ||
      (rc = dev_alloc(c, L, (void**)&s->x, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->b, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->r, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->z, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->d, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->p0, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->p1, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->q, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->r32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->z32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->p32a, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->p32b, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->q32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->d32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->e32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->part, sizeof(double) * 2 * kMaxParts)) ||
      (rc = dev_alloc(c, L, (void**)&s->st, sizeof(CgState)))) {
    free_solver(s);
    return rc;
  }
  if (hipHostMalloc((void**)&s->hst, sizeof(CgState)) != hipSuccess) {
    s->hst = nullptr;
    free_solver(s);
    return fail(c, XG_EHIP, "hipHostMalloc");
  }
  XG_HIP(c, hipMemset(s->x, 0, n * 8));
  XG_HIP(c, hipMemset(s->b, 0, n * 8));
  CgState z0;
  std::memset(&z0, 0, sizeof z0);
  XG_HIP(c, hipMemcpy(s->st, &z0, sizeof z0, hipMemcpyHostToDevice));
  free_solver(c->cg);
  c->cg = s;
  return XG_OK;
}

int xg_cg_init(xg_ctx* c, const double* x, const double* b, const double* neubc,
               int64_t ndir, const size_t* dirnode, const int32_t* dirmask,
               const double* dirval, int apply, int pc) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  if (pc != 0 && pc != 1) return fail(c, XG_EINVAL, "pc must be 0 (none) or 1 (jacobi)");
  if (ndir < 0 || (ndir && (!dirnode || !dirmask || !dirval)))
    return fail(c, XG_EINVAL, "xg_cg_init: bad Dirichlet arrays");
  XG_HIP(c, hipSetDevice(c->device));
  const size_t n = (size_t)s->nrow;
  s->pc = pc;
  s->ready = false;
  if (x) XG_HIP(c, hipMemcpyAsync(s->x, x, n * 8, hipMemcpyHostToDevice, c->stream));
  if (!apply) {
    // ConjugateGradients::init with apply = false: solve with A as stored
    if (b) XG_HIP(c, hipMemcpyAsync(s->b, b, n * 8, hipMemcpyHostToDevice, c->stream));
    XG_HIP(c, hipStreamSynchronize(c->stream));
    if (s->dev_bc && (rc = restore_values(c))) return rc;
    return XG_OK;
  }
  // ConjugateGradients::init/apply/r (ConjugateGradients.cpp:335-556): b +=
  // neubc; Dirichlet rows in decreasing node order; b -= A_col*val; b[dir] =
  // val.  Applied to a copy of the original matrix (A is restored after the
  // solve in the reference).
  std::vector<double> bb(n, 0.0);
  if (b) std::memcpy(bb.data(), b, n * 8);
  else XG_HIP(c, hipMemcpy(bb.data(), s->b, n * 8, hipMemcpyDeviceToHost));
  if (neubc) for (size_t i = 0; i < n; ++i) bb[i] += neubc[i];
  std::vector<std::pair<size_t, int64_t>> order(ndir);
  for (int64_t k = 0; k < ndir; ++k) {
    if (dirnode[k] * (size_t)s->ncomp >= n)
      return fail(c, XG_EINVAL, "xg_cg_init: Dirichlet node out of range");
    order[k] = {dirnode[k], k};
  }
  std::sort(order.begin(), order.end());
  // the BC-modified matrix depends only on which (node, component) pairs are
  // Dirichlet: with the same set as the matrix on the device holds, only b is
  // rebuilt
  std::vector<int64_t> key;
  key.reserve(ndir * (1 + s->ncomp));
  for (int64_t k = 0; k < ndir; ++k) {
    int64_t i = order[k].second;
    key.push_back((int64_t)dirnode[i]);
    for (int j = 0; j < s->ncomp; ++j) key.push_back(dirmask[i * s->ncomp + j] ? 1 : 0);
  }
  const bool same = s->dev_bc && key == s->bc_key;
  std::vector<double> rr(n, 0.0);
  if (same) {
    // b -= A_col * val as CSR::dirichlet accumulates it, in decreasing node
    // order; entries of the original columns equal the progressively zeroed
    // ones in every row that is not itself a Dirichlet row, and those rows of
    // b are overwritten with their values below
    for (int64_t k = ndir - 1; k >= 0; --k) {
      int64_t i = order[k].second;
      for (int j = 0; j < s->ncomp; ++j) {
        if (!dirmask[i * s->ncomp + j]) continue;
        const int64_t col = (int64_t)dirnode[i] * s->ncomp + j;
        const double val = dirval[i * s->ncomp + j];
        for (int64_t e = s->ia[col]; e < s->ia[col + 1]; ++e) {
          const int64_t r = s->ja[e];
          const int32_t* b0 = s->ja.data() + s->ia[r];
          const int32_t* b1 = s->ja.data() + s->ia[r + 1];
          const int32_t* f = std::lower_bound(b0, b1, (int32_t)col);
          if (f != b1 && *f == col) rr[r] += s->a0[f - s->ja.data()] * val;
        }
      }
    }
  } else {
    std::vector<double> a(s->a0);
    for (int64_t k = ndir - 1; k >= 0; --k) {
      int64_t i = order[k].second;
      for (int j = 0; j < s->ncomp; ++j)
        if (dirmask[i * s->ncomp + j])
          xg::csr_dirichlet(s->ncomp, s->ia.data(), s->ja.data(), a.data(),
                            (int64_t)dirnode[i], dirval[i * s->ncomp + j], rr.data(), j);
    }
    std::vector<float> a32(a.begin(), a.end());
    XG_HIP(c, hipMemcpy(s->da, a.data(), a.size() * 8, hipMemcpyHostToDevice));
    XG_HIP(c, hipMemcpy(s->da32, a32.data(), a32.size() * 4, hipMemcpyHostToDevice));
    s->bc_key.swap(key);
  }
  for (size_t i = 0; i < n; ++i) bb[i] -= rr[i];
  for (int64_t k = 0; k < ndir; ++k)
    for (int j = 0; j < s->ncomp; ++j)
      if (dirmask[k * s->ncomp + j])
        bb[dirnode[k] * s->ncomp + j] = dirval[k * s->ncomp + j];
  XG_HIP(c, hipMemcpy(s->b, bb.data(), n * 8, hipMemcpyHostToDevice));
  s->dev_bc = true;
  s->restore = false;
  return XG_OK;
}

int xg_cg_setup(xg_ctx* c, double* normb) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  XG_HIP(c, hipSetDevice(c->device));
  if (s->restore && (rc = restore_values(c))) return rc;
  hipLaunchKernelGGL(k_cg_residual, dim3(grid_rows(s->nrow)), dim3(kCgThreads), 0,
                     c->stream, s->nrow, s->dia, s->dja, s->da, s->x, s->b, s->r,
                     s->d, s->pc);
  hipLaunchKernelGGL(k_cg_init, dim3(grid_vec(s->nrow)), dim3(kCgThreads), 0,
                     c->stream, s->nrow, s->b, s->r, s->d, s->z, s->p0, s->st,
                     s->part, 0.0);
  XG_HIP(c, hipGetLastError());
  s->pbuf = 0;
  if ((rc = read_state(c))) return rc;
  s->ready = true;
  if (normb) *normb = s->hst->normb;
  return XG_OK;
}

int xg_cg_solve(xg_ctx* c, int64_t maxit, double tol, double* normr,
                int64_t* it, int* converged) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  if (!s->ready) return fail(c, XG_ENOMESH, "xg_cg_setup not called");
  if (maxit < 1 || !(tol >= 0.0)) return fail(c, XG_EINVAL, "xg_cg_solve: bad maxit/tol");
  XG_HIP(c, hipSetDevice(c->device));
  // stop test parameters (ConjugateGradients.cpp:785-787)
  CgState* h = s->hst;
  double nb = h->normb > 1.0e-14 ? h->normb : 1.0;
  h->tol_nb = tol * nb;
  h->maxit = maxit;
  h->it = 0;
  h->done = 0;
  h->finished = 0;
  h->converged = 0;
  h->beta = 0.0;
  h->count = 0;
  XG_HIP(c, hipMemcpyAsync(s->st, h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
  if ((rc = run_iterations(c, maxit))) return rc;
  h = s->hst;
  if (normr) *normr = sqrt(h->rr);
  if (it) *it = h->it;
  if (converged) *converged = h->converged;
  s->ready = false;
  s->restore = s->dev_bc;      // the reference restores A after the solve
  return XG_OK;
}

int xg_cg_solution(xg_ctx* c, double* x) {
  int rc = check_cg(c);
  if (rc) return rc;
  if (!x) return fail(c, XG_EINVAL, "xg_cg_solution: null x");
  XG_HIP(c, hipMemcpyAsync(x, c->cg->x, (size_t)c->cg->nrow * 8, hipMemcpyDeviceToHost,
                           c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

int xg_cg_device_x(xg_ctx* c, double** x) {
  int rc = check_cg(c);
  if (rc) return rc;
  if (!x) return fail(c, XG_EINVAL, "xg_cg_device_x: null");
  *x = c->cg->x;
  return XG_OK;
}

// Mixed precision: FP64 residual r = b - A x, FP32 Jacobi-PCG on A e = r to
// a relative tolerance inner_tol (or inner_maxit), x += e, until the true
// FP64 residual satisfies ||b - A x|| < tol * ||b||.
int xg_cg_solve_mixed(xg_ctx* c, int64_t maxit, double tol, double inner_tol,
                      int64_t inner_maxit, double* normr, int64_t* iters,
                      int64_t* nouter, int* converged) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  if (maxit < 1 || inner_maxit < 1 || !(tol > 0.0) || !(inner_tol > 0.0))
    return fail(c, XG_EINVAL, "xg_cg_solve_mixed: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  const int64_t n = s->nrow;
  const int gr = grid_rows(n), gv = grid_vec(n);
  // preconditioner (FP64 diagonal, then FP32 copy) and ||b||
  hipLaunchKernelGGL(k_cg_residual, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                     s->dia, s->dja, s->da, s->x, s->b, s->r, s->d, s->pc);
  hipLaunchKernelGGL(k_f32_from_f64, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n,
                     s->d, s->d32);
  hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->b,
                     s->r, s->d, s->z, s->p0, s->st, s->part, tol);
  if ((rc = read_state(c))) return rc;
  const double normb = s->hst->normb;
  const double target = tol * (normb > 1.0e-14 ? normb : 1.0);
  int64_t total = 0, outer = 0;
  double nr = 0.0;
  int conv = 0;
  for (;;) {
    hipLaunchKernelGGL(k_residual_norm, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                       s->dia, s->dja, s->da, s->x, s->b, s->r, s->st, s->part);
    if ((rc = read_state(c))) return rc;
    nr = sqrt(s->hst->rr);
    if (nr < target) { conv = 1; break; }
    if (total >= maxit) break;
    ++outer;
    // inner FP32 solve of A e = r to |r_e| < inner_tol * |r|
    hipLaunchKernelGGL(k_cg32_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                       s->r, s->d32, s->r32, s->z32, s->p32a, s->e32, s->st, s->part,
                       inner_tol * nr, (long long)inner_maxit);
    int pb = 0;
    const int batch = 8;
    for (int64_t k = 0; k < inner_maxit; k += batch) {
      for (int i = 0; i < batch; ++i) {
        float* pold = pb ? s->p32b : s->p32a;
        float* pnew = pb ? s->p32a : s->p32b;
        hipLaunchKernelGGL(k_cg_spmv<float>, dim3(gr), dim3(kCgThreads), 0,
                           c->stream, n, s->dia, s->dja, s->da32, s->z32, pold,
                           pnew, s->q32, s->st, s->part);
        hipLaunchKernelGGL(k_cg_update<float>, dim3(gv), dim3(kCgThreads), 0,
                           c->stream, n, s->q32, s->d32, pnew, s->r32, s->z32,
                           s->e32, s->st, s->part);
        pb ^= 1;
      }
      XG_HIP(c, hipGetLastError());
      if ((rc = read_state(c))) return rc;
      if (s->hst->done) break;
    }
    total += s->hst->it;
    hipLaunchKernelGGL(k_add_f32, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                       s->e32, s->x);
    if (s->hst->it == 0) break;   // no progress possible in FP32
  }
  if (normr) *normr = nr;
  if (iters) *iters = total;
  if (nouter) *nouter = outer;
  if (converged) *converged = conv;
  s->ready = false;
  s->restore = s->dev_bc;
  return XG_OK;
}

int xg_cg_destroy(xg_ctx* c) {
  if (!c) return XG_EINVAL;
  free_solver(c->cg);
  c->cg = nullptr;
  return XG_OK;
}

}  // extern "C"
