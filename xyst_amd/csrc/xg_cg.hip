// MI355X device path of tk::ConjugateGradients (ConjugateGradients.cpp:
// 105-823) for one GPU-resident partition, plus a mixed-precision variant.
//
// Iteration j of the reference recurrence is two kernels:
//   k_cg_spmv:   p_j = z + beta p_{j-1} (into the other p buffer), q = A p_j,
//                per-workgroup partial sums of (p_j, q);
//   k_cg_update: alpha = rho/(p,q) (0 and "finished" when |(p,q)| is below
//                the breakdown threshold), r -= alpha q, z = r/d,
//                x += alpha p_j, per-workgroup partials of (r,z) and (r,r).
// Each kernel starts by summing the partials the previous kernel left, in
// index order with a fixed tree: every workgroup obtains bitwise the same
// global value, so no workgroup has to publish it and there is no arrival
// counter, atomic or cross-workgroup fence.  The (r,z)/(r,r) partials
// alternate between two buffers by iteration parity, so iteration j still
// sees rho_{j-2} for beta.  The stop test of ConjugateGradients::x runs at
// the start of both kernels of the next iteration; once it holds, the
// remaining launches are no-ops and workgroup 0 records the final state,
// which the host reads every few iterations.  Results are deterministic.
//
// SpMV of the recurrences (CG and flexible CG): tiles of consecutive rows (at
// most 2048 nonzeros in FP64, 4096 in FP32); a workgroup forms a tile's
// products with coalesced loads of the values and column indices into LDS,
// (z, p) gathered as one pair, then one thread per row sums them in CSR
// order.  The setup's residual keeps 8 lanes per CSR row (rows hold ~15
// entries on tet meshes), lane partials combined by a fixed butterfly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "xg_csr.hpp"
#include "xg_ctx.hpp"
#include "xg_amg.hpp"
#include "xg_cg.hpp"
#include "xg_gnn.hpp"

using namespace xgd;

namespace {

constexpr int kCgThreads = 256;
constexpr int kLanesPerRow = 8;
constexpr int kMaxParts = 2048;
// row tiles of the CG SpMV: consecutive rows holding at most tile_nnz<T>
// nonzeros (a longer row is a tile of its own) and at most kCgThreads rows;
// 16 KiB of LDS products per tile in FP64 and in FP32
template <class T>
constexpr int tile_nnz() { return sizeof(T) == 4 ? 4096 : 2048; }

constexpr int kVecParts = 512;    // workgroups of the vector kernels (max)
constexpr int kBufLen = 2 * kVecParts + 8;   // rz[.], rr[.], finished flag

struct CgState {
  double normb, rho0, rr;
  long long it;
  int done, converged;
};

}  // namespace

struct CgSolver {
  int ncomp = 1;
  int64_t nrow = 0, nnz = 0;
  // host copies: pattern and original values (restored semantics of apply)
  std::vector<int64_t> ia;
  std::vector<int32_t> ja;
  std::vector<double> a0;
  std::vector<double> abc;       // host copy of the values on the device
  int64_t vals_version = 0;      // changes whenever the device values do
  bool dev_bc = false;           // device values carry Dirichlet rows
  bool restore = false;          // restore the original values before reuse
  std::vector<int64_t> bc_key;   // Dirichlet (node, mask...) set of the device values
  std::vector<DevArr> allocs;
  int *dia = nullptr, *dja = nullptr;
  int ntile = 0, ntile32 = 0;
  int* dtile = nullptr;          // first row of each FP64 SpMV tile (ntile + 1)
  int* dtile32 = nullptr;        // the same for the FP32 products
  double* da = nullptr;          // values as used by the solve
  float* da32 = nullptr;         // FP32 copy for the mixed-precision inner solve
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *d = nullptr;
  double *p0 = nullptr, *p1 = nullptr, *q = nullptr;
  double* zp = nullptr;          // (z_i, p_i) pairs the SpMV gathers (FP64)
  float* zp32 = nullptr;         // the same for the FP32 recurrence
  double *q2 = nullptr;          // FCG: q of the previous iteration
  double* fb = nullptr;          // FCG partial sums (2 x 4 x kMaxParts)
  double *xs = nullptr, *ebuf = nullptr;   // reference solution, error (training)
  float *r32 = nullptr, *z32 = nullptr, *p32a = nullptr, *p32b = nullptr,
        *q32 = nullptr, *d32 = nullptr, *e32 = nullptr;
  double* pq = nullptr;          // kMaxParts partials of (p,q) (or scratch)
  double* rzb = nullptr;         // 2 x kBufLen partials of (r,z), (r,r)
  CgState* st = nullptr;
  CgState* hst = nullptr;        // pinned host mirror
  DevArr bcbuf;                  // Dirichlet rows/values of xg_cg_init_dev
  int pc = 0;
  bool ready = false;            // setup done
};

namespace {

// ---------------------------------------------------------------- kernels --

// (z_i, p_i) of one row, gathered by the SpMV with one load
template <class T>
struct alignas(2 * sizeof(T)) Pair {
  T z, p;
};

// zp = (z_i, p_i) pairs: the recurrence's state for its first SpMV
template <class T>
__global__ __launch_bounds__(kCgThreads) void k_pack_zp(int64_t n, const T* __restrict__ z,
                                                        const T* __restrict__ p,
                                                        T* __restrict__ zp) {
  const int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x;
  if (i < n) reinterpret_cast<Pair<T>*>(zp)[i] = Pair<T>{z[i], p[i]};
}

template <class T>
__device__ inline T row_sum8(T v) {
  // butterfly over the 8 lanes of a row: fixed order, same result every run
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

// Workgroup sum returned to every thread (fixed tree, so every workgroup that
// sums the same values gets bitwise the same result).
__device__ inline double block_sum_all(double v, double* sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();                       // sm free from a previous use
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kCgThreads / 64; ++k) s += sm[k];
  return s;
}

// Sum of part[0, n) in index order (n <= 4 * kCgThreads... any n): thread t
// adds part[t], part[t + 256], ... then the fixed tree.
__device__ inline double sum_parts(const double* __restrict__ part, int n, double* sm) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kCgThreads) v += part[i];
  return block_sum_all(v, sm);
}

// Writes a workgroup partial: slot `k` of `part`.
__device__ inline void put_part(double* part, int k, double v, double* sm) {
  double s = block_sum_all(v, sm);
  if (threadIdx.x == 0) part[k] = s;
}

// The stop test of ConjugateGradients::x (ConjugateGradients.cpp:772-823)
// after iteration it = j-1, from that iteration's (r,r) and finished flag.
struct StopTest {
  double rz, rr;
  bool stop, converged;
};

__device__ inline StopTest stop_test(const double* __restrict__ buf, int nvp, int j,
                                     long long maxit, double tol_nb, double* sm) {
  StopTest s;
  s.rz = sum_parts(buf, nvp, sm);
  s.rr = sum_parts(buf + kVecParts, nvp, sm);
  const bool finished = buf[2 * kVecParts] != 0.0;
  const double normr = sqrt(s.rr);
  s.stop = j >= 2 && (finished || normr < tol_nb || (long long)(j - 1) >= maxit);
  s.converged = !(normr > tol_nb);
  return s;
}

__device__ inline void record_stop(CgState* st, int j, const StopTest& s) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && !st->done) {
    st->it = j - 1;
    st->rr = s.rr;
    st->converged = s.converged ? 1 : 0;
    st->done = 1;
  }
}

// r = b - A x (ConjugateGradients::residual) and the diagonal preconditioner
// d (pc 1 = jacobi: diag(A); 0 = none: 1), ConjugateGradients.cpp:160-260.
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
    s = row_sum8(s);
    dg = row_sum8(dg);
    if (lane == 0) {
      r[row] = b[row] - s;
      d[row] = pc ? dg : 1.0;
    }
  }
}

// initres (ConjugateGradients.cpp:300-333): z = r/d, p_0 = r; partials of
// (r,z) and (r,r) into buffer 0 and of (b,b) into `bb`.
__global__ __launch_bounds__(kCgThreads) void k_cg_init(
    int64_t n, const double* __restrict__ b, const double* __restrict__ r,
    const double* __restrict__ d, double* __restrict__ z, double* __restrict__ p,
    double* __restrict__ buf0, double* __restrict__ bb) {
  __shared__ double sm[kCgThreads / 64];
  double sbb = 0.0, rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    double zi = r[i] / d[i];
    z[i] = zi;
    p[i] = r[i];
    sbb += b[i] * b[i];
    rz += r[i] * zi;
    rr += r[i] * r[i];
  }
  put_part(bb, blockIdx.x, sbb, sm);
  put_part(buf0, blockIdx.x, rz, sm);
  put_part(buf0 + kVecParts, blockIdx.x, rr, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) buf0[2 * kVecParts] = 0.0;
}

// normb = ||b|| and rho_0 = (r,z) into the state (one workgroup)
__global__ __launch_bounds__(kCgThreads) void k_cg_setup_final(
    const double* __restrict__ bb, const double* __restrict__ buf0, int nvp,
    CgState* st) {
  __shared__ double sm[kCgThreads / 64];
  const double BB = sum_parts(bb, nvp, sm);
  const double RZ = sum_parts(buf0, nvp, sm);
  const double RR = sum_parts(buf0 + kVecParts, nvp, sm);
  if (threadIdx.x == 0) {
    st->normb = sqrt(BB);
    st->rho0 = RZ;
    st->rr = RR;
    st->it = 0;
    st->done = 0;
    st->converged = 0;
  }
}

// q = A p with p = z + bt p_old for the rows of this workgroup's tiles
// (ti = blockIdx.x, + gridDim.x, ...): a tile's products a_k (z_c + bt p_c)
// are formed with coalesced loads of a and ja into LDS, (z_c, p_c) read as
// one pair, then each row is summed by one thread in CSR order (the order of
// ConjugateGradients' CSR::mult); row(r, p_r, q_r) is called once per row,
// by that thread (thread 0 for a row longer than a tile).  Every thread of
// the workgroup calls this (it holds barriers).
template <class T, class Row>
__device__ inline void tiles_spmv(int ntile, const int* __restrict__ tile,
                                  const int* __restrict__ ia, const int* __restrict__ ja,
                                  const T* __restrict__ a, const T* __restrict__ zp, T bt,
                                  T* prod, double* sm, Row row) {
  constexpr int kTileNnz = tile_nnz<T>();
  const Pair<T>* zpp = reinterpret_cast<const Pair<T>*>(zp);
  const int t = threadIdx.x;
  for (int ti = blockIdx.x; ti < ntile; ti += gridDim.x) {
    const int r0 = tile[ti], r1 = tile[ti + 1];
    const int k0 = ia[r0], nz = ia[r1] - k0;
    if (nz <= kTileNnz) {
      for (int k = t; k < nz; k += 4 * kCgThreads) {
        int cc[4];
        T av[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int kk = k + u * kCgThreads;
          cc[u] = kk < nz ? ja[k0 + kk] : 0;
          av[u] = kk < nz ? a[k0 + kk] : T(0);
        }
        T zv[4], pv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const Pair<T> v = zpp[cc[u]];
          zv[u] = v.z;
          pv[u] = v.p;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int kk = k + u * kCgThreads;
          if (kk < nz) prod[kk] = av[u] * (zv[u] + bt * pv[u]);
        }
      }
      __syncthreads();
      const int r = r0 + t;
      if (r < r1) {
        const int e0 = ia[r] - k0, e1 = ia[r + 1] - k0;
        T acc = T(0);
        for (int k = e0; k < e1; ++k) acc += prod[k];
        const Pair<T> v = zpp[r];
        row(r, v.z + bt * v.p, acc);
      }
      __syncthreads();        // prod is the next tile's
    } else {
      // one row longer than a tile: strided partial sums, fixed-order tree
      T acc = T(0);
      for (int k = t; k < nz; k += kCgThreads) {
        const Pair<T> v = zpp[ja[k0 + k]];
        acc += a[k0 + k] * (v.z + bt * v.p);
      }
      const double sum = block_sum_all((double)acc, sm);
      if (t == 0) {
        const Pair<T> v = zpp[r0];
        row(r0, v.z + bt * v.p, (T)sum);
      }
    }
  }
}

// iteration j, first half: p_j = z + beta p_{j-1} (both read from the (z, p)
// pairs the previous update wrote), q = A p_j, partials of
// (p_j, q) (ConjugateGradients::next/qAp, :585-635)
template <class T>
__global__ __launch_bounds__(kCgThreads) void k_cg_spmv(
    int j, int ntile, const int* __restrict__ tile, const int* __restrict__ ia,
    const int* __restrict__ ja, const T* __restrict__ a, const T* __restrict__ zp,
    T* __restrict__ pnew, T* __restrict__ q,
    const double* __restrict__ rzb, int nvp, double* __restrict__ pq, CgState* st,
    long long maxit, double tol_nb) {
  constexpr int kTileNnz = tile_nnz<T>();
  __shared__ double sm[kCgThreads / 64];
  __shared__ T prod[kTileNnz];
  // after the stop the partial sums of the two buffers are one update apart
  // from each other no more: every later launch of the batch is a no-op
  if (st->done) return;
  const StopTest s = stop_test(rzb + ((j - 1) & 1) * kBufLen, nvp, j, maxit, tol_nb, sm);
  if (s.stop) {
    record_stop(st, j, s);
    return;
  }
  double beta = 0.0;   // the reference's m_alpha = rho/rho0 (:590)
  if (j >= 2) beta = s.rz / sum_parts(rzb + (j & 1) * kBufLen, nvp, sm);
  const T bt = (T)beta;
  double spq = 0.0;
  tiles_spmv<T>(ntile, tile, ia, ja, a, zp, bt, prod, sm, [&](int r, T pr, T qr) {
    pnew[r] = pr;
    q[r] = qr;
    spq += (double)pr * (double)qr;
  });
  put_part(pq, blockIdx.x, spq, sm);
}

// iteration j, second half (ConjugateGradients::pq/rz, :672-750): alpha,
// r -= alpha q, z = r/d, x += alpha p_j; partials of (r,z), (r,r) and the
// finished flag into buffer j & 1
template <class T>
__global__ __launch_bounds__(kCgThreads) void k_cg_update(
    int j, int64_t n, const T* __restrict__ q, const T* __restrict__ d,
    const T* __restrict__ p, T* __restrict__ r, T* __restrict__ z, T* __restrict__ x,
    T* __restrict__ zp, double* __restrict__ rzb, int nvp, const double* __restrict__ pq,
    int nsp, const CgState* st, long long maxit, double tol_nb, double brk) {
  __shared__ double sm[kCgThreads / 64];
  if (st->done) return;
  const StopTest s = stop_test(rzb + ((j - 1) & 1) * kBufLen, nvp, j, maxit, tol_nb, sm);
  if (s.stop) return;
  const double PQ = sum_parts(pq, nsp, sm);
  const bool finished = fabs(PQ) < brk;        // ConjugateGradients.cpp:680-687
  const double alpha = finished ? 0.0 : s.rz / PQ;
  const T al = (T)alpha;
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    const T ri = r[i] - al * q[i];
    const T zi = ri / d[i];
    const T pi = p[i];
    r[i] = ri;
    z[i] = zi;
    x[i] += al * pi;
    reinterpret_cast<Pair<T>*>(zp)[i] = Pair<T>{zi, pi};   // the next SpMV's gathers
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  double* out = rzb + (j & 1) * kBufLen;
  put_part(out, blockIdx.x, rz, sm);
  put_part(out + kVecParts, blockIdx.x, rr, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) out[2 * kVecParts] = finished ? 1.0 : 0.0;
}

// ---- mixed precision helpers ----

// reliable update of the FP32 recurrence (the search direction is kept):
// r32 = (float) r64 (the FP64 residual b - A x), z32 = r32 / d32, e32 = 0;
// partials of (r,z) and (r,r) into `buf`, the slot the next iteration's
// stop test and beta read
__global__ __launch_bounds__(kCgThreads) void k_cg32_replace(
    int64_t n, const double* __restrict__ rin, const float* __restrict__ d,
    float* __restrict__ r, float* __restrict__ z, float* __restrict__ e,
    double* __restrict__ buf) {
  __shared__ double sm[kCgThreads / 64];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    const float ri = (float)rin[i];
    const float zi = ri / d[i];
    r[i] = ri;
    z[i] = zi;
    e[i] = 0.0f;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  put_part(buf, blockIdx.x, rz, sm);
  put_part(buf + kVecParts, blockIdx.x, rr, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) buf[2 * kVecParts] = 0.0;
}

// x += (double)e
__global__ __launch_bounds__(kCgThreads) void k_add_f32(
    int64_t n, const float* __restrict__ e, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads)
    x[i] += (double)e[i];
}

// FP32 inner setup: r32 = rin, e = 0, z = r/d, p_0 = r; (r,z), (r,r) -> buffer 0
__global__ __launch_bounds__(kCgThreads) void k_cg32_init(
    int64_t n, const double* __restrict__ rin, const float* __restrict__ d,
    float* __restrict__ r, float* __restrict__ z, float* __restrict__ p,
    float* __restrict__ e, double* __restrict__ buf0) {
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
  put_part(buf0, blockIdx.x, rz, sm);
  put_part(buf0 + kVecParts, blockIdx.x, rr, sm);
  if (blockIdx.x == 0 && threadIdx.x == 0) buf0[2 * kVecParts] = 0.0;
}

// r = b - A x in FP64 with partials of (r,r)
__global__ __launch_bounds__(kCgThreads) void k_residual_norm(
    int64_t nrow, const int* __restrict__ ia, const int* __restrict__ ja,
    const double* __restrict__ a, const double* __restrict__ x,
    const double* __restrict__ b, double* __restrict__ r, double* __restrict__ part) {
  __shared__ double sm[kCgThreads / 64];
  const int lane = threadIdx.x & (kLanesPerRow - 1);
  double rr = 0.0;
  for (int64_t row = ((int64_t)blockIdx.x * kCgThreads + threadIdx.x) / kLanesPerRow;
       row < nrow; row += (int64_t)gridDim.x * kCgThreads / kLanesPerRow) {
    double s = 0.0;
    for (int j = ia[row] + lane; j < ia[row + 1]; j += kLanesPerRow) s += a[j] * x[ja[j]];
    s = row_sum8(s);
    if (lane == 0) {
      double ri = b[row] - s;
      r[row] = ri;
      rr += ri * ri;
    }
  }
  put_part(part, blockIdx.x, rr, sm);
}

// a new solve on the state: iteration count and flags cleared
__global__ void k_state_reset(CgState* st) {
  if (threadIdx.x == 0) {
    st->it = 0;
    st->done = 0;
    st->converged = 0;
  }
}

// st->rr = sum of part[0, n) (one workgroup)
__global__ __launch_bounds__(kCgThreads) void k_sum_to_state(
    const double* __restrict__ part, int n, CgState* st) {
  __shared__ double sm[kCgThreads / 64];
  const double v = sum_parts(part, n, sm);
  if (threadIdx.x == 0) st->rr = v;
}

__global__ void k_f32_from_f64(int64_t n, const double* __restrict__ s,
                               float* __restrict__ d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

// ---- flexible CG (Notay's FCG(1)) ----
//   z_k = M(r_k);  beta = -(z_k, q_{k-1}) / (p_{k-1}, q_{k-1});
//   p_k = z_k + beta p_{k-1};  q_k = A p_k;  alpha = (p_k, r_k) / (p_k, q_k);
//   x += alpha p_k;  r -= alpha q_k.
// Partial-sum slots of the buffer of iteration k (fb + (k & 1) * 4 kMaxParts):
// [0] (z,q_prev)  [1] (p,q)  [2] (p,r)  [3] (r,r), kMaxParts each.

// partials of (z, q_prev)
__global__ __launch_bounds__(kCgThreads) void k_fcg_dot(
    int64_t n, const double* __restrict__ z, const double* __restrict__ qp,
    double* __restrict__ fb) {
  __shared__ double sm[kCgThreads / 64];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads)
    v += z[i] * qp[i];
  put_part(fb, blockIdx.x, v, sm);
}

// FCG: partials of (z, q_{k-1}) (dot != 0) and the (z, p_{k-1}) pairs the
// SpMV gathers
__global__ __launch_bounds__(kCgThreads) void k_fcg_pack_dot(
    int64_t n, int dot, const double* __restrict__ z, const double* __restrict__ qp,
    const double* __restrict__ pold, double* __restrict__ zp, double* __restrict__ fb) {
  __shared__ double sm[kCgThreads / 64];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    const double zi = z[i];
    // the (z, p_{k-1}) pairs of the SpMV (p = 0 in the first iteration)
    reinterpret_cast<Pair<double>*>(zp)[i] = Pair<double>{zi, dot ? pold[i] : 0.0};
    if (dot) v += zi * qp[i];
  }
  put_part(fb, blockIdx.x, v, sm);
}

// p_k = z + beta p_{k-1}, q_k = A p_k; partials of (p,q) and (p,r) into
// the buffer of iteration k.  first != 0: beta = 0.  `cur` holds the
// (z_k, q_{k-1}) partials, `prev` the (p,q) partials of iteration k-1 (the
// buffers alternate by iteration parity, so nothing read is overwritten).
__global__ __launch_bounds__(kCgThreads) void k_fcg_spmv(
    int first, int ntile, const int* __restrict__ tile, const int* __restrict__ ia,
    const int* __restrict__ ja, const double* __restrict__ a, const double* __restrict__ zp,
    double* __restrict__ pnew, double* __restrict__ q, const double* __restrict__ r,
    const double* __restrict__ prev, double* __restrict__ cur, int nvp, int nsp) {
  __shared__ double sm[kCgThreads / 64];
  __shared__ double prod[tile_nnz<double>()];
  double beta = 0.0;
  if (!first) {
    const double zq = sum_parts(cur, nvp, sm);                 // (z, q_prev)
    const double pq = sum_parts(prev + kMaxParts, nsp, sm);    // (p, q)_prev
    beta = pq != 0.0 ? -zq / pq : 0.0;
  }
  double spq = 0.0, spr = 0.0;
  tiles_spmv<double>(ntile, tile, ia, ja, a, zp, beta, prod, sm,
                     [&](int row, double pr, double qr) {
                       pnew[row] = pr;
                       q[row] = qr;
                       spq += pr * qr;
                       spr += pr * r[row];
                     });
  put_part(cur + kMaxParts, blockIdx.x, spq, sm);
  put_part(cur + 2 * kMaxParts, blockIdx.x, spr, sm);
}

// alpha = (p,r)/(p,q); x += alpha p; r -= alpha q; partials of (r,r)
__global__ __launch_bounds__(kCgThreads) void k_fcg_update(
    int64_t n, const double* __restrict__ p, const double* __restrict__ q,
    double* __restrict__ r, double* __restrict__ x, double* __restrict__ fb, int nsp) {
  __shared__ double sm[kCgThreads / 64];
  const double PQ = sum_parts(fb + kMaxParts, nsp, sm);
  const double PR = sum_parts(fb + 2 * kMaxParts, nsp, sm);
  const double alpha = PQ != 0.0 ? PR / PQ : 0.0;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    rr += ri * ri;
  }
  put_part(fb + 3 * kMaxParts, blockIdx.x, rr, sm);
}

// z = r / d (Jacobi preconditioner of the flexible solve)
__global__ __launch_bounds__(kCgThreads) void k_jacobi_apply(
    int64_t n, const double* __restrict__ r, const double* __restrict__ d,
    double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads)
    z[i] = r[i] / d[i];
}

// z = (double)e
__global__ __launch_bounds__(kCgThreads) void k_f64_from_f32(
    int64_t n, const float* __restrict__ e, double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads)
    z[i] = (double)e[i];
}

// e = a - b
__global__ __launch_bounds__(kCgThreads) void k_sub(
    int64_t n, const double* __restrict__ a, const double* __restrict__ b,
    double* __restrict__ e) {
  for (int64_t i = (int64_t)blockIdx.x * kCgThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kCgThreads)
    e[i] = a[i] - b[i];
}

// --------------------------------------------------------------- host side --

int grid_rows(int64_t nrow) {
  int64_t g = (nrow * kLanesPerRow + kCgThreads - 1) / kCgThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kMaxParts));
}

int grid_vec(int64_t n) {
  int64_t g = (n + kCgThreads - 1) / kCgThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, kVecParts));
}

void free_solver(CgSolver* s) {
  if (!s) return;
  dev_free_all(s->allocs);
  if (s->hst) (void)hipHostFree(s->hst);
  if (s->bcbuf.p) (void)hipFree(s->bcbuf.p);
  delete s;
}

// b[rows[i]] -= vals[i] for i < nsub (the rr of CSR::dirichlet), then, in a
// second launch, b[rows[i]] = vals[i] for the Dirichlet rows
__global__ void k_rows_sub(int n, const int* __restrict__ rows,
                           const double* __restrict__ vals, double* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[rows[i]] -= vals[i];
}

__global__ void k_rows_set(int n, const int* __restrict__ rows,
                           const double* __restrict__ vals, double* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[rows[i]] = vals[i];
}

int grow(xg_ctx* c, DevArr& d, size_t bytes) {
  if (d.bytes >= bytes) return XG_OK;
  if (d.p) (void)hipFree(d.p);
  d.p = nullptr;
  d.bytes = 0;
  if (hipMalloc(&d.p, bytes) != hipSuccess) return fail(c, XG_EHIP, "hipMalloc (CG BC buffer)");
  d.bytes = bytes;
  return XG_OK;
}

// Device values back to the matrix as given to xg_cg_matrix.
int restore_values(xg_ctx* c) {
  CgSolver* s = c->cg;
  std::vector<float> a32(s->a0.begin(), s->a0.end());
  XG_HIP(c, hipMemcpy(s->da, s->a0.data(), s->a0.size() * 8, hipMemcpyHostToDevice));
  XG_HIP(c, hipMemcpy(s->da32, a32.data(), a32.size() * 4, hipMemcpyHostToDevice));
  s->abc = s->a0;
  ++s->vals_version;
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

// Iterations j = 1, 2, ... of the recurrence set up in buffer 0 of rzb, until
// the stop test holds (at the latest at j = maxit + 1); the host checks the
// state every `batch` iterations.  T: FP64 solve or the FP32 inner solve.
template <class T>
int run_cg(xg_ctx* c, const T* a, T* z, T* p0, T* p1, T* q, const T* d, T* r, T* x,
           int64_t maxit, double tol_nb, double brk, int64_t jstart = 1) {
  CgSolver* s = c->cg;
  const bool f32 = sizeof(T) == 4;
  const int ntile = f32 ? s->ntile32 : s->ntile;
  const int* dtile = f32 ? s->dtile32 : s->dtile;
  const int gr = std::max(1, std::min(ntile, kMaxParts)), gv = grid_vec(s->nrow);
  const int batch = 8;
  // the (z, p) pairs of the first SpMV; each update writes the next ones
  T* zp = f32 ? reinterpret_cast<T*>(s->zp32) : reinterpret_cast<T*>(s->zp);
  hipLaunchKernelGGL(k_pack_zp<T>, dim3((unsigned)((s->nrow + kCgThreads - 1) / kCgThreads)),
                     dim3(kCgThreads), 0, c->stream, s->nrow, z,
                     (jstart - 1) & 1 ? p1 : p0, zp);
  for (int64_t j0 = jstart;; j0 += batch) {
    if (j0 > maxit + 1 + batch)
      return fail(c, XG_EHIP, "CG: stop test not reached on the device");
    for (int i = 0; i < batch; ++i) {
      const int j = (int)(j0 + i);
      T* pnew = j & 1 ? p1 : p0;
      hipLaunchKernelGGL(k_cg_spmv<T>, dim3(gr), dim3(kCgThreads), 0, c->stream, j,
                         ntile, dtile, s->dia, s->dja, a, zp, pnew, q, s->rzb, gv, s->pq,
                         s->st, (long long)maxit, tol_nb);
      hipLaunchKernelGGL(k_cg_update<T>, dim3(gv), dim3(kCgThreads), 0, c->stream, j,
                         s->nrow, q, d, pnew, r, z, x, zp, s->rzb, gv, s->pq, gr, s->st,
                         (long long)maxit, tol_nb, brk);
    }
    XG_HIP(c, hipGetLastError());
    int rc = read_state(c);
    if (rc) return rc;
    if (s->hst->done) return XG_OK;
  }
}

// ConjugateGradients::apply's Dirichlet rows (CSR::dirichlet in decreasing
// node order): makes the device matrix carry them, rebuilt from the original
// values when the (node, component) set differs from the one it holds; when
// rr is given, rr += A_col * val as CSR::dirichlet accumulates it.
int cg_dirichlet(xg_ctx* c, CgSolver* s, int64_t ndir, const size_t* dirnode,
                 const int32_t* dirmask, const double* dirval, std::vector<double>* rr) {
  const size_t n = (size_t)s->nrow;
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
  if (same) {
    if (!rr) return XG_OK;
    // entries of the original columns equal the progressively zeroed ones in
    // every row that is not itself a Dirichlet row, and those rows of b are
    // overwritten with their values afterwards
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
          if (f != b1 && *f == col) (*rr)[r] += s->a0[f - s->ja.data()] * val;
        }
      }
    }
    return XG_OK;
  }
  std::vector<double> a(s->a0);
  std::vector<double> scratch;
  std::vector<double>* acc = rr;
  if (!acc) { scratch.assign(n, 0.0); acc = &scratch; }
  for (int64_t k = ndir - 1; k >= 0; --k) {
    int64_t i = order[k].second;
    for (int j = 0; j < s->ncomp; ++j)
      if (dirmask[i * s->ncomp + j])
        xg::csr_dirichlet(s->ncomp, s->ia.data(), s->ja.data(), a.data(),
                          (int64_t)dirnode[i], dirval[i * s->ncomp + j], acc->data(), j);
  }
  std::vector<float> a32(a.begin(), a.end());
  XG_HIP(c, hipStreamSynchronize(c->stream));
  XG_HIP(c, hipMemcpy(s->da, a.data(), a.size() * 8, hipMemcpyHostToDevice));
  XG_HIP(c, hipMemcpy(s->da32, a32.data(), a32.size() * 4, hipMemcpyHostToDevice));
  s->bc_key.swap(key);
  s->abc.swap(a);
  ++s->vals_version;
  return XG_OK;
}

}  // namespace

// =================================================================== C-ABI ==

CgSolver* xg_cg_exchange(xg_ctx* c, CgSolver* s) {
  CgSolver* old = c->cg;
  c->cg = s;
  return old;
}

void xg_cg_free_solver(CgSolver* s) { free_solver(s); }

double* xg_cg_values_dev(xg_ctx* c) {
  CgSolver* s = c->cg;
  s->dev_bc = false;
  s->restore = false;
  s->bc_key.clear();
  ++s->vals_version;
  return s->da;
}

double* xg_cg_rhs_dev(xg_ctx* c) { return c->cg->b; }

int xg_cg_raw_init(xg_ctx* c, int pc) {
  int rc = check_cg(c);
  if (rc) return rc;
  if (pc != 0 && pc != 1) return fail(c, XG_EINVAL, "pc must be 0 (none) or 1 (jacobi)");
  CgSolver* s = c->cg;
  s->pc = pc;
  s->ready = false;
  s->restore = false;
  s->dev_bc = false;
  return XG_OK;
}

int64_t xg_cg_rows(const xg_ctx* c) {
  if (!c || !c->cg || c->cg->ncomp != 1) return -1;
  return c->cg->nrow;
}

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
  s->abc = s->a0;
  std::vector<int32_t> ia32(s->ia.begin(), s->ia.end());
  // SpMV tiles: consecutive rows, at most kCgThreads rows and tnz nonzeros
  // each (a longer row alone)
  auto tiles = [&](int64_t tnz) {
    std::vector<int32_t> tile(1, 0);
    for (int64_t r = 0; r < nrow;) {
      int64_t r1 = r, nz = 0;
      while (r1 < nrow && r1 - r < kCgThreads && nz + (s->ia[r1 + 1] - s->ia[r1]) <= tnz) {
        nz += s->ia[r1 + 1] - s->ia[r1];
        ++r1;
      }
      if (r1 == r) r1 = r + 1;
      tile.push_back((int32_t)r1);
      r = r1;
    }
    return tile;
  };
  const std::vector<int32_t> tile = tiles(tile_nnz<double>()), tile32 = tiles(tile_nnz<float>());
  s->ntile = (int)(tile.size() - 1);
  s->ntile32 = (int)(tile32.size() - 1);
  std::vector<float> a32(nnz);
  for (int64_t j = 0; j < nnz; ++j) a32[j] = (float)a[j];
  int rc;
  auto& L = s->allocs;
  const size_t n = (size_t)nrow;
  if ((rc = upload(c, L, &s->dia, ia32)) || (rc = upload(c, L, &s->dja, s->ja)) ||
      (rc = upload(c, L, &s->dtile, tile)) || (rc = upload(c, L, &s->dtile32, tile32)) ||
      (rc = upload(c, L, &s->da, s->a0)) || (rc = upload(c, L, &s->da32, a32)) ||
      (rc = dev_alloc(c, L, (void**)&s->x, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->b, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->r, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->z, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->d, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->p0, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->p1, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->q, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->zp, n * 16)) ||
      (rc = dev_alloc(c, L, (void**)&s->zp32, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->q2, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->fb, sizeof(double) * 8 * kMaxParts)) ||
      (rc = dev_alloc(c, L, (void**)&s->xs, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->ebuf, n * 8)) ||
      (rc = dev_alloc(c, L, (void**)&s->r32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->z32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->p32a, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->p32b, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->q32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->d32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->e32, n * 4)) ||
      (rc = dev_alloc(c, L, (void**)&s->pq, sizeof(double) * kMaxParts)) ||
      (rc = dev_alloc(c, L, (void**)&s->rzb, sizeof(double) * 2 * kBufLen)) ||
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
  std::vector<double> rr(n, 0.0);
  if ((rc = cg_dirichlet(c, s, ndir, dirnode, dirmask, dirval, &rr))) return rc;
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

int xg_cg_init_dev(xg_ctx* c, const double* b, int64_t ndir, const size_t* dirnode,
                   const int32_t* dirmask, const double* dirval, int apply, int pc) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  if (pc != 0 && pc != 1) return fail(c, XG_EINVAL, "pc must be 0 (none) or 1 (jacobi)");
  if (!b) return fail(c, XG_EINVAL, "xg_cg_init_dev: null rhs");
  if (ndir < 0 || (ndir && (!dirnode || !dirmask || !dirval)))
    return fail(c, XG_EINVAL, "xg_cg_init_dev: bad Dirichlet arrays");
  XG_HIP(c, hipSetDevice(c->device));
  const size_t n = (size_t)s->nrow;
  s->pc = pc;
  s->ready = false;
  if (b != s->b)
    XG_HIP(c, hipMemcpyAsync(s->b, b, n * 8, hipMemcpyDeviceToDevice, c->stream));
  if (!apply) {
    if (s->dev_bc && (rc = restore_values(c))) return rc;
    return XG_OK;
  }
  bool nonzero = false;
  for (int64_t k = 0; k < ndir; ++k)
    for (int j = 0; j < s->ncomp; ++j)
      if (dirmask[k * s->ncomp + j] && dirval[k * s->ncomp + j] != 0.0) nonzero = true;
  // b -= A_col * val only matters for nonzero values (b - 0.0 == b)
  std::vector<double> rr;
  if (nonzero) rr.assign(n, 0.0);
  if ((rc = cg_dirichlet(c, s, ndir, dirnode, dirmask, dirval, nonzero ? &rr : nullptr)))
    return rc;
  std::vector<int> rows;
  std::vector<double> vals;
  if (nonzero)
    for (size_t i = 0; i < n; ++i)
      if (rr[i] != 0.0) { rows.push_back((int)i); vals.push_back(rr[i]); }
  const size_t nsub = rows.size();
  for (int64_t k = 0; k < ndir; ++k)
    for (int j = 0; j < s->ncomp; ++j)
      if (dirmask[k * s->ncomp + j]) {
        rows.push_back((int)(dirnode[k] * s->ncomp + j));
        vals.push_back(dirval[k * s->ncomp + j]);
      }
  if (!rows.empty()) {
    if ((rc = grow(c, s->bcbuf, rows.size() * 12 + 16))) return rc;
    int* drow = (int*)s->bcbuf.p;
    double* dval = (double*)((char*)s->bcbuf.p + ((rows.size() * 4 + 7) / 8) * 8);
    XG_HIP(c, hipStreamSynchronize(c->stream));
    XG_HIP(c, hipMemcpy(drow, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
    XG_HIP(c, hipMemcpy(dval, vals.data(), vals.size() * 8, hipMemcpyHostToDevice));
    if (nsub)
      hipLaunchKernelGGL(k_rows_sub, dim3(nblocks((int64_t)nsub)), dim3(kBlock), 0,
                         c->stream, (int)nsub, drow, dval, s->b);
    const int nset = (int)(rows.size() - nsub);
    if (nset)
      hipLaunchKernelGGL(k_rows_set, dim3(nblocks((int64_t)nset)), dim3(kBlock), 0,
                         c->stream, nset, drow + nsub, dval + nsub, s->b);
    XG_HIP(c, hipGetLastError());
  }
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
  const int gv = grid_vec(s->nrow);
  hipLaunchKernelGGL(k_cg_residual, dim3(grid_rows(s->nrow)), dim3(kCgThreads), 0,
                     c->stream, s->nrow, s->dia, s->dja, s->da, s->x, s->b, s->r,
                     s->d, s->pc);
  hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kCgThreads), 0, c->stream, s->nrow,
                     s->b, s->r, s->d, s->z, s->p0, s->rzb, s->pq);
  hipLaunchKernelGGL(k_cg_setup_final, dim3(1), dim3(kCgThreads), 0, c->stream, s->pq,
                     s->rzb, gv, s->st);
  XG_HIP(c, hipGetLastError());
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
  const double nb = s->hst->normb > 1.0e-14 ? s->hst->normb : 1.0;
  rc = run_cg<double>(c, s->da, s->z, s->p0, s->p1, s->q, s->d, s->r, s->x, maxit,
                      tol * nb, 2.220446049250313e-16);
  if (rc) return rc;
  const CgState* h = s->hst;
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
  if (s->restore && (rc = restore_values(c))) return rc;
  const int64_t n = s->nrow;
  const int gr = grid_rows(n), gv = grid_vec(n);
  // preconditioner (FP64 diagonal, then FP32 copy) and ||b||
  hipLaunchKernelGGL(k_cg_residual, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                     s->dia, s->dja, s->da, s->x, s->b, s->r, s->d, s->pc);
  hipLaunchKernelGGL(k_f32_from_f64, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n,
                     s->d, s->d32);
  hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->b,
                     s->r, s->d, s->z, s->p0, s->rzb, s->pq);
  hipLaunchKernelGGL(k_cg_setup_final, dim3(1), dim3(kCgThreads), 0, c->stream, s->pq,
                     s->rzb, gv, s->st);
  if ((rc = read_state(c))) return rc;
  const double normb = s->hst->normb;
  const double target = tol * (normb > 1.0e-14 ? normb : 1.0);
  int64_t total = 0, outer = 0;
  double nr = 0.0;
  int conv = 0;
  for (;;) {
    hipLaunchKernelGGL(k_residual_norm, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                       s->dia, s->dja, s->da, s->x, s->b, s->r, s->pq);
    hipLaunchKernelGGL(k_sum_to_state, dim3(1), dim3(kCgThreads), 0, c->stream, s->pq,
                       gr, s->st);
    if ((rc = read_state(c))) return rc;
    nr = sqrt(s->hst->rr);
    if (nr < target) { conv = 1; break; }
    if (total >= maxit) break;
    ++outer;
    // inner FP32 solve of A e = r to ||r_e|| < inner_tol ||r|| (a relative
    // stop only: the breakdown threshold is a denormal-scale guard)
    hipLaunchKernelGGL(k_cg32_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->r,
                       s->d32, s->r32, s->z32, s->p32a, s->e32, s->rzb);
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, c->stream, s->st);
    rc = run_cg<float>(c, s->da32, s->z32, s->p32a, s->p32b, s->q32, s->d32, s->r32,
                       s->e32, std::min<int64_t>(inner_maxit, maxit - total),
                       inner_tol * nr, 1.0e-300);
    if (rc) return rc;
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

// Mixed precision with reliable updates (no counterpart in the reference):
// the Jacobi-PCG recurrence runs in FP32 (FP32 matrix values, vectors and
// SpMV; the solution increment accumulated in FP32 and added to the FP64
// x).  Whenever the FP32 residual has fallen by `delta` since the last
// update, x takes the increment, the FP64 residual r = b - A x is computed
// and replaces the FP32 one, and the recurrence continues with its search
// direction (Sleijpen and van der Vorst's reliable updates).  Stops when
// the FP64 residual ||b - A x|| < tol ||b||.
int xg_cg_solve_reliable(xg_ctx* c, int64_t maxit, double tol, double delta,
                         double* normr, int64_t* iters, int64_t* nupdates,
                         int* converged) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  if (maxit < 1 || !(tol > 0.0) || !(delta > 0.0 && delta < 1.0))
    return fail(c, XG_EINVAL, "xg_cg_solve_reliable: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  if (s->restore && (rc = restore_values(c))) return rc;
  const int64_t n = s->nrow;
  const int gr = grid_rows(n), gv = grid_vec(n);
  hipLaunchKernelGGL(k_cg_residual, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                     s->dia, s->dja, s->da, s->x, s->b, s->r, s->d, s->pc);
  hipLaunchKernelGGL(k_f32_from_f64, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n,
                     s->d, s->d32);
  hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->b,
                     s->r, s->d, s->z, s->p0, s->rzb, s->pq);
  hipLaunchKernelGGL(k_cg_setup_final, dim3(1), dim3(kCgThreads), 0, c->stream, s->pq,
                     s->rzb, gv, s->st);
  if ((rc = read_state(c))) return rc;
  const double normb = s->hst->normb;
  const double target = tol * (normb > 1.0e-14 ? normb : 1.0);
  int64_t done = 0, nup = 0, j = 1;
  double nr = 0.0;
  int conv = 0;
  for (;;) {
    // the FP64 residual of x
    hipLaunchKernelGGL(k_residual_norm, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                       s->dia, s->dja, s->da, s->x, s->b, s->r, s->pq);
    hipLaunchKernelGGL(k_sum_to_state, dim3(1), dim3(kCgThreads), 0, c->stream, s->pq,
                       gr, s->st);
    if ((rc = read_state(c))) return rc;
    nr = sqrt(s->hst->rr);
    if (nr < target) { conv = 1; break; }
    if (done >= maxit) break;
    if (j == 1) {   // first pass: p_0 = r
      hipLaunchKernelGGL(k_cg32_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->r,
                         s->d32, s->r32, s->z32, s->p32a, s->e32, s->rzb);
    } else {        // reliable update: the recurrence goes on from the true residual
      ++nup;
      hipLaunchKernelGGL(k_cg32_replace, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                         s->r, s->d32, s->r32, s->z32, s->e32,
                         s->rzb + ((j - 1) & 1) * kBufLen);
    }
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, c->stream, s->st);
    rc = run_cg<float>(c, s->da32, s->z32, s->p32a, s->p32b, s->q32, s->d32, s->r32,
                       s->e32, maxit, delta * nr, 1.0e-300, j);
    if (rc) return rc;
    const int64_t it = s->hst->it;       // iterations done in total
    hipLaunchKernelGGL(k_add_f32, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->e32,
                       s->x);
    if (it < j) break;                   // no progress possible in FP32
    done = it;
    j = it + 1;
  }
  if (normr) *normr = nr;
  if (iters) *iters = done;
  if (nupdates) *nupdates = nup;
  if (converged) *converged = conv;
  s->ready = false;
  s->restore = s->dev_bc;
  return XG_OK;
}

// Flexible CG (no counterpart in the reference): Notay's FCG(1) in FP64 with
// a preconditioner that may change from one application to the next.
// pc_kind 0: Jacobi; 1: `inner` iterations of FP32 Jacobi-PCG on A e = r
// from e = 0 (nested Krylov: the FP64 outer recurrence keeps its search
// directions, so the inner solves are not restarts of the outer one).  Stops
// when the true-recurrence residual ||r|| < tol * ||b|| or after maxit outer
// iterations.
int xg_cg_solve_flexible(xg_ctx* c, int pc_kind, int64_t inner, int64_t maxit,
                         double tol, double* normr, int64_t* iters,
                         int64_t* inner_total, int* converged) {
  int rc = check_cg(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  if (maxit < 1 || !(tol > 0.0) || pc_kind < 0 || pc_kind > 3 ||
      (pc_kind == 1 && inner < 1))
    return fail(c, XG_EINVAL, "xg_cg_solve_flexible: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  if (s->restore && (rc = restore_values(c))) return rc;
  const int64_t n = s->nrow;
  const int gr = grid_rows(n), gv = grid_vec(n);
  const int gt = std::max(1, std::min(s->ntile, kMaxParts));   // FCG SpMV tiles
  hipLaunchKernelGGL(k_cg_residual, dim3(gr), dim3(kCgThreads), 0, c->stream, n,
                     s->dia, s->dja, s->da, s->x, s->b, s->r, s->d, s->pc);
  hipLaunchKernelGGL(k_f32_from_f64, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n,
                     s->d, s->d32);
  hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n, s->b,
                     s->r, s->d, s->z, s->p0, s->rzb, s->pq);
  hipLaunchKernelGGL(k_cg_setup_final, dim3(1), dim3(kCgThreads), 0, c->stream, s->pq,
                     s->rzb, gv, s->st);
  if (pc_kind == 2 &&
      (rc = xg_gnn_prepare_internal(c, s->dia, s->dja, s->da, s->d, n)))
    return rc;
  if (pc_kind == 3) {
    if (s->ncomp != 1)
      return fail(c, XG_EUNSUPPORTED, "AMG preconditioner: one component per node only");
    if ((rc = xg_amg_prepare_internal(c, n, s->ia.data(), s->ja.data(), s->abc.data(),
                                      s->vals_version)))
      return rc;
  }
  if ((rc = read_state(c))) return rc;
  const double normb = s->hst->normb;
  const double target = tol * (normb > 1.0e-14 ? normb : 1.0);
  double nr = sqrt(s->hst->rr);
  int64_t k = 0, tin = 0;
  double* P[2] = {s->p0, s->p1};
  double* Q[2] = {s->q, s->q2};
  // the learned preconditioner is judged by the residual: once ||r|| has not
  // improved on its best value for kJudge iterations, the rest of the solve
  // uses Jacobi (flexible CG allows the switch)
  constexpr int64_t kJudge = 20;
  double best = nr;
  int64_t kbest = 0;
  bool learned = pc_kind == 2;
  for (; k < maxit && !(nr < target); ++k) {
    double* Bk = s->fb + (k & 1) * 4 * kMaxParts;
    double* Bp = s->fb + ((k + 1) & 1) * 4 * kMaxParts;
    if (pc_kind == 0 || (pc_kind == 2 && !learned)) {
      if (pc_kind == 2) ++tin;
      hipLaunchKernelGGL(k_jacobi_apply, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                         s->r, s->d, s->z);
    } else if (pc_kind == 3) {
      if ((rc = xg_amg_apply_internal(c, s->r, s->z))) return rc;
      ++tin;
    } else if (pc_kind == 2) {
      if ((rc = xg_gnn_apply_internal(c, s->dia, s->dja, s->r, s->d, s->z))) return rc;
      int used = 0;
      if ((rc = xg_gnn_fallback_flag(c, &used))) return rc;
      tin += used;
    } else {
      hipLaunchKernelGGL(k_cg32_init, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                         s->r, s->d32, s->r32, s->z32, s->p32a, s->e32, s->rzb);
      hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, c->stream, s->st);
      if ((rc = run_cg<float>(c, s->da32, s->z32, s->p32a, s->p32b, s->q32, s->d32,
                              s->r32, s->e32, inner, 0.0, 1.0e-300)))
        return rc;
      tin += s->hst->it;
      hipLaunchKernelGGL(k_f64_from_f32, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                         s->e32, s->z);
    }
    hipLaunchKernelGGL(k_fcg_pack_dot, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                       k > 0 ? 1 : 0,
                       s->z, Q[(k + 1) & 1], P[(k + 1) & 1], s->zp, Bk);
    hipLaunchKernelGGL(k_fcg_spmv, dim3(gt), dim3(kCgThreads), 0, c->stream, k == 0 ? 1 : 0,
                       s->ntile, s->dtile, s->dia, s->dja, s->da, s->zp, P[k & 1], Q[k & 1],
                       s->r, Bp, Bk, gv, gt);
    hipLaunchKernelGGL(k_fcg_update, dim3(gv), dim3(kCgThreads), 0, c->stream, n,
                       P[k & 1], Q[k & 1], s->r, s->x, Bk, gt);
    hipLaunchKernelGGL(k_sum_to_state, dim3(1), dim3(kCgThreads), 0, c->stream,
                       Bk + 3 * kMaxParts, gv, s->st);
    XG_HIP(c, hipGetLastError());
    if ((rc = read_state(c))) return rc;
    nr = sqrt(s->hst->rr);
    if (nr < best) {
      best = nr;
      kbest = k;
    } else if (learned && k - kbest >= kJudge) {
      learned = false;
    }
  }
  if (normr) *normr = nr;
  if (iters) *iters = k;
  if (inner_total) *inner_total = tin;
  if (converged) *converged = nr < target ? 1 : 0;
  s->ready = false;
  s->restore = s->dev_bc;
  return XG_OK;
}

// ---- learned preconditioner entry points that need the CG system ----

static int check_gnn(xg_ctx* c) {
  int rc = check_cg(c);
  if (rc) return rc;
  if (!c->gnn) return fail(c, XG_ENOMESH, "no learned preconditioner (xg_gnn_create)");
  return XG_OK;
}

// e.r on the device, read back
static int dot_dev(xg_ctx* c, const double* a, const double* b, double* out) {
  CgSolver* s = c->cg;
  const int gv = grid_vec(s->nrow);
  hipLaunchKernelGGL(k_fcg_dot, dim3(gv), dim3(kCgThreads), 0, c->stream, s->nrow, a, b,
                     s->fb);
  hipLaunchKernelGGL(k_sum_to_state, dim3(1), dim3(kCgThreads), 0, c->stream, s->fb, gv,
                     s->st);
  int rc = read_state(c);
  if (rc) return rc;
  *out = s->hst->rr;
  return XG_OK;
}

int xg_gnn_prepare(xg_ctx* c) {
  int rc = check_gnn(c);
  if (rc) return rc;
  CgSolver* s = c->cg;
  XG_HIP(c, hipSetDevice(c->device));
  return xg_gnn_prepare_internal(c, s->dia, s->dja, s->da, s->d, s->nrow);
}

int xg_gnn_apply_dev(xg_ctx* c, const double* r, double* z) {
  int rc = check_gnn(c);
  if (rc) return rc;
  if (!r || !z) return fail(c, XG_EINVAL, "xg_gnn_apply_dev: null vector");
  CgSolver* s = c->cg;
  XG_HIP(c, hipSetDevice(c->device));
  return xg_gnn_forward_internal(c, s->dia, s->dja, r, s->d, z);
}

int xg_gnn_loss_grad_dev(xg_ctx* c, const double* r, const double* e, double* z,
                         double* loss) {
  int rc = check_gnn(c);
  if (rc) return rc;
  if (!r || !e || !z) return fail(c, XG_EINVAL, "xg_gnn_loss_grad_dev: null vector");
  CgSolver* s = c->cg;
  XG_HIP(c, hipSetDevice(c->device));
  double etr;
  if ((rc = dot_dev(c, e, r, &etr))) return rc;
  if (etr == 0.0 || !std::isfinite(etr))
    return fail(c, XG_EINVAL, "xg_gnn_loss_grad_dev: e.r must be finite and nonzero");
  if ((rc = xg_gnn_forward_internal(c, s->dia, s->dja, r, s->d, z))) return rc;
  return xg_gnn_backward_internal(c, s->dia, s->dja, s->da, r, etr, z, loss);
}

// Training pairs from the system set up by xg_cg_init (with its Dirichlet
// conditions applied): the reference solution x* by Jacobi-PCG to 1e-12,
// then restarted PCG from x = 0 in runs of `stride` iterations; at every
// restart (r = b - A x, e = x* - x) is one pair: forward, backward, and an
// Adam step every `batch` pairs.  Returns the mean pair loss and the count.
int xg_gnn_train_system(xg_ctx* c, int64_t stride, int64_t max_pairs, int batch,
                        double lr, double* mean_loss, int64_t* npairs) {
  int rc = check_gnn(c);
  if (rc) return rc;
  if (stride < 1 || max_pairs < 1 || batch < 1 || !(lr > 0.0))
    return fail(c, XG_EINVAL, "xg_gnn_train_system: bad arguments");
  CgSolver* s = c->cg;
  XG_HIP(c, hipSetDevice(c->device));
  const size_t n = (size_t)s->nrow;
  const int gv = grid_vec(s->nrow);
  XG_HIP(c, hipMemsetAsync(s->x, 0, n * 8, c->stream));
  if ((rc = xg_cg_setup(c, nullptr))) return rc;
  const double normb = s->hst->normb;
  const double nb = normb > 1.0e-14 ? normb : 1.0;
  if ((rc = run_cg<double>(c, s->da, s->z, s->p0, s->p1, s->q, s->d, s->r, s->x,
                           20 * (int64_t)std::cbrt((double)n) * 10, 1.0e-12 * nb,
                           2.220446049250313e-16)))
    return rc;
  XG_HIP(c, hipMemcpyAsync(s->xs, s->x, n * 8, hipMemcpyDeviceToDevice, c->stream));
  XG_HIP(c, hipMemsetAsync(s->x, 0, n * 8, c->stream));
  if ((rc = xg_gnn_prepare_internal(c, s->dia, s->dja, s->da, s->d, s->nrow))) return rc;
  double lsum = 0.0;
  int64_t np = 0;
  double* z = xg_gnn_zbuf(c);
  while (np < max_pairs) {
    if ((rc = xg_cg_setup(c, nullptr))) return rc;
    if (sqrt(s->hst->rr) < 1.0e-10 * nb) break;
    hipLaunchKernelGGL(k_sub, dim3(gv), dim3(kCgThreads), 0, c->stream, s->nrow, s->xs,
                       s->x, s->ebuf);
    double etr;
    if ((rc = dot_dev(c, s->ebuf, s->r, &etr))) return rc;
    if (etr == 0.0 || !std::isfinite(etr)) break;
    double L = 0.0;
    if ((rc = xg_gnn_forward_internal(c, s->dia, s->dja, s->r, s->d, z)) ||
        (rc = xg_gnn_backward_internal(c, s->dia, s->dja, s->da, s->r, etr, z, &L)))
      return rc;
    lsum += L;
    ++np;
    if (np % batch == 0 && (rc = xg_gnn_adam(c, lr, 0.9, 0.999, 1.0e-8))) return rc;
    if ((rc = run_cg<double>(c, s->da, s->z, s->p0, s->p1, s->q, s->d, s->r, s->x,
                             stride, 0.0, 2.220446049250313e-16)))
      return rc;
  }
  if ((rc = xg_gnn_adam(c, lr, 0.9, 0.999, 1.0e-8))) return rc;
  s->ready = false;
  s->restore = s->dev_bc;
  if (mean_loss) *mean_loss = np ? lsum / np : 0.0;
  if (npairs) *npairs = np;
  return XG_OK;
}

int xg_cg_destroy(xg_ctx* c) {
  if (!c) return XG_EINVAL;
  free_solver(c->cg);
  c->cg = nullptr;
  return XG_OK;
}

}  // extern "C"
