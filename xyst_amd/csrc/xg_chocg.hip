// MI355X device path of the ChoCG edge operators: chorin::div, grad, vgrad,
// flux and rhs (advection with damp2/damp4), Chorin.cpp:34-1044, for one
// GPU-resident partition.
//
// All five are sums over mesh edges of an edge value scattered -/+ to the
// edge's ends, plus a boundary-face term with (6,1,1)/8 corner weights.  They
// share one node-block kernel (the layout of xg_layout.cpp: node blocks of
// consecutive ids, edges/faces owned by the block of their smallest node):
// the block's nodes are staged in LDS with the NV values the operator reads;
// per chunk each thread computes one edge's (or one face corner's) NF values
// into LDS; the thread owning local node t sums the chunk's slots touching it
// in slot order; halo partials go to slots that k_fixup adds in block order.
// No atomics: every result is bitwise reproducible.
//
// div/grad/vgrad/flux add to their output, as the reference's do (the caller
// zeroes it); rhs overwrites R.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "xg_blocks.hpp"
#include "xg_cg.hpp"
#include "xg_csr.hpp"
#include "xg_ctx.hpp"
#include "xg_flux.hpp"
#include "xg_layout.hpp"

using namespace xgd;

struct ChoMesh {
  int npoin = 0;
  int nblk = 0;
  std::vector<DevArr> allocs;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  int4 *blk0 = nullptr, *blk1 = nullptr;
  int* halo = nullptr;
  int4* chk = nullptr;
  int* chk_ent = nullptr;
  uint16_t *tend = nullptr, *entries = nullptr;
  uint32_t* elpq = nullptr;
  double *dx = nullptr, *dy = nullptr, *dz = nullptr, *dw = nullptr;
  uint32_t *fab = nullptr, *fcs = nullptr;
  double *fnx = nullptr, *fny = nullptr, *fnz = nullptr;
  int nslot = 0;
  double* slots = nullptr;     // halo partials, kSlotVals per slot
  int nfx = 0;
  int *fx_node = nullptr, *fx_off = nullptr, *fx_slot = nullptr;
  // staging of the host-pointer calls (grow only)
  std::vector<DevArr> stage;
  // ChoCG stepping (xg_chocg_setup)
  int nc = 3;                  // components of U: velocity + transported scalars
  bool have_step = false;
  std::vector<DevArr> step_allocs;
  xg_chostep cfg{};
  double* vol = nullptr;
  int nsymnode = 0, nnoslip = 0, ndir = 0;
  int *symnode = nullptr, *symoff = nullptr, *noslip = nullptr, *dirnode = nullptr;
  double *symnorm = nullptr, *dirval = nullptr;
  uint8_t* dirmask = nullptr;
  std::vector<size_t> pdir;                 // pressure Dirichlet nodes (host)
  std::vector<int32_t> pmask;
  std::vector<double> pval0, pzero;         // initial values / increments (0)
  double *pgrad = nullptr, *vgrad = nullptr, *sgrad = nullptr, *R = nullptr;
  double *div = nullptr, *F = nullptr, *red = nullptr;
  double* usave = nullptr;     // a stage's starting state (frozen flow)
  std::vector<int> h_dirnode, h_noslip;    // host copies of the BC lists
  std::vector<uint8_t> h_dirmask;
  // theta-implicit momentum solve (xg_chocg_momentum_setup)
  double theta = 0.0;
  int m_pc = 1;
  int64_t m_maxit = 50, m_nnz = 0, m_its = 0;
  double m_tol = 1e-3;
  CgSolver* cgm = nullptr;
  std::vector<DevArr> mom_allocs;
  double *mM = nullptr, *mK = nullptr;    // per-entry mass and stiffness sums
  int nzp = 0, ndp = 0, ndr = 0;
  int *zpos = nullptr, *dpos = nullptr, *drow = nullptr;
};

namespace {

constexpr int kT = xg::kChunkSlots;
constexpr int kCap = xg::kBlockCap;

enum { OP_GRAD = 0, OP_VGRAD, OP_DIV, OP_DIVS, OP_FLUX, OP_ADV2, OP_ADV4 };

// NV staged values per node, NF values per edge; NC: components of U
// (velocity + transported scalars) for vgrad and the advection operators
template <int OP, int NC> struct OpShape;
template <int NC> struct OpShape<OP_GRAD, NC>  { static constexpr int NV = 1,  NF = 3; };
template <int NC> struct OpShape<OP_VGRAD, NC> { static constexpr int NV = NC, NF = 3 * NC; };
template <int NC> struct OpShape<OP_DIV, NC>   { static constexpr int NV = 3,  NF = 1; };
template <int NC> struct OpShape<OP_DIVS, NC>  { static constexpr int NV = 10, NF = 1; };
template <int NC> struct OpShape<OP_FLUX, NC>  { static constexpr int NV = 12, NF = 3; };
template <int NC> struct OpShape<OP_ADV2, NC>  { static constexpr int NV = NC + 1, NF = NC; };
template <int NC> struct OpShape<OP_ADV4, NC>  { static constexpr int NV = 4 * NC + 4, NF = NC; };
// values per halo slot: the widest operator, vgrad of velocity + scalar
constexpr int kSlotVals = OpShape<OP_VGRAD, 4>::NF;

struct ChoArgs {
  int nblk;
  const int4 *blk0, *blk1;
  const int* halo;
  const int4* chk;
  const int* chk_ent;
  const uint16_t *tend, *entries;
  const uint32_t* elpq;
  const double *dx, *dy, *dz, *dw;
  const uint32_t *fab, *fcs;
  const double *fnx, *fny, *fnz;
  const double *x, *y, *z;
  const double* U;   // velocity (columns 0..2) or the scalar (GRAD)
  int ldu;
  const double* P;   // pressure
  const double* G;   // pressure gradient (DIVS, 3 of ldg) or velocity gradients (9 of ldg)
  int ldg;
  double* out;
  double* S;
  int accum;
  double dt, mu, stab2coef, diff;
  int stab, stab2;
};

// staged value c of global node g; layout per operator (NC components):
//   GRAD: u | VGRAD: u0..NC-1 | DIV: u0 u1 u2 | DIVS: u0..2 P G0..2 x y z
//   FLUX: u0..2 G0..8 | ADV2: u0..NC-1 P | ADV4: u0..NC-1 P G0..3NC-1 x y z
template <int OP, int NC>
__device__ inline double node_value(const ChoArgs& a, int g, int c) {
  if (OP == OP_GRAD || OP == OP_VGRAD || OP == OP_DIV) return a.U[(size_t)g * a.ldu + c];
  if (OP == OP_FLUX) return c < 3 ? a.U[(size_t)g * a.ldu + c] : a.G[(size_t)g * a.ldg + (c - 3)];
  if (OP == OP_DIVS) {
    if (c < 3) return a.U[(size_t)g * a.ldu + c];
    if (c == 3) return a.P[g];
    if (c < 7) return a.G[(size_t)g * a.ldg + (c - 4)];
    return c == 7 ? a.x[g] : c == 8 ? a.y[g] : a.z[g];
  }
  // ADV2, ADV4
  if (c < NC) return a.U[(size_t)g * a.ldu + c];
  if (c == NC) return a.P[g];
  if (c < 4 * NC + 1) return a.G[(size_t)g * a.ldg + (c - NC - 1)];
  const int k = c - 4 * NC - 1;
  return k == 0 ? a.x[g] : k == 1 ? a.y[g] : a.z[g];
}

// momentum flux of chorin::flux for edge nodes (P, Q) (Chorin.cpp:451-479),
// from staged u (0..2) and gradients (3..11)
__device__ inline double mflux(const double* sN, int P, int Q, int i, int j, double mu) {
  const double inv = sN[i * kCap + P] * sN[j * kCap + P] + sN[i * kCap + Q] * sN[j * kCap + Q];
  if (mu < 2.220446049250313e-16) return -inv;
  double vis = sN[(3 + i * 3 + j) * kCap + P] + sN[(3 + j * 3 + i) * kCap + P] +
               sN[(3 + i * 3 + j) * kCap + Q] + sN[(3 + j * 3 + i) * kCap + Q];
  if (i == j)
    vis -= 2.0 / 3.0 * (sN[3 * kCap + P] + sN[7 * kCap + P] + sN[11 * kCap + P] +
                        sN[3 * kCap + Q] + sN[7 * kCap + Q] + sN[11 * kCap + Q]);
  return mu * vis - inv;
}

__device__ inline double mflux_point(const double* sN, int P, int i, int j, double mu) {
  const double inv = sN[i * kCap + P] * sN[j * kCap + P];
  if (mu < 2.220446049250313e-16) return -inv;
  double vis = sN[(3 + i * 3 + j) * kCap + P] + sN[(3 + j * 3 + i) * kCap + P];
  if (i == j) vis -= 2.0 / 3.0 * (sN[3 * kCap + P] + sN[7 * kCap + P] + sN[11 * kCap + P]);
  return mu * vis - inv;
}

// advective flux of adv_damp2/adv_damp4 (Chorin.cpp:640-829) on edge (p, q):
// velocity (0..2) and NC - 3 transported scalars
template <int OP, int NC>
__device__ inline void adv_edge(const ChoArgs& a, const double* sN, int p, int q,
                                const double d[3], double dwv, double f[NC]) {
  double uL[NC], uR[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    uL[c] = sN[c * kCap + p];
    uR[c] = sN[c * kCap + q];
  }
  double vdL[NC];  // unreconstructed differences for the viscous/diffusive terms
#pragma unroll
  for (int c = 0; c < NC; ++c) vdL[c] = uR[c] - uL[c];
  if (OP == OP_ADV4) {
    constexpr int X0 = 4 * NC + 1, G0 = NC + 1;
    const double vx = sN[X0 * kCap + q] - sN[X0 * kCap + p];
    const double vy = sN[(X0 + 1) * kCap + q] - sN[(X0 + 1) * kCap + p];
    const double vz = sN[(X0 + 2) * kCap + q] - sN[(X0 + 2) * kCap + p];
    constexpr double eps = 1.0e-9, c1 = 1.0 - 1.0 / 3.0, c2 = 1.0 + 1.0 / 3.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double g1 = sN[(G0 + c * 3) * kCap + p] * vx + sN[(G0 + 1 + c * 3) * kCap + p] * vy +
                        sN[(G0 + 2 + c * 3) * kCap + p] * vz;
      const double g2 = sN[(G0 + c * 3) * kCap + q] * vx + sN[(G0 + 1 + c * 3) * kCap + q] * vy +
                        sN[(G0 + 2 + c * 3) * kCap + q] * vz;
      const double delta2 = uR[c] - uL[c];
      const double delta1 = 2.0 * g1 - delta2;
      const double delta3 = 2.0 * g2 - delta2;
      // van Leer: phi(r) = (|r|+r)/(|r|+1) with r = s/t is 2s/(s+t) when s
      // and t have the same sign and 0 otherwise; phi(1/r) likewise
      const double A = delta2 + eps, BL = delta1 + eps, BR = delta3 + eps;
      const double il = xg::vl_rcp(A, BL);
      const double ir = xg::vl_rcp(A, BR);
      uL[c] += 0.5 * il * (delta1 * c1 * A + delta2 * c2 * BL);
      uR[c] -= 0.5 * ir * (delta3 * c1 * A + delta2 * c2 * BR);
    }
  }
  const double vnL = uL[0] * d[0] + uL[1] * d[1] + uL[2] * d[2];
  const double vnR = uR[0] * d[0] + uR[1] * d[1] + uR[2] * d[2];
  double aw = 0.0;
  if (a.stab) aw = fabs(vnL + vnR) * 0.5;
  if (a.stab2) aw += a.stab2coef * fmax(fabs(vnL), fabs(vnR));
  const double v = dwv * a.mu;
  const double pf = sN[NC * kCap + p] + sN[NC * kCap + q];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (OP == OP_ADV2)
      f[c] = uL[c] * vnL + uR[c] * vnR + pf * d[c] + (aw - v) * (uR[c] - uL[c]);
    else
      f[c] = uL[c] * vnL + uR[c] * vnR + pf * d[c] + aw * (uR[c] - uL[c]) - v * vdL[c];
  }
  if constexpr (NC > 3) {
    const double dd = dwv * a.diff;
#pragma unroll
    for (int c = 3; c < NC; ++c) {
      if (OP == OP_ADV2)
        f[c] = uL[c] * vnL + uR[c] * vnR + (aw - dd) * (uR[c] - uL[c]);
      else
        f[c] = uL[c] * vnL + uR[c] * vnR + aw * (uR[c] - uL[c]) - dd * vdL[c];
    }
  }
}

template <int OP, int NC>
__device__ inline void edge_values(const ChoArgs& a, const double* sN, int p, int q,
                                   const double d[3], double dwv,
                                   double f[OpShape<OP, NC>::NF]) {
  if (OP == OP_GRAD) {
    const double s = sN[q] + sN[p];
#pragma unroll
    for (int j = 0; j < 3; ++j) f[j] = d[j] * s;
  } else if (OP == OP_VGRAD) {
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const double s = sN[i * kCap + q] + sN[i * kCap + p];
#pragma unroll
      for (int j = 0; j < 3; ++j) f[i * 3 + j] = d[j] * s;
    }
  } else if (OP == OP_DIV || OP == OP_DIVS) {
    double dv = d[0] * (sN[p] + sN[q]) + d[1] * (sN[kCap + p] + sN[kCap + q]) +
                d[2] * (sN[2 * kCap + p] + sN[2 * kCap + q]);
    if (OP == OP_DIVS) {
      const double ex = sN[7 * kCap + p] - sN[7 * kCap + q];
      const double ey = sN[8 * kCap + p] - sN[8 * kCap + q];
      const double ez = sN[9 * kCap + p] - sN[9 * kCap + q];
      const double dl = sqrt(ex * ex + ey * ey + ez * ez);
      const double p2 = sN[3 * kCap + q] - sN[3 * kCap + p];
      const double D = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double p4 = 0.5 * (ex * (sN[4 * kCap + p] + sN[4 * kCap + q]) +
                               ey * (sN[5 * kCap + p] + sN[5 * kCap + q]) +
                               ez * (sN[6 * kCap + p] + sN[6 * kCap + q]));
      dv += D * a.dt * xg::rcp(dl) * (p2 + p4);
    }
    f[0] = dv;
  } else if (OP == OP_FLUX) {
    // the reference evaluates flux(U,G,i,j,N[1],N[0]) for edge N[0]-N[1]
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) s += d[j] * mflux(sN, q, p, i, j, a.mu);
      f[i] = s;
    }
  } else {
    adv_edge<OP, NC>(a, sN, p, q, d, dwv, f);
  }
}

// corner `k` (weight 6 on it, 1 on the others, /8) of boundary face (A,B,C)
template <int OP, int NC>
__device__ inline void face_values(const ChoArgs& a, const double* sN, const int N[3],
                                   int k, const double n[3],
                                   double f[OpShape<OP, NC>::NF]) {
  const double w0 = k == 0 ? 6.0 : 1.0, w1 = k == 1 ? 6.0 : 1.0, w2 = k == 2 ? 6.0 : 1.0;
  if (OP == OP_GRAD) {
    const double u = (w0 * sN[N[0]] + w1 * sN[N[1]] + w2 * sN[N[2]]) / 8.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) f[j] = u * n[j];
  } else if (OP == OP_VGRAD) {
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const double u = (w0 * sN[i * kCap + N[0]] + w1 * sN[i * kCap + N[1]] +
                        w2 * sN[i * kCap + N[2]]) / 8.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) f[i * 3 + j] = u * n[j];
    }
  } else if (OP == OP_DIV || OP == OP_DIVS) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      s += (w0 * sN[j * kCap + N[0]] + w1 * sN[j * kCap + N[1]] + w2 * sN[j * kCap + N[2]]) /
           8.0 * n[j];
    f[0] = s;
  } else if (OP == OP_FLUX) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
        s += (w0 * mflux_point(sN, N[0], i, j, a.mu) + w1 * mflux_point(sN, N[1], i, j, a.mu) +
              w2 * mflux_point(sN, N[2], i, j, a.mu)) / 8.0 * n[j];
      f[i] = s;
    }
  } else {
    double fn[3][NC];   // [node][component]
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int P = N[m];
      const double u = sN[P], v = sN[kCap + P], w = sN[2 * kCap + P], pr = sN[NC * kCap + P];
      const double vn = n[0] * u + n[1] * v + n[2] * w;
      fn[m][0] = u * vn + pr * n[0];
      fn[m][1] = v * vn + pr * n[1];
      fn[m][2] = w * vn + pr * n[2];
#pragma unroll
      for (int c = 3; c < NC; ++c) fn[m][c] = sN[c * kCap + P] * vn;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = (w0 * fn[0][c] + w1 * fn[1][c] + w2 * fn[2][c]) / 8.0;
  }
}

template <int OP, int NC>
__global__ __launch_bounds__(kT) void k_cho_blocks(ChoArgs a) {
  constexpr int NV = OpShape<OP, NC>::NV, NF = OpShape<OP, NC>::NF;
  __shared__ double sN[NV * kCap];
  __shared__ double F[kT * NF];
  __shared__ uint16_t sE[2 * kT];
  __shared__ uint16_t sT[kCap];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, a.nblk);
  const int4 B0 = a.blk0[b];   // own0, nown, nloc, halo_off
  const int4 B1 = a.blk1[b];   // slot0, chunk0, chunk1
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  // stage (node, value) pairs over all threads, 4 loads in flight each
  for (int base = 0; base < nloc * NV; base += 4 * kT) {
    double v[4];
    int dst[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = base + t + r * kT;
      dst[r] = -1;
      v[r] = 0.0;
      if (k < nloc * NV) {
        const int i = k / NV, c = k - NV * i;
        const int g = i < nown ? own0 + i : a.halo[B0.w + i - nown];
        v[r] = node_value<OP, NC>(a, g, c);
        dst[r] = c * kCap + i;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (dst[r] >= 0) sN[dst[r]] = v[r];
  }
  double acc[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) acc[k] = 0.0;
  for (int ch = B1.y; ch < B1.z; ++ch) {
    const int4 C = a.chk[ch];   // kind, first, count, tend_off
    const int eoff = a.chk_ent[ch];
    const int nent = C.z * (C.x == 0 ? 2 : 3);
    uint16_t e0 = 0, e1 = 0, te = 0;
    if (t < nent) e0 = a.entries[eoff + t];
    if (t + kT < nent) e1 = a.entries[eoff + kT + t];
    if (t < nloc) te = a.tend[C.w + t];
    __syncthreads();          // staging done / previous chunk consumed
    if (t < nent) sE[t] = e0;
    if (t + kT < nent) sE[kT + t] = e1;
    if (t < nloc) sT[t] = te;
    if (t < C.z) {
      if (C.x == 0) {
        const int e = C.y + t;
        const uint32_t pq = a.elpq[e];
        const double d[3] = {a.dx[e], a.dy[e], a.dz[e]};
        const double dwv = (OP == OP_ADV2 || OP == OP_ADV4) ? a.dw[e] : 0.0;
        double f[NF];
        edge_values<OP, NC>(a, sN, (int)(pq & 0xffff), (int)(pq >> 16), d, dwv, f);
#pragma unroll
        for (int k = 0; k < NF; ++k) F[t * NF + k] = f[k];
      } else {
        const int fi = C.y + t;
        const uint32_t ab = a.fab[fi], cs = a.fcs[fi];
        const int N[3] = {(int)(ab & 0xffff), (int)(ab >> 16), (int)(cs & 0xffff)};
        const double n[3] = {a.fnx[fi], a.fny[fi], a.fnz[fi]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double f[NF];
          face_values<OP, NC>(a, sN, N, k, n, f);
#pragma unroll
          for (int m = 0; m < NF; ++m) F[(3 * t + k) * NF + m] = f[m];
        }
      }
    }
    __syncthreads();          // F, sE, sT complete
    if (t < nloc) {
      const int j0 = t ? sT[t - 1] : 0, j1 = sT[t];
      for (int j = j0; j < j1; ++j) {
        const int v = sE[j];
        const double* fs = F + (v & (kT - 1)) * NF;
        const double sg = (v & 0x8000) ? -1.0 : 1.0;
#pragma unroll
        for (int k = 0; k < NF; ++k) acc[k] = fma(sg, fs[k], acc[k]);
      }
    }
  }
  if (t < nown) {
    double* o = a.out + (size_t)(own0 + t) * NF;
#pragma unroll
    for (int k = 0; k < NF; ++k) o[k] = a.accum ? o[k] + acc[k] : acc[k];
  } else if (t < nloc) {
    double* o = a.S + (size_t)(B1.x + t - nown) * NF;
#pragma unroll
    for (int k = 0; k < NF; ++k) o[k] = acc[k];
  }
}

// Adds halo partials to an output with ldo columns of which NF are written.
template <int NF>
__global__ __launch_bounds__(kBlock) void k_fixup_ld(
    int nfx, const int* __restrict__ fx_node, const int* __restrict__ fx_off,
    const int* __restrict__ fx_slot, const double* __restrict__ S,
    double* __restrict__ R, int ldo) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfx) return;
  double* r = R + (size_t)fx_node[i] * ldo;
  double acc[NF];
#pragma unroll
  for (int c = 0; c < NF; ++c) acc[c] = r[c];
  for (int k = fx_off[i]; k < fx_off[i + 1]; ++k) {
    const double* s = S + (size_t)fx_slot[k] * NF;
#pragma unroll
    for (int c = 0; c < NF; ++c) acc[c] += s[c];
  }
#pragma unroll
  for (int c = 0; c < NF; ++c) r[c] = acc[c];
}

// Writes NF-wide node sums from a dense (npoin, NF) buffer into an output of
// ldo columns (outputs whose leading dimension is wider than NF).
__global__ void k_scatter_cols(int npoin, int nf, const double* __restrict__ src,
                               double* __restrict__ dst, int ldo, int accum) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * nf) return;
  const int64_t n = i / nf, c = i - n * nf;
  dst[n * ldo + c] = accum ? dst[n * ldo + c] + src[i] : src[i];
}

void free_cho(ChoMesh* m) {
  if (!m) return;
  if (m->cgm) xg_cg_free_solver(m->cgm);
  dev_free_all(m->mom_allocs);
  dev_free_all(m->step_allocs);
  dev_free_all(m->allocs);
  dev_free_all(m->stage);
  delete m;
}

ChoArgs base_args(const ChoMesh* m) {
  ChoArgs a{};
  a.nblk = m->nblk;
  a.blk0 = m->blk0; a.blk1 = m->blk1; a.halo = m->halo;
  a.chk = m->chk; a.chk_ent = m->chk_ent; a.tend = m->tend; a.entries = m->entries;
  a.elpq = m->elpq; a.dx = m->dx; a.dy = m->dy; a.dz = m->dz; a.dw = m->dw;
  a.fab = m->fab; a.fcs = m->fcs; a.fnx = m->fnx; a.fny = m->fny; a.fnz = m->fnz;
  a.x = m->x; a.y = m->y; a.z = m->z;
  a.S = m->slots;
  return a;
}

// Runs operator OP (NC components) into `out` (leading dimension ldo >= NF).
// Outputs with ldo == NF are written in place; wider ones go through a dense
// temporary.
template <int OP, int NC = 3>
int run_op(xg_ctx* c, ChoArgs a, double* out, int ldo, int accum, double* tmp) {
  constexpr int NF = OpShape<OP, NC>::NF;
  static_assert(NF <= kSlotVals, "the halo slots hold kSlotVals values each");
  ChoMesh* m = c->cho;
  const bool dense = ldo == NF;
  a.out = dense ? out : tmp;
  a.accum = dense ? accum : 0;
  if (m->nblk)
    hipLaunchKernelGGL((k_cho_blocks<OP, NC>), dim3(m->nblk), dim3(kT), 0, c->stream, a);
  if (m->nfx)
    hipLaunchKernelGGL(k_fixup_ld<NF>, dim3(nblocks(m->nfx)), dim3(kBlock), 0, c->stream,
                       m->nfx, m->fx_node, m->fx_off, m->fx_slot, m->slots, a.out, NF);
  if (!dense)
    hipLaunchKernelGGL(k_scatter_cols, dim3(nblocks((int64_t)m->npoin * NF)), dim3(kBlock),
                       0, c->stream, m->npoin, NF, tmp, out, ldo, accum);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int check_cho(xg_ctx* c) {
  if (!c) return XG_EINVAL;
  if (!c->cho) return fail(c, XG_ENOMESH, "no ChoCG mesh (xg_chocg_upload)");
  return XG_OK;
}

// Device buffer k of the host-pointer entry points (grow only).
int stage_buf(xg_ctx* c, int k, size_t bytes, double** p) {
  ChoMesh* m = c->cho;
  if ((int)m->stage.size() <= k) m->stage.resize(k + 1);
  DevArr& d = m->stage[k];
  if (d.bytes < bytes) {
    if (d.p) (void)hipFree(d.p);
    d.p = nullptr;
    d.bytes = 0;
    if (hipMalloc(&d.p, bytes) != hipSuccess) return fail(c, XG_EHIP, "hipMalloc (staging)");
    d.bytes = bytes;
  }
  *p = (double*)d.p;
  return XG_OK;
}

// in: host or device input -> device pointer (copied when host)
int dev_in(xg_ctx* c, int dev, int k, const double* src, size_t n, const double** out) {
  if (dev || !src) { *out = src; return XG_OK; }
  double* d;
  int rc = stage_buf(c, k, n * 8, &d);
  if (rc) return rc;
  XG_HIP(c, hipMemcpyAsync(d, src, n * 8, hipMemcpyHostToDevice, c->stream));
  *out = d;
  return XG_OK;
}

// ------------------------------------------------------------ stepping --

__global__ void k_fill0(int64_t n, double* __restrict__ p) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ChoCG::fingrad (ChoCG.cpp:841-866): divide by the nodal volume
__global__ void k_divvol_n(int npoin, int nc, double* __restrict__ G,
                           const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * nc) return;
  G[i] /= vol[i / nc];
}

// explicit stage of ChoCG::solve (ChoCG.cpp:1555-1560): U = Un - dt R / vol
__global__ void k_cho_update(int npoin, int nc, double dt, double* __restrict__ U,
                             const double* __restrict__ Un, const double* __restrict__ R,
                             const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * nc) return;
  U[i] = Un[i] - dt * R[i] / vol[i / nc];
}

// physics::dirbc on the nc components (BC.cpp:30-69)
__global__ void k_cho_dirbc(int nd, int nc, const int* __restrict__ node,
                            const uint8_t* __restrict__ mask,
                            const double* __restrict__ val, double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  double* u = U + (size_t)node[i] * nc;
  for (int c = 0; c < nc; ++c)
    if (mask[i * nc + c]) u[c] = val[i * nc + c];
}

// physics::noslipbc (BC.cpp:139-150): the velocity of nodes with nc
// components
__global__ void k_noslip(int n, int nc, const int* __restrict__ node, double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* u = U + (size_t)node[i] * nc;
  u[0] = u[1] = u[2] = 0.0;
}

// the velocity (components 0..2 of nc) of src copied into dst: the frozen
// flow of ChoCG::solve (ChoCG.cpp:1550-1568)
__global__ void k_copy_vel(int npoin, int nc, const double* __restrict__ src,
                           double* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * 3) return;
  const int64_t n = i / 3, c = i - 3 * n;
  dst[n * nc + c] = src[n * nc + c];
}

// chorin::src (Chorin.cpp:985-1008) of slot_cyl, the one ChoCG problem with a
// source (Problems.cpp:640-672): s = (-(x - 0.5), 0.5 - y, 0, 0), the
// acceleration of the prescribed rotation; R -= s vol
__global__ void k_cho_src_slot_cyl(int npoin, int nc, const double* __restrict__ x,
                                   const double* __restrict__ y,
                                   const double* __restrict__ vol, double* __restrict__ R) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npoin) return;
  const double v = vol[i];
  double* r = R + (size_t)i * nc;
  r[0] -= -(x[i] - 0.5) * v;
  r[1] -= (0.5 - y[i]) * v;
}

// ChoCG::pinit's div /= dt (ChoCG.cpp:1044)
__global__ void k_divide(int n, double* __restrict__ x, double dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] /= dt;
}

// ChoCG::psolved's projection (ChoCG.cpp:1208-1213): U -= dt grad(dp) on
// the velocity of nodes with nc components
__global__ void k_project(int npoin, int nc, double dt, double* __restrict__ U,
                          const double* __restrict__ sg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * 3) return;
  const int64_t n = i / 3, c = i - 3 * n;
  U[n * nc + c] -= dt * sg[i];
}

__global__ void k_add_to(int n, double* __restrict__ P, const double* __restrict__ dP) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) P[i] += dP[i];
}

// ChoCG::pinit's problem rhs (ChoCG.cpp:1112-1119, div = pr * vol) and
// ConjugateGradients' Neumann term (ConjugateGradients.cpp:382, 473: b +=
// neubc); either may be absent
__global__ void k_poisson_rhs(int n, const double* __restrict__ pr,
                              const double* __restrict__ vol,
                              const double* __restrict__ neubc, double* __restrict__ div) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = pr ? pr[i] * vol[i] : div[i];
  if (neubc) d += neubc[i];
  div[i] = d;
}

// per-node limits of ChoCG::dt (ChoCG.cpp:1384-1394), block minimum
__global__ __launch_bounds__(kBlock) void k_chodt_partial(
    int npoin, int nc, const double* __restrict__ U, const double* __restrict__ vol,
    double dif, double* __restrict__ part) {
  __shared__ double sm[kBlock / 64];
  constexpr double eps = 2.220446049250313e-16;
  double m = 1.7976931348623157e308;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npoin;
       p += (int64_t)gridDim.x * blockDim.x) {
    const double* s = U + p * nc;
    const double vel = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double L = cbrt(vol[p]);
    m = fmin(m, L / fmax(vel, 1.0e-8));
    if (dif > eps) m = fmin(m, L * L / dif);
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int k = 1; k < kBlock / 64; ++k) b = fmin(b, sm[k]);
    part[blockIdx.x] = b;
  }
}

// the momentum matrix of ChoCG::lhs (ChoCG.cpp:1434-1477) for step dt:
// a = -(M/dt + theta mu K) from the per-entry mass and stiffness sums
__global__ void k_mom_values(int64_t nnz, const double* __restrict__ M,
                             const double* __restrict__ K, double dt, double tmu,
                             double* __restrict__ a) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) a[k] = -(M[k] / dt + tmu * K[k]);
}

__global__ void k_set_at(int n, const int* __restrict__ pos, double v,
                         double* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[pos[i]] = v;
}

// u = un + du (ChoCG::msolved, ChoCG.cpp:1636-1641)
__global__ void k_add_into(int64_t n, const double* __restrict__ un,
                           const double* __restrict__ du, double* __restrict__ u) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = un[i] + du[i];
}

// the momentum solver made the context's solver for the guard's lifetime
struct MomentumSolver {
  xg_ctx* c;
  CgSolver* saved;
  MomentumSolver(xg_ctx* c_, CgSolver* s) : c(c_) { saved = xg_cg_exchange(c, s); }
  ~MomentumSolver() { xg_cg_exchange(c, saved); }
};

// ChoCG::solve's semi-implicit branch, msolve, msolved (ChoCG.cpp:1570-1646):
// A du = R with the Dirichlet masks and noslip nodes set to du = 0, CG from
// the previous du; U = Un + du
int cho_momentum(xg_ctx* c, double dt, double* U, const double* Un) {
  ChoMesh* m = c->cho;
  const int64_t n3 = (int64_t)m->npoin * 3;
  MomentumSolver g(c, m->cgm);
  double* a = xg_cg_values_dev(c);
  hipLaunchKernelGGL(k_mom_values, dim3(nblocks(m->m_nnz)), dim3(kBlock), 0, c->stream,
                     m->m_nnz, m->mM, m->mK, dt, m->theta * m->cfg.flux.mu, a);
  if (m->nzp)
    hipLaunchKernelGGL(k_set_at, dim3(nblocks(m->nzp)), dim3(kBlock), 0, c->stream, m->nzp,
                       m->zpos, 0.0, a);
  if (m->ndp)
    hipLaunchKernelGGL(k_set_at, dim3(nblocks(m->ndp)), dim3(kBlock), 0, c->stream, m->ndp,
                       m->dpos, 1.0, a);
  double* b = xg_cg_rhs_dev(c);
  XG_HIP(c, hipMemcpyAsync(b, m->R, (size_t)n3 * 8, hipMemcpyDeviceToDevice, c->stream));
  if (m->ndr)
    hipLaunchKernelGGL(k_set_at, dim3(nblocks(m->ndr)), dim3(kBlock), 0, c->stream, m->ndr,
                       m->drow, 0.0, b);
  XG_HIP(c, hipGetLastError());
  int rc = xg_cg_raw_init(c, m->m_pc);
  if (rc) return rc;
  double normb = 0.0, normr = 0.0;
  int conv = 0;
  if ((rc = xg_cg_setup(c, &normb))) return rc;
  if ((rc = xg_cg_solve(c, m->m_maxit, m->m_tol, &normr, &m->m_its, &conv))) return rc;
  double* x = nullptr;
  if ((rc = xg_cg_device_x(c, &x))) return rc;
  hipLaunchKernelGGL(k_add_into, dim3(nblocks(n3)), dim3(kBlock), 0, c->stream, n3, Un, x, U);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int check_step(xg_ctx* c) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (!c->cho->have_step) return fail(c, XG_ENOMESH, "xg_chocg_setup not called");
  return XG_OK;
}

void zero(xg_ctx* c, double* p, int64_t n) {
  hipLaunchKernelGGL(k_fill0, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, p);
}

// ChoCG::BC (ChoCG.cpp:1341-1355): dirbc, symbc (pos 0), noslip
int cho_bc(xg_ctx* c, double* U) {
  ChoMesh* m = c->cho;
  if (m->ndir)
    hipLaunchKernelGGL(k_cho_dirbc, dim3(nblocks(m->ndir)), dim3(kBlock), 0, c->stream,
                       m->ndir, m->nc, m->dirnode, m->dirmask, m->dirval, U);
  if (m->nsymnode)
    hipLaunchKernelGGL(k_symbc, dim3(nblocks(m->nsymnode)), dim3(kBlock), 0, c->stream,
                       m->nsymnode, m->symnode, m->symoff, m->symnorm, U, m->nc, 0);
  if (m->nnoslip)
    hipLaunchKernelGGL(k_noslip, dim3(nblocks(m->nnoslip)), dim3(kBlock), 0, c->stream,
                       m->nnoslip, m->nc, m->noslip, U);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

// chorin::grad of a scalar, finalised (zeroed output, divided by vol)
int cho_grad(xg_ctx* c, const double* s, double* G) {
  ChoMesh* m = c->cho;
  zero(c, G, (int64_t)m->npoin * 3);
  int rc = xg_chorin_grad(c, 1, s, G, 3);
  if (rc) return rc;
  hipLaunchKernelGGL(k_divvol_n, dim3(nblocks((int64_t)m->npoin * 3)), dim3(kBlock), 0,
                     c->stream, m->npoin, 3, G, m->vol);
  return XG_OK;
}

int cho_vgrad(xg_ctx* c, const double* U) {
  ChoMesh* m = c->cho;
  const int ng = 3 * m->nc;
  zero(c, m->vgrad, (int64_t)m->npoin * ng);
  int rc = xg_chorin_vgrad(c, 1, m->nc, U, m->vgrad);
  if (rc) return rc;
  hipLaunchKernelGGL(k_divvol_n, dim3(nblocks((int64_t)m->npoin * ng)), dim3(kBlock), 0,
                     c->stream, m->npoin, ng, m->vgrad, m->vol);
  return XG_OK;
}

// ChoCG::pinit + psolve (ChoCG.cpp:1026-1143) on the device rhs m->div
int cho_psolve(xg_ctx* c, const std::vector<double>& pvals, int64_t* it) {
  ChoMesh* m = c->cho;
  const int64_t nd = (int64_t)m->pdir.size();
  // p_pc 2: Jacobi diagonal for the setup, multigrid inside flexible CG
  const int pc = m->cfg.p_pc == 2 ? 1 : m->cfg.p_pc;
  int rc = xg_cg_init_dev(c, m->div, nd, nd ? m->pdir.data() : nullptr,
                          nd ? m->pmask.data() : nullptr, nd ? pvals.data() : nullptr, 1, pc);
  if (rc) return rc;
  double normb = 0.0, normr = 0.0;
  int conv = 0;
  if (m->cfg.p_pc == 2) {
    int64_t vc = 0;
    return xg_cg_solve_flexible(c, 3, 1, m->cfg.p_maxit, m->cfg.p_tol, &normr, it, &vc, &conv);
  }
  if ((rc = xg_cg_setup(c, &normb))) return rc;
  if ((rc = xg_cg_solve(c, m->cfg.p_maxit, m->cfg.p_tol, &normr, it, &conv))) return rc;
  return XG_OK;
}

}  // namespace

void xg_chocg_free(xg_ctx* c) {
  free_cho(c->cho);
  c->cho = nullptr;
}

extern "C" {

int xg_chocg_upload(xg_ctx* c, int64_t npoin, const double* x, const double* y,
                    const double* z, const int64_t nsup[3], const size_t* s0,
                    const double* i0, const size_t* s1, const double* i1,
                    const size_t* s2, const double* i2, int64_t ntri,
                    const size_t* tri) {
  if (!c) return XG_EINVAL;
  if (npoin <= 0 || npoin >= (1ll << 31) || !x || !y || !z || !nsup)
    return fail(c, XG_EINVAL, "xg_chocg_upload: bad mesh size or coordinates");
  if (nsup[0] < 0 || nsup[1] < 0 || nsup[2] < 0 || ntri < 0 ||
      (nsup[0] && (!s0 || !i0)) || (nsup[1] && (!s1 || !i1)) ||
      (nsup[2] && (!s2 || !i2)) || (ntri && !tri))
    return fail(c, XG_EINVAL, "xg_chocg_upload: missing superedge/face arrays");
  XG_HIP(c, hipSetDevice(c->device));
  static const int lpoed[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
  static const int lpoet[3][2] = {{0,1},{1,2},{2,0}};
  const int64_t ne = nsup[0] * 6 + nsup[1] * 3 + nsup[2];
  if (ne >= (1ll << 31)) return fail(c, XG_EINVAL, "xg_chocg_upload: too many edges");
  std::vector<int> p(ne), q(ne);
  std::vector<double> d(ne * 3), w(ne);
  int64_t k = 0;
  auto push = [&](size_t a, size_t b, const double* D) {
    p[k] = (int)std::min<size_t>(a, (size_t)INT32_MAX);
    q[k] = (int)std::min<size_t>(b, (size_t)INT32_MAX);
    d[k * 3 + 0] = D[0]; d[k * 3 + 1] = D[1]; d[k * 3 + 2] = D[2];
    w[k] = D[4];
    ++k;
  };
  for (int64_t e = 0; e < nsup[0]; ++e)
    for (int j = 0; j < 6; ++j)
      push(s0[e * 4 + lpoed[j][0]], s0[e * 4 + lpoed[j][1]], i0 + (e * 6 + j) * 5);
  for (int64_t e = 0; e < nsup[1]; ++e)
    for (int j = 0; j < 3; ++j)
      push(s1[e * 3 + lpoet[j][0]], s1[e * 3 + lpoet[j][1]], i1 + (e * 3 + j) * 5);
  for (int64_t e = 0; e < nsup[2]; ++e) push(s2[e * 2], s2[e * 2 + 1], i2 + e * 5);
  for (int64_t e = 0; e < ne; ++e)
    if (p[e] < 0 || q[e] < 0 || p[e] >= npoin || q[e] >= npoin || p[e] == q[e])
      return fail(c, XG_EINVAL, "xg_chocg_upload: superedge node id out of range");
  for (int64_t f = 0; f < ntri * 3; ++f)
    if (tri[f] >= (size_t)npoin)
      return fail(c, XG_EINVAL, "xg_chocg_upload: boundary face node out of range");
  std::vector<int> tri32(ntri * 3);
  for (int64_t f = 0; f < ntri * 3; ++f) tri32[f] = (int)tri[f];
  std::vector<uint8_t> nosym(ntri * 3 + 1, 0);
  xg::Layout L;
  int rc = xg::build_layout((int)npoin, p, q, d, ntri, tri32.data(), nosym.data(),
                            xg::kBlockNodes, xg::kBlockCap, xg::kChunkSlots, L, c->err);
  if (rc) return rc;
  std::vector<double> dwl(L.eidx.size());
  for (size_t e = 0; e < L.eidx.size(); ++e) dwl[e] = w[L.eidx[e]];
  // boundary normals n = crossdiv(b - a, c - a, 6) (Vector.hpp:68-77)
  std::vector<double> fnx(L.fidx.size()), fny(L.fidx.size()), fnz(L.fidx.size());
  for (size_t i = 0; i < L.fidx.size(); ++i) {
    const size_t* N = tri + L.fidx[i] * 3;
    const double ax = x[N[1]] - x[N[0]], ay = y[N[1]] - y[N[0]], az = z[N[1]] - z[N[0]];
    const double bx = x[N[2]] - x[N[0]], by = y[N[2]] - y[N[0]], bz = z[N[2]] - z[N[0]];
    fnx[i] = (ay * bz - by * az) / 6.0;
    fny[i] = (az * bx - bz * ax) / 6.0;
    fnz[i] = (ax * by - bx * ay) / 6.0;
  }
  std::vector<int4> b0(L.blk.size()), b1(L.blk.size());
  for (size_t b = 0; b < L.blk.size(); ++b) {
    const auto& B = L.blk[b];
    b0[b] = make_int4(B.own0, B.nown, B.nloc, B.halo_off);
    b1[b] = make_int4(B.slot0, B.chunk0, B.chunk1, 0);
  }
  std::vector<int4> c0(L.chunks.size());
  std::vector<int> c1(L.chunks.size());
  for (size_t i = 0; i < L.chunks.size(); ++i) {
    const auto& C = L.chunks[i];
    c0[i] = make_int4(C.kind, C.first, C.count, C.tend_off);
    c1[i] = C.ent_off;
  }
  auto* m = new ChoMesh;
  m->npoin = (int)npoin;
  m->nblk = (int)L.blk.size();
  m->nslot = L.nslot;
  m->nfx = (int)L.fx_node.size();
  auto& M = m->allocs;
  if ((rc = upload(c, M, &m->x, x, npoin)) || (rc = upload(c, M, &m->y, y, npoin)) ||
      (rc = upload(c, M, &m->z, z, npoin)) || (rc = upload(c, M, &m->elpq, L.elpq)) ||
      (rc = upload(c, M, &m->dx, L.dx)) || (rc = upload(c, M, &m->dy, L.dy)) ||
      (rc = upload(c, M, &m->dz, L.dz)) || (rc = upload(c, M, &m->dw, dwl)) ||
      (rc = upload(c, M, &m->blk0, b0)) || (rc = upload(c, M, &m->blk1, b1)) ||
      (rc = upload(c, M, &m->halo, L.halo)) || (rc = upload(c, M, &m->chk, c0)) ||
      (rc = upload(c, M, &m->chk_ent, c1)) || (rc = upload(c, M, &m->tend, L.tend)) ||
      (rc = upload(c, M, &m->entries, L.entries)) || (rc = upload(c, M, &m->fab, L.fab)) ||
      (rc = upload(c, M, &m->fcs, L.fcs)) || (rc = upload(c, M, &m->fnx, fnx)) ||
      (rc = upload(c, M, &m->fny, fny)) || (rc = upload(c, M, &m->fnz, fnz)) ||
      (rc = dev_alloc(c, M, (void**)&m->slots,
                      (size_t)L.nslot * kSlotVals * sizeof(double))) ||
      (rc = upload(c, M, &m->fx_node, L.fx_node)) ||
      (rc = upload(c, M, &m->fx_off, L.fx_off)) ||
      (rc = upload(c, M, &m->fx_slot, L.fx_slot))) {
    free_cho(m);
    return rc;
  }
  free_cho(c->cho);
  c->cho = m;
  return XG_OK;
}

int xg_chorin_div(xg_ctx* c, int dev, double dt, const double* P, const double* G,
                  int ldg, const double* U, int ldu, double* D, int stab) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (!U || !D || ldu < 3 || (stab && (!P || !G || ldg < 3)))
    return fail(c, XG_EINVAL, "xg_chorin_div: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const size_t n = (size_t)m->npoin;
  ChoArgs a = base_args(m);
  a.ldu = ldu; a.ldg = ldg; a.dt = dt;
  if ((rc = dev_in(c, dev, 0, U, n * ldu, &a.U)) ||
      (stab && (rc = dev_in(c, dev, 1, P, n, &a.P))) ||
      (stab && (rc = dev_in(c, dev, 2, G, n * ldg, &a.G))))
    return rc;
  double* out = D;
  if (!dev) {
    double* d;
    if ((rc = stage_buf(c, 3, n * 8, &d))) return rc;
    XG_HIP(c, hipMemcpyAsync(d, D, n * 8, hipMemcpyHostToDevice, c->stream));
    out = d;
  }
  rc = stab ? run_op<OP_DIVS>(c, a, out, 1, 1, nullptr) : run_op<OP_DIV>(c, a, out, 1, 1, nullptr);
  if (rc) return rc;
  if (!dev) {
    XG_HIP(c, hipMemcpyAsync(D, out, n * 8, hipMemcpyDeviceToHost, c->stream));
    XG_HIP(c, hipStreamSynchronize(c->stream));
  }
  return XG_OK;
}

int xg_chorin_grad(xg_ctx* c, int dev, const double* U, double* G, int ldg) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (!U || !G || ldg < 3) return fail(c, XG_EINVAL, "xg_chorin_grad: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const size_t n = (size_t)m->npoin;
  ChoArgs a = base_args(m);
  a.ldu = 1;
  if ((rc = dev_in(c, dev, 0, U, n, &a.U))) return rc;
  double* out = G;
  if (!dev) {
    if ((rc = stage_buf(c, 3, n * ldg * 8, &out))) return rc;
    XG_HIP(c, hipMemcpyAsync(out, G, n * ldg * 8, hipMemcpyHostToDevice, c->stream));
  }
  double* tmp = nullptr;
  if (ldg != 3 && (rc = stage_buf(c, 4, n * 3 * 8, &tmp))) return rc;
  if ((rc = run_op<OP_GRAD>(c, a, out, ldg, 1, tmp))) return rc;
  if (!dev) {
    XG_HIP(c, hipMemcpyAsync(G, out, n * ldg * 8, hipMemcpyDeviceToHost, c->stream));
    XG_HIP(c, hipStreamSynchronize(c->stream));
  }
  return XG_OK;
}

int xg_chorin_vgrad(xg_ctx* c, int dev, int ncomp, const double* U, double* G) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (ncomp != 3 && ncomp != 4)
    return fail(c, XG_EUNSUPPORTED, "chorin::vgrad on the device: ncomp must be 3 or 4");
  if (!U || !G) return fail(c, XG_EINVAL, "xg_chorin_vgrad: null field");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const size_t n = (size_t)m->npoin, nf = 3 * (size_t)ncomp;
  ChoArgs a = base_args(m);
  a.ldu = ncomp;
  if ((rc = dev_in(c, dev, 0, U, n * ncomp, &a.U))) return rc;
  double* out = G;
  if (!dev) {
    if ((rc = stage_buf(c, 3, n * nf * 8, &out))) return rc;
    XG_HIP(c, hipMemcpyAsync(out, G, n * nf * 8, hipMemcpyHostToDevice, c->stream));
  }
  rc = ncomp == 3 ? run_op<OP_VGRAD, 3>(c, a, out, 9, 1, nullptr)
                  : run_op<OP_VGRAD, 4>(c, a, out, 12, 1, nullptr);
  if (rc) return rc;
  if (!dev) {
    XG_HIP(c, hipMemcpyAsync(G, out, n * nf * 8, hipMemcpyDeviceToHost, c->stream));
    XG_HIP(c, hipStreamSynchronize(c->stream));
  }
  return XG_OK;
}

int xg_chorin_flux(xg_ctx* c, int dev, double mu, const double* U, int ldu,
                   const double* G, int ldg, double* F) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (!U || !G || !F || ldu < 3 || ldg < 9)
    return fail(c, XG_EINVAL, "xg_chorin_flux: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const size_t n = (size_t)m->npoin;
  ChoArgs a = base_args(m);
  a.ldu = ldu; a.ldg = ldg; a.mu = mu;
  if ((rc = dev_in(c, dev, 0, U, n * ldu, &a.U)) || (rc = dev_in(c, dev, 2, G, n * ldg, &a.G)))
    return rc;
  double* out = F;
  if (!dev) {
    if ((rc = stage_buf(c, 3, n * 3 * 8, &out))) return rc;
    XG_HIP(c, hipMemcpyAsync(out, F, n * 3 * 8, hipMemcpyHostToDevice, c->stream));
  }
  if ((rc = run_op<OP_FLUX>(c, a, out, 3, 1, nullptr))) return rc;
  if (!dev) {
    XG_HIP(c, hipMemcpyAsync(F, out, n * 3 * 8, hipMemcpyDeviceToHost, c->stream));
    XG_HIP(c, hipStreamSynchronize(c->stream));
  }
  return XG_OK;
}

int xg_chorin_rhs(xg_ctx* c, int dev, const xg_chocfg* cfg, int ncomp,
                  const double* P, const double* U, const double* G, double* R) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (!cfg) return fail(c, XG_EINVAL, "xg_chorin_rhs: null config");
  if (cfg->flux != 0 && cfg->flux != 1) return fail(c, XG_EINVAL, "Flux not correctly configured");
  if (ncomp != 3 && ncomp != 4)
    return fail(c, XG_EUNSUPPORTED, "chorin::rhs on the device: ncomp must be 3 or 4");
  if (!P || !U || !R || (cfg->flux == 1 && !G))
    return fail(c, XG_EINVAL, "xg_chorin_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const size_t n = (size_t)m->npoin, nc = (size_t)ncomp;
  ChoArgs a = base_args(m);
  a.ldu = ncomp; a.ldg = 3 * ncomp; a.mu = cfg->mu; a.diff = cfg->diff;
  a.stab = cfg->stab; a.stab2 = cfg->stab2; a.stab2coef = cfg->stab2coef;
  if ((rc = dev_in(c, dev, 0, U, n * nc, &a.U)) || (rc = dev_in(c, dev, 1, P, n, &a.P)) ||
      (cfg->flux == 1 && (rc = dev_in(c, dev, 2, G, n * 3 * nc, &a.G))))
    return rc;
  double* out = R;
  if (!dev && (rc = stage_buf(c, 3, n * nc * 8, &out))) return rc;
  if (ncomp == 3)
    rc = cfg->flux == 0 ? run_op<OP_ADV2, 3>(c, a, out, 3, 0, nullptr)
                        : run_op<OP_ADV4, 3>(c, a, out, 3, 0, nullptr);
  else
    rc = cfg->flux == 0 ? run_op<OP_ADV2, 4>(c, a, out, 4, 0, nullptr)
                        : run_op<OP_ADV4, 4>(c, a, out, 4, 0, nullptr);
  if (rc) return rc;
  if (!dev) {
    XG_HIP(c, hipMemcpyAsync(R, out, n * nc * 8, hipMemcpyDeviceToHost, c->stream));
    XG_HIP(c, hipStreamSynchronize(c->stream));
  }
  return XG_OK;
}

int xg_chocg_setup(xg_ctx* c, const double* vol, const xg_chostep* cfg,
                   int64_t nsym, const size_t* symnodes, const double* symnorms,
                   int64_t nnoslip, const size_t* noslipnodes, int64_t ndir,
                   const size_t* dirnodes, const int* dirmask, const double* dirvals,
                   int64_t npdir, const size_t* pdirnodes, const double* pdirvals) {
  int rc = check_cho(c);
  if (rc) return rc;
  ChoMesh* m = c->cho;
  if (!vol || !cfg || nsym < 0 || (nsym && (!symnodes || !symnorms)) || nnoslip < 0 ||
      (nnoslip && !noslipnodes) || ndir < 0 || (ndir && (!dirnodes || !dirmask || !dirvals)) ||
      npdir < 0 || (npdir && (!pdirnodes || !pdirvals)))
    return fail(c, XG_EINVAL, "xg_chocg_setup: bad arguments");
  if (cfg->rk < 1 || cfg->rk > 4) return fail(c, XG_EINVAL, "xg_chocg_setup: rk must be 1..4");
  if (cfg->flux.flux != 0 && cfg->flux.flux != 1)
    return fail(c, XG_EINVAL, "Flux not correctly configured");
  if (cfg->p_pc < 0 || cfg->p_pc > 2) return fail(c, XG_EINVAL, "p_pc must be 0, 1 or 2");
  if (cfg->p_maxit < 1 || !(cfg->p_tol >= 0.0) || !(cfg->cfl > 0.0))
    return fail(c, XG_EINVAL, "xg_chocg_setup: bad solver settings");
  if (!c->cg || xg_cg_rows(c) != m->npoin)
    return fail(c, XG_ENOMESH, "xg_chocg_setup: the pressure matrix (xg_cg_matrix, "
                               "one component per node) is not set");
  const int64_t n = m->npoin;
  for (int64_t i = 0; i < n; ++i)
    if (!(vol[i] > 0)) return fail(c, XG_EINVAL, "xg_chocg_setup: non-positive nodal volume");
  auto bad = [n](const size_t* v, int64_t k) {
    for (int64_t i = 0; i < k; ++i)
      if (v[i] >= (size_t)n) return true;
    return false;
  };
  if (bad(symnodes, nsym) || bad(noslipnodes, nnoslip) || bad(dirnodes, ndir) ||
      bad(pdirnodes, npdir))
    return fail(c, XG_EINVAL, "xg_chocg_setup: node id out of range");
  // symmetry (node, normal) pairs grouped by node, list order kept per node
  std::vector<int> node, off(1, 0);
  std::vector<double> nrm;
  {
    std::vector<int64_t> order(nsym);
    for (int64_t i = 0; i < nsym; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      return symnodes[a] < symnodes[b];
    });
    for (int64_t q = 0; q < nsym; ++q) {
      int64_t i = order[q];
      if (q == 0 || symnodes[order[q - 1]] != symnodes[i]) {
        if (q) off.push_back((int)(nrm.size() / 3));
        node.push_back((int)symnodes[i]);
      }
      nrm.insert(nrm.end(), symnorms + i * 3, symnorms + i * 3 + 3);
    }
    if (nsym) off.push_back((int)(nrm.size() / 3));
  }
  const int nc = m->nc;
  std::vector<int> ns(noslipnodes, noslipnodes + nnoslip), dn(ndir);
  std::vector<uint8_t> dm((size_t)ndir * nc);
  for (int64_t i = 0; i < ndir; ++i) {
    dn[i] = (int)dirnodes[i];
    for (int q = 0; q < nc; ++q) dm[i * nc + q] = dirmask[i * nc + q] ? 1 : 0;
  }
  XG_HIP(c, hipSetDevice(c->device));
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(m->step_allocs);
  m->have_step = false;
  auto& L = m->step_allocs;
  const size_t nn = (size_t)n;
  if ((rc = upload(c, L, &m->vol, vol, nn)) || (rc = upload(c, L, &m->symnode, node)) ||
      (rc = upload(c, L, &m->symoff, off)) || (rc = upload(c, L, &m->symnorm, nrm)) ||
      (rc = upload(c, L, &m->noslip, ns)) || (rc = upload(c, L, &m->dirnode, dn)) ||
      (rc = upload(c, L, &m->dirmask, dm)) ||
      (rc = upload(c, L, &m->dirval, dirvals, (size_t)ndir * nc)) ||
      (rc = dev_alloc(c, L, (void**)&m->pgrad, nn * 3 * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->vgrad, nn * 3 * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->sgrad, nn * 3 * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->R, nn * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->usave, nn * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->F, nn * 3 * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->div, nn * 8)) ||
      (rc = dev_alloc(c, L, (void**)&m->red, sizeof(double) * (kRedBlocks + 1)))) {
    dev_free_all(m->step_allocs);
    return rc;
  }
  zero(c, m->pgrad, (int64_t)nn * 3);
  zero(c, m->vgrad, (int64_t)nn * 3 * nc);
  m->nsymnode = (int)node.size();
  m->nnoslip = (int)nnoslip;
  m->ndir = (int)ndir;
  m->h_dirnode = dn;
  m->h_dirmask = dm;
  m->h_noslip = ns;
  m->pdir.assign(pdirnodes, pdirnodes + npdir);
  m->pmask.assign(npdir, 1);
  m->pval0.assign(pdirvals, pdirvals + npdir);
  m->pzero.assign(npdir, 0.0);
  m->cfg = *cfg;
  XG_HIP(c, hipGetLastError());
  m->have_step = true;
  return XG_OK;
}

int xg_chocg_initial_dev(xg_ctx* c, double* U, double* P, int64_t* pit) {
  int rc = check_step(c);
  if (rc) return rc;
  if (!U || !P) return fail(c, XG_EINVAL, "xg_chocg_initial: null field");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const int n = m->npoin;
  int64_t it0 = 0, it1 = 0;
  double* x = nullptr;
  if ((rc = xg_cg_device_x(c, &x))) return rc;
  // m_np 0: project the initial velocity (dt = 1)
  if ((rc = cho_bc(c, U))) return rc;
  zero(c, m->div, n);
  if ((rc = xg_chorin_div(c, 1, 0.0, P, m->pgrad, 3, U, m->nc, m->div, 0))) return rc;
  if ((rc = cho_psolve(c, m->pval0, &it0))) return rc;
  if ((rc = cho_grad(c, x, m->sgrad))) return rc;
  hipLaunchKernelGGL(k_project, dim3(nblocks((int64_t)n * 3)), dim3(kBlock), 0, c->stream,
                     n, m->nc, 1.0, U, m->sgrad);
  if ((rc = cho_bc(c, U))) return rc;
  // m_np 1: velocity gradients, momentum flux, its divergence -> pressure
  if ((rc = cho_vgrad(c, U))) return rc;
  zero(c, m->F, (int64_t)n * 3);
  if ((rc = xg_chorin_flux(c, 1, m->cfg.flux.mu, U, m->nc, m->vgrad, 3 * m->nc, m->F)))
    return rc;
  hipLaunchKernelGGL(k_divvol_n, dim3(nblocks((int64_t)n * 3)), dim3(kBlock), 0, c->stream,
                     n, 3, m->F, m->vol);
  if (m->nsymnode)
    hipLaunchKernelGGL(k_symbc, dim3(nblocks(m->nsymnode)), dim3(kBlock), 0, c->stream,
                       m->nsymnode, m->symnode, m->symoff, m->symnorm, m->F, 3, 0);
  zero(c, m->div, n);
  if ((rc = xg_chorin_div(c, 1, 0.0, P, m->pgrad, 3, m->F, 3, m->div, 0))) return rc;
  if ((rc = cho_psolve(c, m->pval0, &it1))) return rc;
  XG_HIP(c, hipMemcpyAsync(P, x, (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
  if ((rc = cho_grad(c, P, m->pgrad))) return rc;
  XG_HIP(c, hipGetLastError());
  if (pit) { pit[0] = it0; pit[1] = it1; }
  return XG_OK;
}

int xg_chocg_poisson_dev(xg_ctx* c, const double* pr, const double* neubc, double* U,
                         double* P, int64_t* pit) {
  int rc = check_step(c);
  if (rc) return rc;
  if (!U || !P) return fail(c, XG_EINVAL, "xg_chocg_poisson: null field");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  const int n = m->npoin;
  int64_t it = 0;
  double* x = nullptr;
  if ((rc = xg_cg_device_x(c, &x))) return rc;
  // m_np 0: div(u), replaced by the problem's rhs where it has one
  if ((rc = cho_bc(c, U))) return rc;
  zero(c, m->div, n);
  if ((rc = xg_chorin_div(c, 1, 0.0, P, m->pgrad, 3, U, m->nc, m->div, 0))) return rc;
  if (pr || neubc)
    hipLaunchKernelGGL(k_poisson_rhs, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, pr,
                       m->vol, neubc, m->div);
  if ((rc = cho_psolve(c, m->pval0, &it))) return rc;
  // psolved with m_np 0: project u with dt = 1, BCs; nstep = 1: p = solution
  if ((rc = cho_grad(c, x, m->sgrad))) return rc;
  hipLaunchKernelGGL(k_project, dim3(nblocks((int64_t)n * 3)), dim3(kBlock), 0, c->stream,
                     n, m->nc, 1.0, U, m->sgrad);
  if ((rc = cho_bc(c, U))) return rc;
  XG_HIP(c, hipMemcpyAsync(P, x, (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (pit) *pit = it;
  return XG_OK;
}

int xg_chocg_dt_dev(xg_ctx* c, const double* U, double* dt) {
  int rc = check_step(c);
  if (rc) return rc;
  if (!U || !dt) return fail(c, XG_EINVAL, "xg_chocg_dt: null argument");
  XG_HIP(c, hipSetDevice(c->device));
  ChoMesh* m = c->cho;
  int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(m->npoin)));
  hipLaunchKernelGGL(k_chodt_partial, dim3(nb), dim3(kBlock), 0, c->stream, m->npoin, m->nc,
                     U, m->vol, std::max(m->cfg.flux.mu, m->cfg.flux.diff), m->red);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kRedBlocks), 0, c->stream, nb, m->red,
                     m->cfg.cfl, m->red + kRedBlocks);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipMemcpyAsync(dt, m->red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (*dt < 2.220446049250313e-16)
    return fail(c, XG_EBLOWUP, "time step below machine epsilon");
  return XG_OK;
}

int xg_chocg_step2_dev(xg_ctx* c, double dt, int freeze, double* U, double* Un, double* P,
                       double* dP, int64_t* pit) {
  static const double rkcoef[4][4] = {{1.0}, {1.0 / 2.0, 1.0},
                                      {1.0 / 3.0, 1.0 / 2.0, 1.0},
                                      {1.0 / 4.0, 1.0 / 3.0, 1.0 / 2.0, 1.0}};
  int rc = check_step(c);
  if (rc) return rc;
  if (!U || !Un || !P || !dP || U == Un) return fail(c, XG_EINVAL, "xg_chocg_step: bad fields");
  if (!(dt > 0.0)) return fail(c, XG_EINVAL, "xg_chocg_step: dt must be positive");
  ChoMesh* m = c->cho;
  const int kind = c->problem.kind;
  if (kind != XG_PROBLEM_NONE && kind != XG_PROBLEM_SLOT_CYL)
    return fail(c, XG_EUNSUPPORTED, "xg_chocg_step: no ChoCG source for this problem");
  if (kind == XG_PROBLEM_SLOT_CYL && m->nc != 4)
    return fail(c, XG_EINVAL, "xg_chocg_step: slot_cyl has 4 components (xg_chocg_set_ncomp)");
  XG_HIP(c, hipSetDevice(c->device));
  const int n = m->npoin, nc = m->nc;
  const int64_t n3 = (int64_t)n * 3, nn = (int64_t)n * nc;
  const int nst = m->cfg.rk;
  const bool implicit = m->theta > 2.220446049250313e-16;
  bool divdone = false;
  for (int st = 0; st < nst; ++st) {
    const bool last = st + 1 == nst;
    // ChoCG::rhs (chorin::rhs with the problem's source), solve (explicit,
    // or semi-implicit at the last stage when theta > 0), pred (BC,
    // velocity gradients for damp4)
    if ((rc = xg_chorin_rhs(c, 1, &m->cfg.flux, nc, P, U, m->vgrad, m->R))) return rc;
    if (kind == XG_PROBLEM_SLOT_CYL)
      hipLaunchKernelGGL(k_cho_src_slot_cyl, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n,
                         nc, m->x, m->y, m->vol, m->R);
    if (st == 0)
      XG_HIP(c, hipMemcpyAsync(Un, U, (size_t)nn * 8, hipMemcpyDeviceToDevice, c->stream));
    const bool restore = freeze && !(implicit && last);
    if (restore)
      hipLaunchKernelGGL(k_copy_vel, dim3(nblocks(n3)), dim3(kBlock), 0, c->stream, n, nc, U,
                         m->usave);
    if (implicit && last) {
      if ((rc = cho_momentum(c, dt, U, Un))) return rc;
    } else {
      hipLaunchKernelGGL(k_cho_update, dim3(nblocks(nn)), dim3(kBlock), 0, c->stream, n, nc,
                         rkcoef[nst - 1][st] * dt, U, Un, m->R, m->vol);
    }
    if ((rc = cho_bc(c, U))) return rc;
    if (m->cfg.flux.flux == 1) {
      if ((rc = cho_vgrad(c, U))) return rc;
    } else if (restore && last) {
      // damp2: corr and its div run inside pred, before the flow is put back
      zero(c, m->div, n);
      if ((rc = xg_chorin_div(c, 1, dt, P, m->pgrad, 3, U, nc, m->div, 1))) return rc;
      divdone = true;
    }
    if (restore)
      hipLaunchKernelGGL(k_copy_vel, dim3(nblocks(n3)), dim3(kBlock), 0, c->stream, n, nc,
                         m->usave, U);
  }
  // ChoCG::div, pinit, psolve, sgrad, psolved, pgrad
  if (!divdone) {
    zero(c, m->div, n);
    if ((rc = xg_chorin_div(c, 1, dt, P, m->pgrad, 3, U, nc, m->div, 1))) return rc;
  }
  hipLaunchKernelGGL(k_divide, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, m->div, dt);
  int64_t it = 0;
  if ((rc = cho_psolve(c, m->pzero, &it))) return rc;
  double* x = nullptr;
  if ((rc = xg_cg_device_x(c, &x))) return rc;
  XG_HIP(c, hipMemcpyAsync(dP, x, (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
  if ((rc = cho_grad(c, dP, m->sgrad))) return rc;
  hipLaunchKernelGGL(k_project, dim3(nblocks(n3)), dim3(kBlock), 0, c->stream, n, nc, dt, U,
                     m->sgrad);
  if ((rc = cho_bc(c, U))) return rc;
  hipLaunchKernelGGL(k_add_to, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, P, dP);
  if ((rc = cho_grad(c, P, m->pgrad))) return rc;
  XG_HIP(c, hipGetLastError());
  if (pit) *pit = it;
  return XG_OK;
}

int xg_chocg_step_dev(xg_ctx* c, double dt, double* U, double* Un, double* P, double* dP,
                      int64_t* pit) {
  return xg_chocg_step2_dev(c, dt, 0, U, Un, P, dP, pit);
}

int xg_chocg_set_ncomp(xg_ctx* c, int ncomp) {
  int rc = check_cho(c);
  if (rc) return rc;
  if (ncomp != 3 && ncomp != 4)
    return fail(c, XG_EUNSUPPORTED, "xg_chocg_set_ncomp: ncomp must be 3 or 4");
  ChoMesh* m = c->cho;
  if (m->nc != ncomp) {
    XG_HIP(c, hipSetDevice(c->device));
    (void)hipStreamSynchronize(c->stream);
    dev_free_all(m->step_allocs);
    m->have_step = false;     // xg_chocg_setup sizes its buffers by ncomp
    m->theta = 0.0;           // and the momentum system is 3-component only
    m->nc = ncomp;
  }
  return XG_OK;
}

int xg_chocg_momentum_setup(xg_ctx* c, int64_t ntet, const size_t* inpoel, double theta,
                            int m_pc, int64_t m_maxit, double m_tol) {
  int rc = check_step(c);
  if (rc) return rc;
  ChoMesh* m = c->cho;
  if (ntet <= 0 || !inpoel || !(theta >= 0.0) || (m_pc != 0 && m_pc != 1) || m_maxit < 1 ||
      !(m_tol >= 0.0))
    return fail(c, XG_EINVAL, "xg_chocg_momentum_setup: bad arguments");
  if (m->nc != 3)
    return fail(c, XG_EUNSUPPORTED, "xg_chocg_momentum_setup: the momentum system has 3 "
                                    "components (no transported scalars)");
  const int64_t n = m->npoin;
  for (int64_t i = 0; i < 4 * ntet; ++i)
    if (inpoel[i] >= (size_t)n) return fail(c, XG_EINVAL, "xg_chocg_momentum_setup: node id out of range");
  // ChoCG::momlhs: the 3-component pattern
  std::vector<int64_t> ia;
  std::vector<int32_t> ja;
  if ((rc = xg::csr_pattern(n, ntet, inpoel, 3, ia, ja, c->err))) return rc;
  const int64_t nnz = ia[3 * n];
  auto find = [&](int64_t row, int64_t col) -> int64_t {
    const int32_t* b0 = ja.data() + ia[row];
    const int32_t* b1 = ja.data() + ia[row + 1];
    const int32_t* f = std::lower_bound(b0, b1, (int32_t)col);
    return (f != b1 && *f == col) ? (int64_t)(f - ja.data()) : -1;
  };
  // per-entry sums of the mass and stiffness terms of ChoCG::lhs, tet order
  std::vector<double> M(nnz, 0.0), K(nnz, 0.0);
  std::vector<double> hx(n), hy(n), hz(n);
  XG_HIP(c, hipMemcpy(hx.data(), m->x, n * 8, hipMemcpyDeviceToHost));
  XG_HIP(c, hipMemcpy(hy.data(), m->y, n * 8, hipMemcpyDeviceToHost));
  XG_HIP(c, hipMemcpy(hz.data(), m->z, n * 8, hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < ntet; ++e) {
    const size_t* N = inpoel + e * 4;
    const double ba[3] = {hx[N[1]] - hx[N[0]], hy[N[1]] - hy[N[0]], hz[N[1]] - hz[N[0]]};
    const double ca[3] = {hx[N[2]] - hx[N[0]], hy[N[2]] - hy[N[0]], hz[N[2]] - hz[N[0]]};
    const double da[3] = {hx[N[3]] - hx[N[0]], hy[N[3]] - hy[N[0]], hz[N[3]] - hz[N[0]]};
    double g[4][3];
    g[1][0] = ca[1] * da[2] - da[1] * ca[2];
    g[1][1] = ca[2] * da[0] - da[2] * ca[0];
    g[1][2] = ca[0] * da[1] - da[0] * ca[1];
    g[2][0] = da[1] * ba[2] - ba[1] * da[2];
    g[2][1] = da[2] * ba[0] - ba[2] * da[0];
    g[2][2] = da[0] * ba[1] - ba[0] * da[1];
    g[3][0] = ba[1] * ca[2] - ca[1] * ba[2];
    g[3][1] = ba[2] * ca[0] - ca[2] * ba[0];
    g[3][2] = ba[0] * ca[1] - ca[0] * ba[1];
    for (int i = 0; i < 3; ++i) g[0][i] = -g[1][i] - g[2][i] - g[3][i];
    const double J = ba[0] * g[1][0] + ba[1] * g[1][1] + ba[2] * g[1][2];
    if (!(J > 0.0)) return fail(c, XG_EINVAL, "xg_chocg_momentum_setup: non-positive tet Jacobian");
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        const double vm = J / 120.0 * (a == b ? 2.0 : 1.0);
        const double vk = (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]) / J / 6.0;
        for (int cc = 0; cc < 3; ++cc) {
          const int64_t p = find((int64_t)N[a] * 3 + cc, (int64_t)N[b] * 3 + cc);
          if (p < 0) return fail(c, XG_EINVAL, "xg_chocg_momentum_setup: entry not in pattern");
          M[p] += vm;
          K[p] += vk;
        }
      }
  }
  // the momentum solve's Dirichlet rows (ChoCG.cpp:1577-1600): masked
  // components of the dirbc nodes and all components of noslip nodes, each
  // set to du = 0 as CSR::dirichlet does: row and column zeroed, diagonal 1
  std::vector<char> isdir(3 * n, 0);
  for (size_t k = 0; k < m->h_dirnode.size(); ++k)
    for (int cc = 0; cc < 3; ++cc)
      if (m->h_dirmask[k * 3 + cc]) isdir[(size_t)m->h_dirnode[k] * 3 + cc] = 1;
  for (int p : m->h_noslip)
    for (int cc = 0; cc < 3; ++cc) isdir[(size_t)p * 3 + cc] = 1;
  std::vector<int> zpos, dpos, drow;
  for (int64_t r = 0; r < 3 * n; ++r) {
    if (!isdir[r]) continue;
    drow.push_back((int)r);
    for (int64_t k = ia[r]; k < ia[r + 1]; ++k) {
      zpos.push_back((int)k);
      if (ja[k] == r) dpos.push_back((int)k);
      const int64_t q = find(ja[k], r);      // the column entry in row ja[k]
      if (q >= 0) zpos.push_back((int)q);
    }
  }
  XG_HIP(c, hipSetDevice(c->device));
  (void)hipStreamSynchronize(c->stream);
  if (m->cgm) { xg_cg_free_solver(m->cgm); m->cgm = nullptr; }
  dev_free_all(m->mom_allocs);
  {
    std::vector<size_t> iz(ia.begin(), ia.end()), jz(ja.begin(), ja.end());
    std::vector<double> zeros(nnz, 0.0);
    CgSolver* pressure = xg_cg_exchange(c, nullptr);
    rc = xg_cg_matrix(c, 3, n, iz.data(), jz.data(), zeros.data(), 0);
    m->cgm = xg_cg_exchange(c, pressure);
    if (rc) { m->cgm = nullptr; return rc; }
  }
  auto& L = m->mom_allocs;
  if ((rc = upload(c, L, &m->mM, M)) || (rc = upload(c, L, &m->mK, K)) ||
      (rc = upload(c, L, &m->zpos, zpos)) || (rc = upload(c, L, &m->dpos, dpos)) ||
      (rc = upload(c, L, &m->drow, drow)))
    return rc;
  m->m_nnz = nnz;
  m->nzp = (int)zpos.size();
  m->ndp = (int)dpos.size();
  m->ndr = (int)drow.size();
  m->theta = theta;
  m->m_pc = m_pc;
  m->m_maxit = m_maxit;
  m->m_tol = m_tol;
  return XG_OK;
}

int xg_chocg_momentum_its(xg_ctx* c, int64_t* its) {
  int rc = check_step(c);
  if (rc) return rc;
  if (!its) return fail(c, XG_EINVAL, "xg_chocg_momentum_its: null");
  *its = c->cho->m_its;
  return XG_OK;
}

}  // extern "C"
