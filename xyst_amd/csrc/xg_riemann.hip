// MI355X (gfx950) device path of the RieCG edge-based Riemann/MUSCL scheme:
// context, device mesh layout, riemann::grad/rhs kernels, RK stage, CFL
// minimum.  All FP64.  Deterministic: no atomics, every nodal sum is taken in
// an order fixed by the layout (xg_layout.cpp).
//
//   rhs:  k_rhs_blocks -- one 512-thread workgroup per node block (two per
//         CU); the block's node data staged in LDS, one node sum per thread;
//         k_fixup<5> adds halo partials.
//   grad: k_grad_blocks -- a gather layout of its own (edges between two
//         blocks evaluated in both, so no halo partials and no fix-up), also
//         512 threads and two workgroups per CU, each node thread gathering
//         its edges' terms from LDS-staged primitive states; k_grad_bnd adds
//         boundary-face terms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "xg_flux.hpp"
#include "xg_layout.hpp"
#include "xg_problems.hpp"

using xg::Cfg;

#include "xg_blocks.hpp"
#include "xg_amg.hpp"
#include "xg_ctx.hpp"
#include "xg_gnn.hpp"

namespace {

using namespace xgd;
constexpr int kGradThreads = xg::kGradSlots;
constexpr int kRhsThreads = xg::kRhsSlots;

// ----------------------------------------------------------------- kernels --

// Per-chunk bookkeeping shared by the block kernels: the chunk's entry list
// (slot | sign bit, grouped by local node) and each local node's entry end
// are prefetched into registers one chunk ahead with coalesced loads, then
// placed in LDS, so the per-node accumulation loop reads LDS only.
// (NT threads, chunks of at most NT edges: up to 2 NT entries)
template <int NT>
struct ChunkPrefetch {
  uint16_t e0 = 0, e1 = 0, te = 0;
  __device__ inline void load(const int4& C, int ent_off, int nloc, int t,
                              const uint16_t* __restrict__ tend,
                              const uint16_t* __restrict__ entries) {
    int nent = C.z * (C.x == 0 ? 2 : 3);
    if (t < nent) e0 = entries[ent_off + t];
    if (t + NT < nent) e1 = entries[ent_off + NT + t];
    if (t < nloc) te = tend[C.w + t];
  }
  __device__ inline void store(const int4& C, int nloc, int t, uint16_t* sE,
                               uint16_t* sT) const {
    int nent = C.z * (C.x == 0 ? 2 : 3);
    if (t < nent) sE[t] = e0;
    if (t + NT < nent) sE[NT + t] = e1;
    if (t < nloc) sT[t] = te;
  }
};

// riemann::grad domain term (Riemann.cpp:263-328) on the grad (gather)
// layout: a block holds every edge touching one of its nodes, so each owned
// node's sum is complete here.  The block's primitive states are staged in
// LDS.  Per chunk: each thread writes one edge's (u_p + u_q, D) to LDS; then
// the thread owning local node t adds -/+ D (u_p + u_q) for the chunk's edges
// touching it, in slot order.
// DIV: owned nodes without boundary terms (gflag 0) are final here and are
// divided by their volume (RieCG.cpp:935-939).  NC components per node
// (5 + transported scalars).
template <bool DIV, int NC>
__global__ __launch_bounds__(kGradThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_grad_blocks(
    int nblk, const int4* __restrict__ blk0, const int4* __restrict__ blk1,
    const int* __restrict__ halo, const int4* __restrict__ chk,
    const int* __restrict__ chk_ent, const uint16_t* __restrict__ tend,
    const uint16_t* __restrict__ entries, const uint32_t* __restrict__ elpq,
    const double* __restrict__ dx, const double* __restrict__ dy,
    const double* __restrict__ dz, const double* __restrict__ U,
    double* __restrict__ G, const uint8_t* __restrict__ gflag,
    const double* __restrict__ vol) {
  constexpr int CAP = xg::kGradCap;
  constexpr int FW = NC + 3;   // an edge's record: u_p + u_q, then D
  static_assert(CAP <= kGradThreads, "one local node per thread");
  __shared__ double sP[NC * CAP];
  __shared__ double F[kGradThreads * FW];
  __shared__ uint16_t sE[2 * kGradThreads];
  __shared__ uint16_t sT[CAP];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 B0 = blk0[b];   // own0, nown, nloc, halo_off
  const int4 B1 = blk1[b];   // slot0, chunk0, chunk1
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  if (t < nloc) {
    int g = t < nown ? own0 + t : halo[B0.w + t - nown];
    double s[NC], u[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = U[(size_t)g * NC + c];
    xg::primitive_fast(s, u);
#pragma unroll
    for (int c = 5; c < NC; ++c) u[c] = s[c];
#pragma unroll
    for (int c = 0; c < NC; ++c) sP[c * CAP + t] = u[c];
  }
  // prefetch of the next edge chunk (a gather layout has no boundary-face
  // chunks, and any would be skipped: k_grad_bnd adds those terms)
  int ch = B1.y;
  while (ch < B1.z && chk[ch].x != 0) ++ch;
  ChunkPrefetch<kGradThreads> pf;
  uint32_t pq_n = 0;
  double dx_n = 0.0, dy_n = 0.0, dz_n = 0.0;
  auto fetch = [&](int c) {
    if (c < B1.z) {
      const int4 C = chk[c];
      pf.load(C, chk_ent[c], nloc, t, tend, entries);
      if (t < C.z) {
        int e = C.y + t;
        pq_n = elpq[e];
        dx_n = dx[e];
        dy_n = dy[e];
        dz_n = dz[e];
      }
    }
  };
  fetch(ch);
  double acc[3 * NC];
#pragma unroll
  for (int k = 0; k < 3 * NC; ++k) acc[k] = 0.0;
  while (ch < B1.z) {
    const int4 C = chk[ch];
    int nx = ch + 1;
    while (nx < B1.z && chk[nx].x != 0) ++nx;
    const uint32_t pq = pq_n;
    const double d0 = dx_n, d1 = dy_n, d2 = dz_n;
    __syncthreads();          // staging done / previous chunk consumed
    pf.store(C, nloc, t, sE, sT);
    if (t < C.z) {
      int lp = pq & 0xffff, lq = pq >> 16;
      double* f = F + t * FW;
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = sP[c * CAP + lq] + sP[c * CAP + lp];
      f[NC] = d0;
      f[NC + 1] = d1;
      f[NC + 2] = d2;
    }
    fetch(nx);                // in flight during the accumulation below
    __syncthreads();          // F, sE, sT complete
    if (t < nown) {           // halo nodes have no entries here
      int e0 = t ? sT[t - 1] : 0, e1 = sT[t];
      for (int j = e0; j < e1; ++j) {
        int v = sE[j];
        const double* f = F + (v & 0x3ff) * FW;
        double sg = (v & 0x8000) ? -1.0 : 1.0;
        double D0 = sg * f[NC], D1 = sg * f[NC + 1], D2 = sg * f[NC + 2];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          acc[c * 3 + 0] += D0 * f[c];
          acc[c * 3 + 1] += D1 * f[c];
          acc[c * 3 + 2] += D2 * f[c];
        }
      }
    }
    ch = nx;
  }
  if (t < nown) {
    double* o = G + (size_t)(own0 + t) * (3 * NC);
    if (DIV && gflag[own0 + t] == 0) {
      const double v = vol[own0 + t];
#pragma unroll
      for (int k = 0; k < 3 * NC; ++k) o[k] = acc[k] / v;
    } else {
#pragma unroll
      for (int k = 0; k < 3 * NC; ++k) o[k] = acc[k];
    }
  }
}

// gflag: the nodes with boundary-face gradient terms
__global__ void k_gflag(int n, const int* __restrict__ node, uint8_t bit,
                        uint8_t* __restrict__ gflag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) gflag[node[i]] |= bit;
}

__global__ void k_fill_u8(int64_t n, uint8_t* __restrict__ a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// riemann::grad boundary integral (Riemann.cpp:332-359) gathered per node.
// As in the reference, every node of a face receives g[j]*n[j] for direction
// j, with g[j] the 4:1:1-weighted value belonging to the face's j-th node.
// DIV: boundary nodes are final here and are divided by their volume
template <bool DIV, int NC>
__global__ __launch_bounds__(kBlock) void k_grad_bnd(
    int nb, const int* __restrict__ bnode, const int* __restrict__ boff,
    const int* __restrict__ bface, const int* __restrict__ tri,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ U,
    double* __restrict__ G, const double* __restrict__ vol) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int n = bnode[i];
  double acc[3 * NC];
#pragma unroll
  for (int k = 0; k < 3 * NC; ++k) acc[k] = G[(size_t)n * (3 * NC) + k];
  for (int k = boff[i]; k < boff[i + 1]; ++k) {
    int f = bface[k] >> 2;
    int A = tri[f * 3 + 0], B = tri[f * 3 + 1], C = tri[f * 3 + 2];
    double u[3][NC];
    xg::primitive<NC>(U, A, u[0]);
    xg::primitive<NC>(U, B, u[1]);
    xg::primitive<NC>(U, C, u[2]);
    double ba[3] = {x[B] - x[A], y[B] - y[A], z[B] - z[A]};
    double ca[3] = {x[C] - x[A], y[C] - y[A], z[C] - z[A]};
    double nr[3] = {(ba[1] * ca[2] - ca[1] * ba[2]) / 12.0,
                    (ba[2] * ca[0] - ca[2] * ba[0]) / 12.0,
                    (ba[0] * ca[1] - ca[0] * ba[1]) / 12.0};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double uab = (u[0][c] + u[1][c]) / 4.0;
      double ubc = (u[1][c] + u[2][c]) / 4.0;
      double uca = (u[2][c] + u[0][c]) / 4.0;
      double g[3] = {uab + uca + u[0][c], uab + ubc + u[1][c], ubc + uca + u[2][c]};
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[c * 3 + j] += g[j] * nr[j];
    }
  }
  if (DIV) {
    const double v = vol[n];
#pragma unroll
    for (int k = 0; k < 3 * NC; ++k) acc[k] /= v;
  }
#pragma unroll
  for (int k = 0; k < 3 * NC; ++k) G[(size_t)n * (3 * NC) + k] = acc[k];
}

// advbnd (Riemann.cpp:768-878) for one boundary face from the staged
// primitive states: the three corner contributions (fab + fca + f_a style
// weights).  The conserved values are rebuilt from the primitive ones; a
// transported scalar's corner flux is its value times the normal velocity.
template <int NC>
__device__ inline void face_flux_lds(int la, int lb, int lc, int sym,
                                     const double* __restrict__ sP,
                                     const double* __restrict__ sX,
                                     double gm1, double out[3][NC]) {
  constexpr int CAP = xg::kRhsCap;   // the rhs blocks' LDS stride
  int N[3] = {la, lb, lc};
  double ba[3], ca[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    ba[j] = sX[j * CAP + lb] - sX[j * CAP + la];
    ca[j] = sX[j * CAP + lc] - sX[j * CAP + la];
  }
  double nx = (ba[1] * ca[2] - ca[1] * ba[2]) / 12.0;
  double ny = (ba[2] * ca[0] - ca[2] * ba[0]) / 12.0;
  double nz = (ba[0] * ca[1] - ca[0] * ba[1]) / 12.0;
  double fl[NC][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double r = sP[0 * CAP + N[a]], u = sP[1 * CAP + N[a]], v = sP[2 * CAP + N[a]];
    double w = sP[3 * CAP + N[a]], e = sP[4 * CAP + N[a]];
    double p = (r * e) * gm1;
    double vn = ((sym >> a) & 1) ? 0.0 : (nx * u + ny * v + nz * w);
    double re = r * (e + 0.5 * (u * u + v * v + w * w));
    fl[0][a] = r * vn;
    fl[1][a] = r * u * vn + p * nx;
    fl[2][a] = r * v * vn + p * ny;
    fl[3][a] = r * w * vn + p * nz;
    fl[4][a] = (re + p) * vn;
#pragma unroll
    for (int c = 5; c < NC; ++c) fl[c][a] = sP[c * CAP + N[a]] * vn;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double fab = (fl[c][0] + fl[c][1]) / 4.0;
    double fbc = (fl[c][1] + fl[c][2]) / 4.0;
    double fca = (fl[c][2] + fl[c][0]) / 4.0;
    out[0][c] = fab + fca + fl[c][0];
    out[1][c] = fab + fbc + fl[c][1];
    out[2][c] = fbc + fca + fl[c][2];
  }
}

// riemann::rhs (advdom + advbnd, Riemann.cpp:652-878) for one node block.
// The block's nodes (owned range + halo) are staged in LDS as primitive
// state, gradient and coordinates.  Thread t owns local node t and keeps its
// nodal sum in registers.  Per chunk: each thread evaluates one edge flux (or
// one face's three corner terms) into LDS, then every node thread adds the
// chunk's slots that touch it, in slot order.  NC components per node
// (5 + transported scalars; with scalars the LDS of one workgroup exceeds
// half a CU's, so one workgroup runs per CU).
template <int FLUX, int NC>
__global__ __launch_bounds__(kRhsThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_rhs_blocks(
    int nblk, const int4* __restrict__ blk0, const int4* __restrict__ blk1,
    const int* __restrict__ halo, const int4* __restrict__ chk,
    const int* __restrict__ chk_ent, const uint16_t* __restrict__ tend,
    const uint16_t* __restrict__ entries, const uint32_t* __restrict__ elpq,
    const double* __restrict__ dx, const double* __restrict__ dy,
    const double* __restrict__ dz, const uint32_t* __restrict__ fab,
    const uint32_t* __restrict__ fcs, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ G, const double* __restrict__ U,
    double* __restrict__ R, double* __restrict__ S, Cfg cfg) {
  constexpr int CAP = xg::kRhsCap;
  __shared__ double sN[(4 * NC + 3) * CAP];
  double* const sP = sN;                  // primitive state, NC x CAP
  double* const sG = sN + NC * CAP;       // gradients, 3 NC x CAP
  double* const sX = sN + 4 * NC * CAP;   // coordinates, 3 x CAP
  __shared__ double F[kRhsThreads * NC];
  __shared__ uint16_t sE[2 * kRhsThreads];
  __shared__ uint16_t sT[CAP];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 B0 = blk0[b];   // own0, nown, nloc, halo_off
  const int4 B1 = blk1[b];   // slot0, chunk0, chunk1
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  // stage the block's nodes (one per thread, nloc <= threads): all 4 NC + 3
  // loads of a node are issued before the first is used
  if (t < nloc) {
    int g = t < nown ? own0 + t : halo[B0.w + t - nown];
    double s[NC], u[NC], gg[3 * NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = U[(size_t)g * NC + c];
#pragma unroll
    for (int c = 0; c < 3 * NC; ++c) gg[c] = G[(size_t)g * (3 * NC) + c];
    double xg_ = x[g], yg = y[g], zg = z[g];
    xg::primitive_fast(s, u);
#pragma unroll
    for (int c = 5; c < NC; ++c) u[c] = s[c];
#pragma unroll
    for (int c = 0; c < NC; ++c) sP[c * CAP + t] = u[c];
#pragma unroll
    for (int c = 0; c < 3 * NC; ++c) sG[c * CAP + t] = gg[c];
    sX[0 * CAP + t] = xg_;
    sX[1 * CAP + t] = yg;
    sX[2 * CAP + t] = zg;
  }
  // the next chunk's edge record and entry lists are loaded while the
  // current chunk computes
  ChunkPrefetch<kRhsThreads> pf;
  uint32_t pq_n = 0;
  double dx_n = 0.0, dy_n = 0.0, dz_n = 0.0;
  auto fetch = [&](int c, int tl) {
    if (c < B1.z) {
      const int4 C = chk[c];
      pf.load(C, chk_ent[c], nloc, tl, tend, entries);
      if (C.x == 0 && tl < C.z) {
        int e = C.y + tl;
        pq_n = elpq[e];
        dx_n = dx[e];
        dy_n = dy[e];
        dz_n = dz[e];
      }
    }
  };
  fetch(B1.y, t);
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (int ch = B1.y; ch < B1.z; ++ch) {
    // the thread index, opaque to the optimizer: thread-dependent LDS
    // addresses are recomputed per chunk instead of being held (and spilled)
    // across the whole loop
    int tl = t;
    asm volatile("" : "+v"(tl));
    const int4 C = chk[ch];   // kind, first, count, tend_off
    const uint32_t pq = pq_n;
    double d[3] = {dx_n, dy_n, dz_n};
    __syncthreads();          // staging done / previous chunk's F consumed
    pf.store(C, nloc, tl, sE, sT);
    if (tl < C.z) {
      if (C.x == 0) {
        int lp = pq & 0xffff, lq = pq >> 16;
        double l[NC], r[NC], f[NC], g1[NC], g2[NC];
        double vw[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) vw[j] = sX[j * CAP + lq] - sX[j * CAP + lp];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          l[c] = sP[c * CAP + lp];
          r[c] = sP[c * CAP + lq];
          g1[c] = sG[(c * 3 + 0) * CAP + lp] * vw[0] + sG[(c * 3 + 1) * CAP + lp] * vw[1] +
                  sG[(c * 3 + 2) * CAP + lp] * vw[2];
          g2[c] = sG[(c * 3 + 0) * CAP + lq] * vw[0] + sG[(c * 3 + 1) * CAP + lq] * vw[1] +
                  sG[(c * 3 + 2) * CAP + lq] * vw[2];
        }
        xg::muscl_fast(g1, g2, l, r);
        xg::ScalarFlux sf;
        if (FLUX == 0)
          xg::rusanov_fast(cfg, d, l, r, f, sf);
        else
          xg::hllc_fast(cfg, d, l, r, f, sf);
        if constexpr (NC > 5) {
          xg::muscl_scalars<NC>(g1, g2, l, r);
#pragma unroll
          for (int c = 5; c < NC; ++c)
            f[c] = (l[c] * sf.a + r[c] * sf.b) * sf.m + sf.sw * (r[c] - l[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) F[tl * NC + c] = f[c];
      } else {
        int f = C.y + tl;
        uint32_t ab = fab[f], cs = fcs[f];
        double o[3][NC];
        face_flux_lds<NC>(ab & 0xffff, ab >> 16, cs & 0xffff, cs >> 16, sP, sX, cfg.gm1, o);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int c = 0; c < NC; ++c) F[(3 * tl + a) * NC + c] = o[a][c];
      }
    }
    fetch(ch + 1, tl);        // in flight during the accumulation below
    __syncthreads();          // F, sE, sT complete
    if (tl < nloc) {
      int e0 = tl ? sT[tl - 1] : 0, e1 = sT[tl];
      for (int j = e0; j < e1; ++j) {
        int v = sE[j];
        const double* fs = F + (v & 0x3ff) * NC;
        double sg = (v & 0x8000) ? -1.0 : 1.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = fma(sg, fs[c], acc[c]);
      }
    }
  }
  if (t < nown) {
    double* o = R + (size_t)(own0 + t) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = acc[c];
  } else if (t < nloc) {
    double* o = S + (size_t)(B1.x + t - nown) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = acc[c];
  }
}

// U = Un - rkcoef*dt*R/vol (RieCG.cpp:1016-1021); STAGE0 first saves
// Un = U (RieCG.cpp:1000-1003); NC components per node
template <bool STAGE0, int NC>
__global__ void k_update(int npoin, double coef_dt, double* __restrict__ U,
                         double* __restrict__ Un, const double* __restrict__ R,
                         const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * NC) return;
  double un;
  if (STAGE0) {
    un = U[i];
    Un[i] = un;
  } else {
    un = Un[i];
  }
  U[i] = un - coef_dt * R[i] / vol[i / NC];
}

// U = Un - rkcoef*dtp[i]*R/vol with per-node time steps (steady runs,
// RieCG.cpp:1017-1018)
template <bool STAGE0, int NC>
__global__ void k_update_steady(int npoin, double rk, const double* __restrict__ dtp,
                                double* __restrict__ U, double* __restrict__ Un,
                                const double* __restrict__ R,
                                const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * NC) return;
  double un;
  if (STAGE0) {
    un = U[i];
    Un[i] = un;
  } else {
    un = Un[i];
  }
  U[i] = un - rk * dtp[i / NC] * R[i] / vol[i / NC];
}

// F(i, :) /= vol[i] for NC components (RieCG.cpp:935-939)
template <int NC>
__global__ void k_div_vol(int npoin, double* __restrict__ F, const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * NC) return;
  F[i] /= vol[i / NC];
}

// riemann::src (Riemann.cpp:880-907): R(p,c) -= s_c(x_p) v[p] for the flow
// components of the nc per node (the scalars have no source)
__global__ void k_src(int npoin, int nc, xg::Prob P, const double* __restrict__ x,
                      const double* __restrict__ y, const double* __restrict__ z,
                      const double* __restrict__ v, double* __restrict__ R) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npoin) return;
  double s[5];
  xg::source(P, x[p], y[p], z[p], s);
  double* r = R + (size_t)p * nc;
  const double vp = v[p];
#pragma unroll
  for (int c = 0; c < 5; ++c) r[c] -= s[c] * vp;
}

// physics::dirbc (BC.cpp:29-72) as the RieCG chare calls it: components
// with mask 1 take the node's initial-condition value (nc per node)
__global__ void k_dirbc_ic(int nd, int nc, const int* __restrict__ node,
                           const uint8_t* __restrict__ mask,
                           const double* __restrict__ val, double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  double* u = U + (size_t)node[i] * nc;
  for (int c = 0; c < nc; ++c)
    if (mask[i * nc + c] == 1) u[c] = val[i * nc + c];
}

// physics::farbc (BC.cpp:152-221), the (node, normal) pairs of a node in
// list order
__global__ void k_farbc(int nf, int nc, const int* __restrict__ node,
                        const int* __restrict__ off, const double* __restrict__ nrm,
                        double fr, double fu, double fv, double fw, double fp, double g,
                        double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  double* s = U + (size_t)node[i] * nc;
  double r = s[0], ru = s[1], rv = s[2], rw = s[3], re = s[4];
  const double a = sqrt(g * fp / fr);
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const double* n = nrm + (size_t)k * 3;
    const double M = (fu * n[0] + fv * n[1] + fw * n[2]) / a;
    if (M <= -1.0) {              // supersonic inflow
      r = fr; ru = fr * fu; rv = fr * fv; rw = fr * fw;
      re = xg::total_energy(g, fr, fu, fv, fw, fp);
    } else if (M > -1.0 && M < 0.0) {   // subsonic inflow: inside pressure
      const double pr = (re - 0.5 * (ru * ru + rv * rv + rw * rw) / r) * (g - 1.0);
      r = fr; ru = fr * fu; rv = fr * fv; rw = fr * fw;
      re = xg::total_energy(g, fr, fu, fv, fw, pr);
    } else if (M >= 0.0 && M < 1.0) {   // subsonic outflow: outside pressure
      re = xg::total_energy(g, r, ru / r, rv / r, rw / r, fp);
    }
  }
  s[0] = r; s[1] = ru; s[2] = rv; s[3] = rw; s[4] = re;
}

// physics::prebc (BC.cpp:223-241), a node's (density, pressure) records in
// list order
__global__ void k_prebc(int np, int nc, const int* __restrict__ node,
                        const int* __restrict__ off, const double* __restrict__ val,
                        double g, double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double* s = U + (size_t)node[i] * nc;
  for (int k = off[i]; k < off[i + 1]; ++k) {
    s[0] = val[k * 2 + 0];
    s[4] = xg::total_energy(g, s[0], s[1] / s[0], s[2] / s[0], s[3] / s[0], val[k * 2 + 1]);
  }
}

// ------------------------------------------------------------- launchers --

int check_ncomp(xg_ctx* c, int ncomp) {
  if (ncomp < 5 || ncomp > XG_MAX_NCOMP)
    return fail(c, XG_EUNSUPPORTED,
                "device path supports ncomp = 5 + at most " +
                    std::to_string(XG_MAX_NCOMP - 5) + " transported scalars");
  return XG_OK;
}

int check_mesh(xg_ctx* c, int ncomp) {
  if (!c) return XG_EINVAL;
  if (!c->have_mesh) return fail(c, XG_ENOMESH, "no mesh uploaded");
  return check_ncomp(c, ncomp);
}

template <int NC>
void launch_grad_nc(xg_ctx* c, const double* U, double* G, bool div) {
  if (div)
    hipLaunchKernelGGL((k_grad_blocks<true, NC>), dim3(c->nblk), dim3(kGradThreads), 0,
                       c->stream, c->nblk, c->blk0, c->blk1, c->halo, c->chk,
                       c->chk_ent, c->tend, c->entries, c->elpq, c->dx, c->dy,
                       c->dz, U, G, c->gflag, c->vol);
  else
    hipLaunchKernelGGL((k_grad_blocks<false, NC>), dim3(c->nblk), dim3(kGradThreads), 0,
                       c->stream, c->nblk, c->blk0, c->blk1, c->halo, c->chk,
                       c->chk_ent, c->tend, c->entries, c->elpq, c->dx, c->dy,
                       c->dz, U, G, nullptr, nullptr);
  if (c->nbnode) {
    if (div)
      hipLaunchKernelGGL((k_grad_bnd<true, NC>), dim3(nblocks(c->nbnode)), dim3(kBlock), 0,
                         c->stream, c->nbnode, c->bnode, c->boff, c->bface,
                         c->tri, c->x, c->y, c->z, U, G, c->vol);
    else
      hipLaunchKernelGGL((k_grad_bnd<false, NC>), dim3(nblocks(c->nbnode)), dim3(kBlock), 0,
                         c->stream, c->nbnode, c->bnode, c->boff, c->bface,
                         c->tri, c->x, c->y, c->z, U, G, nullptr);
  }
}

// div: G = grad(U)/vol (the RK stage; needs xg_riecg_setup), each node
// divided by the last kernel that adds to it; otherwise grad(U) undivided.
// nc components per node (checked by the caller).
int launch_grad(xg_ctx* c, int nc, const double* U, double* G, bool div = false) {
  static_assert(XG_MAX_NCOMP == 7, "one instantiation per component count");
  if (nc == 5) launch_grad_nc<5>(c, U, G, div);
  else if (nc == 6) launch_grad_nc<6>(c, U, G, div);
  else launch_grad_nc<7>(c, U, G, div);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

xg::Prob prob_of(const xg_ctx* c) {
  return xg::Prob{c->problem.kind, c->problem.alpha, c->problem.kappa, c->problem.p0,
                  c->cfg.gamma};
}

template <int NC>
void launch_rhs_nc(xg_ctx* c, const double* G, const double* U, double* R) {
  const double g = c->cfg.gamma;
  Cfg cfg{g, c->cfg.stab2, c->cfg.stab2coef, g - 1.0, g * (g - 1.0)};
  auto kern = c->cfg.flux == 0 ? k_rhs_blocks<0, NC> : k_rhs_blocks<1, NC>;
  const auto& B = c->rb;
  if (B.nblk)
    hipLaunchKernelGGL(kern, dim3(B.nblk), dim3(kRhsThreads), 0, c->stream,
                       B.nblk, B.blk0, B.blk1, B.halo, B.chk, B.chk_ent, B.tend,
                       B.entries, B.elpq, B.dx, B.dy, B.dz, B.fab, B.fcs, c->x, c->y,
                       c->z, G, U, R, B.slots, cfg);
  if (B.nfx)
    hipLaunchKernelGGL(k_fixup<NC>, dim3(nblocks(B.nfx)), dim3(kBlock), 0,
                       c->stream, B.nfx, B.fx_node, B.fx_off, B.fx_slot, B.slots, R);
}

// riemann::rhs with nc components per node (checked by the caller), then
// the problem's source on the flow components
int launch_rhs(xg_ctx* c, int nc, const double* G, const double* U, double* R,
               const double* v) {
  if (c->cfg.flux != 0 && c->cfg.flux != 1)
    return fail(c, XG_EINVAL, "Flux not configured");
  if (nc == 5) launch_rhs_nc<5>(c, G, U, R);
  else if (nc == 6) launch_rhs_nc<6>(c, G, U, R);
  else launch_rhs_nc<7>(c, G, U, R);
  const xg::Prob P = prob_of(c);
  if (xg::has_source(P))
    hipLaunchKernelGGL(k_src, dim3(nblocks(c->npoin)), dim3(kBlock), 0, c->stream,
                       c->npoin, nc, P, c->x, c->y, c->z, v, R);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

// RieCG::BC (RieCG.cpp:764-785): dirbc, symbc, farbc, prebc
int launch_bc(xg_ctx* c, double* U) {
  const int nc = c->ncomp;
  if (c->ndirnode)
    hipLaunchKernelGGL(k_dirbc_ic, dim3(nblocks(c->ndirnode)), dim3(kBlock), 0,
                       c->stream, c->ndirnode, nc, c->dirnode, c->dirmask, c->dirval, U);
  if (c->nsymnode)
    hipLaunchKernelGGL(k_symbc, dim3(nblocks(c->nsymnode)), dim3(kBlock), 0,
                       c->stream, c->nsymnode, c->symnode, c->symoff,
                       c->symnorm, U, nc, 1);
  if (c->nfarnode)
    hipLaunchKernelGGL(k_farbc, dim3(nblocks(c->nfarnode)), dim3(kBlock), 0,
                       c->stream, c->nfarnode, nc, c->farnode, c->faroff, c->farnorm,
                       c->far[0], c->far[1], c->far[2], c->far[3], c->far[4],
                       c->cfg.gamma, U);
  if (c->nprenode)
    hipLaunchKernelGGL(k_prebc, dim3(nblocks(c->nprenode)), dim3(kBlock), 0,
                       c->stream, c->nprenode, nc, c->prenode, c->preoff, c->preval,
                       c->cfg.gamma, U);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

template <int NC>
void launch_update_nc(xg_ctx* c, int stage, double rk, double dt, const double* dtp,
                      double* U, double* Un, const double* R) {
  const int64_t n = (int64_t)c->npoin * NC;
  if (dtp) {
    if (stage == 0)
      hipLaunchKernelGGL((k_update_steady<true, NC>), dim3(nblocks(n)), dim3(kBlock), 0,
                         c->stream, c->npoin, rk, dtp, U, Un, R, c->vol);
    else
      hipLaunchKernelGGL((k_update_steady<false, NC>), dim3(nblocks(n)), dim3(kBlock), 0,
                         c->stream, c->npoin, rk, dtp, U, Un, R, c->vol);
  } else if (stage == 0) {
    hipLaunchKernelGGL((k_update<true, NC>), dim3(nblocks(n)), dim3(kBlock), 0, c->stream,
                       c->npoin, rk * dt, U, Un, R, c->vol);
  } else {
    hipLaunchKernelGGL((k_update<false, NC>), dim3(nblocks(n)), dim3(kBlock), 0, c->stream,
                       c->npoin, rk * dt, U, Un, R, c->vol);
  }
}

// RieCG::solve (RieCG.cpp:992-1057) after the rhs is complete
int launch_solve(xg_ctx* c, int stage, double dt, const double* dtp, double* tp,
                 double* U, double* Un, const double* R) {
  static const double rkcoef[3] = {1.0 / 3.0, 1.0 / 2.0, 1.0};
  const double rk = rkcoef[stage];
  if (c->ncomp == 5) launch_update_nc<5>(c, stage, rk, dt, dtp, U, Un, R);
  else if (c->ncomp == 6) launch_update_nc<6>(c, stage, rk, dt, dtp, U, Un, R);
  else launch_update_nc<7>(c, stage, rk, dt, dtp, U, Un, R);
  int rc = launch_bc(c, U);
  if (rc) return rc;
  if (dtp && stage == 2)
    hipLaunchKernelGGL(k_axpy1, dim3(nblocks(c->npoin)), dim3(kBlock), 0, c->stream,
                       (int64_t)c->npoin, dtp, tp);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

// (node, record) lists grouped by node, records of a node in list order
template <int W>
int group_by_node(xg_ctx* c, int64_t n, const size_t* nodes, const double* rec,
                  std::vector<int>& node, std::vector<int>& off,
                  std::vector<double>& out, const char* what) {
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) {
    if (nodes[i] >= (size_t)c->npoin)
      return fail(c, XG_EINVAL, std::string("xg_riecg_bc: ") + what + " node out of range");
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return nodes[a] < nodes[b]; });
  off.assign(1, 0);
  node.clear();
  out.clear();
  for (int64_t k = 0; k < n; ++k) {
    int64_t i = order[k];
    if (k == 0 || nodes[order[k - 1]] != nodes[i]) {
      if (k) off.push_back((int)(out.size() / W));
      node.push_back((int)nodes[i]);
    }
    out.insert(out.end(), rec + i * W, rec + i * W + W);
  }
  if (n) off.push_back((int)(out.size() / W));
  return XG_OK;
}

// uploads the riemann::rhs layout (its block tables, edge records, chunk
// entry lists, boundary faces, halo slots and fix-up lists)
int upload_blocks(xg_ctx* c, std::vector<DevArr>& M, const xg::Layout& L,
                  xg_ctx::EdgeBlocks& B) {
  std::vector<int4> b0(L.blk.size()), b1(L.blk.size());
  for (size_t b = 0; b < L.blk.size(); ++b) {
    const auto& K = L.blk[b];
    b0[b] = make_int4(K.own0, K.nown, K.nloc, K.halo_off);
    b1[b] = make_int4(K.slot0, K.chunk0, K.chunk1, 0);
  }
  std::vector<int4> c0(L.chunks.size());
  std::vector<int> c1(L.chunks.size());
  for (size_t i = 0; i < L.chunks.size(); ++i) {
    const auto& C = L.chunks[i];
    c0[i] = make_int4(C.kind, C.first, C.count, C.tend_off);
    c1[i] = C.ent_off;
  }
  int rc;
  if ((rc = upload(c, M, &B.blk0, b0)) || (rc = upload(c, M, &B.blk1, b1)) ||
      (rc = upload(c, M, &B.halo, L.halo)) || (rc = upload(c, M, &B.chk, c0)) ||
      (rc = upload(c, M, &B.chk_ent, c1)) || (rc = upload(c, M, &B.tend, L.tend)) ||
      (rc = upload(c, M, &B.entries, L.entries)) || (rc = upload(c, M, &B.elpq, L.elpq)) ||
      (rc = upload(c, M, &B.dx, L.dx)) || (rc = upload(c, M, &B.dy, L.dy)) ||
      (rc = upload(c, M, &B.dz, L.dz)) || (rc = upload(c, M, &B.fab, L.fab)) ||
      (rc = upload(c, M, &B.fcs, L.fcs)) ||
      (rc = dev_alloc(c, M, (void**)&B.slots,
                      (size_t)L.nslot * XG_MAX_NCOMP * sizeof(double))) ||
      (rc = upload(c, M, &B.fx_node, L.fx_node)) || (rc = upload(c, M, &B.fx_off, L.fx_off)) ||
      (rc = upload(c, M, &B.fx_slot, L.fx_slot)))
    return rc;
  B.nblk = (int)L.blk.size();
  B.nfx = (int)L.fx_node.size();
  return XG_OK;
}

// device staging of the host-pointer entries for nc components per node
int stage_buffers(xg_ctx* c, int nc) {
  size_t need = (size_t)c->npoin * 8;
  if (c->sbytes >= need && c->snc >= nc) return XG_OK;
  if (c->sU) { (void)hipFree(c->sU); (void)hipFree(c->sG); (void)hipFree(c->sR); }
  c->sU = c->sG = c->sR = nullptr; c->sbytes = 0; c->snc = 0;
  XG_HIP(c, hipMalloc((void**)&c->sU, need * nc));
  XG_HIP(c, hipMalloc((void**)&c->sG, need * (3 * nc + 1)));   // G and, after it, v
  XG_HIP(c, hipMalloc((void**)&c->sR, need * nc));
  c->sbytes = need;
  c->snc = nc;
  return XG_OK;
}

}  // namespace

// =================================================================== C-ABI ==

extern "C" {

int xg_ctx_create(int device, xg_ctx** out) {
  if (!out) return XG_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return XG_EHIP;
  if (device < 0 || device >= n) return XG_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return XG_EHIP;
  if (xg_diag_scratch_init(device) != XG_OK) return XG_EHIP;
  xg_ctx* c = new xg_ctx;
  c->device = device;
  *out = c;
  return XG_OK;
}

int xg_ctx_destroy(xg_ctx* c) {
  if (!c) return XG_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(c->mesh_allocs);
  dev_free_all(c->riecg_allocs);
  dev_free_all(c->bc_allocs);
  if (c->sU) { (void)hipFree(c->sU); (void)hipFree(c->sG); (void)hipFree(c->sR); }
  if (c->cg) (void)xg_cg_destroy(c);
  if (c->koz) xg_kozak_free(c);
  if (c->cho) xg_chocg_free(c);
  if (c->gnn) xg_gnn_free(c);
  if (c->diagbuf.p) (void)hipFree(c->diagbuf.p);
  if (c->amg) xg_amg_free(c);
  delete c;
  return XG_OK;
}

const char* xg_last_error(const xg_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

int xg_set_stream(xg_ctx* c, void* s) {
  if (!c) return XG_EINVAL;
  c->stream = (hipStream_t)s;
  return XG_OK;
}

int xg_set_config(xg_ctx* c, const xg_config* cfg) {
  if (!c || !cfg) return XG_EINVAL;
  if (cfg->flux != 0 && cfg->flux != 1)
    return fail(c, XG_EINVAL, "Flux not configured");
  if (!(cfg->gamma > 1.0)) return fail(c, XG_EINVAL, "gamma must exceed 1");
  c->cfg = *cfg;
  return XG_OK;
}

int xg_mesh_upload(xg_ctx* c, int64_t npoin, const double* x, const double* y,
                   const double* z, const int64_t nsup[3],
                   const size_t* s0, const double* i0, const size_t* s1,
                   const double* i1, const size_t* s2, const double* i2,
                   int64_t ntri, const size_t* tri, const uint8_t* besym) {
  if (!c) return XG_EINVAL;
  if (npoin <= 0 || npoin >= (1ll << 31) || !x || !y || !z || !nsup)
    return fail(c, XG_EINVAL, "xg_mesh_upload: bad mesh size or coordinates");
  if (nsup[0] < 0 || nsup[1] < 0 || nsup[2] < 0 || ntri < 0 ||
      (nsup[0] && (!s0 || !i0)) || (nsup[1] && (!s1 || !i1)) ||
      (nsup[2] && (!s2 || !i2)) || (ntri && (!tri || !besym)))
    return fail(c, XG_EINVAL, "xg_mesh_upload: missing superedge/face arrays");
  XG_HIP(c, hipSetDevice(c->device));
  // expand superedges into oriented edges, in the reference's loop order
  static const int lpoed[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
  static const int lpoet[3][2] = {{0,1},{1,2},{2,0}};
  int64_t ne = nsup[0] * 6 + nsup[1] * 3 + nsup[2];
  if (ne >= (1ll << 31))
    return fail(c, XG_EINVAL, "xg_mesh_upload: too many edges");
  std::vector<int> p(ne), q(ne);
  std::vector<double> d(ne * 3);
  int64_t k = 0;
  auto push = [&](size_t a, size_t b, const double* D) {
    p[k] = (int)std::min<size_t>(a, (size_t)INT32_MAX);
    q[k] = (int)std::min<size_t>(b, (size_t)INT32_MAX);
    d[k * 3 + 0] = D[0]; d[k * 3 + 1] = D[1]; d[k * 3 + 2] = D[2];
    ++k;
  };
  for (int64_t e = 0; e < nsup[0]; ++e)
    for (int j = 0; j < 6; ++j)
      push(s0[e * 4 + lpoed[j][0]], s0[e * 4 + lpoed[j][1]], i0 + (e * 6 + j) * 3);
  for (int64_t e = 0; e < nsup[1]; ++e)
    for (int j = 0; j < 3; ++j)
      push(s1[e * 3 + lpoet[j][0]], s1[e * 3 + lpoet[j][1]], i1 + (e * 3 + j) * 3);
  for (int64_t e = 0; e < nsup[2]; ++e) push(s2[e * 2], s2[e * 2 + 1], i2 + e * 3);
  for (int64_t e = 0; e < ne; ++e)
    if (p[e] < 0 || q[e] < 0 || p[e] >= npoin || q[e] >= npoin || p[e] == q[e])
      return fail(c, XG_EINVAL, "xg_mesh_upload: superedge node id out of range");
  for (int64_t f = 0; f < ntri * 3; ++f)
    if (tri[f] >= (size_t)npoin)
      return fail(c, XG_EINVAL, "xg_mesh_upload: boundary face node out of range");
  std::vector<int> tri32(ntri * 3);
  for (int64_t f = 0; f < ntri * 3; ++f) tri32[f] = (int)tri[f];
  xg::Layout L, LR;
  int rc = xg::build_layout((int)npoin, p, q, d, ntri, tri32.data(), besym,
                            xg::kGradNodes, xg::kGradCap, xg::kGradSlots, L, c->err,
                            true);
  if (rc) return rc;
  rc = xg::build_layout((int)npoin, p, q, d, ntri, tri32.data(), besym, xg::kRhsNodes,
                        xg::kRhsCap, xg::kRhsSlots, LR, c->err);
  if (rc) return rc;
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
  // boundary gather lists for the gradient: node -> (face*4 + position)
  std::vector<int> bcnt((size_t)npoin + 1, 0);
  for (int64_t f = 0; f < ntri * 3; ++f) ++bcnt[tri[f] + 1];
  std::vector<int> bnode, boff(1, 0), bface;
  {
    std::vector<int> start((size_t)npoin + 1, 0);
    for (int64_t i = 0; i < npoin; ++i) start[i + 1] = start[i] + bcnt[i + 1];
    std::vector<int> fill(ntri * 3);
    std::vector<int> pos(start.begin(), start.end() - 1);
    for (int64_t f = 0; f < ntri; ++f)
      for (int a = 0; a < 3; ++a) fill[pos[tri[f * 3 + a]]++] = (int)(f * 4 + a);
    for (int64_t i = 0; i < npoin; ++i) {
      if (start[i + 1] == start[i]) continue;
      bnode.push_back((int)i);
      bface.insert(bface.end(), fill.begin() + start[i], fill.begin() + start[i + 1]);
      boff.push_back((int)bface.size());
    }
  }
  // device arrays
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(c->mesh_allocs);
  c->have_mesh = false;
  auto& M = c->mesh_allocs;
  if ((rc = upload(c, M, &c->x, x, npoin)) || (rc = upload(c, M, &c->y, y, npoin)) ||
      (rc = upload(c, M, &c->z, z, npoin)) || (rc = upload(c, M, &c->elpq, L.elpq)) ||
      (rc = upload(c, M, &c->dx, L.dx)) ||
      (rc = upload(c, M, &c->dy, L.dy)) || (rc = upload(c, M, &c->dz, L.dz)) ||
      (rc = upload(c, M, &c->blk0, b0)) || (rc = upload(c, M, &c->blk1, b1)) ||
      (rc = upload(c, M, &c->halo, L.halo)) ||
      (rc = upload(c, M, &c->chk, c0)) || (rc = upload(c, M, &c->chk_ent, c1)) ||
      (rc = upload(c, M, &c->tend, L.tend)) ||
      (rc = upload(c, M, &c->entries, L.entries)) ||
      (rc = upload(c, M, &c->tri, tri32)) || (rc = upload(c, M, &c->bnode, bnode)) ||
      (rc = upload(c, M, &c->boff, boff)) || (rc = upload(c, M, &c->bface, bface)))
    return rc;
  if ((rc = upload_blocks(c, M, LR, c->rb))) return rc;
  c->npoin = (int)npoin;
  c->nedge = ne;
  c->nblk = (int)L.blk.size();
  c->nchunk = (int)L.chunks.size();
  c->nbnode = (int)bnode.size();
  c->have_mesh = true;
  c->have_riecg = false;
  return XG_OK;
}

int xg_mesh_device_bytes(const xg_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return XG_EINVAL;
  int64_t b = 0;
  for (auto& a : c->mesh_allocs) b += (int64_t)a.bytes;
  *bytes = b;
  return XG_OK;
}

int xg_riemann_grad_dev(xg_ctx* c, int ncomp, const double* U, double* G) {
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G) return fail(c, XG_EINVAL, "xg_riemann_grad: null field");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_grad(c, ncomp, U, G);
}

int xg_riemann_rhs_dev(xg_ctx* c, int ncomp, const double* G, const double* U,
                       const double* v, double t, const double* tp, double* R) {
  (void)t;   // the sources served do not depend on time
  (void)tp;
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G || !R) return fail(c, XG_EINVAL, "xg_riemann_rhs: null field");
  if (xg::has_source(prob_of(c)) && !v)
    return fail(c, XG_EINVAL, "xg_riemann_rhs: the problem's source needs the nodal volumes v");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_rhs(c, ncomp, G, U, R, v);
}

int xg_riemann_grad(xg_ctx* c, int ncomp, const double* U, double* G) {
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G) return fail(c, XG_EINVAL, "xg_riemann_grad: null field");
  XG_HIP(c, hipSetDevice(c->device));
  if ((rc = stage_buffers(c, ncomp))) return rc;
  size_t n = (size_t)c->npoin * 8;
  XG_HIP(c, hipMemcpyAsync(c->sU, U, n * ncomp, hipMemcpyHostToDevice, c->stream));
  if ((rc = launch_grad(c, ncomp, c->sU, c->sG))) return rc;
  XG_HIP(c, hipMemcpyAsync(G, c->sG, n * 3 * ncomp, hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

int xg_riemann_rhs(xg_ctx* c, int ncomp, const double* G, const double* U,
                   const double* v, double t, const double* tp, double* R) {
  (void)t;   // the sources served do not depend on time
  (void)tp;
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G || !R) return fail(c, XG_EINVAL, "xg_riemann_rhs: null field");
  const bool src = xg::has_source(prob_of(c));
  if (src && !v)
    return fail(c, XG_EINVAL, "xg_riemann_rhs: the problem's source needs the nodal volumes v");
  XG_HIP(c, hipSetDevice(c->device));
  if ((rc = stage_buffers(c, ncomp))) return rc;
  size_t n = (size_t)c->npoin;
  XG_HIP(c, hipMemcpyAsync(c->sU, U, n * 8 * ncomp, hipMemcpyHostToDevice, c->stream));
  XG_HIP(c, hipMemcpyAsync(c->sG, G, n * 24 * ncomp, hipMemcpyHostToDevice, c->stream));
  // the volumes are staged after G in the G staging buffer (3 ncomp + 1
  // doubles per node)
  double* sv = nullptr;
  if (src) {
    sv = c->sG + n * 3 * ncomp;
    XG_HIP(c, hipMemcpyAsync(sv, v, n * 8, hipMemcpyHostToDevice, c->stream));
  }
  if ((rc = launch_rhs(c, ncomp, c->sG, c->sU, c->sR, sv))) return rc;
  XG_HIP(c, hipMemcpyAsync(R, c->sR, n * 8 * ncomp, hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

int xg_riecg_setup(xg_ctx* c, const double* vol, int64_t nsym,
                   const size_t* symnodes, const double* symnorms) {
  if (!c) return XG_EINVAL;
  if (!c->have_mesh) return fail(c, XG_ENOMESH, "no mesh uploaded");
  if (!vol || nsym < 0 || (nsym && (!symnodes || !symnorms)))
    return fail(c, XG_EINVAL, "xg_riecg_setup: bad arguments");
  for (int64_t i = 0; i < c->npoin; ++i)
    if (!(vol[i] > 0)) return fail(c, XG_EINVAL, "xg_riecg_setup: non-positive nodal volume");
  // group the (node, normal) pairs by node, keeping list order per node
  std::vector<int> node, off(1, 0);
  std::vector<double> nrm;
  {
    std::vector<int64_t> order(nsym);
    for (int64_t i = 0; i < nsym; ++i) {
      if (symnodes[i] >= (size_t)c->npoin)
        return fail(c, XG_EINVAL, "xg_riecg_setup: symmetry node out of range");
      order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      return symnodes[a] < symnodes[b];
    });
    for (int64_t k = 0; k < nsym; ++k) {
      int64_t i = order[k];
      if (k == 0 || symnodes[order[k - 1]] != symnodes[i]) {
        if (k) off.push_back((int)(nrm.size() / 3));
        node.push_back((int)symnodes[i]);
      }
      nrm.insert(nrm.end(), symnorms + i * 3, symnorms + i * 3 + 3);
    }
    if (nsym) off.push_back((int)(nrm.size() / 3));
  }
  XG_HIP(c, hipSetDevice(c->device));
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(c->riecg_allocs);
  dev_free_all(c->bc_allocs);
  c->ndirnode = c->nfarnode = c->nprenode = 0;
  c->have_riecg = false;
  c->ncomp = 5;
  int rc;
  auto& L = c->riecg_allocs;
  if ((rc = upload(c, L, &c->vol, vol, (size_t)c->npoin)) ||
      (rc = upload(c, L, &c->symnode, node)) || (rc = upload(c, L, &c->symoff, off)) ||
      (rc = upload(c, L, &c->symnorm, nrm)) ||
      (rc = dev_alloc(c, L, (void**)&c->red, sizeof(double) * (kRedBlocks + 1))) ||
      (rc = dev_alloc(c, L, (void**)&c->gflag, (size_t)c->npoin)))
    return rc;
  hipLaunchKernelGGL(k_fill_u8, dim3(nblocks(c->npoin)), dim3(kBlock), 0, c->stream,
                     (int64_t)c->npoin, c->gflag);
  if (c->nbnode)
    hipLaunchKernelGGL(k_gflag, dim3(nblocks(c->nbnode)), dim3(kBlock), 0, c->stream,
                       c->nbnode, c->bnode, (uint8_t)1, c->gflag);
  XG_HIP(c, hipGetLastError());
  c->nsymnode = (int)node.size();
  c->have_riecg = true;
  return XG_OK;
}

int xg_riecg_set_ncomp(xg_ctx* c, int ncomp) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  int rc = check_ncomp(c, ncomp);
  if (rc) return rc;
  if (ncomp != c->ncomp) {
    // the Dirichlet masks and values are per component: drop the lists
    XG_HIP(c, hipSetDevice(c->device));
    (void)hipStreamSynchronize(c->stream);
    dev_free_all(c->bc_allocs);
    c->ndirnode = c->nfarnode = c->nprenode = 0;
  }
  c->ncomp = ncomp;
  return XG_OK;
}

int xg_riecg_dt_dev(xg_ctx* c, const double* U, double cfl, double* dt) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (!U || !dt) return fail(c, XG_EINVAL, "xg_riecg_dt: null argument");
  XG_HIP(c, hipSetDevice(c->device));
  int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(c->npoin)));
  hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(kBlock), 0, c->stream,
                     c->npoin, c->ncomp, U, c->vol, c->cfg.gamma, c->red);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kRedBlocks), 0, c->stream, nb,
                     c->red, cfl, c->red + kRedBlocks);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipMemcpyAsync(dt, c->red + kRedBlocks, sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (*dt < 2.220446049250313e-16)
    return fail(c, XG_EBLOWUP, "time step below machine epsilon");
  return XG_OK;
}

int xg_riecg_dt_steady_dev(xg_ctx* c, const double* U, double cfl, double* dtp,
                           double* dt) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (!U || !dtp || !dt) return fail(c, XG_EINVAL, "xg_riecg_dt_steady: null argument");
  XG_HIP(c, hipSetDevice(c->device));
  int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(c->npoin)));
  hipLaunchKernelGGL(k_dt_steady, dim3(nb), dim3(kBlock), 0, c->stream, c->npoin,
                     c->ncomp, U, c->vol, c->cfg.gamma, cfl, dtp, c->red);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kRedBlocks), 0, c->stream, nb,
                     c->red, 1.0, c->red + kRedBlocks);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipMemcpyAsync(dt, c->red + kRedBlocks, sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (*dt < 2.220446049250313e-16)
    return fail(c, XG_EBLOWUP, "time step below machine epsilon");
  return XG_OK;
}

int xg_riecg_bc(xg_ctx* c, int64_t ndir, const size_t* dirnodes, const int32_t* dirmask,
                const double* dirvals, int64_t nfar, const size_t* farnodes,
                const double* farnorms, const double far[5], int64_t npre,
                const size_t* prenodes, const double* prevals) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (ndir < 0 || nfar < 0 || npre < 0 ||
      (ndir && (!dirnodes || !dirmask || !dirvals)) ||
      (nfar && (!farnodes || !farnorms || !far)) || (npre && (!prenodes || !prevals)))
    return fail(c, XG_EINVAL, "xg_riecg_bc: bad arguments");
  const int nc = c->ncomp;
  std::vector<int> dn(ndir);
  std::vector<uint8_t> dm(ndir * nc);
  for (int64_t i = 0; i < ndir; ++i) {
    if (dirnodes[i] >= (size_t)c->npoin)
      return fail(c, XG_EINVAL, "xg_riecg_bc: Dirichlet node out of range");
    dn[i] = (int)dirnodes[i];
    for (int k = 0; k < nc; ++k) dm[i * nc + k] = (uint8_t)(dirmask[i * nc + k] == 1);
  }
  std::vector<int> fnode, foff, pnode, poff;
  std::vector<double> fnrm, pval;
  int rc;
  if ((rc = group_by_node<3>(c, nfar, farnodes, farnorms, fnode, foff, fnrm, "farfield")) ||
      (rc = group_by_node<2>(c, npre, prenodes, prevals, pnode, poff, pval, "pressure")))
    return rc;
  if (nfar && !(far[0] > 0.0 && far[4] > 0.0))
    return fail(c, XG_EINVAL, "xg_riecg_bc: farfield density and pressure must be positive");
  XG_HIP(c, hipSetDevice(c->device));
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(c->bc_allocs);
  c->ndirnode = c->nfarnode = c->nprenode = 0;
  auto& L = c->bc_allocs;
  if ((rc = upload(c, L, &c->dirnode, dn)) || (rc = upload(c, L, &c->dirmask, dm)) ||
      (rc = upload(c, L, &c->dirval, dirvals, (size_t)ndir * nc)) ||
      (rc = upload(c, L, &c->farnode, fnode)) || (rc = upload(c, L, &c->faroff, foff)) ||
      (rc = upload(c, L, &c->farnorm, fnrm)) || (rc = upload(c, L, &c->prenode, pnode)) ||
      (rc = upload(c, L, &c->preoff, poff)) || (rc = upload(c, L, &c->preval, pval)))
    return rc;
  if (nfar)
    for (int k = 0; k < 5; ++k) c->far[k] = far[k];
  c->ndirnode = (int)ndir;
  c->nfarnode = (int)fnode.size();
  c->nprenode = (int)pnode.size();
  return XG_OK;
}

int xg_riecg_apply_bc_dev(xg_ctx* c, double* U) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (!U) return fail(c, XG_EINVAL, "xg_riecg_apply_bc: null field");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_bc(c, U);
}

int xg_riecg_stage_dev(xg_ctx* c, int stage, double t, double dt, const double* dtp,
                       double* tp, double* U, double* Un, double* G, double* R) {
  (void)t;   // BC values and sources served are independent of time
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (stage < 0 || stage > 2) return fail(c, XG_EINVAL, "stage must be 0, 1 or 2");
  if (!U || !Un || !G || !R) return fail(c, XG_EINVAL, "xg_riecg_stage: null field");
  if ((dtp == nullptr) != (tp == nullptr))
    return fail(c, XG_EINVAL, "xg_riecg_stage: dtp and tp go together");
  XG_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = launch_grad(c, c->ncomp, U, G, true))) return rc;
  if ((rc = launch_rhs(c, c->ncomp, G, U, R, c->vol))) return rc;
  return launch_solve(c, stage, dt, dtp, tp, U, Un, R);
}

int xg_riecg_grad_div_dev(xg_ctx* c, double* G) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (!G) return fail(c, XG_EINVAL, "xg_riecg_grad_div: null field");
  XG_HIP(c, hipSetDevice(c->device));
  const int64_t n = (int64_t)c->npoin * 3 * c->ncomp;
  if (c->ncomp == 5)
    hipLaunchKernelGGL(k_div_vol<15>, dim3(nblocks(n)), dim3(kBlock), 0, c->stream,
                       c->npoin, G, c->vol);
  else if (c->ncomp == 6)
    hipLaunchKernelGGL(k_div_vol<18>, dim3(nblocks(n)), dim3(kBlock), 0, c->stream,
                       c->npoin, G, c->vol);
  else
    hipLaunchKernelGGL(k_div_vol<21>, dim3(nblocks(n)), dim3(kBlock), 0, c->stream,
                       c->npoin, G, c->vol);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int xg_riecg_solve_dev(xg_ctx* c, int stage, double t, double dt, const double* dtp,
                       double* tp, double* U, double* Un, const double* R) {
  (void)t;
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (stage < 0 || stage > 2) return fail(c, XG_EINVAL, "stage must be 0, 1 or 2");
  if (!U || !Un || !R) return fail(c, XG_EINVAL, "xg_riecg_solve: null field");
  if ((dtp == nullptr) != (tp == nullptr))
    return fail(c, XG_EINVAL, "xg_riecg_solve: dtp and tp go together");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_solve(c, stage, dt, dtp, tp, U, Un, R);
}

int xg_set_problem(xg_ctx* c, const xg_problem* p) {
  if (!c || !p) return XG_EINVAL;
  if (p->kind != XG_PROBLEM_NONE && p->kind != XG_PROBLEM_VORTICAL_FLOW &&
      p->kind != XG_PROBLEM_TAYLOR_GREEN && p->kind != XG_PROBLEM_SLOT_CYL)
    return fail(c, XG_EUNSUPPORTED, "xg_set_problem: no device source for this problem");
  c->problem = *p;
  return XG_OK;
}

int xg_get_problem(const xg_ctx* c, xg_problem* p) {
  if (!c || !p) return XG_EINVAL;
  *p = c->problem;
  return XG_OK;
}

}  // extern "C"
