// MI355X device path of kozak::rhs (Kozak.cpp:29-180): the KozCG
// Taylor-Galerkin right-hand side of one GPU-resident partition.
//
// Layout (xg_layout.cpp, build_tet_layout): node blocks of consecutive ids;
// a block owns the tets whose smallest node it holds, in chunks of 256 tets.
// One 512-thread workgroup per block (two share a CU) stages the conserved
// state and coordinates of the block's nodes (owned + halo) in LDS and walks
// the block's tets chunk by chunk, the next chunk's tet records and entry
// lists loaded while the current one computes.  Thread t handles corners 2h
// and 2h + 1 (h = t % 2) of chunk tet t / 2; the two threads of a tet are
// neighbouring lanes and combine their partial values with one DPP lane
// exchange.
//  * kozak::rhs (k_kozak_pairs): each thread evaluates the tet geometry and
//    two corners' shares of the half-step state, the exchange completes the
//    state, and the thread writes its two corners' second-step fluxes; in
//    the FCT step also their antidiffusive contributions (second pass).
//  * the FCT bounds (k_kozcg_alw) and limiter (k_kozcg_lim) on the same
//    chunks.
// Then the thread owning local node t adds (or, for the bounds, takes the
// max/min of) the chunk's corner values that touch it, in slot order; halo
// partials go to slots that k_fixup adds to their owners in block order.
// No atomics: every sum has a fixed order and the result is bitwise
// reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "xg_blocks.hpp"
#include "xg_ctx.hpp"
#include "xg_flux.hpp"
#include "xg_layout.hpp"
#include "xg_problems.hpp"

using namespace xgd;

struct KozakMesh {
  int npoin = 0;
  int64_t ntet = 0;
  int nblk = 0;
  std::vector<DevArr> allocs;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  // node blocks, their tets in chunks of kTetsPerPairChunk
  int4 *blk0 = nullptr, *blk1 = nullptr;
  int* halo = nullptr;
  int4* chk = nullptr;
  int* chk_ent = nullptr;
  uint16_t *tend = nullptr, *entries = nullptr;
  uint32_t *tab = nullptr, *tcd = nullptr;
  double* slots = nullptr;
  int nfx = 0;
  int *fx_node = nullptr, *fx_off = nullptr, *fx_slot = nullptr;
  double *sU = nullptr, *sR = nullptr;   // staging of the host-pointer entry
  double* sD = nullptr;                  // (XG_MAX_NCOMP and dtp per node)
  int nslot = 0;
  int ncomp = 5;                         // components of the KozCG step's fields
  // KozCG stepping (xg_kozcg_setup)
  bool have_step = false;
  std::vector<DevArr> step_allocs;
  double* vol = nullptr;
  int nsymnode = 0, ndir = 0;
  int *symnode = nullptr, *symoff = nullptr, *dirnode = nullptr;
  double *symnorm = nullptr, *dirval = nullptr;
  uint8_t* dirmask = nullptr;
  xg_fctcfg fct{};
  double *W = nullptr, *UL = nullptr, *Q = nullptr, *C = nullptr, *A = nullptr;
  double *S15 = nullptr, *red = nullptr;
};

namespace {

constexpr int kKozThreads = xg::kKozPairSlots / 2;   // two corners per thread
static_assert(xg::kKozCap <= kKozThreads, "one local node per thread");

// v from lane (lane ^ 1) of the same quad: two v_mov_b32_dpp with
// quad_perm [1,0,3,2] (0xB1).  Every lane of the wave is active where this
// is used.
template <int CTRL>
__device__ inline double quad_swap(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double quad_xor1(double v) { return quad_swap<0xB1>(v); }

// One chunk of the pair layout, loaded one chunk ahead so that its
// global-memory latency overlaps the current chunk's work: the chunk record,
// thread t's tet (t >> 1; idle threads read the chunk's first tet, so the
// lane exchanges run with every lane), two entries of the chunk's entry
// list (4 per tet) and local node t's entry end.
struct PairPrefetch {
  int4 C = make_int4(0, 0, 0, 0);   // kind, first, count, tend_off
  uint32_t ab = 0, cd = 0;
  uint16_t e0 = 0, e1 = 0, te = 0;
  __device__ inline void load(int ch, int chunk1, const int4* __restrict__ chk,
                              const int* __restrict__ chk_ent,
                              const uint16_t* __restrict__ tend,
                              const uint16_t* __restrict__ entries,
                              const uint32_t* __restrict__ tab,
                              const uint32_t* __restrict__ tcd, int t, int nloc) {
    if (ch >= chunk1) return;
    C = chk[ch];
    const int i = t >> 1;
    const int e = C.y + (i < C.z ? i : 0);
    ab = tab[e];
    cd = tcd[e];
    const int off = chk_ent[ch], ne = 4 * C.z;
    e0 = t < ne ? entries[off + t] : 0;
    e1 = kKozThreads + t < ne ? entries[off + kKozThreads + t] : 0;
    te = t < nloc ? tend[C.w + t] : 0;
  }
  // the chunk's entry list and entry ends into LDS
  __device__ inline void store(int t, int nloc, uint16_t* sE, uint16_t* sT) const {
    if (t < 4 * C.z) sE[t] = e0;
    if (kKozThreads + t < 4 * C.z) sE[kKozThreads + t] = e1;
    if (t < nloc) sT[t] = te;
  }
};

__device__ inline void cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - b[1] * a[2];
  c[1] = a[2] * b[0] - b[2] * a[0];
  c[2] = a[0] * b[1] - b[0] * a[1];
}

// The tet's Jacobian and the antidiffusive contributions of its corners
// 2h and 2h + 1 (KozCG::aec, KozCG.cpp:794-812) for NC components, in the
// reference's operation order and without fused multiply-adds: the limiter
// branches on their signs, so they are bit-identical.
template <int NC>
__device__ inline void pair_aec(const double X[4][3], const double* sU, int cap,
                                const int N[4], int h, double ctau, double out[2][NC]) {
#pragma clang fp contract(off)
  double ba[3], ca[3], da[3], t[3];
  for (int j = 0; j < 3; ++j) {
    ba[j] = X[1][j] - X[0][j];
    ca[j] = X[2][j] - X[0][j];
    da[j] = X[3][j] - X[0][j];
  }
  t[0] = ca[1] * da[2] - da[1] * ca[2];
  t[1] = ca[2] * da[0] - da[2] * ca[0];
  t[2] = ca[0] * da[1] - da[0] * ca[1];
  const double J = ba[0] * t[0] + ba[1] * t[1] + ba[2] * t[2];
  const double j120 = J / 120.0;
  const double mo = j120 * -1.0, md = j120 * 3.0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double u[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) u[b] = sU[c * cap + N[b]];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int a = 2 * h + k;
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) s += (b == a ? md : mo) * ctau * u[b];
      out[k][c] = s;
    }
  }
}

// kozak::rhs (Kozak.cpp:29-180) on the pair layout: chunks of
// kTetsPerPairChunk tets, thread t handles corners 2h and 2h + 1 (h = t & 1)
// of chunk tet t >> 1.  Each thread evaluates the tet geometry and its two
// corners' shares of the half-step state; one lane exchange with its partner
// completes the state ((h0 + h1) + (h2 + h3) in both lanes), and each thread
// writes its two corners' second-step fluxes into the chunk's corner slots
// (4 i + a).  The thread owning local node t then adds the chunk's slots of
// its node, in slot order; halo partials go to slots that k_fixup adds to
// their owners.
// SRC: the problem's source term (Kozak.cpp:124-136, 163-173).
// AEC: also the antidiffusive contributions of KozCG::aec, written into the
// same corner slots after the fluxes are summed and summed per node by sign
// (P+, P-), the output record of a node being R (NC), P+ (NC), P- (NC).
// NC: components per node, 5 + transported scalars (advected with the flow
// velocity, Kozak.cpp:119-121, 155-157).  dtp (nullable): the nodes' local
// time steps of a steady run, each tet taking their mean (Kozak.cpp:105).
template <bool SRC, bool AEC, int NC>
__global__ __launch_bounds__(kKozThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_kozak_pairs(
    int nblk, const int4* __restrict__ blk0, const int4* __restrict__ blk1,
    const int* __restrict__ halo, const int4* __restrict__ chk,
    const int* __restrict__ chk_ent, const uint16_t* __restrict__ tend,
    const uint16_t* __restrict__ entries, const uint32_t* __restrict__ tab,
    const uint32_t* __restrict__ tcd, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ U, double* __restrict__ R,
    double* __restrict__ S, double gm1, double dt, const double* __restrict__ dtp,
    double ctau, xg::Prob P) {
  constexpr int CAP = xg::kKozCap;
  constexpr int NS = xg::kKozPairSlots;
  constexpr int NO = AEC ? 3 * NC : NC;
  static_assert(NS == 2 * kKozThreads, "two corner slots per thread");
  __shared__ double sU[NC * CAP];
  __shared__ double sX[3 * CAP];
  __shared__ double F[NS * NC];
  __shared__ double sD[CAP];           // the nodes' local time steps (steady)
  __shared__ uint16_t sE[NS];
  __shared__ uint16_t sT[CAP];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 B0 = blk0[b];   // own0, nown, nloc, halo_off
  const int4 B1 = blk1[b];   // slot0, chunk0, chunk1
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  if (t < nloc) {
    int g = t < nown ? own0 + t : halo[B0.w + t - nown];
    double s[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) s[c] = U[(size_t)g * NC + c];
    double xg_ = x[g], yg = y[g], zg = z[g];
#pragma unroll
    for (int c = 0; c < NC; ++c) sU[c * CAP + t] = s[c];
    sX[0 * CAP + t] = xg_;
    sX[1 * CAP + t] = yg;
    sX[2 * CAP + t] = zg;
    if (dtp) sD[t] = dtp[g];
  }
  const int i = t >> 1, h = t & 1;
  PairPrefetch pf;
  pf.load(B1.y, B1.z, chk, chk_ent, tend, entries, tab, tcd, t, nloc);
  double acc[NO];
#pragma unroll
  for (int c = 0; c < NO; ++c) acc[c] = 0.0;
  for (int ch = B1.y; ch < B1.z; ++ch) {
    const int4 C = pf.C;
    const uint32_t ab = pf.ab, cd = pf.cd;
    __syncthreads();          // staging done / previous chunk's F consumed
    pf.store(t, nloc, sE, sT);
    pf.load(ch + 1, B1.z, chk, chk_ent, tend, entries, tab, tcd, t, nloc);
    double aec[2][NC];
    {
      const int N[4] = {(int)(ab & 0xffff), (int)(ab >> 16), (int)(cd & 0xffff),
                        (int)(cd >> 16)};
      double X[4][3];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[k][j] = sX[j * CAP + N[k]];
      double ba[3], ca[3], da[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        ba[j] = X[1][j] - X[0][j];
        ca[j] = X[2][j] - X[0][j];
        da[j] = X[3][j] - X[0][j];
      }
      double g1[3], g2[3], g3[3];
      cross(ca, da, g1);
      cross(da, ba, g2);
      cross(ba, ca, g3);
      const double J = ba[0] * g1[0] + ba[1] * g1[1] + ba[2] * g1[2];
      // this thread's corners: (0, 1) or (2, 3)
      double gA[3], gB[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        gA[j] = h ? g2[j] : -g1[j] - g2[j] - g3[j];
        gB[j] = h ? g3[j] : g1[j];
      }
      const int nA = h ? N[2] : N[0], nB = h ? N[3] : N[1];
      // steady: the mean of the corners' local time steps
      const double dte =
          dtp ? (((sD[N[0]] + sD[N[1]]) + sD[N[2]]) + sD[N[3]]) / 4.0 : dt;
      const double coef = dte * xg::rcp(J) * 0.5;
      // the two corners' shares of the half-step state (Kozak.cpp:92-124)
      double hs[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) hs[c] = 0.0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int na = k ? nB : nA;
        const double* ga = k ? gB : gA;
        double u[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) u[c] = sU[c * CAP + na];
        const double ir = xg::rcp(u[0]);
        const double p = (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * ir) * gm1;
        double hh[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) hh[c] = 0.25 * u[c];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double cg = coef * ga[j];
          const double uj = u[j + 1] * ir;
          hh[0] -= cg * u[j + 1];
          hh[1] -= cg * u[1] * uj;
          hh[2] -= cg * u[2] * uj;
          hh[3] -= cg * u[3] * uj;
          hh[j + 1] -= cg * p;
          hh[4] -= cg * (u[4] + p) * uj;
#pragma unroll
          for (int c = 5; c < NC; ++c) hh[c] -= cg * u[c] * uj;
        }
        if constexpr (SRC) {
          double sa[5];
          xg::source(P, sX[0 * CAP + na], sX[1 * CAP + na], sX[2 * CAP + na], sa);
          const double c8 = dte / 8.0;
#pragma unroll
          for (int c = 0; c < 5; ++c) hh[c] += c8 * sa[c];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) hs[c] = k ? hs[c] + hh[c] : hh[c];
      }
      // the partner holds the other two corners' sum
      double ue[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) ue[c] = hs[c] + quad_xor1(hs[c]);
      // second half step (Kozak.cpp:138-161): the two corners' fluxes
      const double ie = xg::rcp(ue[0]);
      const double pr = (ue[4] - 0.5 * (ue[1] * ue[1] + ue[2] * ue[2] + ue[3] * ue[3]) * ie) * gm1;
      double fl[3][NC];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double uj = ue[j + 1] * ie;
        fl[j][0] = ue[j + 1];
        fl[j][1] = ue[1] * uj;
        fl[j][2] = ue[2] * uj;
        fl[j][3] = ue[3] * uj;
        fl[j][j + 1] += pr;
        fl[j][4] = (ue[4] + pr) * uj;
#pragma unroll
        for (int c = 5; c < NC; ++c) fl[j][c] = ue[c] * uj;
      }
      double se[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      double cj = 0.0;
      if constexpr (SRC) {
        xg::source(P, (X[0][0] + X[1][0] + X[2][0] + X[3][0]) / 4.0,
                   (X[0][1] + X[1][1] + X[2][1] + X[3][1]) / 4.0,
                   (X[0][2] + X[1][2] + X[2][2] + X[3][2]) / 4.0, se);
        cj = J / 24.0;
      }
      if (i < C.z) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const double* ga = k ? gB : gA;
          const int slot = 4 * i + 2 * h + k;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            double f = (ga[0] * (1.0 / 6.0)) * fl[0][c] + (ga[1] * (1.0 / 6.0)) * fl[1][c] +
                       (ga[2] * (1.0 / 6.0)) * fl[2][c];
            if constexpr (SRC) {
              if (c < 5) f += cj * se[c];
            }
            F[slot * NC + c] = f;
          }
        }
      }
      if constexpr (AEC) pair_aec<NC>(X, sU, CAP, N, h, ctau, aec);
    }
    __syncthreads();          // F, sE, sT complete
    if (t < nloc) {
      int e0 = t ? sT[t - 1] : 0, e1 = sT[t];
      for (int j = e0; j < e1; ++j) {
        const double* fs = F + (int)sE[j] * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += fs[c];
      }
    }
    if constexpr (AEC) {
      __syncthreads();        // the fluxes are summed: the slots take the aec
      if (i < C.z) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int c = 0; c < NC; ++c) F[(4 * i + 2 * h + k) * NC + c] = aec[k][c];
      }
      __syncthreads();
      if (t < nloc) {
        int e0 = t ? sT[t - 1] : 0, e1 = sT[t];
        for (int j = e0; j < e1; ++j) {
          const double* fs = F + (int)sE[j] * NC;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            // std::max(0.0, aec), std::min(0.0, aec) up to the sign of a
            // zero, which no later sum or comparison can tell apart
            const double v = fs[c];
            acc[NC + c] += fmax(v, 0.0);
            acc[2 * NC + c] += fmin(v, 0.0);
          }
        }
      }
    }
  }
  if (t < nown) {
    double* o = R + (size_t)(own0 + t) * NO;
#pragma unroll
    for (int c = 0; c < NO; ++c) o[c] = acc[c];
  } else if (t < nloc) {
    double* o = S + (size_t)(B1.x + t - nown) * NO;
#pragma unroll
    for (int c = 0; c < NO; ++c) o[c] = acc[c];
  }
}

// ---------------------------------------------------------------- FCT --
// The KozCG step after the rhs (KozCG.cpp:866-1197).  W holds per node the
// summed rhs and P+/- (stride 3 NC for NC components); the kernels below
// follow the reference's arithmetic order without fused multiply-adds
// wherever a comparison (limiter branch, bound) reads the result.

// std::max / std::min: which argument a tie or a NaN returns
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }

// low-order solution of node g (KozCG.cpp:883-899; dt the node's own in a
// steady run)
template <int NC>
__device__ inline void low_order(const double* __restrict__ W,
                                 const double* __restrict__ U, double vol,
                                 double dt, int g, double ul[NC]) {
#pragma clang fp contract(off)
  const double* w = W + (size_t)g * (3 * NC);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double pp = w[NC + c] / vol, pn = w[2 * NC + c] / vol;
    ul[c] = U[(size_t)g * NC + c] + dt * w[c] / vol - pp - pn;
  }
}

__device__ inline int local_to_global(int t, int own0, int nown,
                                      const int* __restrict__ halo, int hoff) {
  return t < nown ? own0 + t : halo[hoff + t - nown];
}

// Allowed limits Q+/- (KozCG.cpp:901-935): per tet the max/min over its
// corners of the low-order solution (with CLIP = false also of U), taken
// into every corner's running max/min.  Each local node's own bound pair
// (max and min of its low-order solution and, without clipping, its U) is
// staged in LDS once.  Per chunk of the pair layout, thread t takes the
// max/min over its two corners, one lane exchange with its partner gives the
// tet's, and the pair writes the tet's bounds (10 doubles) to LDS; then the
// thread owning local node t merges the bounds of the chunk's tets that
// touch it (its entry list) into its running max/min in registers -- max
// and min are exact and independent of order, so the result is
// deterministic.  Halo partials go to slots that k_fixup_minmax merges.
// Owners also store the low-order solution UL.  NC components; dtp
// (nullable) the nodes' local time steps of a steady run.
template <bool CLIP, int NC>
__global__ __launch_bounds__(kKozThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_kozcg_alw(
    int nblk, const int4* __restrict__ blk0, const int4* __restrict__ blk1,
    const int* __restrict__ halo, const int4* __restrict__ chk,
    const int* __restrict__ chk_ent, const uint16_t* __restrict__ tend,
    const uint16_t* __restrict__ entries, const uint32_t* __restrict__ tab,
    const uint32_t* __restrict__ tcd, const double* __restrict__ W,
    const double* __restrict__ U, const double* __restrict__ vol, double dt,
    const double* __restrict__ dtp, double* __restrict__ UL, double* __restrict__ Q,
    double* __restrict__ S) {
  constexpr int CAP = xg::kKozCap;
  constexpr int NT = xg::kTetsPerPairChunk;
  constexpr int NB = 2 * NC;
  constexpr double kLarge = 1.7976931348623157e308;
  __shared__ double sB[NB * CAP];    // each node's own max (NC) and min (NC)
  __shared__ double sQ[NT * NB];     // the chunk's tets' max (NC) and min (NC)
  __shared__ uint16_t sE[xg::kKozPairSlots];
  __shared__ uint16_t sT[CAP];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 B0 = blk0[b];
  const int4 B1 = blk1[b];
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  const int i = t >> 1, h = t & 1;
  PairPrefetch pf;
  pf.load(B1.y, B1.z, chk, chk_ent, tend, entries, tab, tcd, t, nloc);
  if (t < nloc) {
    const int g = local_to_global(t, own0, nown, halo, B0.w);
    double ul[NC];
    low_order<NC>(W, U, vol[g], dtp ? dtp[g] : dt, g, ul);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double hi = ul[c], lo = ul[c];
      if (!CLIP) {
        const double u = U[(size_t)g * NC + c];
        hi = smax(hi, u);
        lo = smin(lo, u);
      }
      sB[c * CAP + t] = hi;
      sB[(NC + c) * CAP + t] = lo;
      if (t < nown) UL[(size_t)g * NC + c] = ul[c];
    }
  }
  double qh[NC], ql[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    qh[c] = -kLarge;
    ql[c] = kLarge;
  }
  for (int ch = B1.y; ch < B1.z; ++ch) {
    const int4 C = pf.C;
    const uint32_t w = h ? pf.cd : pf.ab;
    __syncthreads();          // staging done / previous chunk's sQ consumed
    pf.store(t, nloc, sE, sT);
    pf.load(ch + 1, B1.z, chk, chk_ent, tend, entries, tab, tcd, t, nloc);
    const int nA = (int)(w & 0xffff), nB = (int)(w >> 16);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double hi = fmax(sB[c * CAP + nA], sB[c * CAP + nB]);
      double lo = fmin(sB[(NC + c) * CAP + nA], sB[(NC + c) * CAP + nB]);
      hi = fmax(hi, quad_xor1(hi));
      lo = fmin(lo, quad_xor1(lo));
      if (i < C.z) sQ[i * NB + NC * h + c] = h ? lo : hi;
    }
    __syncthreads();          // sQ, sE, sT complete
    if (t < nloc) {
      int e0 = t ? sT[t - 1] : 0, e1 = sT[t];
      for (int j = e0; j < e1; ++j) {
        const double* q = sQ + ((int)sE[j] >> 2) * NB;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          qh[c] = fmax(qh[c], q[c]);
          ql[c] = fmin(ql[c], q[NC + c]);
        }
      }
    }
  }
  if (t < nloc) {
    double* o = t < nown ? Q + (size_t)(own0 + t) * NB
                         : S + (size_t)(B1.x + t - nown) * NB;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      o[c] = qh[c];
      o[NC + c] = ql[c];
    }
  }
}

// merges halo bounds into their owners (max for Q+, min for Q-), NC
// components
template <int NC>
__global__ __launch_bounds__(kBlock) void k_fixup_minmax(
    int nfx, const int* __restrict__ fx_node, const int* __restrict__ fx_off,
    const int* __restrict__ fx_slot, const double* __restrict__ S,
    double* __restrict__ Q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfx) return;
  double* q = Q + (size_t)fx_node[i] * (2 * NC);
  double v[2 * NC];
#pragma unroll
  for (int c = 0; c < 2 * NC; ++c) v[c] = q[c];
  for (int k = fx_off[i]; k < fx_off[i + 1]; ++k) {
    const double* s = S + (size_t)fx_slot[k] * (2 * NC);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      v[c] = smax(v[c], s[c]);
      v[NC + c] = smin(v[NC + c], s[NC + c]);
    }
  }
#pragma unroll
  for (int c = 0; c < 2 * NC; ++c) q[c] = v[c];
}

// limit coefficients C+/- (KozCG.cpp:1005-1024), one thread per node and
// component (neighbouring threads read neighbouring words)
template <int NC>
__global__ __launch_bounds__(kBlock) void k_kozcg_coef(
    int npoin, const double* __restrict__ W, const double* __restrict__ vol,
    const double* __restrict__ UL, const double* __restrict__ Q,
    double* __restrict__ Cf) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)npoin * NC) return;
  const int64_t g = k / NC;
  const int c = (int)(k - g * NC);
  constexpr double eps = 2.220446049250313e-16;
  const double v = vol[g];
  const double pp = W[g * (3 * NC) + NC + c] / v, pn = W[g * (3 * NC) + 2 * NC + c] / v;
  const double ul = UL[k];
  const double qp = Q[g * (2 * NC) + c] - ul, qn = Q[g * (2 * NC) + NC + c] - ul;
  Cf[g * (2 * NC) + c] = pp < eps ? 0.0 : smin(1.0, qp / pp);
  Cf[g * (2 * NC) + NC + c] = pn > -eps ? 0.0 : smin(1.0, qn / pn);
}

// limited antidiffusive contributions (KozCG.cpp:1049-1078): per tet the
// coefficient of each component is the min over the corners of C+ or C-
// (by the sign of the corner's contribution), the fctsys components share
// their min; the corner adds coef * aec.  Pair layout as kozak::rhs: thread
// t recomputes the antidiffusive contributions of its two corners, takes
// the min over them and, with one lane exchange, over its partner's; the
// corner slots are summed per node as in the rhs.  NC components.
template <int NC>
__global__ __launch_bounds__(kKozThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_kozcg_lim(
    int nblk, const int4* __restrict__ blk0, const int4* __restrict__ blk1,
    const int* __restrict__ halo, const int4* __restrict__ chk,
    const int* __restrict__ chk_ent, const uint16_t* __restrict__ tend,
    const uint16_t* __restrict__ entries, const uint32_t* __restrict__ tab,
    const uint32_t* __restrict__ tcd, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ U, const double* __restrict__ Cf,
    double* __restrict__ A, double* __restrict__ S, double ctau, int sysmask) {
  constexpr int CAP = xg::kKozCap;
  constexpr int NS = xg::kKozPairSlots;
  constexpr int NB = 2 * NC;
  __shared__ double sU[NC * CAP];
  __shared__ double sX[3 * CAP];
  __shared__ double F[NS * NC];
  __shared__ int sG[CAP];
  __shared__ uint16_t sE[NS];
  __shared__ uint16_t sT[CAP];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 B0 = blk0[b];
  const int4 B1 = blk1[b];
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  if (t < nloc) {
    const int g = local_to_global(t, own0, nown, halo, B0.w);
    sG[t] = g;
#pragma unroll
    for (int c = 0; c < NC; ++c) sU[c * CAP + t] = U[(size_t)g * NC + c];
    sX[0 * CAP + t] = x[g];
    sX[1 * CAP + t] = y[g];
    sX[2 * CAP + t] = z[g];
  }
  const int i = t >> 1, h = t & 1;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  PairPrefetch pf;
  pf.load(B1.y, B1.z, chk, chk_ent, tend, entries, tab, tcd, t, nloc);
  __syncthreads();            // staging done
  // the two corner nodes' limit coefficients C+ (NC) and C- (NC): the next
  // chunk's are loaded as soon as the current chunk's coefficients are
  // formed, so the loads overlap the slot writes and the accumulation
  double cfn[2][NB];
  auto fetch_cf = [&](const PairPrefetch& Pf) {
    const uint32_t w = h ? Pf.cd : Pf.ab;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int na = k ? (int)(w >> 16) : (int)(w & 0xffff);
      const double* cf = Cf + (size_t)sG[na] * NB;
#pragma unroll
      for (int c = 0; c < NB; ++c) cfn[k][c] = cf[c];
    }
  };
  if (B1.y < B1.z) fetch_cf(pf);
  for (int ch = B1.y; ch < B1.z; ++ch) {
    const int4 C = pf.C;
    const uint32_t ab = pf.ab, cd = pf.cd;
    __syncthreads();          // previous chunk's F consumed
    pf.store(t, nloc, sE, sT);
    pf.load(ch + 1, B1.z, chk, chk_ent, tend, entries, tab, tcd, t, nloc);
    {
      const int N[4] = {(int)(ab & 0xffff), (int)(ab >> 16), (int)(cd & 0xffff),
                        (int)(cd >> 16)};
      double X[4][3];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) X[k][j] = sX[j * CAP + N[k]];
      double aec[2][NC], coef[NC];
      pair_aec<NC>(X, sU, CAP, N, h, ctau, aec);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double qa = aec[0][c] > 0.0 ? cfn[0][c] : cfn[0][NC + c];
        const double qb = aec[1][c] > 0.0 ? cfn[1][c] : cfn[1][NC + c];
        // min over the four corners, starting from 1 (exact, order-free)
        double q = fmin(qa, qb);
        q = fmin(q, quad_xor1(q));
        coef[c] = smin(1.0, q);
      }
      if (ch + 1 < B1.z) fetch_cf(pf);
      if (sysmask) {
        double cs = 1.0;
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (sysmask >> c & 1) cs = smin(cs, coef[c]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (sysmask >> c & 1) coef[c] = cs;
      }
      if (i < C.z) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int c = 0; c < NC; ++c)
            F[(4 * i + 2 * h + k) * NC + c] = coef[c] * aec[k][c];
      }
    }
    __syncthreads();
    if (t < nloc) {
      int e0 = t ? sT[t - 1] : 0, e1 = sT[t];
      for (int j = e0; j < e1; ++j) {
        const double* fs = F + (int)sE[j] * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += fs[c];
      }
    }
  }
  if (t < nown) {
    double* o = A + (size_t)(own0 + t) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = acc[c];
  } else if (t < nloc) {
    double* o = S + (size_t)(B1.x + t - nown) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = acc[c];
  }
}

// new solution (KozCG.cpp:1144-1160), nc components per node; Un receives
// the old one.  dtp (nullable): the nodes' local time steps without FCT.
template <bool FCT>
__global__ __launch_bounds__(kBlock) void k_kozcg_solve(
    int npoin, int nc, double dt, const double* __restrict__ dtp,
    const double* __restrict__ UL, const double* __restrict__ A,
    const double* __restrict__ R, const double* __restrict__ vol,
    double* __restrict__ U, double* __restrict__ Un) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)npoin * nc) return;
  const int64_t g = k / nc;
  const double u = U[k];
  Un[k] = u;
  if (FCT)
    U[k] = UL[k] + A[k] / vol[g];
  else
    U[k] = u + (dtp ? dtp[g] : dt) * R[k] / vol[g];
}

// Dirichlet BCs (physics::dirbc, BC.cpp:30-69): masked components of the
// listed nodes take their values (nc per node)
__global__ __launch_bounds__(kBlock) void k_dirbc(
    int nd, int nc, const int* __restrict__ node, const uint8_t* __restrict__ mask,
    const double* __restrict__ val, double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  double* u = U + (size_t)node[i] * nc;
  for (int c = 0; c < nc; ++c)
    if (mask[i * nc + c]) u[c] = val[i * nc + c];
}

// freezeflow (KozCG.cpp:1168-1175): the flow components (0..4) of U take
// their values of the step's start
__global__ __launch_bounds__(kBlock) void k_freeze_flow(
    int npoin, int nc, const double* __restrict__ Un, double* __restrict__ U) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)npoin * 5) return;
  const int64_t g = k / 5;
  const int c = (int)(k - g * 5);
  U[g * nc + c] = Un[g * nc + c];
}

void free_kozak(KozakMesh* k) {
  if (!k) return;
  dev_free_all(k->step_allocs);
  dev_free_all(k->allocs);
  delete k;
}

xg::Prob prob_of(const xg_ctx* c) {
  return xg::Prob{c->problem.kind, c->problem.alpha, c->problem.kappa, c->problem.p0,
                  c->cfg.gamma};
}

// the pair kernel with or without the antidiffusive contributions (ctau
// their coefficient) and the problem's source term; out and slots take 3 NC
// or NC values per node; dtp (nullable) the local time steps
template <bool AEC, int NC>
void launch_kozak_pairs(xg_ctx* c, const double* U, double* out, double* slots,
                        double dt, const double* dtp, double ctau) {
  KozakMesh* k = c->koz;
  const xg::Prob P = prob_of(c);
  auto kern =
      xg::has_source(P) ? k_kozak_pairs<true, AEC, NC> : k_kozak_pairs<false, AEC, NC>;
  hipLaunchKernelGGL(kern, dim3(k->nblk), dim3(kKozThreads), 0, c->stream, k->nblk,
                     k->blk0, k->blk1, k->halo, k->chk, k->chk_ent, k->tend, k->entries,
                     k->tab, k->tcd, k->x, k->y, k->z, U, out, slots,
                     c->cfg.gamma - 1.0, dt, dtp, ctau, P);
}

template <int NC>
void launch_kozak_nc(xg_ctx* c, double dt, const double* dtp, const double* U,
                     double* R) {
  KozakMesh* k = c->koz;
  if (k->nblk) launch_kozak_pairs<false, NC>(c, U, R, k->slots, dt, dtp, 0.0);
  if (k->nfx)
    hipLaunchKernelGGL(k_fixup<NC>, dim3(nblocks(k->nfx)), dim3(kBlock), 0,
                       c->stream, k->nfx, k->fx_node, k->fx_off, k->fx_slot,
                       k->slots, R);
}

// kozak::rhs with nc components per node (checked by the caller)
int launch_kozak(xg_ctx* c, int nc, double dt, const double* dtp, const double* U,
                 double* R) {
  static_assert(XG_MAX_NCOMP == 7, "one instantiation per component count");
  if (nc == 5) launch_kozak_nc<5>(c, dt, dtp, U, R);
  else if (nc == 6) launch_kozak_nc<6>(c, dt, dtp, U, R);
  else launch_kozak_nc<7>(c, dt, dtp, U, R);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int check_kozak(xg_ctx* c, int ncomp) {
  if (!c) return XG_EINVAL;
  if (!c->koz) return fail(c, XG_ENOMESH, "no tet mesh (xg_kozak_upload)");
  if (ncomp < 5 || ncomp > XG_MAX_NCOMP)
    return fail(c, XG_EUNSUPPORTED,
                "kozak::rhs on the device: ncomp = 5 + at most " +
                    std::to_string(XG_MAX_NCOMP - 5) + " transported scalars");
  return XG_OK;
}

// One KozCG step (KozCG::rhs ... solve, BC) for NC components; dtp, tp
// (nullable) local time steps, freeze: the flow is frozen
template <int NC>
void kozcg_step_nc(xg_ctx* c, double dt, const double* dtp, double* tp, int freeze,
                   double* U, double* Un) {
  KozakMesh* k = c->koz;
  const int n = k->npoin;
  const int64_t nn = (int64_t)n * NC;
  const auto& f = k->fct;
  if (!f.fct) {
    if (k->nblk) launch_kozak_pairs<false, NC>(c, U, k->W, k->S15, dt, dtp, 0.0);
    if (k->nfx)
      hipLaunchKernelGGL(k_fixup<NC>, dim3(nblocks(k->nfx)), dim3(kBlock), 0,
                         c->stream, k->nfx, k->fx_node, k->fx_off, k->fx_slot,
                         k->S15, k->W);
    hipLaunchKernelGGL(k_kozcg_solve<false>, dim3(nblocks(nn)), dim3(kBlock), 0,
                       c->stream, n, NC, dt, dtp, nullptr, nullptr, k->W, k->vol, U, Un);
  } else {
    int sysmask = 0;
    for (int s = 0; s < f.nsys; ++s) sysmask |= 1 << f.sys[s];
    // rhs and P+/- (KozCG::rhs, aec)
    if (k->nblk) launch_kozak_pairs<true, NC>(c, U, k->W, k->S15, dt, dtp, f.fctdif);
    if (k->nfx)
      hipLaunchKernelGGL(k_fixup<3 * NC>, dim3(nblocks(k->nfx)), dim3(kBlock), 0,
                         c->stream, k->nfx, k->fx_node, k->fx_off, k->fx_slot,
                         k->S15, k->W);
    // symmetry BCs on the antidiffusive contributions (KozCG.cpp:816-830)
    if (k->nsymnode) {
      hipLaunchKernelGGL(k_symbc, dim3(nblocks(k->nsymnode)), dim3(kBlock), 0,
                         c->stream, k->nsymnode, k->symnode, k->symoff,
                         k->symnorm, k->W, 3 * NC, NC + 1);
      hipLaunchKernelGGL(k_symbc, dim3(nblocks(k->nsymnode)), dim3(kBlock), 0,
                         c->stream, k->nsymnode, k->symnode, k->symoff,
                         k->symnorm, k->W, 3 * NC, 2 * NC + 1);
    }
    // low-order solution and allowed bounds (KozCG::alw)
    if (k->nblk) {
      auto alw = f.fctclip ? k_kozcg_alw<true, NC> : k_kozcg_alw<false, NC>;
      hipLaunchKernelGGL(alw, dim3(k->nblk), dim3(kKozThreads), 0, c->stream, k->nblk,
                         k->blk0, k->blk1, k->halo, k->chk, k->chk_ent, k->tend,
                         k->entries, k->tab, k->tcd, k->W, U, k->vol, dt, dtp, k->UL,
                         k->Q, k->S15);
    }
    if (k->nfx)
      hipLaunchKernelGGL(k_fixup_minmax<NC>, dim3(nblocks(k->nfx)), dim3(kBlock), 0,
                         c->stream, k->nfx, k->fx_node, k->fx_off, k->fx_slot,
                         k->S15, k->Q);
    // limit coefficients and limited contributions (KozCG::lim)
    hipLaunchKernelGGL(k_kozcg_coef<NC>, dim3(nblocks(nn)), dim3(kBlock), 0,
                       c->stream, n, k->W, k->vol, k->UL, k->Q, k->C);
    if (k->nblk)
      hipLaunchKernelGGL(k_kozcg_lim<NC>, dim3(k->nblk), dim3(kKozThreads), 0,
                         c->stream, k->nblk, k->blk0, k->blk1, k->halo, k->chk,
                         k->chk_ent, k->tend, k->entries, k->tab, k->tcd, k->x,
                         k->y, k->z, U, k->C, k->A, k->S15, f.fctdif, sysmask);
    if (k->nfx)
      hipLaunchKernelGGL(k_fixup<NC>, dim3(nblocks(k->nfx)), dim3(kBlock), 0,
                         c->stream, k->nfx, k->fx_node, k->fx_off, k->fx_slot,
                         k->S15, k->A);
    // new solution (KozCG::solve)
    hipLaunchKernelGGL(k_kozcg_solve<true>, dim3(nblocks(nn)), dim3(kBlock), 0,
                       c->stream, n, NC, dt, nullptr, k->UL, k->A, nullptr, k->vol, U,
                       Un);
  }
  // KozCG::BC: Dirichlet, then symmetry
  if (k->ndir)
    hipLaunchKernelGGL(k_dirbc, dim3(nblocks(k->ndir)), dim3(kBlock), 0, c->stream,
                       k->ndir, NC, k->dirnode, k->dirmask, k->dirval, U);
  if (k->nsymnode)
    hipLaunchKernelGGL(k_symbc, dim3(nblocks(k->nsymnode)), dim3(kBlock), 0,
                       c->stream, k->nsymnode, k->symnode, k->symoff, k->symnorm,
                       U, NC, 1);
  if (freeze)
    hipLaunchKernelGGL(k_freeze_flow, dim3(nblocks((int64_t)n * 5)), dim3(kBlock), 0,
                       c->stream, n, NC, Un, U);
  // local time stepping: tp += dtp (KozCG.cpp:1189-1193)
  if (dtp && tp)
    hipLaunchKernelGGL(k_axpy1, dim3(nblocks(n)), dim3(kBlock), 0, c->stream,
                       (int64_t)n, dtp, tp);
}

}  // namespace

void xg_kozak_free(xg_ctx* c) {
  free_kozak(c->koz);
  c->koz = nullptr;
}

extern "C" {

int xg_kozak_upload(xg_ctx* c, int64_t npoin, const double* x, const double* y,
                    const double* z, int64_t ntet, const size_t* inpoel) {
  if (!c) return XG_EINVAL;
  if (npoin <= 0 || npoin >= (1ll << 31) || !x || !y || !z || ntet < 0 ||
      (ntet && !inpoel))
    return fail(c, XG_EINVAL, "xg_kozak_upload: bad arguments");
  XG_HIP(c, hipSetDevice(c->device));
  xg::TetLayout L;
  int rc = xg::build_tet_layout((int)npoin, ntet, inpoel, xg::kKozNodes, xg::kKozCap,
                                L, c->err, xg::kTetsPerPairChunk);
  if (rc) return rc;
  auto* k = new KozakMesh;
  k->npoin = (int)npoin;
  k->ntet = ntet;
  k->nblk = (int)L.blk.size();
  std::vector<int4> b0(k->nblk), b1(k->nblk);
  for (int b = 0; b < k->nblk; ++b) {
    const auto& B = L.blk[b];
    b0[b] = make_int4(B.own0, B.nown, B.nloc, B.halo_off);
    b1[b] = make_int4(B.slot0, B.chunk0, B.chunk1, 0);
  }
  std::vector<int4> ch(L.chunks.size());
  std::vector<int> ce(L.chunks.size());
  for (size_t i = 0; i < L.chunks.size(); ++i) {
    const auto& C = L.chunks[i];
    ch[i] = make_int4(C.kind, C.first, C.count, C.tend_off);
    ce[i] = C.ent_off;
  }
  auto& A = k->allocs;
  const size_t n = (size_t)npoin;
  if ((rc = upload(c, A, &k->x, x, n)) || (rc = upload(c, A, &k->y, y, n)) ||
      (rc = upload(c, A, &k->z, z, n)) || (rc = upload(c, A, &k->blk0, b0)) ||
      (rc = upload(c, A, &k->blk1, b1)) || (rc = upload(c, A, &k->halo, L.halo)) ||
      (rc = upload(c, A, &k->chk, ch)) || (rc = upload(c, A, &k->chk_ent, ce)) ||
      (rc = upload(c, A, &k->tend, L.tend)) ||
      (rc = upload(c, A, &k->entries, L.entries)) ||
      (rc = upload(c, A, &k->tab, L.tab)) || (rc = upload(c, A, &k->tcd, L.tcd)) ||
      (rc = dev_alloc(c, A, (void**)&k->slots,
                      (size_t)L.nslot * XG_MAX_NCOMP * sizeof(double))) ||
      (rc = upload(c, A, &k->fx_node, L.fx_node)) ||
      (rc = upload(c, A, &k->fx_off, L.fx_off)) ||
      (rc = upload(c, A, &k->fx_slot, L.fx_slot))) {
    free_kozak(k);
    return rc;
  }
  k->nfx = (int)L.fx_node.size();
  k->nslot = L.nslot;
  free_kozak(c->koz);
  c->koz = k;
  return XG_OK;
}

int xg_kozak_rhs_dev(xg_ctx* c, int ncomp, double t, double dt, const double* tp,
                     const double* dtp, const double* U, double* R) {
  (void)t;    // the sources served are independent of time,
  (void)tp;   // so neither t nor the steady per-node times change them
  int rc = check_kozak(c, ncomp);
  if (rc) return rc;
  if (!U || !R) return fail(c, XG_EINVAL, "xg_kozak_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_kozak(c, ncomp, dt, dtp, U, R);
}

int xg_kozak_rhs(xg_ctx* c, int ncomp, double t, double dt, const double* tp,
                 const double* dtp, const double* U, double* R) {
  (void)t;
  (void)tp;
  int rc = check_kozak(c, ncomp);
  if (rc) return rc;
  if (!U || !R) return fail(c, XG_EINVAL, "xg_kozak_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  KozakMesh* k = c->koz;
  const size_t n = (size_t)k->npoin;
  if (!k->sU) {
    const size_t most = n * XG_MAX_NCOMP * sizeof(double);
    if ((rc = dev_alloc(c, k->allocs, (void**)&k->sU, most)) ||
        (rc = dev_alloc(c, k->allocs, (void**)&k->sR, most)) ||
        (rc = dev_alloc(c, k->allocs, (void**)&k->sD, n * sizeof(double))))
      return rc;
  }
  const size_t bytes = n * (size_t)ncomp * sizeof(double);
  XG_HIP(c, hipMemcpyAsync(k->sU, U, bytes, hipMemcpyHostToDevice, c->stream));
  if (dtp)
    XG_HIP(c, hipMemcpyAsync(k->sD, dtp, n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  if ((rc = launch_kozak(c, ncomp, dt, dtp ? k->sD : nullptr, k->sU, k->sR))) return rc;
  XG_HIP(c, hipMemcpyAsync(R, k->sR, bytes, hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

int xg_kozcg_set_ncomp(xg_ctx* c, int ncomp) {
  if (!c) return XG_EINVAL;
  if (!c->koz) return fail(c, XG_ENOMESH, "no tet mesh (xg_kozak_upload)");
  int rc = check_kozak(c, ncomp);
  if (rc) return rc;
  if (ncomp != c->koz->ncomp) {
    // the step's buffers and Dirichlet lists are per component: setup again
    XG_HIP(c, hipSetDevice(c->device));
    (void)hipStreamSynchronize(c->stream);
    dev_free_all(c->koz->step_allocs);
    c->koz->have_step = false;
  }
  c->koz->ncomp = ncomp;
  return XG_OK;
}

int xg_kozcg_setup(xg_ctx* c, const double* vol, int64_t nsym,
                   const size_t* symnodes, const double* symnorms, int64_t ndir,
                   const size_t* dirnodes, const int* dirmask,
                   const double* dirvals, const xg_fctcfg* fct) {
  if (!c) return XG_EINVAL;
  if (!c->koz) return fail(c, XG_ENOMESH, "no tet mesh (xg_kozak_upload)");
  KozakMesh* k = c->koz;
  if (!vol || nsym < 0 || (nsym && (!symnodes || !symnorms)) || ndir < 0 ||
      (ndir && (!dirnodes || !dirmask || !dirvals)) || !fct)
    return fail(c, XG_EINVAL, "xg_kozcg_setup: bad arguments");
  const int nc = k->ncomp;
  if (fct->nsys < 0 || fct->nsys > 5)
    return fail(c, XG_EINVAL, "xg_kozcg_setup: fctsys holds at most 5 components");
  for (int s = 0; s < fct->nsys; ++s)
    if (fct->sys[s] < 0 || fct->sys[s] >= nc)
      return fail(c, XG_EINVAL, "xg_kozcg_setup: fctsys component out of range");
  for (int64_t i = 0; i < k->npoin; ++i)
    if (!(vol[i] > 0)) return fail(c, XG_EINVAL, "xg_kozcg_setup: non-positive nodal volume");
  // symmetry (node, normal) pairs grouped by node, list order kept per node
  std::vector<int> node, off(1, 0);
  std::vector<double> nrm;
  {
    std::vector<int64_t> order(nsym);
    for (int64_t i = 0; i < nsym; ++i) {
      if (symnodes[i] >= (size_t)k->npoin)
        return fail(c, XG_EINVAL, "xg_kozcg_setup: symmetry node out of range");
      order[i] = i;
    }
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
  std::vector<int> dn(ndir);
  std::vector<uint8_t> dm((size_t)ndir * nc);
  for (int64_t i = 0; i < ndir; ++i) {
    if (dirnodes[i] >= (size_t)k->npoin)
      return fail(c, XG_EINVAL, "xg_kozcg_setup: Dirichlet node out of range");
    dn[i] = (int)dirnodes[i];
    for (int q = 0; q < nc; ++q) dm[i * nc + q] = dirmask[i * nc + q] ? 1 : 0;
  }
  XG_HIP(c, hipSetDevice(c->device));
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(k->step_allocs);
  k->have_step = false;
  const size_t n = (size_t)k->npoin;
  const size_t nsl = std::max<size_t>(1, (size_t)k->nslot);
  auto& L = k->step_allocs;
  int rc;
  if ((rc = upload(c, L, &k->vol, vol, n)) || (rc = upload(c, L, &k->symnode, node)) ||
      (rc = upload(c, L, &k->symoff, off)) || (rc = upload(c, L, &k->symnorm, nrm)) ||
      (rc = upload(c, L, &k->dirnode, dn)) || (rc = upload(c, L, &k->dirmask, dm)) ||
      (rc = upload(c, L, &k->dirval, dirvals, (size_t)ndir * nc)) ||
      (rc = dev_alloc(c, L, (void**)&k->W, n * 3 * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&k->UL, n * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&k->Q, n * 2 * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&k->C, n * 2 * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&k->A, n * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&k->S15, nsl * 3 * nc * 8)) ||
      (rc = dev_alloc(c, L, (void**)&k->red, sizeof(double) * (kRedBlocks + 1)))) {
    dev_free_all(k->step_allocs);
    return rc;
  }
  k->nsymnode = (int)node.size();
  k->ndir = (int)ndir;
  k->fct = *fct;
  k->have_step = true;
  return XG_OK;
}

int xg_kozcg_dt_dev(xg_ctx* c, const double* U, double cfl, double* dt) {
  if (!c) return XG_EINVAL;
  if (!c->koz || !c->koz->have_step) return fail(c, XG_ENOMESH, "xg_kozcg_setup not called");
  if (!U || !dt) return fail(c, XG_EINVAL, "xg_kozcg_dt: null argument");
  XG_HIP(c, hipSetDevice(c->device));
  KozakMesh* k = c->koz;
  int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(k->npoin)));
  hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(kBlock), 0, c->stream,
                     k->npoin, k->ncomp, U, k->vol, c->cfg.gamma, k->red);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kRedBlocks), 0, c->stream, nb,
                     k->red, cfl, k->red + kRedBlocks);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipMemcpyAsync(dt, k->red + kRedBlocks, sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (*dt < 2.220446049250313e-16)
    return fail(c, XG_EBLOWUP, "time step below machine epsilon");
  return XG_OK;
}

int xg_kozcg_dt_steady_dev(xg_ctx* c, const double* U, double cfl, double* dtp,
                           double* dt) {
  if (!c) return XG_EINVAL;
  if (!c->koz || !c->koz->have_step) return fail(c, XG_ENOMESH, "xg_kozcg_setup not called");
  if (!U || !dtp || !dt) return fail(c, XG_EINVAL, "xg_kozcg_dt_steady: null argument");
  XG_HIP(c, hipSetDevice(c->device));
  KozakMesh* k = c->koz;
  int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(k->npoin)));
  hipLaunchKernelGGL(k_dt_steady, dim3(nb), dim3(kBlock), 0, c->stream, k->npoin,
                     k->ncomp, U, k->vol, c->cfg.gamma, cfl, dtp, k->red);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kRedBlocks), 0, c->stream, nb,
                     k->red, 1.0, k->red + kRedBlocks);
  XG_HIP(c, hipGetLastError());
  XG_HIP(c, hipMemcpyAsync(dt, k->red + kRedBlocks, sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  if (*dt < 2.220446049250313e-16)
    return fail(c, XG_EBLOWUP, "time step below machine epsilon");
  return XG_OK;
}

int xg_kozcg_step2_dev(xg_ctx* c, double t, double dt, const double* dtp, double* tp,
                       int freeze, double* U, double* Un) {
  (void)t;   // the sources served are independent of time
  if (!c) return XG_EINVAL;
  if (!c->koz || !c->koz->have_step) return fail(c, XG_ENOMESH, "xg_kozcg_setup not called");
  if (!U || !Un || U == Un) return fail(c, XG_EINVAL, "xg_kozcg_step: bad field pointers");
  if ((dtp == nullptr) != (tp == nullptr))
    return fail(c, XG_EINVAL, "xg_kozcg_step: dtp and tp go together");
  XG_HIP(c, hipSetDevice(c->device));
  const int nc = c->koz->ncomp;
  if (nc == 5) kozcg_step_nc<5>(c, dt, dtp, tp, freeze, U, Un);
  else if (nc == 6) kozcg_step_nc<6>(c, dt, dtp, tp, freeze, U, Un);
  else kozcg_step_nc<7>(c, dt, dtp, tp, freeze, U, Un);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int xg_kozcg_step_dev(xg_ctx* c, double t, double dt, double* U, double* Un) {
  return xg_kozcg_step2_dev(c, t, dt, nullptr, nullptr, 0, U, Un);
}

}  // extern "C"
