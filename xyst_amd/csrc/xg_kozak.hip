// MI355X device path of kozak::rhs (Kozak.cpp:29-180): the KozCG
// Taylor-Galerkin right-hand side of one GPU-resident partition.
//
// Layout (xg_layout.cpp, build_tet_layout): node blocks of consecutive ids;
// a block owns the tets whose smallest node it holds.  One 1024-thread
// workgroup per block stages the conserved state and coordinates of the
// block's nodes (owned + halo) in LDS.  Per chunk of 256 tets, thread t
// handles corner a = t % 4 of tet t / 4: it evaluates its node's share of the
// half-step element state, the four corners' shares are combined with two
// lane exchanges (every corner sees the same sum), and the thread writes its
// corner's second-step flux to LDS.  Then the thread owning local node t
// adds the chunk's corner fluxes that touch it, in slot order; halo partials
// go to slots that k_fixup adds to their owners in block order.  No atomics:
// every sum has a fixed order and the result is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "xg_blocks.hpp"
#include "xg_ctx.hpp"
#include "xg_flux.hpp"
#include "xg_layout.hpp"

using namespace xgd;

struct KozakMesh {
  int npoin = 0;
  int64_t ntet = 0;
  int nblk = 0;
  std::vector<DevArr> allocs;
  double *x = nullptr, *y = nullptr, *z = nullptr;
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
};

namespace {

constexpr int kKozThreads = xg::kChunkSlots;
static_assert(xg::kKozCap <= kKozThreads, "one local node per thread");

__device__ inline void cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - b[1] * a[2];
  c[1] = a[2] * b[0] - b[2] * a[0];
  c[2] = a[0] * b[1] - b[0] * a[1];
}

__global__ __launch_bounds__(kKozThreads) void k_kozak_blocks(
    int nblk, const int4* __restrict__ blk0, const int4* __restrict__ blk1,
    const int* __restrict__ halo, const int4* __restrict__ chk,
    const int* __restrict__ chk_ent, const uint16_t* __restrict__ tend,
    const uint16_t* __restrict__ entries, const uint32_t* __restrict__ tab,
    const uint32_t* __restrict__ tcd, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ U, double* __restrict__ R,
    double* __restrict__ S, double gm1, double dt) {
  constexpr int CAP = xg::kKozCap;
  __shared__ double sU[5 * CAP];
  __shared__ double sX[3 * CAP];
  __shared__ double F[kKozThreads * 5];
  __shared__ uint16_t sE[kKozThreads];
  __shared__ uint16_t sT[CAP];
  const int t = threadIdx.x;
  const int b = xcd_block(blockIdx.x, nblk);
  const int4 B0 = blk0[b];   // own0, nown, nloc, halo_off
  const int4 B1 = blk1[b];   // slot0, chunk0, chunk1
  const int own0 = B0.x, nown = B0.y, nloc = B0.z;
  if (t < nloc) {
    int g = t < nown ? own0 + t : halo[B0.w + t - nown];
    double s[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) s[c] = U[(size_t)g * 5 + c];
    double xg_ = x[g], yg = y[g], zg = z[g];
#pragma unroll
    for (int c = 0; c < 5; ++c) sU[c * CAP + t] = s[c];
    sX[0 * CAP + t] = xg_;
    sX[1 * CAP + t] = yg;
    sX[2 * CAP + t] = zg;
  }
  const int i = t >> 2, a = t & 3;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int ch = B1.y; ch < B1.z; ++ch) {
    const int4 C = chk[ch];   // kind, first, count, tend_off
    const int eoff = chk_ent[ch];
    // this chunk's tet (a valid one for idle lanes, so the lane exchanges
    // below run with every lane active)
    const int e = C.y + (i < C.z ? i : 0);
    const uint32_t ab = tab[e], cd = tcd[e];
    uint16_t ent = 0, te = 0;
    if (t < 4 * C.z) ent = entries[eoff + t];
    if (t < nloc) te = tend[C.w + t];
    __syncthreads();          // staging done / previous chunk's F consumed
    if (t < 4 * C.z) sE[t] = ent;
    if (t < nloc) sT[t] = te;
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
      double g1[3], g2[3], g3[3], ga[3];
      cross(ca, da, g1);
      cross(da, ba, g2);
      cross(ba, ca, g3);
      const double J = ba[0] * g1[0] + ba[1] * g1[1] + ba[2] * g1[2];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        ga[j] = a == 0 ? -g1[j] - g2[j] - g3[j] : a == 1 ? g1[j] : a == 2 ? g2[j] : g3[j];
      const int na = N[a];
      double u[5];
#pragma unroll
      for (int c = 0; c < 5; ++c) u[c] = sU[c * CAP + na];
      // this corner's share of the half-step state (Kozak.cpp:92-124)
      const double ir = xg::rcp(u[0]);
      const double p = (u[4] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * ir) * gm1;
      const double coef = dt * xg::rcp(J) * 0.5;
      double h[5];
#pragma unroll
      for (int c = 0; c < 5; ++c) h[c] = 0.25 * u[c];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double cg = coef * ga[j];
        const double uj = u[j + 1] * ir;
        h[0] -= cg * u[j + 1];
        h[1] -= cg * u[1] * uj;
        h[2] -= cg * u[2] * uj;
        h[3] -= cg * u[3] * uj;
        h[j + 1] -= cg * p;
        h[4] -= cg * (u[4] + p) * uj;
      }
      // sum over the tet's four corners (lanes 4i..4i+3); (h0+h1)+(h2+h3) and
      // (h2+h3)+(h0+h1) round identically, so all four lanes agree
      double ue[5];
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double v = h[c] + __shfl_xor(h[c], 1, 64);
        ue[c] = v + __shfl_xor(v, 2, 64);
      }
      // second half step (Kozak.cpp:138-161): corner a's flux
      const double ie = xg::rcp(ue[0]);
      const double pr = (ue[4] - 0.5 * (ue[1] * ue[1] + ue[2] * ue[2] + ue[3] * ue[3]) * ie) * gm1;
      double f[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double uj = ue[j + 1] * ie;
        const double cg = ga[j] * (1.0 / 6.0);
        f[0] += cg * ue[j + 1];
        f[1] += cg * ue[1] * uj;
        f[2] += cg * ue[2] * uj;
        f[3] += cg * ue[3] * uj;
        f[j + 1] += cg * pr;
        f[4] += cg * (ue[4] + pr) * uj;
      }
      if (i < C.z) {
#pragma unroll
        for (int c = 0; c < 5; ++c) F[t * 5 + c] = f[c];
      }
    }
    __syncthreads();          // F, sE, sT complete
    if (t < nloc) {
      int e0 = t ? sT[t - 1] : 0, e1 = sT[t];
      for (int j = e0; j < e1; ++j) {
        const double* fs = F + (int)sE[j] * 5;
#pragma unroll
        for (int c = 0; c < 5; ++c) acc[c] += fs[c];
      }
    }
  }
  if (t < nown) {
    double* o = R + (size_t)(own0 + t) * 5;
#pragma unroll
    for (int c = 0; c < 5; ++c) o[c] = acc[c];
  } else if (t < nloc) {
    double* o = S + (size_t)(B1.x + t - nown) * 5;
#pragma unroll
    for (int c = 0; c < 5; ++c) o[c] = acc[c];
  }
}

void free_kozak(KozakMesh* k) {
  if (!k) return;
  dev_free_all(k->allocs);
  delete k;
}

int launch_kozak(xg_ctx* c, double dt, const double* U, double* R) {
  KozakMesh* k = c->koz;
  if (k->nblk)
    hipLaunchKernelGGL(k_kozak_blocks, dim3(k->nblk), dim3(kKozThreads), 0,
                       c->stream, k->nblk, k->blk0, k->blk1, k->halo, k->chk,
                       k->chk_ent, k->tend, k->entries, k->tab, k->tcd, k->x,
                       k->y, k->z, U, R, k->slots, c->cfg.gamma - 1.0, dt);
  if (k->nfx)
    hipLaunchKernelGGL(k_fixup<5>, dim3(nblocks(k->nfx)), dim3(kBlock), 0,
                       c->stream, k->nfx, k->fx_node, k->fx_off, k->fx_slot,
                       k->slots, R);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int check_kozak(xg_ctx* c, int ncomp) {
  if (!c) return XG_EINVAL;
  if (!c->koz) return fail(c, XG_ENOMESH, "no tet mesh (xg_kozak_upload)");
  if (ncomp != 5)
    return fail(c, XG_EUNSUPPORTED, "kozak::rhs on the device: ncomp must be 5");
  return XG_OK;
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
                                L, c->err);
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
      (rc = dev_alloc(c, A, (void**)&k->slots, (size_t)L.nslot * 40)) ||
      (rc = upload(c, A, &k->fx_node, L.fx_node)) ||
      (rc = upload(c, A, &k->fx_off, L.fx_off)) ||
      (rc = upload(c, A, &k->fx_slot, L.fx_slot))) {
    free_kozak(k);
    return rc;
  }
  k->nfx = (int)L.fx_node.size();
  free_kozak(c->koz);
  c->koz = k;
  return XG_OK;
}

int xg_kozak_rhs_dev(xg_ctx* c, int ncomp, double dt, const double* U, double* R) {
  int rc = check_kozak(c, ncomp);
  if (rc) return rc;
  if (!U || !R) return fail(c, XG_EINVAL, "xg_kozak_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_kozak(c, dt, U, R);
}

int xg_kozak_rhs(xg_ctx* c, int ncomp, double dt, const double* U, double* R) {
  int rc = check_kozak(c, ncomp);
  if (rc) return rc;
  if (!U || !R) return fail(c, XG_EINVAL, "xg_kozak_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  KozakMesh* k = c->koz;
  const size_t bytes = (size_t)k->npoin * 40;
  if (!k->sU) {
    if ((rc = dev_alloc(c, k->allocs, (void**)&k->sU, bytes)) ||
        (rc = dev_alloc(c, k->allocs, (void**)&k->sR, bytes)))
      return rc;
  }
  XG_HIP(c, hipMemcpyAsync(k->sU, U, bytes, hipMemcpyHostToDevice, c->stream));
  if ((rc = launch_kozak(c, dt, k->sU, k->sR))) return rc;
  XG_HIP(c, hipMemcpyAsync(R, k->sR, bytes, hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

}  // extern "C"
