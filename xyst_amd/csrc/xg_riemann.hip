// MI355X (gfx950) device path of the RieCG edge-based Riemann/MUSCL scheme:
// context, device mesh layout, riemann::grad/rhs kernels, RK stage, CFL
// minimum.  All FP64.  Deterministic: no atomics, every nodal sum is taken in
// an order fixed by the layout.
//
// Scatter: edges are greedily coloured on the host so that no two edges of a
// colour share a node; one launch per colour then updates both end nodes with
// plain read-modify-writes.  Boundary-face terms are gathered per node
// (node -> incident boundary faces, in face order) by a separate kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "xg_internal.hpp"

// ----------------------------------------------------------------- context --

struct DevArr {
  void* p = nullptr;
  size_t bytes = 0;
};

struct xg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  xg_config cfg{1.4, 0, 0, 0.0};
  std::string err;
  std::vector<DevArr> mesh_allocs;     // freed on re-upload / destroy
  std::vector<DevArr> riecg_allocs;
  // mesh
  bool have_mesh = false;
  int npoin = 0;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  int64_t nedge = 0;
  int *ep = nullptr, *eq = nullptr;
  double *dx = nullptr, *dy = nullptr, *dz = nullptr;
  std::vector<int64_t> coloff;         // host: edge offsets of each colour
  int nbnode = 0;
  int *bnode = nullptr, *boff = nullptr, *bface = nullptr;
  int *tri = nullptr;
  unsigned char* besym = nullptr;
  int64_t ntri = 0;
  // RieCG stepping
  bool have_riecg = false;
  double* vol = nullptr;
  int nsymnode = 0;
  int *symnode = nullptr, *symoff = nullptr;
  double* symnorm = nullptr;
  double* red = nullptr;               // reduction partials + result
  // staging for the host-pointer entry points
  double *sU = nullptr, *sG = nullptr, *sR = nullptr;
  size_t sbytes = 0;
};

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;

struct Cfg {
  double g;
  int stab2;
  double stab2coef;
};

int fail(xg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define XG_HIP(ctx, call)                                                    \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(ctx, XG_EHIP, std::string(#call) + ": " +                  \
                                    hipGetErrorString(e_));                  \
  } while (0)

int dev_alloc(xg_ctx* c, std::vector<DevArr>& list, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(c, XG_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  list.push_back({*p, bytes});
  return XG_OK;
}

void dev_free_all(std::vector<DevArr>& list) {
  for (auto& a : list) (void)hipFree(a.p);
  list.clear();
}

template <class T>
int upload(xg_ctx* c, std::vector<DevArr>& list, T** dst, const T* src,
           size_t n) {
  int rc = dev_alloc(c, list, (void**)dst, n * sizeof(T));
  if (rc) return rc;
  if (n) XG_HIP(c, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return XG_OK;
}

inline unsigned nblocks(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

// ------------------------------------------------------------ device math --

// riemann::primitive, Riemann.cpp:211-227
__device__ inline void primitive(const double* __restrict__ U, int i,
                                 double u[5]) {
  const double* s = U + (size_t)i * 5;
  u[0] = s[0];
  u[1] = s[1] / u[0];
  u[2] = s[2] / u[0];
  u[3] = s[3] / u[0];
  u[4] = s[4] / u[0] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
}

// MUSCL / van Leer reconstruction of the flow variables, Riemann.cpp:34-143
__device__ inline void muscl(const double vw[3], const double* __restrict__ Gp,
                             const double* __restrict__ Gq, double l[5],
                             double r[5]) {
  constexpr double eps = 1.0e-9;
  constexpr double mc = 1.0 / 3.0;
  double url[5], urr[5], d1[5], d3[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double g1 = Gp[c * 3 + 0] * vw[0] + Gp[c * 3 + 1] * vw[1] + Gp[c * 3 + 2] * vw[2];
    double g2 = Gq[c * 3 + 0] * vw[0] + Gq[c * 3 + 1] * vw[1] + Gq[c * 3 + 2] * vw[2];
    double delta2 = r[c] - l[c];
    double delta1 = 2.0 * g1 - delta2;
    double delta3 = 2.0 * g2 - delta2;
    d1[c] = delta1;
    d3[c] = delta3;
    double rcL = (delta2 + eps) / (delta1 + eps);
    double rcR = (delta2 + eps) / (delta3 + eps);
    double rLinv = (delta1 + eps) / (delta2 + eps);
    double rRinv = (delta3 + eps) / (delta2 + eps);
    double phiL = (fabs(rcL) + rcL) / (fabs(rcL) + 1.0);
    double phiR = (fabs(rcR) + rcR) / (fabs(rcR) + 1.0);
    double phiLi = (fabs(rLinv) + rLinv) / (fabs(rLinv) + 1.0);
    double phiRi = (fabs(rRinv) + rRinv) / (fabs(rRinv) + 1.0);
    url[c] = l[c] + 0.25 * (delta1 * (1.0 - mc) * phiL + delta2 * (1.0 + mc) * phiLi);
    urr[c] = r[c] - 0.25 * (delta3 * (1.0 - mc) * phiR + delta2 * (1.0 + mc) * phiRi);
  }
  // first order where density or internal energy would turn negative
  bool keepL = (l[0] < d1[0] || l[4] < d1[4]);
  bool keepR = (r[0] < -d3[0] || r[4] < -d3[4]);
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    if (!keepL) l[c] = url[c];
    if (!keepR) r[c] = urr[c];
  }
}

// Rusanov flux, Riemann.cpp:369-478 (domain edges: no symmetry masking)
__device__ inline void rusanov(const Cfg& cfg, const double n[3], double l[5],
                               double r[5], double f[5]) {
  const double g = cfg.g;
  double pL = (l[0] * l[4]) * (g - 1.0);
  double pR = (r[0] * r[4]) * (g - 1.0);
  double vnL = l[1] * n[0] + l[2] * n[1] + l[3] * n[2];
  double vnR = r[1] * n[0] + r[2] * n[1] + r[3] * n[2];
  l[4] = (l[4] + 0.5 * (l[1] * l[1] + l[2] * l[2] + l[3] * l[3])) * l[0];
  l[1] *= l[0]; l[2] *= l[0]; l[3] *= l[0];
  r[4] = (r[4] + 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3])) * r[0];
  r[1] *= r[0]; r[2] *= r[0]; r[3] *= r[0];
  double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double sl = fabs(vnL) + sqrt(g * pL / l[0]) * len;
  double sr = fabs(vnR) + sqrt(g * pR / r[0]) * len;
  double fw = fmax(sl, sr);
  f[0] = l[0] * vnL + r[0] * vnR + fw * (r[0] - l[0]);
  f[1] = l[1] * vnL + r[1] * vnR + (pL + pR) * n[0] + fw * (r[1] - l[1]);
  f[2] = l[2] * vnL + r[2] * vnR + (pL + pR) * n[1] + fw * (r[2] - l[2]);
  f[3] = l[3] * vnL + r[3] * vnR + (pL + pR) * n[2] + fw * (r[3] - l[3]);
  f[4] = (l[4] + pL) * vnL + (r[4] + pR) * vnR + fw * (r[4] - l[4]);
  if (cfg.stab2) {
    double fws = cfg.stab2coef * fw;
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c] -= fws * (l[c] - r[c]);
  }
}

// HLLC flux, Riemann.cpp:480-650 (domain edges: no symmetry masking)
__device__ inline void hllc(const Cfg& cfg, const double d[3], double l[5],
                            double r[5], double f[5]) {
  const double g = cfg.g;
  double nx = -d[0], ny = -d[1], nz = -d[2];
  double len = sqrt(nx * nx + ny * ny + nz * nz);
  nx /= len; ny /= len; nz /= len;
  double qL = l[1] * nx + l[2] * ny + l[3] * nz;
  double qR = r[1] * nx + r[2] * ny + r[3] * nz;
  double pL = (l[0] * l[4]) * (g - 1.0);
  double pR = (r[0] * r[4]) * (g - 1.0);
  l[4] = (l[4] + 0.5 * (l[1] * l[1] + l[2] * l[2] + l[3] * l[3])) * l[0];
  l[1] *= l[0]; l[2] *= l[0]; l[3] *= l[0];
  r[4] = (r[4] + 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3])) * r[0];
  r[1] *= r[0]; r[2] *= r[0]; r[3] *= r[0];
  double cL = sqrt(g * pL / l[0]);
  double cR = sqrt(g * pR / r[0]);
  double sL = fmin(qL - cL, qR - cR);
  double sR = fmax(qL + cL, qR + cR);
  double tL = sL - qL;
  double tR = sR - qR;
  double sM = (r[0] * qR * tR - l[0] * qL * tL + pL - pR) / (r[0] * tR - l[0] * tL);
  double pS = pL - l[0] * tL * (qL - sM);
  double L2 = -2.0 * len;
  double mx = nx * L2, my = ny * L2, mz = nz * L2;
  if (sL > 0.0) {
    double qL2 = qL * L2;
    f[0] = l[0] * qL2;
    f[1] = l[1] * qL2 + pL * mx;
    f[2] = l[2] * qL2 + pL * my;
    f[3] = l[3] * qL2 + pL * mz;
    f[4] = (l[4] + pL) * qL2;
  } else if (sL <= 0.0 && sM > 0.0) {
    double s = sL - sM;
    double u0 = tL * l[0] / s;
    double u1 = (tL * l[1] + (pS - pL) * nx) / s;
    double u2 = (tL * l[2] + (pS - pL) * ny) / s;
    double u3 = (tL * l[3] + (pS - pL) * nz) / s;
    double u4 = (tL * l[4] - pL * qL + pS * sM) / s;
    double qL2 = qL * L2;
    double sL2 = sL * L2;
    f[0] = l[0] * qL2 + sL2 * (u0 - l[0]);
    f[1] = l[1] * qL2 + pL * mx + sL2 * (u1 - l[1]);
    f[2] = l[2] * qL2 + pL * my + sL2 * (u2 - l[2]);
    f[3] = l[3] * qL2 + pL * mz + sL2 * (u3 - l[3]);
    f[4] = (l[4] + pL) * qL2 + sL2 * (u4 - l[4]);
  } else if (sM <= 0.0 && sR >= 0.0) {
    double s = sR - sM;
    double u0 = tR * r[0] / s;
    double u1 = (tR * r[1] + (pS - pR) * nx) / s;
    double u2 = (tR * r[2] + (pS - pR) * ny) / s;
    double u3 = (tR * r[3] + (pS - pR) * nz) / s;
    double u4 = (tR * r[4] - pR * qR + pS * sM) / s;
    double qR2 = qR * L2;
    double sR2 = sR * L2;
    f[0] = r[0] * qR2 + sR2 * (u0 - r[0]);
    f[1] = r[1] * qR2 + pR * mx + sR2 * (u1 - r[1]);
    f[2] = r[2] * qR2 + pR * my + sR2 * (u2 - r[2]);
    f[3] = r[3] * qR2 + pR * mz + sR2 * (u3 - r[3]);
    f[4] = (r[4] + pR) * qR2 + sR2 * (u4 - r[4]);
  } else {
    double qR2 = qR * L2;
    f[0] = r[0] * qR2;
    f[1] = r[1] * qR2 + pR * mx;
    f[2] = r[2] * qR2 + pR * my;
    f[3] = r[3] * qR2 + pR * mz;
    f[4] = (r[4] + pR) * qR2;
  }
  if (cfg.stab2) {
    double sl = fabs(qL) + cL;
    double sr = fabs(qR) + cR;
    double fws = cfg.stab2coef * fmax(sl, sr) * len;
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c] -= fws * (l[c] - r[c]);
  }
}

// ----------------------------------------------------------------- kernels --

__global__ void k_fill(int64_t n, double* __restrict__ a, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// riemann::grad domain edges (Riemann.cpp:263-328), one colour: the edges of
// a colour share no node, so both end updates are race free.
__global__ __launch_bounds__(kBlock) void k_grad_edges(
    int64_t beg, int64_t end, const int* __restrict__ ep,
    const int* __restrict__ eq, const double* __restrict__ dx,
    const double* __restrict__ dy, const double* __restrict__ dz,
    const double* __restrict__ U, double* __restrict__ G) {
  int64_t e = beg + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= end) return;
  int p = ep[e], q = eq[e];
  double d[3] = {dx[e], dy[e], dz[e]};
  double up[5], uq[5];
  primitive(U, p, up);
  primitive(U, q, uq);
  double* Gp = G + (size_t)p * 15;
  double* Gq = G + (size_t)q * 15;
#pragma unroll
  for (int c = 0; c < 5; ++c)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double f = d[j] * (uq[c] + up[c]);
      Gp[c * 3 + j] -= f;
      Gq[c * 3 + j] += f;
    }
}

// riemann::grad boundary integral (Riemann.cpp:332-359) gathered per node.
// As in the reference, every node of a face receives g[j]*n[j] for direction
// j, with g[j] the 4:1:1-weighted value belonging to the face's j-th node.
__global__ __launch_bounds__(kBlock) void k_grad_bnd(
    int nb, const int* __restrict__ bnode, const int* __restrict__ boff,
    const int* __restrict__ bface, const int* __restrict__ tri,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ U,
    double* __restrict__ G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int n = bnode[i];
  double acc[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) acc[k] = G[(size_t)n * 15 + k];
  for (int k = boff[i]; k < boff[i + 1]; ++k) {
    int f = bface[k] >> 2;
    int A = tri[f * 3 + 0], B = tri[f * 3 + 1], C = tri[f * 3 + 2];
    double u[3][5];
    primitive(U, A, u[0]);
    primitive(U, B, u[1]);
    primitive(U, C, u[2]);
    double ba[3] = {x[B] - x[A], y[B] - y[A], z[B] - z[A]};
    double ca[3] = {x[C] - x[A], y[C] - y[A], z[C] - z[A]};
    double nr[3] = {(ba[1] * ca[2] - ca[1] * ba[2]) / 12.0,
                    (ba[2] * ca[0] - ca[2] * ba[0]) / 12.0,
                    (ba[0] * ca[1] - ca[0] * ba[1]) / 12.0};
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double uab = (u[0][c] + u[1][c]) / 4.0;
      double ubc = (u[1][c] + u[2][c]) / 4.0;
      double uca = (u[2][c] + u[0][c]) / 4.0;
      double g[3] = {uab + uca + u[0][c], uab + ubc + u[1][c], ubc + uca + u[2][c]};
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[c * 3 + j] += g[j] * nr[j];
    }
  }
#pragma unroll
  for (int k = 0; k < 15; ++k) G[(size_t)n * 15 + k] = acc[k];
}

// advdom (Riemann.cpp:652-766) for one colour of edges.
template <int FLUX>
__global__ __launch_bounds__(kBlock) void k_rhs_edges(
    int64_t beg, int64_t end, const int* __restrict__ ep,
    const int* __restrict__ eq, const double* __restrict__ dx,
    const double* __restrict__ dy, const double* __restrict__ dz,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ G,
    const double* __restrict__ U, double* __restrict__ R, Cfg cfg) {
  int64_t e = beg + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= end) return;
  int p = ep[e], q = eq[e];
  double d[3] = {dx[e], dy[e], dz[e]};
  double l[5], r[5], f[5];
  primitive(U, p, l);
  primitive(U, q, r);
  double vw[3] = {x[q] - x[p], y[q] - y[p], z[q] - z[p]};
  muscl(vw, G + (size_t)p * 15, G + (size_t)q * 15, l, r);
  if (FLUX == 0)
    rusanov(cfg, d, l, r, f);
  else
    hllc(cfg, d, l, r, f);
  double* Rp = R + (size_t)p * 5;
  double* Rq = R + (size_t)q * 5;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    Rp[c] -= f[c];
    Rq[c] += f[c];
  }
}

// advbnd (Riemann.cpp:768-878) gathered per node; bface packs face*4 + the
// node's position within the face.
__global__ __launch_bounds__(kBlock) void k_rhs_bnd(
    int nb, const int* __restrict__ bnode, const int* __restrict__ boff,
    const int* __restrict__ bface, const int* __restrict__ tri,
    const unsigned char* __restrict__ besym, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ U, double* __restrict__ R, double gamma) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int n = bnode[i];
  double acc[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) acc[c] = R[(size_t)n * 5 + c];
  for (int k = boff[i]; k < boff[i + 1]; ++k) {
    int f = bface[k] >> 2, pos = bface[k] & 3;
    int N[3] = {tri[f * 3 + 0], tri[f * 3 + 1], tri[f * 3 + 2]};
    double ba[3] = {x[N[1]] - x[N[0]], y[N[1]] - y[N[0]], z[N[1]] - z[N[0]]};
    double ca[3] = {x[N[2]] - x[N[0]], y[N[2]] - y[N[0]], z[N[2]] - z[N[0]]};
    double nx = (ba[1] * ca[2] - ca[1] * ba[2]) / 12.0;
    double ny = (ba[2] * ca[0] - ca[2] * ba[0]) / 12.0;
    double nz = (ba[0] * ca[1] - ca[0] * ba[1]) / 12.0;
    double fl[5][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double* s = U + (size_t)N[a] * 5;
      double r = s[0], ru = s[1], rv = s[2], rw = s[3], re = s[4];
      double p = (re - 0.5 * (ru * ru + rv * rv + rw * rw) / r) * (gamma - 1.0);
      double vn = besym[f * 3 + a] ? 0.0 : (nx * ru + ny * rv + nz * rw) / r;
      fl[0][a] = r * vn;
      fl[1][a] = ru * vn + p * nx;
      fl[2][a] = rv * vn + p * ny;
      fl[3][a] = rw * vn + p * nz;
      fl[4][a] = (re + p) * vn;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double fab = (fl[c][0] + fl[c][1]) / 4.0;
      double fbc = (fl[c][1] + fl[c][2]) / 4.0;
      double fca = (fl[c][2] + fl[c][0]) / 4.0;
      double v = pos == 0 ? fab + fca + fl[c][0]
               : pos == 1 ? fab + fbc + fl[c][1]
                          : fbc + fca + fl[c][2];
      acc[c] += v;
    }
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) R[(size_t)n * 5 + c] = acc[c];
}

// G /= vol (RieCG.cpp:935-939)
__global__ void k_divvol(int npoin, double* __restrict__ G,
                         const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * 15) return;
  G[i] /= vol[i / 15];
}

// U = Un - rkcoef*dt*R/vol (RieCG.cpp:1016-1021)
__global__ void k_update(int npoin, double coef_dt, double* __restrict__ U,
                         const double* __restrict__ Un,
                         const double* __restrict__ R,
                         const double* __restrict__ vol) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)npoin * 5) return;
  U[i] = Un[i] - coef_dt * R[i] / vol[i / 5];
}

// physics::symbc (BC.cpp:110-136): the normals of a node are applied in list
// order by one thread.
__global__ void k_symbc(int ns, const int* __restrict__ node,
                        const int* __restrict__ off,
                        const double* __restrict__ nrm, double* __restrict__ U) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  double* u = U + (size_t)node[i] * 5 + 1;
  double a = u[0], b = u[1], c = u[2];
  for (int k = off[i]; k < off[i + 1]; ++k) {
    const double* n = nrm + (size_t)k * 3;
    double vn = a * n[0] + b * n[1] + c * n[2];
    a -= vn * n[0];
    b -= vn * n[1];
    c -= vn * n[2];
  }
  u[0] = a; u[1] = b; u[2] = c;
}

__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// per-node CFL limit of RieCG::dt (RieCG.cpp:827-839), block minimum
__global__ __launch_bounds__(kBlock) void k_dt_partial(
    int npoin, const double* __restrict__ U, const double* __restrict__ vol,
    double g, double* __restrict__ part) {
  __shared__ double sm[kBlock / 64];
  double m = 1.7976931348623157e308;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npoin;
       p += (int64_t)gridDim.x * blockDim.x) {
    const double* s = U + p * 5;
    double r = s[0];
    double u = s[1] / r, v = s[2] / r, w = s[3] / r;
    double pr = (s[4] - 0.5 * r * (u * u + v * v + w * w)) * (g - 1.0);
    double c = sqrt(g * fmax(pr, 0.0) / r);
    double L = cbrt(vol[p]);
    double vel = sqrt(u * u + v * v + w * w);
    m = fmin(m, L / fmax(vel + c, 1.0e-8));
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

__global__ __launch_bounds__(kRedBlocks) void k_min_final(
    int n, const double* __restrict__ part, double scale,
    double* __restrict__ out) {
  __shared__ double sm[kRedBlocks / 64];
  double m = 1.7976931348623157e308;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, part[i]);
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int k = 1; k < kRedBlocks / 64; ++k) b = fmin(b, sm[k]);
    out[0] = b * scale;
  }
}

// ------------------------------------------------------------- launchers --

int check_mesh(xg_ctx* c, int ncomp) {
  if (!c) return XG_EINVAL;
  if (!c->have_mesh) return fail(c, XG_ENOMESH, "no mesh uploaded");
  if (ncomp != 5)
    return fail(c, XG_EUNSUPPORTED,
                "device path supports ncomp = 5 (no transported scalars)");
  return XG_OK;
}

int launch_grad(xg_ctx* c, const double* U, double* G) {
  int64_t n = (int64_t)c->npoin * 15;
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, G, 0.0);
  for (size_t k = 0; k + 1 < c->coloff.size(); ++k) {
    int64_t b = c->coloff[k], e = c->coloff[k + 1];
    if (e == b) continue;
    hipLaunchKernelGGL(k_grad_edges, dim3(nblocks(e - b)), dim3(kBlock), 0,
                       c->stream, b, e, c->ep, c->eq, c->dx, c->dy, c->dz, U, G);
  }
  if (c->nbnode)
    hipLaunchKernelGGL(k_grad_bnd, dim3(nblocks(c->nbnode)), dim3(kBlock), 0,
                       c->stream, c->nbnode, c->bnode, c->boff, c->bface,
                       c->tri, c->x, c->y, c->z, U, G);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int launch_rhs(xg_ctx* c, const double* G, const double* U, double* R) {
  if (c->cfg.flux != 0 && c->cfg.flux != 1)
    return fail(c, XG_EINVAL, "Flux not configured");
  Cfg cfg{c->cfg.gamma, c->cfg.stab2, c->cfg.stab2coef};
  int64_t n = (int64_t)c->npoin * 5;
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, R, 0.0);
  for (size_t k = 0; k + 1 < c->coloff.size(); ++k) {
    int64_t b = c->coloff[k], e = c->coloff[k + 1];
    if (e == b) continue;
    if (c->cfg.flux == 0)
      hipLaunchKernelGGL(k_rhs_edges<0>, dim3(nblocks(e - b)), dim3(kBlock), 0,
                         c->stream, b, e, c->ep, c->eq, c->dx, c->dy, c->dz,
                         c->x, c->y, c->z, G, U, R, cfg);
    else
      hipLaunchKernelGGL(k_rhs_edges<1>, dim3(nblocks(e - b)), dim3(kBlock), 0,
                         c->stream, b, e, c->ep, c->eq, c->dx, c->dy, c->dz,
                         c->x, c->y, c->z, G, U, R, cfg);
  }
  if (c->nbnode)
    hipLaunchKernelGGL(k_rhs_bnd, dim3(nblocks(c->nbnode)), dim3(kBlock), 0,
                       c->stream, c->nbnode, c->bnode, c->boff, c->bface,
                       c->tri, c->besym, c->x, c->y, c->z, U, R, c->cfg.gamma);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

int stage_buffers(xg_ctx* c) {
  size_t need = (size_t)c->npoin * 8;
  if (c->sbytes >= need) return XG_OK;
  if (c->sU) { (void)hipFree(c->sU); (void)hipFree(c->sG); (void)hipFree(c->sR); }
  c->sU = c->sG = c->sR = nullptr; c->sbytes = 0;
  XG_HIP(c, hipMalloc((void**)&c->sU, need * 5));
  XG_HIP(c, hipMalloc((void**)&c->sG, need * 15));
  XG_HIP(c, hipMalloc((void**)&c->sR, need * 5));
  c->sbytes = need;
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
  if (c->sU) { (void)hipFree(c->sU); (void)hipFree(c->sG); (void)hipFree(c->sR); }
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
    p[k] = (int)a; q[k] = (int)b;
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
  // greedy colouring: smallest colour free at both ends
  std::vector<uint64_t> used((size_t)npoin, 0);
  std::vector<uint8_t> col(ne);
  int ncol = 0;
  for (int64_t e = 0; e < ne; ++e) {
    uint64_t m = used[p[e]] | used[q[e]];
    if (~m == 0) return fail(c, XG_EUNSUPPORTED, "edge colouring needs > 64 colours");
    int cc = __builtin_ctzll(~m);
    col[e] = (uint8_t)cc;
    used[p[e]] |= 1ull << cc;
    used[q[e]] |= 1ull << cc;
    ncol = std::max(ncol, cc + 1);
  }
  std::vector<int64_t> coloff(ncol + 1, 0);
  for (int64_t e = 0; e < ne; ++e) ++coloff[col[e] + 1];
  for (int cc = 0; cc < ncol; ++cc) coloff[cc + 1] += coloff[cc];
  std::vector<int> sp(ne), sq(ne);
  std::vector<double> sx(ne), sy(ne), sz(ne);
  {
    std::vector<int64_t> pos(coloff.begin(), coloff.end() - 1);
    for (int64_t e = 0; e < ne; ++e) {
      int64_t t = pos[col[e]]++;
      sp[t] = p[e]; sq[t] = q[e];
      sx[t] = d[e * 3]; sy[t] = d[e * 3 + 1]; sz[t] = d[e * 3 + 2];
    }
  }
  // boundary gather lists: node -> (face*4 + position), faces in order
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
  std::vector<int> tri32(ntri * 3);
  for (int64_t f = 0; f < ntri * 3; ++f) tri32[f] = (int)tri[f];
  // device arrays
  (void)hipStreamSynchronize(c->stream);
  dev_free_all(c->mesh_allocs);
  c->have_mesh = false;
  int rc;
  auto& L = c->mesh_allocs;
  if ((rc = upload(c, L, &c->x, x, npoin)) || (rc = upload(c, L, &c->y, y, npoin)) ||
      (rc = upload(c, L, &c->z, z, npoin)) || (rc = upload(c, L, &c->ep, sp.data(), ne)) ||
      (rc = upload(c, L, &c->eq, sq.data(), ne)) ||
      (rc = upload(c, L, &c->dx, sx.data(), ne)) ||
      (rc = upload(c, L, &c->dy, sy.data(), ne)) ||
      (rc = upload(c, L, &c->dz, sz.data(), ne)) ||
      (rc = upload(c, L, &c->tri, tri32.data(), tri32.size())) ||
      (rc = upload(c, L, &c->besym, (const unsigned char*)besym, (size_t)ntri * 3)) ||
      (rc = upload(c, L, &c->bnode, bnode.data(), bnode.size())) ||
      (rc = upload(c, L, &c->boff, boff.data(), boff.size())) ||
      (rc = upload(c, L, &c->bface, bface.data(), bface.size())))
    return rc;
  c->npoin = (int)npoin;
  c->nedge = ne;
  c->coloff = coloff;
  c->nbnode = (int)bnode.size();
  c->ntri = ntri;
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
  return launch_grad(c, U, G);
}

int xg_riemann_rhs_dev(xg_ctx* c, int ncomp, const double* G, const double* U,
                       double* R) {
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G || !R) return fail(c, XG_EINVAL, "xg_riemann_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  return launch_rhs(c, G, U, R);
}

int xg_riemann_grad(xg_ctx* c, int ncomp, const double* U, double* G) {
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G) return fail(c, XG_EINVAL, "xg_riemann_grad: null field");
  XG_HIP(c, hipSetDevice(c->device));
  if ((rc = stage_buffers(c))) return rc;
  size_t n = (size_t)c->npoin;
  XG_HIP(c, hipMemcpyAsync(c->sU, U, n * 40, hipMemcpyHostToDevice, c->stream));
  if ((rc = launch_grad(c, c->sU, c->sG))) return rc;
  XG_HIP(c, hipMemcpyAsync(G, c->sG, n * 120, hipMemcpyDeviceToHost, c->stream));
  XG_HIP(c, hipStreamSynchronize(c->stream));
  return XG_OK;
}

int xg_riemann_rhs(xg_ctx* c, int ncomp, const double* G, const double* U,
                   double* R) {
  int rc = check_mesh(c, ncomp);
  if (rc) return rc;
  if (!U || !G || !R) return fail(c, XG_EINVAL, "xg_riemann_rhs: null field");
  XG_HIP(c, hipSetDevice(c->device));
  if ((rc = stage_buffers(c))) return rc;
  size_t n = (size_t)c->npoin;
  XG_HIP(c, hipMemcpyAsync(c->sU, U, n * 40, hipMemcpyHostToDevice, c->stream));
  XG_HIP(c, hipMemcpyAsync(c->sG, G, n * 120, hipMemcpyHostToDevice, c->stream));
  if ((rc = launch_rhs(c, c->sG, c->sU, c->sR))) return rc;
  XG_HIP(c, hipMemcpyAsync(R, c->sR, n * 40, hipMemcpyDeviceToHost, c->stream));
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
  c->have_riecg = false;
  int rc;
  auto& L = c->riecg_allocs;
  if ((rc = upload(c, L, &c->vol, vol, (size_t)c->npoin)) ||
      (rc = upload(c, L, &c->symnode, node.data(), node.size())) ||
      (rc = upload(c, L, &c->symoff, off.data(), off.size())) ||
      (rc = upload(c, L, &c->symnorm, nrm.data(), nrm.size())) ||
      (rc = dev_alloc(c, L, (void**)&c->red, sizeof(double) * (kRedBlocks + 1))))
    return rc;
  c->nsymnode = (int)node.size();
  c->have_riecg = true;
  return XG_OK;
}

int xg_riecg_dt_dev(xg_ctx* c, const double* U, double cfl, double* dt) {
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (!U || !dt) return fail(c, XG_EINVAL, "xg_riecg_dt: null argument");
  XG_HIP(c, hipSetDevice(c->device));
  int nb = std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, nblocks(c->npoin)));
  hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(kBlock), 0, c->stream,
                     c->npoin, U, c->vol, c->cfg.gamma, c->red);
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

int xg_riecg_stage_dev(xg_ctx* c, int stage, double dt, double* U, double* Un,
                       double* G, double* R) {
  static const double rkcoef[3] = {1.0 / 3.0, 1.0 / 2.0, 1.0};
  if (!c) return XG_EINVAL;
  if (!c->have_riecg) return fail(c, XG_ENOMESH, "xg_riecg_setup not called");
  if (stage < 0 || stage > 2) return fail(c, XG_EINVAL, "stage must be 0, 1 or 2");
  if (!U || !Un || !G || !R) return fail(c, XG_EINVAL, "xg_riecg_stage: null field");
  XG_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = launch_grad(c, U, G))) return rc;
  int64_t n15 = (int64_t)c->npoin * 15, n5 = (int64_t)c->npoin * 5;
  hipLaunchKernelGGL(k_divvol, dim3(nblocks(n15)), dim3(kBlock), 0, c->stream,
                     c->npoin, G, c->vol);
  if ((rc = launch_rhs(c, G, U, R))) return rc;
  if (stage == 0)
    XG_HIP(c, hipMemcpyAsync(Un, U, (size_t)n5 * 8, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_update, dim3(nblocks(n5)), dim3(kBlock), 0, c->stream,
                     c->npoin, rkcoef[stage] * dt, U, Un, R, c->vol);
  if (c->nsymnode)
    hipLaunchKernelGGL(k_symbc, dim3(nblocks(c->nsymnode)), dim3(kBlock), 0,
                       c->stream, c->nsymnode, c->symnode, c->symoff,
                       c->symnorm, U);
  XG_HIP(c, hipGetLastError());
  return XG_OK;
}

}  // extern "C"
