// Device arithmetic of the RieCG edge flux (FP64): primitive variables,
// MUSCL/van Leer reconstruction, Rusanov and HLLC fluxes, in divide-free
// forms of the reference's arithmetic (each function names the reference
// line range whose result it reproduces).
#pragma once
#include <hip/hip_runtime.h>

namespace xg {

struct Cfg {
  double g;
  int stab2;
  double stab2coef;
  double gm1;   // g - 1, computed on the host (keeps it a scalar operand)
  double gg1;   // g * (g - 1)
};

// riemann::primitive, Riemann.cpp:211-227: NC components per node, the
// transported scalars (c >= 5) as they are
template <int NC>
__device__ inline void primitive(const double* __restrict__ U, int i,
                                 double u[NC]) {
  const double* s = U + (size_t)i * NC;
  u[0] = s[0];
  u[1] = s[1] / u[0];
  u[2] = s[2] / u[0];
  u[3] = s[3] / u[0];
  u[4] = s[4] / u[0] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
#pragma unroll
  for (int c = 5; c < NC; ++c) u[c] = s[c];
}

// The transported scalars' flux of an edge, f_c = (l_c a + r_c b) m +
// sw (r_c - l_c), in the reference's association (Riemann.cpp:460-466 for
// Rusanov: a, b the dual-face normal velocities, m = 1; :637-643 for HLLC:
// a, b the unit-normal velocities, m the dual-face area).
struct ScalarFlux {
  double a, b, m, sw;
};

// ---------------------------------------------------------------------------
// Divide-free forms of the reference arithmetic.  Each rewrite is an identity in
// exact arithmetic; in FP64 the results differ from the reference's by a few
// units in the last place.
//
//  * 1/x: hardware reciprocal refined by two Newton steps (x > 0 or the
//    operand of a product whose sign is checked first).
//  * van Leer limiter (Riemann.cpp:93-104): with a = d2+eps, b = d1+eps,
//      phi(a/b)   = (|r|+r)/(|r|+1) = 2a/(a+b)  if a and b share a sign, else 0
//      phi(b/a)   =                   2b/(a+b)  likewise,
//    so both limiter values of an edge end share one reciprocal 1/(a+b)
//    (vl_rcp: sign test as a*b > 0, one Newton step).  (The reference
//    evaluates 0/0 -> NaN when a or b is exactly zero; here that
//    measure-zero case limits to 0.)
//  * sound speed sqrt(g p/rho) with p = (g-1) rho e equals sqrt(g (g-1) e);
//    Rusanov's c |n| is taken as one root, sqrt(g (g-1) e n.n).
// ---------------------------------------------------------------------------

__device__ inline double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

__device__ inline void primitive_fast(const double s[5], double u[5]) {
  double ir = rcp(s[0]);
  u[0] = s[0];
  u[1] = s[1] * ir;
  u[2] = s[2] * ir;
  u[3] = s[3] * ir;
  u[4] = s[4] * ir - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
}

// the van Leer limiter's shared reciprocal 1/(a + b) when a and b have the
// same sign, else 0: the sign test as one product (a, b are slope
// differences plus 1e-9, far from under- or overflow), the hardware
// reciprocal refined by one Newton step (the limiter value carries ~1e-16
// relative error against 2e-16 for two steps)
__device__ inline double vl_rcp(double a, double b) {
  if (!(a * b > 0.0)) return 0.0;
  const double x = a + b;
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// MUSCL / van Leer reconstruction (Riemann.cpp:34-143), one shared
// reciprocal per edge end and component.  g1[c] = G_p,c . vw and
// g2[c] = G_q,c . vw are the gradient projections on the edge vector.
__device__ inline void muscl_fast(const double g1[5], const double g2[5],
                                  double l[5], double r[5]) {
  constexpr double eps = 1.0e-9;
  constexpr double c1 = 2.0 / 3.0;   // 1 - muscl_const
  constexpr double c2 = 4.0 / 3.0;   // 1 + muscl_const
  double url[5], urr[5], d1[5], d3[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double delta2 = r[c] - l[c];
    double delta1 = 2.0 * g1[c] - delta2;
    double delta3 = 2.0 * g2[c] - delta2;
    d1[c] = delta1;
    d3[c] = delta3;
    double a = delta2 + eps, bl = delta1 + eps, br = delta3 + eps;
    double il = vl_rcp(a, bl);
    double ir = vl_rcp(a, br);
    url[c] = l[c] + 0.5 * il * (delta1 * c1 * a + delta2 * c2 * bl);
    urr[c] = r[c] - 0.5 * ir * (delta3 * c1 * a + delta2 * c2 * br);
  }
  bool keepL = (l[0] < d1[0] || l[4] < d1[4]);
  bool keepR = (r[0] < -d3[0] || r[4] < -d3[4]);
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    if (!keepL) l[c] = url[c];
    if (!keepR) r[c] = urr[c];
  }
}

// MUSCL / van Leer reconstruction of the transported scalars c = 5..NC-1
// (Riemann.cpp:145-209): the flow variables' limiter, without the
// positivity fallback.
template <int NC>
__device__ inline void muscl_scalars(const double g1[NC], const double g2[NC],
                                     double l[NC], double r[NC]) {
  constexpr double eps = 1.0e-9;
  constexpr double c1 = 2.0 / 3.0;
  constexpr double c2 = 4.0 / 3.0;
#pragma unroll
  for (int c = 5; c < NC; ++c) {
    double delta2 = r[c] - l[c];
    double delta1 = 2.0 * g1[c] - delta2;
    double delta3 = 2.0 * g2[c] - delta2;
    double a = delta2 + eps, bl = delta1 + eps, br = delta3 + eps;
    double il = vl_rcp(a, bl);
    double ir = vl_rcp(a, br);
    l[c] += 0.5 * il * (delta1 * c1 * a + delta2 * c2 * bl);
    r[c] -= 0.5 * ir * (delta3 * c1 * a + delta2 * c2 * br);
  }
}

// Rusanov flux (Riemann.cpp:369-478) on primitive end states (l, r hold
// rho, u, v, w, e), divide-free.
__device__ inline void rusanov_fast(const Cfg& cfg, const double n[3],
                                    double l[5], double r[5], double f[5],
                                    ScalarFlux& sf) {
  const double gg1 = cfg.gg1;
  double pL = (l[0] * l[4]) * cfg.gm1;
  double pR = (r[0] * r[4]) * cfg.gm1;
  // the wave speeds times the dual-face area, c len = sqrt(g (g-1) e n.n)
  const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  double cL = sqrt(gg1 * l[4] * nn);
  double cR = sqrt(gg1 * r[4] * nn);
  double vnL = l[1] * n[0] + l[2] * n[1] + l[3] * n[2];
  double vnR = r[1] * n[0] + r[2] * n[1] + r[3] * n[2];
  l[4] = (l[4] + 0.5 * (l[1] * l[1] + l[2] * l[2] + l[3] * l[3])) * l[0];
  l[1] *= l[0]; l[2] *= l[0]; l[3] *= l[0];
  r[4] = (r[4] + 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3])) * r[0];
  r[1] *= r[0]; r[2] *= r[0]; r[3] *= r[0];
  double fw = fmax(fabs(vnL) + cL, fabs(vnR) + cR);
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
  sf = ScalarFlux{vnL, vnR, 1.0, fmax(fabs(vnL), fabs(vnR))};
}

// HLLC flux (Riemann.cpp:480-650), reciprocals shared where the reference
// divides several terms by the same quantity.
__device__ inline void hllc_fast(const Cfg& cfg, const double d[3], double l[5],
                                 double r[5], double f[5], ScalarFlux& sf) {
  const double gg1 = cfg.gg1;
  double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double il = rcp(len);
  double nx = -d[0] * il, ny = -d[1] * il, nz = -d[2] * il;
  double qL = l[1] * nx + l[2] * ny + l[3] * nz;
  double qR = r[1] * nx + r[2] * ny + r[3] * nz;
  double pL = (l[0] * l[4]) * cfg.gm1;
  double pR = (r[0] * r[4]) * cfg.gm1;
  double cL = sqrt(gg1 * l[4]);
  double cR = sqrt(gg1 * r[4]);
  l[4] = (l[4] + 0.5 * (l[1] * l[1] + l[2] * l[2] + l[3] * l[3])) * l[0];
  l[1] *= l[0]; l[2] *= l[0]; l[3] *= l[0];
  r[4] = (r[4] + 0.5 * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3])) * r[0];
  r[1] *= r[0]; r[2] *= r[0]; r[3] *= r[0];
  double sL = fmin(qL - cL, qR - cR);
  double sR = fmax(qL + cL, qR + cR);
  double tL = sL - qL;
  double tR = sR - qR;
  double sM = (r[0] * qR * tR - l[0] * qL * tL + pL - pR) * rcp(r[0] * tR - l[0] * tL);
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
    double is = rcp(sL - sM);
    double u0 = tL * l[0] * is;
    double u1 = (tL * l[1] + (pS - pL) * nx) * is;
    double u2 = (tL * l[2] + (pS - pL) * ny) * is;
    double u3 = (tL * l[3] + (pS - pL) * nz) * is;
    double u4 = (tL * l[4] - pL * qL + pS * sM) * is;
    double qL2 = qL * L2;
    double sL2 = sL * L2;
    f[0] = l[0] * qL2 + sL2 * (u0 - l[0]);
    f[1] = l[1] * qL2 + pL * mx + sL2 * (u1 - l[1]);
    f[2] = l[2] * qL2 + pL * my + sL2 * (u2 - l[2]);
    f[3] = l[3] * qL2 + pL * mz + sL2 * (u3 - l[3]);
    f[4] = (l[4] + pL) * qL2 + sL2 * (u4 - l[4]);
  } else if (sM <= 0.0 && sR >= 0.0) {
    double is = rcp(sR - sM);
    double u0 = tR * r[0] * is;
    double u1 = (tR * r[1] + (pS - pR) * nx) * is;
    double u2 = (tR * r[2] + (pS - pR) * ny) * is;
    double u3 = (tR * r[3] + (pS - pR) * nz) * is;
    double u4 = (tR * r[4] - pR * qR + pS * sM) * is;
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
    double fws = cfg.stab2coef * fmax(fabs(qL) + cL, fabs(qR) + cR) * len;
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c] -= fws * (l[c] - r[c]);
  }
  sf = ScalarFlux{qL, qR, len, fmax(fabs(qL), fabs(qR)) * len};
}

}  // namespace xg
