/* CPU oracle of kozak::rhs -- TEST INFRASTRUCTURE ONLY (see xyst_oracle.h).
 *
 * kozak::rhs, Kozak.cpp:29-180: per tet, the Taylor-Galerkin half-step
 * element state from the four nodal states and the tet's shape-function
 * gradients, then the second-step fluxes scattered to the four nodes.
 * No source term (problems::SRC() is null for the problems served).
 *
 * orc_kozak_rhs_mag repeats the same operations on (value, magnitude) pairs:
 * for every intermediate r it carries m(r) with |r_computed - r| <= eps*m(r)
 * to first order (m(a+b) = m(a)+m(b)+|a+b|, m(ab) = m(a)|b|+|a|m(b)+|ab|,
 * m(a/b) = (m(a)+|a/b|m(b))/|b| + |a/b|), and sums m over the terms added to
 * each R entry.  That is the round-off scale against which another
 * evaluation order of the same arithmetic is compared.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xyst_oracle.h"

/* tk::cross / tk::triple (Vector.hpp) on 3-vectors */
static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1]*b[2] - b[1]*a[2];
  c[1] = a[2]*b[0] - b[2]*a[0];
  c[2] = a[0]*b[1] - b[0]*a[1];
}

void orc_kozak_rhs(const orc_cfg* cfg, int64_t npoin, int ncomp,
                   const double* x, const double* y, const double* z,
                   int64_t ntet, const int64_t* inpoel, double dt,
                   const double* dtp, const double* U, double* R) {
  orc_kozak_rhs_src(cfg, NULL, npoin, ncomp, x, y, z, ntet, inpoel, 0.0, dt,
                    NULL, dtp, U, R);
}

void orc_kozak_rhs_src(const orc_cfg* cfg, const orc_problem* P,
                       int64_t npoin, int ncomp, const double* x,
                       const double* y, const double* z, int64_t ntet,
                       const int64_t* inpoel, double t0, double dt,
                       const double* tp, const double* dtp, const double* U,
                       double* R) {
  const double g = cfg->gamma;
  const int src = P && P->kind >= 1 && P->kind <= 3;
  memset(R, 0, sizeof(double) * (size_t)npoin * (size_t)ncomp);   /* :54 */
  double ue[64];
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    double ba[3] = { x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]] };
    double ca[3] = { x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]] };
    double da[3] = { x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]] };
    double grad[4][3], t[3];
    cross3(ca, da, grad[1]);
    cross3(da, ba, grad[2]);
    cross3(ba, ca, grad[3]);
    cross3(ca, da, t);
    const double J = ba[0]*t[0] + ba[1]*t[1] + ba[2]*t[2];          /* triple */
    for (int i = 0; i < 3; ++i) grad[0][i] = -grad[1][i]-grad[2][i]-grad[3][i];
    /* first half step, Kozak.cpp:92-124 */
    double p[4];
    for (int a = 0; a < 4; ++a) {
      const double* u = U + N[a]*ncomp;
      double r = u[0], ru = u[1], rv = u[2], rw = u[3];
      p[a] = (u[4] - 0.5*(ru*ru + rv*rv + rw*rw)/r) * (g - 1.0);
    }
    for (int c = 0; c < ncomp; ++c)
      ue[c] = (U[N[0]*ncomp+c] + U[N[1]*ncomp+c] + U[N[2]*ncomp+c] +
               U[N[3]*ncomp+c])/4.0;
    double dte = dt;
    if (dtp) dte = (dtp[N[0]] + dtp[N[1]] + dtp[N[2]] + dtp[N[3]])/4.0;
    double coef = dte/J/2.0;
    for (int j = 0; j < 3; ++j)
      for (int a = 0; a < 4; ++a) {
        const double* u = U + N[a]*ncomp;
        double cg = coef * grad[a][j];
        double uj = u[j+1] / u[0];
        ue[0] -= cg * u[j+1];
        ue[1] -= cg * u[1] * uj;
        ue[2] -= cg * u[2] * uj;
        ue[3] -= cg * u[3] * uj;
        ue[j+1] -= cg * p[a];
        ue[4] -= cg * (u[4] + p[a]) * uj;
        for (int c = 5; c < ncomp; ++c) ue[c] -= cg * u[c] * uj;
      }
    double te = t0;
    if (src) {                                          /* Kozak.cpp:124-136 */
      coef = dte/8.0;
      if (tp) te = (tp[N[0]] + tp[N[1]] + tp[N[2]] + tp[N[3]])/4.0;
      for (int a = 0; a < 4; ++a) {
        double s[5];
        orc_problem_src(P, x[N[a]], y[N[a]], z[N[a]], te, s);
        for (int c = 0; c < ncomp && c < 5; ++c) ue[c] += coef * s[c];
      }
    }
    /* second half step, Kozak.cpp:138-161 */
    double r = ue[0], ru = ue[1], rv = ue[2], rw = ue[3];
    double pr = (ue[4] - 0.5*(ru*ru + rv*rv + rw*rw)/r) * (g - 1.0);
    coef = 1.0/6.0;
    for (int j = 0; j < 3; ++j) {
      double uj = ue[j+1] / ue[0];
      for (int a = 0; a < 4; ++a) {
        double* o = R + N[a]*ncomp;
        double cg = coef * grad[a][j];
        o[0] += cg * ue[j+1];
        o[1] += cg * ue[1] * uj;
        o[2] += cg * ue[2] * uj;
        o[3] += cg * ue[3] * uj;
        o[j+1] += cg * pr;
        o[4] += cg * (ue[4] + pr) * uj;
        for (int c = 5; c < ncomp; ++c) o[c] += cg * ue[c] * uj;
      }
    }
    if (src) {                                          /* Kozak.cpp:163-173 */
      double xe = (x[N[0]] + x[N[1]] + x[N[2]] + x[N[3]])/4.0;
      double ye = (y[N[0]] + y[N[1]] + y[N[2]] + y[N[3]])/4.0;
      double ze = (z[N[0]] + z[N[1]] + z[N[2]] + z[N[3]])/4.0;
      double se[5];
      orc_problem_src(P, xe, ye, ze, te + dte/2.0, se);
      coef = J/24.0;
      for (int a = 0; a < 4; ++a)
        for (int c = 0; c < ncomp && c < 5; ++c) R[N[a]*ncomp+c] += coef * se[c];
    }
  }
}

/* ---- magnitude-tracking restatement ---- */

typedef struct { double v, m; } em;

static em E(double v) { em r = { v, 0.0 }; return r; }          /* exact input */
static em K(double v) { em r = { v, fabs(v) }; return r; }      /* rounded constant */
static em add(em a, em b) { em r = { a.v + b.v, a.m + b.m }; r.m += fabs(r.v); return r; }
static em sub(em a, em b) { em r = { a.v - b.v, a.m + b.m }; r.m += fabs(r.v); return r; }
static em mul(em a, em b) {
  em r = { a.v * b.v, a.m*fabs(b.v) + fabs(a.v)*b.m };
  r.m += fabs(r.v);
  return r;
}
static em dvd(em a, em b) {
  em r = { a.v / b.v, 0.0 };
  r.m = (a.m + fabs(r.v)*b.m)/fabs(b.v) + fabs(r.v);
  return r;
}

void orc_kozak_rhs_mag(const orc_cfg* cfg, int64_t npoin, int ncomp,
                       const double* x, const double* y, const double* z,
                       int64_t ntet, const int64_t* inpoel, double dt,
                       const double* U, double* R, double* A) {
  const em g1 = sub(E(cfg->gamma), E(1.0));
  memset(R, 0, sizeof(double) * (size_t)npoin * (size_t)ncomp);
  memset(A, 0, sizeof(double) * (size_t)npoin * (size_t)ncomp);
  em ue[64];
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    em X[4][3];
    for (int a = 0; a < 4; ++a) {
      X[a][0] = E(x[N[a]]); X[a][1] = E(y[N[a]]); X[a][2] = E(z[N[a]]);
    }
    em ba[3], ca[3], da[3];
    for (int i = 0; i < 3; ++i) {
      ba[i] = sub(X[1][i], X[0][i]);
      ca[i] = sub(X[2][i], X[0][i]);
      da[i] = sub(X[3][i], X[0][i]);
    }
    em grad[4][3];
#define CROSS(A_, B_, C_) \
    C_[0] = sub(mul(A_[1], B_[2]), mul(B_[1], A_[2])); \
    C_[1] = sub(mul(A_[2], B_[0]), mul(B_[2], A_[0])); \
    C_[2] = sub(mul(A_[0], B_[1]), mul(B_[0], A_[1]));
    CROSS(ca, da, grad[1]);
    CROSS(da, ba, grad[2]);
    CROSS(ba, ca, grad[3]);
#undef CROSS
    em J = add(add(mul(ba[0], grad[1][0]), mul(ba[1], grad[1][1])), mul(ba[2], grad[1][2]));
    for (int i = 0; i < 3; ++i)
      grad[0][i] = sub(sub(sub(E(0.0), grad[1][i]), grad[2][i]), grad[3][i]);
    em u[4][64], p[4];
    for (int a = 0; a < 4; ++a) {
      for (int c = 0; c < ncomp; ++c) u[a][c] = E(U[N[a]*ncomp + c]);
      em ke = add(add(mul(u[a][1], u[a][1]), mul(u[a][2], u[a][2])), mul(u[a][3], u[a][3]));
      p[a] = mul(sub(u[a][4], dvd(mul(K(0.5), ke), u[a][0])), g1);
    }
    for (int c = 0; c < ncomp; ++c)
      ue[c] = dvd(add(add(add(u[0][c], u[1][c]), u[2][c]), u[3][c]), E(4.0));
    em coef = dvd(dvd(E(dt), J), E(2.0));
    for (int j = 0; j < 3; ++j)
      for (int a = 0; a < 4; ++a) {
        em cg = mul(coef, grad[a][j]);
        em uj = dvd(u[a][j+1], u[a][0]);
        ue[0] = sub(ue[0], mul(cg, u[a][j+1]));
        ue[1] = sub(ue[1], mul(mul(cg, u[a][1]), uj));
        ue[2] = sub(ue[2], mul(mul(cg, u[a][2]), uj));
        ue[3] = sub(ue[3], mul(mul(cg, u[a][3]), uj));
        ue[j+1] = sub(ue[j+1], mul(cg, p[a]));
        ue[4] = sub(ue[4], mul(mul(cg, add(u[a][4], p[a])), uj));
        for (int c = 5; c < ncomp; ++c) ue[c] = sub(ue[c], mul(mul(cg, u[a][c]), uj));
      }
    em ke = add(add(mul(ue[1], ue[1]), mul(ue[2], ue[2])), mul(ue[3], ue[3]));
    em pr = mul(sub(ue[4], dvd(mul(K(0.5), ke), ue[0])), g1);
    em c6 = K(1.0/6.0);
    for (int j = 0; j < 3; ++j) {
      em uj = dvd(ue[j+1], ue[0]);
      for (int a = 0; a < 4; ++a) {
        double* o = R + N[a]*ncomp;
        double* m = A + N[a]*ncomp;
        em cg = mul(c6, grad[a][j]);
        em t0 = mul(cg, ue[j+1]);
        em t1 = mul(mul(cg, ue[1]), uj);
        em t2 = mul(mul(cg, ue[2]), uj);
        em t3 = mul(mul(cg, ue[3]), uj);
        em tp = mul(cg, pr);
        em t4 = mul(mul(cg, add(ue[4], pr)), uj);
        o[0] += t0.v; m[0] += t0.m;
        o[1] += t1.v; m[1] += t1.m;
        o[2] += t2.v; m[2] += t2.m;
        o[3] += t3.v; m[3] += t3.m;
        o[j+1] += tp.v; m[j+1] += tp.m;
        o[4] += t4.v; m[4] += t4.m;
        for (int c = 5; c < ncomp; ++c) {
          em tc = mul(mul(cg, ue[c]), uj);
          o[c] += tc.v; m[c] += tc.m;
        }
      }
    }
  }
}

/* ---- KozCG step with flux-corrected transport ---- */

/* std::max / std::min, including which argument a tie or NaN returns */
static double kz_max(double a, double b) { return (a < b) ? b : a; }
static double kz_min(double a, double b) { return (b < a) ? b : a; }

/* element antidiffusive contributions of component c, KozCG.cpp:799-806 */
static void aec_tet(const double* x, const double* y, const double* z,
                    const int64_t* N, const double* U, int ncomp, int c,
                    double ctau, double aec[4]) {
  const double ba[3] = { x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]] };
  const double ca[3] = { x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]] };
  const double da[3] = { x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]] };
  double t[3];
  cross3(ca, da, t);
  const double J = ba[0]*t[0] + ba[1]*t[1] + ba[2]*t[2];
  for (int a = 0; a < 4; ++a) {
    aec[a] = 0.0;
    for (int b = 0; b < 4; ++b) {
      double m = J/120.0 * ((a == b) ? 3.0 : -1.0);
      aec[a] += m * ctau * U[N[b]*ncomp + c];
    }
  }
}

int orc_kozcg_step(const orc_cfg* cfg, const orc_fctcfg* f,
                   const orc_problem* P, double t, int64_t npoin,
                   int ncomp, const double* x, const double* y,
                   const double* z, int64_t ntet, const int64_t* inpoel,
                   const double* vol, double dt, const double* dtp, double* tp,
                   int freeze, int64_t nsym,
                   const int64_t* symnodes, const double* symnorms,
                   int64_t ndir, const int64_t* dirnodes, const int32_t* dirmask,
                   const double* dirvals, double* U, double* Un) {
  if (ncomp < 5 || ncomp > 16) return -1;
  const size_t nn = (size_t)npoin * (size_t)ncomp;
  double* R = (double*)malloc(sizeof(double) * nn);
  double* A = (double*)calloc(nn, sizeof(double));
  orc_kozak_rhs_src(cfg, P, npoin, ncomp, x, y, z, ntet, inpoel, t, dt, tp,
                    dtp, U, R);
  if (f->fct) {
    const double ctau = f->fctdif;
    double* P = (double*)calloc(2*nn, sizeof(double));
    double* Q = (double*)malloc(sizeof(double) * 2*nn);
    /* aec: P+/- (KozCG.cpp:774-838) */
    for (int64_t e = 0; e < ntet; ++e) {
      const int64_t* N = inpoel + e*4;
      for (int c = 0; c < ncomp; ++c) {
        double aec[4];
        aec_tet(x, y, z, N, U, ncomp, c, ctau, aec);
        for (int a = 0; a < 4; ++a) {
          P[N[a]*2*ncomp + 2*c] += kz_max(0.0, aec[a]);
          P[N[a]*2*ncomp + 2*c + 1] += kz_min(0.0, aec[a]);
        }
      }
    }
    for (int64_t i = 0; i < nsym; ++i) {           /* :817-830 */
      double* p = P + symnodes[i]*2*ncomp;
      const double nx = symnorms[i*3], ny = symnorms[i*3+1], nz = symnorms[i*3+2];
      double rvnp = p[2]*nx + p[4]*ny + p[6]*nz;
      double rvnn = p[3]*nx + p[5]*ny + p[7]*nz;
      p[2] -= rvnp * nx;  p[3] -= rvnn * nx;
      p[4] -= rvnp * ny;  p[5] -= rvnn * ny;
      p[6] -= rvnp * nz;  p[7] -= rvnn * nz;
    }
    /* alw: low-order solution (into R) and Q+/- (:866-948); steady: the
     * node's own time step (:889-890) */
    for (int64_t i = 0; i < npoin; ++i) {
      const double dti = dtp ? dtp[i] : dt;
      for (int c = 0; c < ncomp; ++c) {
        double* p = P + i*2*ncomp + 2*c;
        p[0] /= vol[i];
        p[1] /= vol[i];
        R[i*ncomp+c] = U[i*ncomp+c] + dti*R[i*ncomp+c]/vol[i] - p[0] - p[1];
      }
    }
    const double large = 1.7976931348623157e308;
    for (size_t k = 0; k < nn; ++k) { Q[2*k] = -large; Q[2*k+1] = large; }
    for (int64_t e = 0; e < ntet; ++e) {
      const int64_t* N = inpoel + e*4;
      for (int c = 0; c < ncomp; ++c) {
        double alwp = -large, alwn = large;
        for (int a = 0; a < 4; ++a) {
          const double ul = R[N[a]*ncomp+c];
          if (f->fctclip) {
            alwp = kz_max(alwp, ul);
            alwn = kz_min(alwn, ul);
          } else {
            alwp = kz_max(alwp, kz_max(ul, U[N[a]*ncomp+c]));
            alwn = kz_min(alwn, kz_min(ul, U[N[a]*ncomp+c]));
          }
        }
        for (int a = 0; a < 4; ++a) {
          double* q = Q + N[a]*2*ncomp + 2*c;
          q[0] = kz_max(q[0], alwp);
          q[1] = kz_min(q[1], alwn);
        }
      }
    }
    /* lim: limit coefficients (:979-1023) */
    const double eps = 2.220446049250313e-16;
    for (int64_t i = 0; i < npoin; ++i)
      for (int c = 0; c < ncomp; ++c) {
        double* q = Q + i*2*ncomp + 2*c;
        const double* p = P + i*2*ncomp + 2*c;
        q[0] -= R[i*ncomp+c];
        q[1] -= R[i*ncomp+c];
        q[0] = p[0] <  eps ? 0.0 : kz_min(1.0, q[0]/p[0]);
        q[1] = p[1] > -eps ? 0.0 : kz_min(1.0, q[1]/p[1]);
      }
    /* limited antidiffusive contributions (:1025-1070) */
    for (int64_t e = 0; e < ntet; ++e) {
      const int64_t* N = inpoel + e*4;
      double coef[16], aec[16][4];
      for (int c = 0; c < ncomp; ++c) {
        coef[c] = 1.0;
        aec_tet(x, y, z, N, U, ncomp, c, ctau, aec[c]);
        for (int a = 0; a < 4; ++a)
          coef[c] = kz_min(coef[c], aec[c][a] > 0.0 ? Q[N[a]*2*ncomp + 2*c]
                                                  : Q[N[a]*2*ncomp + 2*c + 1]);
      }
      double cs = 1.0;
      for (int k = 0; k < f->nsys; ++k) cs = kz_min(cs, coef[f->sys[k]]);
      for (int k = 0; k < f->nsys; ++k) coef[f->sys[k]] = cs;
      for (int c = 0; c < ncomp; ++c)
        for (int a = 0; a < 4; ++a) A[N[a]*ncomp+c] += coef[c] * aec[c][a];
    }
    /* solve: low-order + limited contributions (:1140-1146) */
    for (int64_t i = 0; i < npoin; ++i)
      for (int c = 0; c < ncomp; ++c)
        A[i*ncomp+c] = R[i*ncomp+c] + A[i*ncomp+c]/vol[i];
    free(P);
    free(Q);
  } else {
    for (int64_t i = 0; i < npoin; ++i) {                  /* :1148-1154 */
      const double dti = dtp ? dtp[i] : dt;
      for (int c = 0; c < ncomp; ++c)
        A[i*ncomp+c] = U[i*ncomp+c] + dti*R[i*ncomp+c]/vol[i];
    }
  }
  /* BC: dirbc (components with a nonzero mask; all without masks) then
   * symbc on momentum (KozCG.cpp:590-605) */
  for (int64_t k = 0; k < ndir; ++k)
    for (int c = 0; c < ncomp; ++c)
      if (!dirmask || dirmask[k*ncomp+c]) A[dirnodes[k]*ncomp+c] = dirvals[k*ncomp+c];
  orc_symbc(npoin, ncomp, A, nsym, symnodes, symnorms);
  /* freezeflow: the flow components keep their old values (:1168-1175) */
  if (freeze)
    for (int64_t i = 0; i < npoin; ++i)
      for (int c = 0; c < 5; ++c) A[i*ncomp+c] = U[i*ncomp+c];
  memcpy(Un, U, sizeof(double) * nn);
  memcpy(U, A, sizeof(double) * nn);
  /* local time stepping: tp += dtp (:1189-1193) */
  if (dtp && tp)
    for (int64_t i = 0; i < npoin; ++i) tp[i] += dtp[i];
  free(R);
  free(A);
  return 0;
}
