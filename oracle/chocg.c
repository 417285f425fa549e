/* CPU oracle of the ChoCG edge operators -- TEST INFRASTRUCTURE ONLY (see
 * xyst_oracle.h).
 *
 * chorin::div, vgrad, grad, flux and rhs (adv with damp2/damp4), Chorin.cpp
 * :34-1044, over superedges with 5 values per edge integral (ChoCG::domint:
 * normal components, J/120, dot(grad_p,grad_q)/J/6).  div/vgrad/grad/flux
 * accumulate into their output (the caller zeroes, ChoCG.cpp:885,932,984);
 * rhs zeroes R.  Each output entry also receives, in A (when given), the sum
 * of the magnitudes of the products its added terms are made of: the round-off
 * scale used to compare another evaluation order of the same arithmetic.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xyst_oracle.h"

static const double muscl_eps = 1.0e-9;          /* Chorin.cpp:29-30 */
static const double muscl_const = 1.0/3.0;

/* tk::crossdiv (Vector.hpp:68-77): cross(v1, v2) / j */
static void crossdiv(double ax, double ay, double az, double bx, double by,
                     double bz, double j, double* n) {
  n[0] = (ay*bz - by*az) / j;
  n[1] = (az*bx - bz*ax) / j;
  n[2] = (ax*by - bx*ay) / j;
}

/* The superedge iteration shared by the operators: for tet, triangle and
 * single-edge groups, the per-edge values f[c] of edge (p,q) with integral
 * d[5] are combined as the reference scatters them (Chorin.cpp:123-158). */
typedef void (*edge_fn)(void* ctx, const double* d, int64_t p, int64_t q,
                        double* f, double* m);

static void superedges(const int64_t* nsup, const int64_t* sup0,
                       const double* int0, const int64_t* sup1,
                       const double* int1, const int64_t* sup2,
                       const double* int2, int nc, edge_fn fn, void* ctx,
                       double* R, int ldr, double* A) {
  double f[6][16], m[6][16];
  static const int tp[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
  for (int64_t e = 0; e < nsup[0]; ++e) {
    const int64_t* N = sup0 + e*4;
    for (int k = 0; k < 6; ++k) {
      int64_t p = N[tp[k][0]], q = N[tp[k][1]];
      fn(ctx, int0 + (e*6+k)*5, p, q, f[k], m[k]);
    }
    for (int c = 0; c < nc; ++c) {
      R[N[0]*ldr+c] = R[N[0]*ldr+c] - f[0][c] + f[2][c] - f[3][c];
      R[N[1]*ldr+c] = R[N[1]*ldr+c] + f[0][c] - f[1][c] - f[4][c];
      R[N[2]*ldr+c] = R[N[2]*ldr+c] + f[1][c] - f[2][c] - f[5][c];
      R[N[3]*ldr+c] = R[N[3]*ldr+c] + f[3][c] + f[4][c] + f[5][c];
      if (A) {
        A[N[0]*ldr+c] += m[0][c] + m[2][c] + m[3][c];
        A[N[1]*ldr+c] += m[0][c] + m[1][c] + m[4][c];
        A[N[2]*ldr+c] += m[1][c] + m[2][c] + m[5][c];
        A[N[3]*ldr+c] += m[3][c] + m[4][c] + m[5][c];
      }
    }
  }
  static const int tt[3][2] = {{0,1},{1,2},{2,0}};
  for (int64_t e = 0; e < nsup[1]; ++e) {
    const int64_t* N = sup1 + e*3;
    for (int k = 0; k < 3; ++k) {
      int64_t p = N[tt[k][0]], q = N[tt[k][1]];
      fn(ctx, int1 + (e*3+k)*5, p, q, f[k], m[k]);
    }
    for (int c = 0; c < nc; ++c) {
      R[N[0]*ldr+c] = R[N[0]*ldr+c] - f[0][c] + f[2][c];
      R[N[1]*ldr+c] = R[N[1]*ldr+c] + f[0][c] - f[1][c];
      R[N[2]*ldr+c] = R[N[2]*ldr+c] + f[1][c] - f[2][c];
      if (A) {
        A[N[0]*ldr+c] += m[0][c] + m[2][c];
        A[N[1]*ldr+c] += m[0][c] + m[1][c];
        A[N[2]*ldr+c] += m[1][c] + m[2][c];
      }
    }
  }
  for (int64_t e = 0; e < nsup[2]; ++e) {
    const int64_t* N = sup2 + e*2;
    int64_t p = N[0], q = N[1];
    fn(ctx, int2 + e*5, p, q, f[0], m[0]);
    for (int c = 0; c < nc; ++c) {
      R[N[0]*ldr+c] -= f[0][c];
      R[N[1]*ldr+c] += f[0][c];
      if (A) { A[N[0]*ldr+c] += m[0][c]; A[N[1]*ldr+c] += m[0][c]; }
    }
  }
}

/* ------------------------------------------------------------------ div -- */

typedef struct {
  const double *x, *y, *z, *P, *G, *U;
  int ldg, ldu, stab;
  double dt;
} div_ctx;

/* chorin::div edge term, Chorin.cpp:34-79 */
static void div_edge(void* vc, const double* d, int64_t p, int64_t q,
                     double* f, double* m) {
  const div_ctx* k = (const div_ctx*)vc;
  const double* U = k->U;
  const int lu = k->ldu;
  double dv = d[0] * (U[p*lu+0] + U[q*lu+0]) +
              d[1] * (U[p*lu+1] + U[q*lu+1]) +
              d[2] * (U[p*lu+2] + U[q*lu+2]);
  double mv = fabs(d[0]) * (fabs(U[p*lu+0]) + fabs(U[q*lu+0])) +
              fabs(d[1]) * (fabs(U[p*lu+1]) + fabs(U[q*lu+1])) +
              fabs(d[2]) * (fabs(U[p*lu+2]) + fabs(U[q*lu+2]));
  if (k->stab) {
    const double* G = k->G;
    const int lg = k->ldg;
    double dx = k->x[p] - k->x[q];
    double dy = k->y[p] - k->y[q];
    double dz = k->z[p] - k->z[q];
    double dl = sqrt(dx*dx + dy*dy + dz*dz);
    double p2 = k->P[q] - k->P[p];
    double D = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
    double dpx = G[p*lg+0] + G[q*lg+0];
    double dpy = G[p*lg+1] + G[q*lg+1];
    double dpz = G[p*lg+2] + G[q*lg+2];
    double p4 = 0.5 * (dx*dpx + dy*dpy + dz*dpz);
    dv += D*k->dt/dl*(p2 + p4);
    mv += fabs(D*k->dt/dl) * (fabs(k->P[q]) + fabs(k->P[p]) +
          0.5*(fabs(dx)*(fabs(G[p*lg+0]) + fabs(G[q*lg+0])) +
               fabs(dy)*(fabs(G[p*lg+1]) + fabs(G[q*lg+1])) +
               fabs(dz)*(fabs(G[p*lg+2]) + fabs(G[q*lg+2]))));
  }
  f[0] = dv;
  m[0] = mv;
}

void orc_chorin_div(const int64_t* nsup, const int64_t* sup0, const double* int0,
                    const int64_t* sup1, const double* int1, const int64_t* sup2,
                    const double* int2, const double* x, const double* y,
                    const double* z, int64_t ntri, const int64_t* tri, double dt,
                    const double* P, const double* G, int ldg, const double* U,
                    int ldu, double* D, int stab, double* A) {
  div_ctx k = { x, y, z, P, G, U, ldg, ldu, stab, dt };
  superedges(nsup, sup0, int0, sup1, int1, sup2, int2, 1, div_edge, &k, D, 1, A);
  /* boundary integral, Chorin.cpp:165-200 */
  for (int64_t e = 0; e < ntri; ++e) {
    const int64_t* N = tri + e*3;
    double n[3];
    crossdiv(x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]],
             x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]], 6.0, n);
    double uA[3], uB[3], uC[3];
    for (int j = 0; j < 3; ++j) {
      uA[j] = U[N[0]*ldu+j]; uB[j] = U[N[1]*ldu+j]; uC[j] = U[N[2]*ldu+j];
    }
    double w[3][3] = {{6.0,1.0,1.0},{1.0,6.0,1.0},{1.0,1.0,6.0}};
    for (int a = 0; a < 3; ++a) {
      double ux = (w[a][0]*uA[0] + w[a][1]*uB[0] + w[a][2]*uC[0])/8.0;
      double uy = (w[a][0]*uA[1] + w[a][1]*uB[1] + w[a][2]*uC[1])/8.0;
      double uz = (w[a][0]*uA[2] + w[a][1]*uB[2] + w[a][2]*uC[2])/8.0;
      D[N[a]] += ux*n[0] + uy*n[1] + uz*n[2];
      if (A) {
        double m = 0.0;
        for (int j = 0; j < 3; ++j)
          m += fabs(n[j]) * (w[a][0]*fabs(uA[j]) + w[a][1]*fabs(uB[j]) +
                             w[a][2]*fabs(uC[j])) / 8.0;
        A[N[a]] += m;
      }
    }
  }
}

/* --------------------------------------------------------- vgrad, grad -- */

typedef struct { const double* U; int ldu, ncomp; } grad_ctx;

/* d[j]*(u_q + u_p) per component and direction, Chorin.cpp:245-262 */
static void grad_edge(void* vc, const double* d, int64_t p, int64_t q,
                      double* f, double* m) {
  const grad_ctx* k = (const grad_ctx*)vc;
  for (int i = 0; i < k->ncomp; ++i) {
    double up = k->U[p*k->ldu+i], uq = k->U[q*k->ldu+i];
    for (int j = 0; j < 3; ++j) {
      f[i*3+j] = d[j] * (uq + up);
      m[i*3+j] = fabs(d[j]) * (fabs(uq) + fabs(up));
    }
  }
}

static void grad_bnd(const double* x, const double* y, const double* z,
                     int64_t ntri, const int64_t* tri, const double* U, int ldu,
                     int ncomp, double* G, int ldg, double* A) {
  for (int64_t e = 0; e < ntri; ++e) {
    const int64_t* N = tri + e*3;
    double n[3];
    crossdiv(x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]],
             x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]], 6.0, n);
    for (int i = 0; i < ncomp; ++i) {
      double u[3] = { U[N[0]*ldu+i], U[N[1]*ldu+i], U[N[2]*ldu+i] };
      double fa[3] = { (6.0*u[0] + u[1] + u[2])/8.0, (u[0] + 6.0*u[1] + u[2])/8.0,
                       (u[0] + u[1] + 6.0*u[2])/8.0 };
      double ma[3] = { (6.0*fabs(u[0]) + fabs(u[1]) + fabs(u[2]))/8.0,
                       (fabs(u[0]) + 6.0*fabs(u[1]) + fabs(u[2]))/8.0,
                       (fabs(u[0]) + fabs(u[1]) + 6.0*fabs(u[2]))/8.0 };
      for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 3; ++j) {
          G[N[a]*ldg + i*3+j] += fa[a] * n[j];
          if (A) A[N[a]*ldg + i*3+j] += ma[a] * fabs(n[j]);
        }
    }
  }
}

void orc_chorin_vgrad(const int64_t* nsup, const int64_t* sup0, const double* int0,
                      const int64_t* sup1, const double* int1, const int64_t* sup2,
                      const double* int2, const double* x, const double* y,
                      const double* z, int64_t ntri, const int64_t* tri,
                      const double* U, int ncomp, double* G, double* A) {
  if (ncomp < 1 || ncomp > 5) return;
  grad_ctx k = { U, ncomp, ncomp };
  superedges(nsup, sup0, int0, sup1, int1, sup2, int2, 3*ncomp, grad_edge, &k,
             G, 3*ncomp, A);
  grad_bnd(x, y, z, ntri, tri, U, ncomp, ncomp, G, 3*ncomp, A);
}

void orc_chorin_grad(const int64_t* nsup, const int64_t* sup0, const double* int0,
                     const int64_t* sup1, const double* int1, const int64_t* sup2,
                     const double* int2, const double* x, const double* y,
                     const double* z, int64_t ntri, const int64_t* tri,
                     const double* U, double* G, int ldg, double* A) {
  /* G has ldg >= 3 columns; the gradient adds to columns 0..2 */
  grad_ctx k = { U, 1, 1 };
  superedges(nsup, sup0, int0, sup1, int1, sup2, int2, 3, grad_edge, &k, G, ldg, A);
  grad_bnd(x, y, z, ntri, tri, U, 1, 1, G, ldg, A);
}

/* ----------------------------------------------------------------- flux -- */

typedef struct { const double *U, *G; int ldu, ldg; double mu; } flux_ctx;

/* momentum flux over edge p-q, Chorin.cpp:451-479 (value and magnitude) */
static double flux_pq(const flux_ctx* k, int i, int j, int64_t p, int64_t q,
                      double* m) {
  const double* U = k->U;
  const double* G = k->G;
  const int lu = k->ldu, lg = k->ldg;
  double inv = U[p*lu+i]*U[p*lu+j] + U[q*lu+i]*U[q*lu+j];
  double mi = fabs(U[p*lu+i]*U[p*lu+j]) + fabs(U[q*lu+i]*U[q*lu+j]);
  const double eps = 2.220446049250313e-16;
  if (k->mu < eps) { *m = mi; return -inv; }
  double vis = G[p*lg+i*3+j] + G[p*lg+j*3+i] + G[q*lg+i*3+j] + G[q*lg+j*3+i];
  double mv = fabs(G[p*lg+i*3+j]) + fabs(G[p*lg+j*3+i]) + fabs(G[q*lg+i*3+j]) +
              fabs(G[q*lg+j*3+i]);
  if (i == j) {
    vis -= 2.0/3.0 * (G[p*lg+0] + G[p*lg+4] + G[p*lg+8] + G[q*lg+0] + G[q*lg+4] +
                      G[q*lg+8]);
    mv += 2.0/3.0 * (fabs(G[p*lg+0]) + fabs(G[p*lg+4]) + fabs(G[p*lg+8]) +
                     fabs(G[q*lg+0]) + fabs(G[q*lg+4]) + fabs(G[q*lg+8]));
  }
  *m = k->mu*mv + mi;
  return k->mu*vis - inv;
}

/* momentum flux in point p, Chorin.cpp:481-506 */
static double flux_p(const flux_ctx* k, int i, int j, int64_t p, double* m) {
  const double* U = k->U;
  const double* G = k->G;
  const int lu = k->ldu, lg = k->ldg;
  double inv = U[p*lu+i]*U[p*lu+j];
  double mi = fabs(inv);
  const double eps = 2.220446049250313e-16;
  if (k->mu < eps) { *m = mi; return -inv; }
  double vis = G[p*lg+i*3+j] + G[p*lg+j*3+i];
  double mv = fabs(G[p*lg+i*3+j]) + fabs(G[p*lg+j*3+i]);
  if (i == j) {
    vis -= 2.0/3.0 * (G[p*lg+0] + G[p*lg+4] + G[p*lg+8]);
    mv += 2.0/3.0 * (fabs(G[p*lg+0]) + fabs(G[p*lg+4]) + fabs(G[p*lg+8]));
  }
  *m = k->mu*mv + mi;
  return k->mu*vis - inv;
}

void orc_chorin_flux(const int64_t* nsup, const int64_t* sup0, const double* int0,
                     const int64_t* sup1, const double* int1, const int64_t* sup2,
                     const double* int2, const double* x, const double* y,
                     const double* z, int64_t ntri, const int64_t* tri,
                     int64_t npoin, const double* U, int ldu, const double* G,
                     int ldg, double mu, double* F, double* A) {
  (void)npoin;
  flux_ctx k = { U, G, ldu, ldg, mu };
  /* per superedge the (i, j) loops are outermost, as in the reference, so
   * every F(N,i) receives its terms in the same order */
  static const int tp[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
  for (int64_t e = 0; e < nsup[0]; ++e) {
    const int64_t* N = sup0 + e*4;
    const double* d = int0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double f[6], m[6];
        for (int s = 0; s < 6; ++s) {
          double mm;
          f[s] = d[(e*6+s)*5+j] * flux_pq(&k, i, j, N[tp[s][1]], N[tp[s][0]], &mm);
          m[s] = fabs(d[(e*6+s)*5+j]) * mm;
        }
        F[N[0]*3+i] = F[N[0]*3+i] - f[0] + f[2] - f[3];
        F[N[1]*3+i] = F[N[1]*3+i] + f[0] - f[1] - f[4];
        F[N[2]*3+i] = F[N[2]*3+i] + f[1] - f[2] - f[5];
        F[N[3]*3+i] = F[N[3]*3+i] + f[3] + f[4] + f[5];
        if (A) {
          A[N[0]*3+i] += m[0] + m[2] + m[3];
          A[N[1]*3+i] += m[0] + m[1] + m[4];
          A[N[2]*3+i] += m[1] + m[2] + m[5];
          A[N[3]*3+i] += m[3] + m[4] + m[5];
        }
      }
  }
  static const int tt[3][2] = {{0,1},{1,2},{2,0}};
  for (int64_t e = 0; e < nsup[1]; ++e) {
    const int64_t* N = sup1 + e*3;
    const double* d = int1;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double f[3], m[3];
        for (int s = 0; s < 3; ++s) {
          double mm;
          f[s] = d[(e*3+s)*5+j] * flux_pq(&k, i, j, N[tt[s][1]], N[tt[s][0]], &mm);
          m[s] = fabs(d[(e*3+s)*5+j]) * mm;
        }
        F[N[0]*3+i] = F[N[0]*3+i] - f[0] + f[2];
        F[N[1]*3+i] = F[N[1]*3+i] + f[0] - f[1];
        F[N[2]*3+i] = F[N[2]*3+i] + f[1] - f[2];
        if (A) {
          A[N[0]*3+i] += m[0] + m[2];
          A[N[1]*3+i] += m[0] + m[1];
          A[N[2]*3+i] += m[1] + m[2];
        }
      }
  }
  for (int64_t e = 0; e < nsup[2]; ++e) {
    const int64_t* N = sup2 + e*2;
    const double* d = int2 + e*5;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double mm;
        double f = d[j] * flux_pq(&k, i, j, N[1], N[0], &mm);
        F[N[0]*3+i] -= f;
        F[N[1]*3+i] += f;
        if (A) { A[N[0]*3+i] += fabs(d[j])*mm; A[N[1]*3+i] += fabs(d[j])*mm; }
      }
  }
  /* boundary integral, Chorin.cpp:585-627 */
  for (int64_t e = 0; e < ntri; ++e) {
    const int64_t* N = tri + e*3;
    double n[3];
    crossdiv(x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]],
             x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]], 6.0, n);
    for (int i = 0; i < 3; ++i) {
      double fv[3][3], fm[3][3];   /* [node][direction] */
      for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 3; ++j) fv[a][j] = flux_p(&k, i, j, N[a], &fm[a][j]);
      double w[3][3] = {{6.0,1.0,1.0},{1.0,6.0,1.0},{1.0,1.0,6.0}};
      for (int a = 0; a < 3; ++a) {
        double fx = (w[a][0]*fv[0][0] + w[a][1]*fv[1][0] + w[a][2]*fv[2][0])/8.0;
        double fy = (w[a][0]*fv[0][1] + w[a][1]*fv[1][1] + w[a][2]*fv[2][1])/8.0;
        double fz = (w[a][0]*fv[0][2] + w[a][1]*fv[1][2] + w[a][2]*fv[2][2])/8.0;
        F[N[a]*3+i] += fx*n[0] + fy*n[1] + fz*n[2];
        if (A) {
          double m = 0.0;
          for (int j = 0; j < 3; ++j)
            m += fabs(n[j]) * (w[a][0]*fm[0][j] + w[a][1]*fm[1][j] + w[a][2]*fm[2][j]) / 8.0;
          A[N[a]*3+i] += m;
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ adv -- */

typedef struct {
  const orc_chocfg* cfg;
  const double *U, *G, *P, *x, *y, *z;
  int ncomp;
} adv_ctx;

/* adv_damp2, Chorin.cpp:640-708 */
static void adv_damp2(void* vc, const double* s, int64_t p, int64_t q,
                      double* f, double* m) {
  const adv_ctx* k = (const adv_ctx*)vc;
  const orc_chocfg* cf = k->cfg;
  const double* U = k->U;
  const int nc = k->ncomp;
  double nx = s[0], ny = s[1], nz = s[2];
  double uL = U[p*nc+0], vL = U[p*nc+1], wL = U[p*nc+2];
  double vnL = uL*nx + vL*ny + wL*nz;
  double uR = U[q*nc+0], vR = U[q*nc+1], wR = U[q*nc+2];
  double vnR = uR*nx + vR*ny + wR*nz;
  double mnL = fabs(uL*nx) + fabs(vL*ny) + fabs(wL*nz);
  double mnR = fabs(uR*nx) + fabs(vR*ny) + fabs(wR*nz);
  double aw = 0.0, maw = 0.0;
  if (cf->stab) { aw = fabs(vnL + vnR) / 2.0; maw = (mnL + mnR) / 2.0; }
  if (cf->stab2) {
    double sl = fabs(vnL), sr = fabs(vnR);
    aw += cf->stab2coef * (sl > sr ? sl : sr);
    maw += cf->stab2coef * (mnL > mnR ? mnL : mnR);
  }
  double v = s[4] * cf->mu;
  double pf = k->P[p] + k->P[q];
  double mpf = fabs(k->P[p]) + fabs(k->P[q]);
  double L[3] = { uL, vL, wL }, Rr[3] = { uR, vR, wR }, n[3] = { nx, ny, nz };
  for (int c = 0; c < 3; ++c) {
    f[c] = L[c]*vnL + Rr[c]*vnR + pf*n[c] + (aw-v)*(Rr[c]-L[c]);
    m[c] = fabs(L[c])*mnL + fabs(Rr[c])*mnR + mpf*fabs(n[c]) +
           (maw + fabs(v))*(fabs(Rr[c]) + fabs(L[c]));
  }
  if (nc == 3) return;
  double dd = s[4] * cf->diff;
  for (int c = 3; c < nc; ++c) {
    f[c] = U[p*nc+c]*vnL + U[q*nc+c]*vnR + (aw-dd)*(U[q*nc+c]-U[p*nc+c]);
    m[c] = fabs(U[p*nc+c])*mnL + fabs(U[q*nc+c])*mnR +
           (maw + fabs(dd))*(fabs(U[q*nc+c]) + fabs(U[p*nc+c]));
  }
}

/* adv_damp4, Chorin.cpp:710-829 */
static void adv_damp4(void* vc, const double* s, int64_t p, int64_t q,
                      double* f, double* m) {
  const adv_ctx* k = (const adv_ctx*)vc;
  const orc_chocfg* cf = k->cfg;
  const double* U = k->U;
  const double* G = k->G;
  const int nc = k->ncomp;
  double dx = k->x[q] - k->x[p];
  double dy = k->y[q] - k->y[p];
  double dz = k->z[q] - k->z[p];
  double uL[16], uR[16], mL[16], mR[16];
  for (int i = 0; i < nc; ++i) { uL[i] = U[p*nc+i]; uR[i] = U[q*nc+i]; }
  for (int c = 0; c < nc; ++c) {
    double g1 = G[p*nc*3+c*3+0]*dx + G[p*nc*3+c*3+1]*dy + G[p*nc*3+c*3+2]*dz;
    double g2 = G[q*nc*3+c*3+0]*dx + G[q*nc*3+c*3+1]*dy + G[q*nc*3+c*3+2]*dz;
    double mg1 = fabs(G[p*nc*3+c*3+0]*dx) + fabs(G[p*nc*3+c*3+1]*dy) +
                 fabs(G[p*nc*3+c*3+2]*dz);
    double mg2 = fabs(G[q*nc*3+c*3+0]*dx) + fabs(G[q*nc*3+c*3+1]*dy) +
                 fabs(G[q*nc*3+c*3+2]*dz);
    double delta2 = uR[c] - uL[c];
    double delta1 = 2.0 * g1 - delta2;
    double delta3 = 2.0 * g2 - delta2;
    double rL = (delta2 + muscl_eps) / (delta1 + muscl_eps);
    double rR = (delta2 + muscl_eps) / (delta3 + muscl_eps);
    double rLinv = (delta1 + muscl_eps) / (delta2 + muscl_eps);
    double rRinv = (delta3 + muscl_eps) / (delta2 + muscl_eps);
    double phiL = (fabs(rL) + rL) / (fabs(rL) + 1.0);
    double phiR = (fabs(rR) + rR) / (fabs(rR) + 1.0);
    double phi_L_inv = (fabs(rLinv) + rLinv) / (fabs(rLinv) + 1.0);
    double phi_R_inv = (fabs(rRinv) + rRinv) / (fabs(rRinv) + 1.0);
    double md2 = fabs(uR[c]) + fabs(uL[c]);
    mL[c] = fabs(uL[c]) + 0.5*(2.0*mg1 + md2) + 0.5*md2;
    mR[c] = fabs(uR[c]) + 0.5*(2.0*mg2 + md2) + 0.5*md2;
    uL[c] += 0.25*(delta1*(1.0-muscl_const)*phiL + delta2*(1.0+muscl_const)*phi_L_inv);
    uR[c] -= 0.25*(delta3*(1.0-muscl_const)*phiR + delta2*(1.0+muscl_const)*phi_R_inv);
  }
  double nx = s[0], ny = s[1], nz = s[2];
  double vnL = uL[0]*nx + uL[1]*ny + uL[2]*nz;
  double vnR = uR[0]*nx + uR[1]*ny + uR[2]*nz;
  double mnL = mL[0]*fabs(nx) + mL[1]*fabs(ny) + mL[2]*fabs(nz);
  double mnR = mR[0]*fabs(nx) + mR[1]*fabs(ny) + mR[2]*fabs(nz);
  double aw = 0.0, maw = 0.0;
  if (cf->stab) { aw = fabs(vnL + vnR) / 2.0; maw = (mnL + mnR) / 2.0; }
  if (cf->stab2) {
    double sl = fabs(vnL), sr = fabs(vnR);
    aw += cf->stab2coef * (sl > sr ? sl : sr);
    maw += cf->stab2coef * (mnL > mnR ? mnL : mnR);
  }
  double v = s[4] * cf->mu;
  double pf = k->P[p] + k->P[q];
  double mpf = fabs(k->P[p]) + fabs(k->P[q]);
  double n[3] = { nx, ny, nz };
  for (int c = 0; c < 3; ++c) {
    f[c] = uL[c]*vnL + uR[c]*vnR + pf*n[c] + aw*(uR[c]-uL[c]) -
           v*(U[q*nc+c]-U[p*nc+c]);
    m[c] = mL[c]*mnL + mR[c]*mnR + mpf*fabs(n[c]) + maw*(mR[c] + mL[c]) +
           fabs(v)*(fabs(U[q*nc+c]) + fabs(U[p*nc+c]));
  }
  if (nc == 3) return;
  double dd = s[4] * cf->diff;
  for (int c = 3; c < nc; ++c) {
    f[c] = uL[c]*vnL + uR[c]*vnR + aw*(uR[c]-uL[c]) - dd*(U[q*nc+c]-U[p*nc+c]);
    m[c] = mL[c]*mnL + mR[c]*mnR + maw*(mR[c] + mL[c]) +
           fabs(dd)*(fabs(U[q*nc+c]) + fabs(U[p*nc+c]));
  }
}

int orc_chorin_rhs(const orc_chocfg* cfg, int64_t npoin, int ncomp,
                   const int64_t* nsup, const int64_t* sup0, const double* int0,
                   const int64_t* sup1, const double* int1, const int64_t* sup2,
                   const double* int2, const double* x, const double* y,
                   const double* z, int64_t ntri, const int64_t* tri,
                   const double* P, const double* U, const double* G, double* R,
                   double* A) {
  if (cfg->flux != 0 && cfg->flux != 1) return -1;   /* Chorin.cpp:862-866 */
  if (ncomp < 3 || ncomp > 16) return -2;
  memset(R, 0, sizeof(double)*(size_t)npoin*(size_t)ncomp);     /* :1041 */
  if (A) memset(A, 0, sizeof(double)*(size_t)npoin*(size_t)ncomp);
  adv_ctx k = { cfg, U, G, P, x, y, z, ncomp };
  superedges(nsup, sup0, int0, sup1, int1, sup2, int2, ncomp,
             cfg->flux == 0 ? adv_damp2 : adv_damp4, &k, R, ncomp, A);
  /* boundary integral, Chorin.cpp:930-976 */
  for (int64_t e = 0; e < ntri; ++e) {
    const int64_t* N = tri + e*3;
    double n[3];
    crossdiv(x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]],
             x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]], 6.0, n);
    double f[16][3], fm[16][3];
    for (int a = 0; a < 3; ++a) {
      const double* u = U + N[a]*ncomp;
      double pr = P[N[a]];
      double vn = n[0]*u[0] + n[1]*u[1] + n[2]*u[2];
      double mvn = fabs(n[0]*u[0]) + fabs(n[1]*u[1]) + fabs(n[2]*u[2]);
      for (int c = 0; c < 3; ++c) {
        f[c][a] = u[c]*vn + pr*n[c];
        fm[c][a] = fabs(u[c])*mvn + fabs(pr*n[c]);
      }
      for (int c = 3; c < ncomp; ++c) {
        f[c][a] = u[c]*vn;
        fm[c][a] = fabs(u[c])*mvn;
      }
    }
    for (int c = 0; c < ncomp; ++c) {
      R[N[0]*ncomp+c] += (6.0*f[c][0] + f[c][1] + f[c][2])/8.0;
      R[N[1]*ncomp+c] += (f[c][0] + 6.0*f[c][1] + f[c][2])/8.0;
      R[N[2]*ncomp+c] += (f[c][0] + f[c][1] + 6.0*f[c][2])/8.0;
      if (A) {
        A[N[0]*ncomp+c] += (6.0*fm[c][0] + fm[c][1] + fm[c][2])/8.0;
        A[N[1]*ncomp+c] += (fm[c][0] + 6.0*fm[c][1] + fm[c][2])/8.0;
        A[N[2]*ncomp+c] += (fm[c][0] + fm[c][1] + 6.0*fm[c][2])/8.0;
      }
    }
  }
  return 0;
}

/* ChoCG::pinit's Neumann vector (ChoCG.cpp:1070-1095): every boundary face
 * flagged in besym subtracts, at each of its three nodes, the dot product of
 * its third of the area vector, crossdiv(x1 - x0, x2 - x0, 6), with the
 * pressure gradient pg (npoin x 3) at that node.  neubc is zeroed first. */
void orc_neumann(int64_t npoin, const double* x, const double* y, const double* z,
                 int64_t ntri, const int64_t* tri, const unsigned char* besym,
                 const double* pg, double* neubc) {
  for (int64_t i = 0; i < npoin; ++i) neubc[i] = 0.0;
  for (int64_t e = 0; e < ntri; ++e) {
    if (!besym[e]) continue;
    const int64_t* N = tri + 3*e;
    double n[3];
    crossdiv(x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]],
             x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]], 6.0, n);
    for (int k = 0; k < 3; ++k) {
      const double* g = pg + 3*N[k];
      neubc[N[k]] -= n[0]*g[0] + n[1]*g[1] + n[2]*g[2];
    }
  }
}
