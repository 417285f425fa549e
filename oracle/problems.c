/* CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * The analytic problems the path evaluates (initial conditions and source
 * terms of Problems.cpp), riemann::src, the RieCG boundary conditions other
 * than symmetry (BC.cpp dirbc/farbc/prebc), the steady-state (local time
 * stepping) branch of RieCG::dt and RieCG::rhs/solve, and the analytic-error
 * columns of NodeDiagnostics::rhocompute.  Each function names the reference
 * lines it restates.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xyst_oracle.h"

static const double kPi = 3.14159265358979323846;

/* eos::totalenergy, EOS.hpp:53-56 */
static double totalenergy(double g, double r, double u, double v, double w,
                          double p) {
  return p / (g - 1.0) + 0.5 * r * (u*u + v*v + w*w);
}

/* vortical_flow::ic, Problems.cpp:455-478 */
static void vf_ic(const orc_problem* P, double x, double y, double z,
                  double* s) {
  const double a = P->alpha, k = P->kappa, p0 = P->p0, g = P->gamma;
  double ru = a*x - k*y;
  double rv = k*x + a*y;
  double rw = -2.0*a*z;
  double rE = (ru*ru + rv*rv + rw*rw)/2.0 + (p0 - 2.0*a*a*z*z) / (g - 1.0);
  s[0] = 1.0; s[1] = ru; s[2] = rv; s[3] = rw; s[4] = rE;
}

/* taylor_green::ic, Problems.cpp:410-430 */
static void tg_ic(const orc_problem* P, double x, double y, double* s) {
  double r = 1.0;
  double p = 10.0 + r/4.0*(cos(2.0*kPi*x) + cos(2.0*kPi*y));
  double u =  sin(kPi*x) * cos(kPi*y);
  double v = -cos(kPi*x) * sin(kPi*y);
  double w = 0.0;
  double rE = totalenergy(P->gamma, r, u, v, w, p);
  s[0] = r; s[1] = r*u; s[2] = r*v; s[3] = r*w; s[4] = rE;
}

/* slot_cyl::ic, Problems.cpp:513-623, RieCG/KozCG form (ncomp 6): unit
 * density, the rigid rotation u = 0.5 - y, v = x - 0.5 about (0.5, 0.5),
 * pressure 1, and one scalar c0 holding a cone, a hump and a slotted
 * cylinder (radius 0.15) rotated by the angle t from their initial centres
 * (0.5, 0.25), (0.25, 0.5), (0.5, 0.75); later shapes overwrite earlier
 * ones where they overlap. */
static void sc_ic(const orc_problem* P, double x, double y, double t,
                  double* s) {
  const double R0 = 0.15;
  double r = 1.0, u = 0.5 - y, v = x - 0.5;
  s[0] = r; s[1] = r*u; s[2] = r*v; s[3] = 0.0;
  s[4] = totalenergy(P->gamma, r, u, v, 0.0, 1.0);
  s[5] = 0.0;
  /* a centre (x0, y0) at angle t: distance rc from the rotation axis */
  double rc = sqrt((0.5-0.5)*(0.5-0.5) + (0.25-0.5)*(0.25-0.5));
  double kx = 0.5 + rc*sin(t), ky = 0.5 - rc*cos(t);               /* cone */
  rc = sqrt((0.25-0.5)*(0.25-0.5) + (0.5-0.5)*(0.5-0.5));
  double hx = 0.5 + rc*sin(t - kPi/2.0), hy = 0.5 - rc*cos(t - kPi/2.0);   /* hump */
  rc = sqrt((0.5-0.5)*(0.5-0.5) + (0.75-0.5)*(0.75-0.5));
  double cx = 0.5 + rc*sin(t + kPi), cy = 0.5 - rc*cos(t + kPi);   /* cylinder */
  /* the slot: three corner points, rotated by t about (0.5, 0.5) */
  double px[3] = {0.525, 0.525, 0.475};
  double py[3] = {cy - rc*cos(asin(0.025/rc)), 0.8, 0.8};
  double qx[3], qy[3];
  for (int k = 0; k < 3; ++k) {
    qx[k] = 0.5 + cos(t)*(px[k]-0.5) - sin(t)*(py[k]-0.5);
    qy[k] = 0.5 + sin(t)*(px[k]-0.5) + cos(t)*(py[k]-0.5);
  }
  double e1x = qx[1]-qx[0], e1y = qy[1]-qy[0];
  double e2x = qx[2]-qx[1], e2y = qy[2]-qy[1];
  double l1 = sqrt(e1x*e1x + e1y*e1y), l2 = sqrt(e2x*e2x + e2y*e2y);
  double d = sqrt((x-kx)*(x-kx) + (y-ky)*(y-ky)) / R0;
  if (d < 1.0) s[5] = 0.6*(1.0 - d);
  d = sqrt((x-hx)*(x-hx) + (y-hy)*(y-hy)) / R0;
  if (d < 1.0) s[5] = 0.2*(1.0 + cos(kPi*(d < 1.0 ? d : 1.0)));
  d = sqrt((x-cx)*(x-cx) + (y-cy)*(y-cy)) / R0;
  /* signed distances from the slot's long side and its end */
  double s1 = (e1x*(y-qy[0]) - (x-qx[0])*e1y) / l1;
  double s2 = (e2x*(y-qy[1]) - (x-qx[1])*e2y) / l2;
  if (d < 1.0 && (s1 > 0.05 || s1 < 0.0 || s2 < 0.0)) s[5] = 0.6;
}

int orc_problem_ic(const orc_problem* P, double x, double y, double z,
                   double t, double* s) {
  if (P->kind == 1) { vf_ic(P, x, y, z, s); return 5; }
  if (P->kind == 2) { tg_ic(P, x, y, s); return 5; }
  if (P->kind == 3) { sc_ic(P, x, y, t, s); return 6; }
  return -1;
}

int orc_problem_src(const orc_problem* P, double x, double y, double z,
                    double t, double* s) {
  for (int c = 0; c < 5; ++c) s[c] = 0.0;
  if (P->kind == 1) {                 /* vortical_flow::src, Problems.cpp:481-507 */
    const double a = P->alpha, k = P->kappa, g = P->gamma;
    double u[5];
    vf_ic(P, x, y, z, u);
    s[1] = a*u[1]/u[0] - k*u[2]/u[0];
    s[2] = k*u[1]/u[0] + a*u[2]/u[0];
    s[4] = (s[1]*u[1] + s[2]*u[2])/u[0] + 8.0*a*a*a*z*z/(g-1.0);
    return 1;
  }
  if (P->kind == 2) {                 /* taylor_green::src, Problems.cpp:433-448 */
    s[4] = 3.0*kPi/8.0*( cos(3.0*kPi*x)*cos(kPi*y)
                       - cos(3.0*kPi*y)*cos(kPi*x) );
    return 1;
  }
  if (P->kind == 3) {                 /* slot_cyl::src, Problems.cpp:625-670 */
    double u[6];
    sc_ic(P, x, y, t, u);
    s[1] = -u[2];
    s[2] = u[1];
    return 1;
  }
  return 0;
}

/* riemann::src, Riemann.cpp:880-907: R(p,c) -= s[c] * v[p] with s at
 * (x_p, t) or, in steady mode, (x_p, tp[p]). */
void orc_src(const orc_problem* P, int64_t npoin, int ncomp, const double* x,
             const double* y, const double* z, const double* v, double t,
             const double* tp, double* R) {
  if (P->kind < 1 || P->kind > 3) return;
  double s[6];
  for (int64_t p = 0; p < npoin; ++p) {
    double tt = tp ? tp[p] : t;
    orc_problem_src(P, x[p], y[p], z[p], tt, s);
    for (int c = 0; c < 5 && c < ncomp; ++c) R[p*ncomp+c] -= s[c] * v[p];
  }
}

/* physics::dirbc, BC.cpp:29-72: mask 1 sets the initial condition at time t
 * (IC(), or vals where the problem has no analytic form), mask 2 the given
 * BC value vals. */
void orc_dirbc(const orc_problem* P, int64_t npoin, int ncomp, double* U,
               double t, const double* x, const double* y, const double* z,
               int64_t ndir, const int64_t* nodes, const int32_t* mask,
               const double* vals) {
  (void)npoin;
  double u[16];
  for (int64_t i = 0; i < ndir; ++i) {
    int64_t p = nodes[i];
    int nic = orc_problem_ic(P, x[p], y[p], z[p], t, u);
    for (int c = 0; c < ncomp; ++c) {
      int m = mask[i*ncomp+c];
      if (m == 1) U[p*ncomp+c] = c < nic ? u[c] : vals[i*ncomp+c];
      else if (m == 2) U[p*ncomp+c] = vals[i*ncomp+c];
    }
  }
}

/* physics::farbc, BC.cpp:152-221: characteristic farfield from the free
 * stream far[] = {density, u, v, w, pressure}. */
void orc_farbc(double g, int ncomp, double* U, int64_t nfar,
               const int64_t* nodes, const double* norms, const double* far) {
  const double fr = far[0], fu = far[1], fv = far[2], fw = far[3], fp = far[4];
  for (int64_t i = 0; i < nfar; ++i) {
    double* s = U + nodes[i]*ncomp;
    double nx = norms[i*3+0], ny = norms[i*3+1], nz = norms[i*3+2];
    double r = s[0], ru = s[1], rv = s[2], rw = s[3], re = s[4];
    double vn = fu*nx + fv*ny + fw*nz;
    double a = sqrt(g * fp / fr);                       /* eos::soundspeed */
    double M = vn / a;
    if (M <= -1.0) {
      s[0] = fr; s[1] = fr*fu; s[2] = fr*fv; s[3] = fr*fw;
      s[4] = totalenergy(g, fr, fu, fv, fw, fp);
    } else if (M > -1.0 && M < 0.0) {
      double pr = (re - 0.5*(ru*ru + rv*rv + rw*rw)/r) * (g - 1.0);
      s[0] = fr; s[1] = fr*fu; s[2] = fr*fv; s[3] = fr*fw;
      s[4] = totalenergy(g, fr, fu, fv, fw, pr);
    } else if (M >= 0.0 && M < 1.0) {
      s[4] = totalenergy(g, r, ru/r, rv/r, rw/r, fp);
    }
  }
}

/* physics::prebc, BC.cpp:223-241: density and pressure set at the nodes,
 * momentum kept. */
void orc_prebc(double g, int ncomp, double* U, int64_t npre,
               const int64_t* nodes, const double* vals) {
  for (int64_t i = 0; i < npre; ++i) {
    double* s = U + nodes[i]*ncomp;
    s[0] = vals[i*2+0];
    s[4] = totalenergy(g, s[0], s[1]/s[0], s[2]/s[0], s[3]/s[0], vals[i*2+1]);
  }
}

/* RieCG::dt, RieCG.cpp:810-824, steady branch: per-node dtp and its minimum. */
double orc_dt_steady(const orc_cfg* cfg, int64_t npoin, int ncomp,
                     const double* U, const double* vol, double cfl,
                     double* dtp) {
  const double g = cfg->gamma;
  double mindt = DBL_MAX;
  for (int64_t p = 0; p < npoin; ++p) {
    const double* s = U + p*ncomp;
    double r = s[0];
    double u = s[1]/r, v = s[2]/r, w = s[3]/r;
    double pr = (s[4] - 0.5*r*(u*u + v*v + w*w)) * (g-1.0);
    double c = sqrt(g * (pr > 0.0 ? pr : 0.0) / r);
    double L = cbrt(vol[p]);
    double vel = sqrt(u*u + v*v + w*w);
    double den = vel + c > 1.0e-8 ? vel + c : 1.0e-8;
    dtp[p] = L / den * cfl;
    if (dtp[p] < mindt) mindt = dtp[p];
  }
  return mindt;
}

/* One RieCG time step with every boundary condition, the source term and
 * optional local time stepping (RieCG::rhs/solve/BC, RieCG.cpp:764-785,
 * 918-1057): per stage grad/vol, rhs (+ src at t, or at tp + prev_rkcoef dtp
 * when steady), Un = U at stage 0, U = Un - rkcoef dt R/vol (dtp[i] when
 * steady), then dirbc, symbc, farbc, prebc at t + rkcoef dt.  With steady,
 * tp += dtp at the end. */
int orc_riecg_step2(const orc_cfg* cfg, const orc_problem* P, int64_t npoin,
                    int ncomp, const int64_t* nsup, const int64_t* sup0,
                    const double* int0, const int64_t* sup1, const double* int1,
                    const int64_t* sup2, const double* int2, const double* x,
                    const double* y, const double* z, int64_t nbtri,
                    const int64_t* triinpoel, const uint8_t* besym,
                    const double* vol, const double* v, const orc_riecg_bc* bc,
                    double t, double dt, const double* dtp, double* tp,
                    double* U, double* Un, double* G, double* R) {
  static const double rkcoef[3] = {1.0/3.0, 1.0/2.0, 1.0};
  const int ng = 3*ncomp;
  for (int stage = 0; stage < 3; ++stage) {
    orc_grad(npoin, ncomp, nsup, sup0, int0, sup1, int1, sup2, int2, x, y, z,
             nbtri, triinpoel, U, G, NULL);
    for (int64_t p = 0; p < npoin; ++p)
      for (int c = 0; c < ng; ++c) G[p*ng+c] /= vol[p];
    double prev = stage == 0 ? 0.0 : rkcoef[stage-1];
    if (dtp) for (int64_t p = 0; p < npoin; ++p) tp[p] += prev * dtp[p];
    int rc = orc_rhs(cfg, npoin, ncomp, nsup, sup0, int0, sup1, int1, sup2,
                     int2, x, y, z, nbtri, triinpoel, besym, G, U, R, NULL);
    if (rc) return rc;
    orc_src(P, npoin, ncomp, x, y, z, v, t, dtp ? tp : NULL, R);
    if (dtp) for (int64_t p = 0; p < npoin; ++p) tp[p] -= prev * dtp[p];
    if (stage == 0) memcpy(Un, U, sizeof(double)*(size_t)(npoin*ncomp));
    for (int64_t i = 0; i < npoin; ++i) {
      double dti = dtp ? dtp[i] : dt;
      for (int c = 0; c < ncomp; ++c)
        U[i*ncomp+c] = Un[i*ncomp+c] - rkcoef[stage] * dti * R[i*ncomp+c] / vol[i];
    }
    double tb = t + rkcoef[stage] * dt;
    orc_dirbc(P, npoin, ncomp, U, tb, x, y, z, bc->ndir, bc->dirnodes,
              bc->dirmask, bc->dirvals);
    orc_symbc(npoin, ncomp, U, bc->nsym, bc->symnodes, bc->symnorms);
    orc_farbc(cfg->gamma, ncomp, U, bc->nfar, bc->farnodes, bc->farnorms, bc->far);
    orc_prebc(cfg->gamma, ncomp, U, bc->npre, bc->prenodes, bc->prevals);
  }
  if (dtp) for (int64_t p = 0; p < npoin; ++p) tp[p] += dtp[p];
  return 0;
}

/* NodeDiagnostics::rhocompute, NodeDiagnostics.cpp:84-128, the analytic-error
 * sums: An holds the analytic solution in the primitive form the reference
 * compares (rho, u, v, w, specific internal energy, then the scalars);
 * out[c] = sum (nu_c - An_c)^2 v, out[ncomp+c] = sum |nu_c - An_c| v with nu
 * the numerical solution in the same form. */
void orc_diag_err(int64_t npoin, int ncomp, const double* U, const double* v,
                  const double* An, double* out) {
  memset(out, 0, sizeof(double) * (size_t)(2*ncomp));
  double nu[16];
  for (int64_t i = 0; i < npoin; ++i) {
    for (int c = 0; c < ncomp; ++c) nu[c] = U[i*ncomp+c];
    nu[1] /= nu[0];
    nu[2] /= nu[0];
    nu[3] /= nu[0];
    nu[4] = nu[4] / nu[0] - 0.5*(nu[1]*nu[1] + nu[2]*nu[2] + nu[3]*nu[3]);
    for (int c = 0; c < ncomp; ++c) {
      double du = nu[c] - An[i*ncomp+c];
      out[c] += du * du * v[i];
      out[ncomp+c] += fabs(du) * v[i];
    }
  }
}
