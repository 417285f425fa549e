/* CPU oracle: RieCG edge-based Riemann/MUSCL path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates src/Physics/Riemann.cpp, src/Physics/BC.cpp (symbc) and the RieCG
 * chare loops (src/Inciter/RieCG.cpp) in plain C, loop for loop, with the same
 * FP64 arithmetic and AoS fields.  Only tests/, the smoke check and bench.py's
 * CPU-baseline leg use it.
 *
 * Pinning (the checks that keep this restatement honest):
 *   - orc_grad/orc_rhs/orc_dt/orc_symbc_setup/orc_riecg_step: the reference
 *     RieCG/Sod regression fixture (tests/golden/riecg_sod.npz, extracted from
 *     tests/regression/inciter/RieCG/Sod/{sod.std.exo,diag.std}) at the
 *     reference's exodiff tolerances, plus free-stream preservation, linear
 *     exactness of grad, conservation and brute-force dt.
 *   - orc_domint/orc_domsuped: D antisymmetry/closure identities and a
 *     brute-force quadrature of the dual-face normal on tiny meshes; every edge
 *     appears exactly once in the superedges.
 */
#include "xyst_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

static const double muscl_eps = 1.0e-9;        /* Riemann.cpp:29 */
static const double muscl_const = 1.0 / 3.0;   /* Riemann.cpp:30 */

/* tet edge and face tables, DerivedData.hpp:36-44 */
static const int lpoed[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
static const int lpofa[4][3] = {{1,2,3},{2,0,3},{3,0,1},{0,2,1}};
static const int lpoet[3][2] = {{0,1},{1,2},{2,0}};

/* tk::cross, Vector.hpp:36-58 */
static void cross(const double* a, const double* b, double* r) {
  r[0] = a[1]*b[2] - b[1]*a[2];
  r[1] = a[2]*b[0] - b[2]*a[0];
  r[2] = a[0]*b[1] - b[0]*a[1];
}

int orc_vol(int64_t npoin, const double* x, const double* y, const double* z,
            int64_t ntet, const int64_t* inpoel, double* vol) {
  memset(vol, 0, sizeof(double) * (size_t)npoin);
  int bad = 0;
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
    double c[3];
    cross(ca, da, c);   /* triple(ba,ca,da) = dot(ba, cross(ca,da)) */
    double J = (ba[0]*c[0] + ba[1]*c[1] + ba[2]*c[2]) / 24.0;
    if (!(J > 0)) bad = 1;
    for (int j = 0; j < 4; ++j) vol[N[j]] += J;
  }
  return bad ? -1 : 0;
}

/* ---------------------------------------------------------------- domint -- */

typedef struct { int64_t p, q, rec; } edrec;

static int edrec_cmp(const void* a, const void* b) {
  const edrec* u = (const edrec*)a; const edrec* v = (const edrec*)b;
  if (u->p != v->p) return u->p < v->p ? -1 : 1;
  if (u->q != v->q) return u->q < v->q ? -1 : 1;
  return (u->rec > v->rec) - (u->rec < v->rec);
}

/* stride 3: RieCG::domint (RieCG.cpp:339-382); stride 5: ChoCG::domint
 * (ChoCG.cpp:398-445), which also sums J/120 and dot(grad_p,grad_q)/J/6 */
static int64_t domint_stride(int stride, const double* x, const double* y,
                             const double* z, int64_t ntet,
                             const int64_t* inpoel, int64_t cap, int64_t* edge,
                             double* dint) {
  int64_t nrec = ntet * 6;
  edrec* r = (edrec*)malloc(sizeof(edrec) * (size_t)nrec);
  double* c = (double*)malloc(sizeof(double) * (size_t)stride * (size_t)nrec);
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
    double grad[4][3];
    cross(ca, da, grad[1]);
    cross(da, ba, grad[2]);
    cross(ba, ca, grad[3]);
    for (int i = 0; i < 3; ++i)
      grad[0][i] = -grad[1][i]-grad[2][i]-grad[3][i];
    const double J = ba[0]*grad[1][0] + ba[1]*grad[1][1] + ba[2]*grad[1][2];
    for (int k = 0; k < 6; ++k) {
      int p = lpoed[k][0], q = lpoed[k][1];
      int64_t a = N[p], b = N[q];
      double sig = 1.0;
      if (a > b) { int64_t t = a; a = b; b = t; sig = -1.0; }
      int64_t i = e*6 + k;
      r[i].p = a; r[i].q = b; r[i].rec = i;
      for (int j = 0; j < 3; ++j)
        c[i*stride+j] = sig * (grad[p][j] - grad[q][j]) / 48.0;
      if (stride == 5) {
        c[i*5+3] = J / 120.0;
        c[i*5+4] = (grad[p][0]*grad[q][0] + grad[p][1]*grad[q][1] +
                    grad[p][2]*grad[q][2]) / J / 6.0;
      }
    }
  }
  qsort(r, (size_t)nrec, sizeof(edrec), edrec_cmp);
  int64_t ne = 0;
  for (int64_t i = 0; i < nrec; ++i) {
    if (i == 0 || r[i].p != r[i-1].p || r[i].q != r[i-1].q) {
      if (ne == cap) { free(r); free(c); return -1; }
      edge[ne*2+0] = r[i].p; edge[ne*2+1] = r[i].q;
      for (int j = 0; j < stride; ++j) dint[ne*stride+j] = 0.0;
      ++ne;
    }
    /* accumulation in tet order, as into the hash map at RieCG.cpp:377-379 */
    for (int j = 0; j < stride; ++j) dint[(ne-1)*stride+j] += c[r[i].rec*stride+j];
  }
  free(r); free(c);
  return ne;
}

int64_t orc_domint(int64_t npoin, const double* x, const double* y,
                   const double* z, int64_t ntet, const int64_t* inpoel,
                   int64_t cap, int64_t* edge, double* dint) {
  (void)npoin;
  return domint_stride(3, x, y, z, ntet, inpoel, cap, edge, dint);
}

int64_t orc_chocg_domint(int64_t npoin, const double* x, const double* y,
                         const double* z, int64_t ntet, const int64_t* inpoel,
                         int64_t cap, int64_t* edge, double* dint) {
  (void)npoin;
  return domint_stride(5, x, y, z, ntet, inpoel, cap, edge, dint);
}

/* ------------------------------------------------------------- domsuped -- */

/* Edge id of (a, b), a < b, in the (p, q)-sorted unique edge list, found by
 * a binary search within row a (row offsets from edge_rows). */
static int64_t edge_find(const int64_t* row, const int64_t* edge, int64_t a,
                         int64_t b) {
  int64_t lo = row[a], hi = row[a+1] - 1;
  while (lo <= hi) {
    int64_t m = (lo + hi) / 2;
    int64_t q = edge[m*2+1];
    if (q == b) return m;
    if (q < b) lo = m + 1; else hi = m - 1;
  }
  return -1;
}

/* row[a] = first edge with p >= a, for a = 0..npoin (edges sorted by p) */
static int64_t* edge_rows(int64_t nedge, const int64_t* edge, int64_t* npoin) {
  int64_t np = 0;
  for (int64_t m = 0; m < nedge; ++m) {
    if (edge[m*2] + 1 > np) np = edge[m*2] + 1;
    if (edge[m*2+1] + 1 > np) np = edge[m*2+1] + 1;
  }
  int64_t* row = (int64_t*)calloc((size_t)np + 1, sizeof(int64_t));
  for (int64_t m = 0; m < nedge; ++m) ++row[edge[m*2] + 1];
  for (int64_t a = 0; a < np; ++a) row[a+1] += row[a];
  *npoin = np;
  return row;
}

/* Superedge grouping for edge integrals of `stride` values per edge, the
 * first three of which (the normal components) change sign with the edge
 * orientation; RieCG uses stride 3, ChoCG stride 5 (ChoCG.cpp:687-813). */
static int domsuped_stride(int stride, int64_t ntet, const int64_t* inpoel,
                           int64_t nedge, const int64_t* edge,
                           const double* dint, int64_t* nsup, int64_t* sup0,
                           double* int0, int64_t* sup1, double* int1,
                           int64_t* sup2, double* int2) {
  char* used = (char*)calloc((size_t)nedge, 1);
  int64_t npoin = 0;
  int64_t* row = edge_rows(nedge, edge, &npoin);
  int64_t n0 = 0, n1 = 0, n2 = 0;
  /* tet superedges, RieCG.cpp:657-680 */
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    int64_t id[6]; double sig[6]; int f = 0;
    for (int k = 0; k < 6; ++k) {
      int64_t a = N[lpoed[k][0]], b = N[lpoed[k][1]];
      sig[k] = a < b ? 1.0 : -1.0;
      int64_t m = a < b ? edge_find(row, edge, a, b) : edge_find(row, edge, b, a);
      if (m < 0 || used[m]) break;
      id[k] = m; ++f;
    }
    if (f == 6) {
      for (int i = 0; i < 4; ++i) sup0[n0*4+i] = N[i];
      for (int k = 0; k < 6; ++k) {
        for (int j = 0; j < stride; ++j)
          int0[(n0*6+k)*stride+j] = (j < 3 ? sig[k] : 1.0) * dint[id[k]*stride+j];
        used[id[k]] = 1;
      }
      ++n0;
    }
  }
  /* triangle superedges over the unique tet faces in first-seen order (the
   * FaceSet of RieCG.cpp:646-651, node order of the first tet holding the
   * face), RieCG.cpp:683-702.  Walking every (tet, face) in order visits each
   * face first at its first-seen place; a later copy of it never qualifies
   * (taken: its edges are used; not taken: an edge was already used, and
   * edges are never released), so no deduplication is needed. */
  for (int64_t e = 0; e < ntet; ++e)
    for (int fa = 0; fa < 4; ++fa) {
      int64_t N[3];
      for (int i = 0; i < 3; ++i) N[i] = inpoel[e*4 + lpofa[fa][i]];
      int64_t id[3]; double sig[3]; int f = 0;
      for (int k = 0; k < 3; ++k) {
        int64_t a = N[lpoet[k][0]], b = N[lpoet[k][1]];
        sig[k] = a < b ? 1.0 : -1.0;
        int64_t m = a < b ? edge_find(row, edge, a, b) : edge_find(row, edge, b, a);
        if (m < 0 || used[m]) break;
        id[k] = m; ++f;
      }
      if (f == 3) {
        for (int j = 0; j < 3; ++j) sup1[n1*3+j] = N[j];
        for (int k = 0; k < 3; ++k) {
          for (int j = 0; j < stride; ++j)
            int1[(n1*3+k)*stride+j] = (j < 3 ? sig[k] : 1.0) * dint[id[k]*stride+j];
          used[id[k]] = 1;
        }
        ++n1;
      }
    }
  free(row);
  /* remaining edges, RieCG.cpp:704-716 */
  for (int64_t m = 0; m < nedge; ++m) {
    if (used[m]) continue;
    sup2[n2*2+0] = edge[m*2+0];
    sup2[n2*2+1] = edge[m*2+1];
    for (int j = 0; j < stride; ++j) int2[n2*stride+j] = dint[m*stride+j];
    ++n2;
  }
  free(used);
  nsup[0] = n0; nsup[1] = n1; nsup[2] = n2;
  return (n0*6 + n1*3 + n2 == nedge) ? 0 : -1;   /* RieCG.cpp:733-735 */
}

int orc_domsuped(int64_t ntet, const int64_t* inpoel, int64_t nedge,
                 const int64_t* edge, const double* dint, int64_t* nsup,
                 int64_t* sup0, double* int0, int64_t* sup1, double* int1,
                 int64_t* sup2, double* int2) {
  return domsuped_stride(3, ntet, inpoel, nedge, edge, dint, nsup, sup0, int0,
                         sup1, int1, sup2, int2);
}

int orc_domsuped5(int64_t ntet, const int64_t* inpoel, int64_t nedge,
                  const int64_t* edge, const double* dint, int64_t* nsup,
                  int64_t* sup0, double* int0, int64_t* sup1, double* int1,
                  int64_t* sup2, double* int2) {
  return domsuped_stride(5, ntet, inpoel, nedge, edge, dint, nsup, sup0, int0,
                         sup1, int1, sup2, int2);
}

/* ------------------------------------------------------------ symmetry BC -- */

int64_t orc_symbc_setup(int64_t npoin, const double* x, const double* y,
                        const double* z, int64_t nfaces, const int64_t* tri,
                        const int32_t* faceset, int32_t nsym,
                        const int32_t* symsets, int64_t cap,
                        int64_t* symnodes, double* symnorms, uint8_t* besym) {
  /* per symmetry set: accumulated r*n (3) and r (1), RieCG.cpp:294-333 */
  double* bn = (double*)calloc((size_t)(nsym * npoin * 4), sizeof(double));
  char* has = (char*)calloc((size_t)(nsym * npoin), 1);
  char* issym = (char*)calloc((size_t)npoin, 1);
  for (int32_t s = 0; s < nsym; ++s) {
    for (int64_t f = 0; f < nfaces; ++f) {
      if (faceset[f] != symsets[s]) continue;
      const int64_t* N = tri + f*3;
      double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
      double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
      double n[3];
      cross(ba, ca, n);
      double A2 = sqrt(n[0]*n[0] + n[1]*n[1] + n[2]*n[2]);
      n[0] /= A2; n[1] /= A2; n[2] /= A2;
      double cen[3] = {(x[N[0]] + x[N[1]] + x[N[2]]) / 3.0,
                       (y[N[0]] + y[N[1]] + y[N[2]]) / 3.0,
                       (z[N[0]] + z[N[1]] + z[N[2]]) / 3.0};
      for (int k = 0; k < 3; ++k) {
        int64_t p = N[lpoet[k][0]];
        double r = 1.0 / ((cen[0]-x[p])*(cen[0]-x[p]) +
                          (cen[1]-y[p])*(cen[1]-y[p]) +
                          (cen[2]-z[p])*(cen[2]-z[p]));
        double* b = bn + ((int64_t)s*npoin + p)*4;
        b[0] += r*n[0]; b[1] += r*n[1]; b[2] += r*n[2]; b[3] += r;
        has[(int64_t)s*npoin + p] = 1;
        issym[p] = 1;
      }
    }
  }
  /* bnorm division, RieCG.cpp:503-513; streamable lists, RieCG.cpp:574-589 */
  int64_t ns = 0;
  for (int64_t p = 0; p < npoin; ++p) {
    if (!issym[p]) continue;
    for (int32_t s = 0; s < nsym; ++s) {
      if (!has[(int64_t)s*npoin + p]) continue;
      if (ns == cap) { ns = -1; goto done; }
      double* b = bn + ((int64_t)s*npoin + p)*4;
      symnodes[ns] = p;
      symnorms[ns*3+0] = b[0] / b[3];
      symnorms[ns*3+1] = b[1] / b[3];
      symnorms[ns*3+2] = b[2] / b[3];
      ++ns;
    }
  }
  /* besym per boundary-triangle node, RieCG.cpp:533-537 */
  for (int64_t i = 0; i < nfaces*3; ++i) besym[i] = (uint8_t)issym[tri[i]];
done:
  free(bn); free(has); free(issym);
  return ns;
}

void orc_symbc(int64_t npoin, int ncomp, double* U, int64_t nsym,
               const int64_t* symnodes, const double* symnorms) {
  (void)npoin;
  for (int64_t i = 0; i < nsym; ++i) {
    int64_t p = symnodes[i];
    const double* n = symnorms + i*3;
    double* u = U + p*ncomp + 1;
    double vn = u[0]*n[0] + u[1]*n[1] + u[2]*n[2];
    u[0] -= vn * n[0];
    u[1] -= vn * n[1];
    u[2] -= vn * n[2];
  }
}

/* ------------------------------------------------------------------ grad -- */

/* riemann::primitive, Riemann.cpp:211-227 */
static void primitive(int ncomp, int64_t i, const double* U, double* u) {
  const double* s = U + i*ncomp;
  u[0] = s[0];
  u[1] = s[1] / u[0];
  u[2] = s[2] / u[0];
  u[3] = s[3] / u[0];
  u[4] = s[4] / u[0] - 0.5*(u[1]*u[1] + u[2]*u[2] + u[3]*u[3]);
  for (int c = 5; c < ncomp; ++c) u[c] = s[c];
}

#define MAXC 16

void orc_grad(int64_t npoin, int ncomp, const int64_t* nsup,
              const int64_t* sup0, const double* int0,
              const int64_t* sup1, const double* int1,
              const int64_t* sup2, const double* int2,
              const double* x, const double* y, const double* z,
              int64_t nbtri, const int64_t* triinpoel,
              const double* U, double* G, double* A) {
  const int ng = 3*ncomp;
  memset(G, 0, sizeof(double) * (size_t)(npoin*ng));
  if (A) memset(A, 0, sizeof(double) * (size_t)(npoin*ng));
#define GG(n,k) G[(n)*ng + (k)]
#define GA(n,k,v) do { if (A) A[(n)*ng + (k)] += fabs(v); } while (0)
  /* tet superedges, Riemann.cpp:263-289 */
  for (int64_t e = 0; e < nsup[0]; ++e) {
    const int64_t* N = sup0 + e*4;
    double u[4][MAXC];
    for (int a = 0; a < 4; ++a) primitive(ncomp, N[a], U, u[a]);
    for (int c = 0; c < ncomp; ++c)
      for (int j = 0; j < 3; ++j) {
        const double* d = int0;
        double f[6];
        f[0] = d[(e*6+0)*3+j] * (u[1][c] + u[0][c]);
        f[1] = d[(e*6+1)*3+j] * (u[2][c] + u[1][c]);
        f[2] = d[(e*6+2)*3+j] * (u[0][c] + u[2][c]);
        f[3] = d[(e*6+3)*3+j] * (u[3][c] + u[0][c]);
        f[4] = d[(e*6+4)*3+j] * (u[3][c] + u[1][c]);
        f[5] = d[(e*6+5)*3+j] * (u[3][c] + u[2][c]);
        GG(N[0],c*3+j) = GG(N[0],c*3+j) - f[0] + f[2] - f[3];
        GG(N[1],c*3+j) = GG(N[1],c*3+j) + f[0] - f[1] - f[4];
        GG(N[2],c*3+j) = GG(N[2],c*3+j) + f[1] - f[2] - f[5];
        GG(N[3],c*3+j) = GG(N[3],c*3+j) + f[3] + f[4] + f[5];
        double s[6];
        s[0] = d[(e*6+0)*3+j] * (fabs(u[1][c]) + fabs(u[0][c]));
        s[1] = d[(e*6+1)*3+j] * (fabs(u[2][c]) + fabs(u[1][c]));
        s[2] = d[(e*6+2)*3+j] * (fabs(u[0][c]) + fabs(u[2][c]));
        s[3] = d[(e*6+3)*3+j] * (fabs(u[3][c]) + fabs(u[0][c]));
        s[4] = d[(e*6+4)*3+j] * (fabs(u[3][c]) + fabs(u[1][c]));
        s[5] = d[(e*6+5)*3+j] * (fabs(u[3][c]) + fabs(u[2][c]));
        GA(N[0],c*3+j, s[0]); GA(N[0],c*3+j, s[2]); GA(N[0],c*3+j, s[3]);
        GA(N[1],c*3+j, s[0]); GA(N[1],c*3+j, s[1]); GA(N[1],c*3+j, s[4]);
        GA(N[2],c*3+j, s[1]); GA(N[2],c*3+j, s[2]); GA(N[2],c*3+j, s[5]);
        GA(N[3],c*3+j, s[3]); GA(N[3],c*3+j, s[4]); GA(N[3],c*3+j, s[5]);
      }
  }
  /* triangle superedges, Riemann.cpp:292-311 */
  for (int64_t e = 0; e < nsup[1]; ++e) {
    const int64_t* N = sup1 + e*3;
    double u[3][MAXC];
    for (int a = 0; a < 3; ++a) primitive(ncomp, N[a], U, u[a]);
    for (int c = 0; c < ncomp; ++c)
      for (int j = 0; j < 3; ++j) {
        const double* d = int1;
        double f[3];
        f[0] = d[(e*3+0)*3+j] * (u[1][c] + u[0][c]);
        f[1] = d[(e*3+1)*3+j] * (u[2][c] + u[1][c]);
        f[2] = d[(e*3+2)*3+j] * (u[0][c] + u[2][c]);
        GG(N[0],c*3+j) = GG(N[0],c*3+j) - f[0] + f[2];
        GG(N[1],c*3+j) = GG(N[1],c*3+j) + f[0] - f[1];
        GG(N[2],c*3+j) = GG(N[2],c*3+j) + f[1] - f[2];
        double s[3];
        s[0] = d[(e*3+0)*3+j] * (fabs(u[1][c]) + fabs(u[0][c]));
        s[1] = d[(e*3+1)*3+j] * (fabs(u[2][c]) + fabs(u[1][c]));
        s[2] = d[(e*3+2)*3+j] * (fabs(u[0][c]) + fabs(u[2][c]));
        GA(N[0],c*3+j, s[0]); GA(N[0],c*3+j, s[2]);
        GA(N[1],c*3+j, s[0]); GA(N[1],c*3+j, s[1]);
        GA(N[2],c*3+j, s[1]); GA(N[2],c*3+j, s[2]);
      }
  }
  /* edges, Riemann.cpp:314-328 */
  for (int64_t e = 0; e < nsup[2]; ++e) {
    const int64_t* N = sup2 + e*2;
    double u[2][MAXC];
    primitive(ncomp, N[0], U, u[0]);
    primitive(ncomp, N[1], U, u[1]);
    const double* d = int2 + e*3;
    for (int c = 0; c < ncomp; ++c)
      for (int j = 0; j < 3; ++j) {
        double f = d[j] * (u[1][c] + u[0][c]);
        GG(N[0],c*3+j) -= f;
        GG(N[1],c*3+j) += f;
        double sf = d[j] * (fabs(u[1][c]) + fabs(u[0][c]));
        GA(N[0],c*3+j, sf); GA(N[1],c*3+j, sf);
      }
  }
  /* boundary integral, Riemann.cpp:332-359 */
  for (int64_t e = 0; e < nbtri; ++e) {
    const int64_t* N = triinpoel + e*3;
    double u[3][MAXC];
    for (int a = 0; a < 3; ++a) primitive(ncomp, N[a], U, u[a]);
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double n[3];
    cross(ba, ca, n);
    n[0] /= 12.0; n[1] /= 12.0; n[2] /= 12.0;
    for (int c = 0; c < ncomp; ++c) {
      double uab = (u[0][c] + u[1][c])/4.0;
      double ubc = (u[1][c] + u[2][c])/4.0;
      double uca = (u[2][c] + u[0][c])/4.0;
      double g[3] = {uab + uca + u[0][c], uab + ubc + u[1][c],
                     ubc + uca + u[2][c]};
      for (int j = 0; j < 3; ++j) {
        GG(N[0],c*3+j) += g[j] * n[j];
        GG(N[1],c*3+j) += g[j] * n[j];
        GG(N[2],c*3+j) += g[j] * n[j];
        double sg = 1.5*(fabs(u[0][c]) + fabs(u[1][c]) + fabs(u[2][c]))*n[j];
        GA(N[0],c*3+j, sg); GA(N[1],c*3+j, sg); GA(N[2],c*3+j, sg);
      }
    }
  }
#undef GG
#undef GA
}

/* ------------------------------------------------------------------- rhs -- */

/* MUSCL reconstruction of the flow variables, Riemann.cpp:34-143 */
static void muscl(int ncomp, int64_t p, int64_t q, const double* x,
                  const double* y, const double* z, const double* G,
                  double* l, double* r) {
  const int ng = 3*ncomp;
  double vw[3] = {x[q]-x[p], y[q]-y[p], z[q]-z[p]};
  double delta1[5], delta2[5], delta3[5];
  double ls[5], rs[5], url[5], urr[5];
  memcpy(ls, l, sizeof ls); memcpy(rs, r, sizeof rs);
  memcpy(url, ls, sizeof ls); memcpy(urr, rs, sizeof rs);
  for (int c = 0; c < 5; ++c) {
    const double* gp = G + p*ng + c*3;
    const double* gq = G + q*ng + c*3;
    double g1 = gp[0]*vw[0] + gp[1]*vw[1] + gp[2]*vw[2];
    double g2 = gq[0]*vw[0] + gq[1]*vw[1] + gq[2]*vw[2];
    delta2[c] = rs[c] - ls[c];
    delta1[c] = 2.0 * g1 - delta2[c];
    delta3[c] = 2.0 * g2 - delta2[c];
    /* van Leer limiter */
    double rcL = (delta2[c] + muscl_eps) / (delta1[c] + muscl_eps);
    double rcR = (delta2[c] + muscl_eps) / (delta3[c] + muscl_eps);
    double rLinv = (delta1[c] + muscl_eps) / (delta2[c] + muscl_eps);
    double rRinv = (delta3[c] + muscl_eps) / (delta2[c] + muscl_eps);
    double phiL = (fabs(rcL) + rcL) / (fabs(rcL) + 1.0);
    double phiR = (fabs(rcR) + rcR) / (fabs(rcR) + 1.0);
    double phi_L_inv = (fabs(rLinv) + rLinv) / (fabs(rLinv) + 1.0);
    double phi_R_inv = (fabs(rRinv) + rRinv) / (fabs(rRinv) + 1.0);
    url[c] += 0.25*(delta1[c]*(1.0-muscl_const)*phiL +
                    delta2[c]*(1.0+muscl_const)*phi_L_inv);
    urr[c] -= 0.25*(delta3[c]*(1.0-muscl_const)*phiR +
                    delta2[c]*(1.0+muscl_const)*phi_R_inv);
  }
  /* first order if density or internal energy would go negative */
  if (ls[0] < delta1[0] || ls[4] < delta1[4]) memcpy(url, ls, sizeof ls);
  if (rs[0] < -delta3[0] || rs[4] < -delta3[4]) memcpy(urr, rs, sizeof rs);
  memcpy(l, url, sizeof url);
  memcpy(r, urr, sizeof urr);
}

/* MUSCL reconstruction of transported scalars, Riemann.cpp:145-209 */
static void muscl_scalar(int ncomp, int64_t p, int64_t q, const double* x,
                         const double* y, const double* z, const double* G,
                         double* uL, double* uR) {
  const int ng = 3*ncomp;
  double vw[3] = {x[q]-x[p], y[q]-y[p], z[q]-z[p]};
  for (int c = 5; c < ncomp; ++c) {
    const double* gp = G + p*ng + c*3;
    const double* gq = G + q*ng + c*3;
    double g1 = gp[0]*vw[0] + gp[1]*vw[1] + gp[2]*vw[2];
    double g2 = gq[0]*vw[0] + gq[1]*vw[1] + gq[2]*vw[2];
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
    uL[c] += 0.25*(delta1*(1.0-muscl_const)*phiL +
                   delta2*(1.0+muscl_const)*phi_L_inv);
    uR[c] -= 0.25*(delta3*(1.0-muscl_const)*phiR +
                   delta2*(1.0+muscl_const)*phi_R_inv);
  }
}

/* Magnitude of the products each flux component of the last flux call was
 * built from: the scale of its round-off, used as the parity yardstick (one
 * per thread: orc_rhs_mt runs orc_rhs on several threads at once). */
static _Thread_local double fsc[MAXC];

/* Rusanov flux, Riemann.cpp:369-478 (symL = symR = 0 in the domain) */
static void rusanov(const orc_cfg* cfg, int ncomp, const double* x,
                    const double* y, const double* z, const double* G,
                    const double* d, int64_t p, int64_t q, const double* L,
                    const double* R, double* f) {
  const double g = cfg->gamma;
  double l[MAXC], r[MAXC];
  memcpy(l, L, sizeof(double)*(size_t)ncomp);
  memcpy(r, R, sizeof(double)*(size_t)ncomp);
  muscl(ncomp, p, q, x, y, z, G, l, r);
  double pL = (l[0]*l[4]) * (g-1.0);
  double pR = (r[0]*r[4]) * (g-1.0);
  double nx = d[0], ny = d[1], nz = d[2];
  double vnL = l[1]*nx + l[2]*ny + l[3]*nz;
  double vnR = r[1]*nx + r[2]*ny + r[3]*nz;
  l[4] = (l[4] + 0.5*(l[1]*l[1] + l[2]*l[2] + l[3]*l[3])) * l[0];
  l[1] *= l[0]; l[2] *= l[0]; l[3] *= l[0];
  r[4] = (r[4] + 0.5*(r[1]*r[1] + r[2]*r[2] + r[3]*r[3])) * r[0];
  r[1] *= r[0]; r[2] *= r[0]; r[3] *= r[0];
  double len = sqrt(nx*nx + ny*ny + nz*nz);
  double sl = fabs(vnL) + sqrt(g * pL / l[0])*len;
  double sr = fabs(vnR) + sqrt(g * pR / r[0])*len;
  double fw = sl > sr ? sl : sr;
  f[0] = l[0]*vnL + r[0]*vnR + fw*(r[0] - l[0]);
  f[1] = l[1]*vnL + r[1]*vnR + (pL + pR)*nx + fw*(r[1] - l[1]);
  f[2] = l[2]*vnL + r[2]*vnR + (pL + pR)*ny + fw*(r[2] - l[2]);
  f[3] = l[3]*vnL + r[3]*vnR + (pL + pR)*nz + fw*(r[3] - l[3]);
  f[4] = (l[4] + pL)*vnL + (r[4] + pR)*vnR + fw*(r[4] - l[4]);
  double fwt = fw * (1.0 + (cfg->stab2 ? fabs(cfg->stab2coef) : 0.0));
  for (int c = 0; c < 5; ++c)
    fsc[c] = fabs(l[c]*vnL) + fabs(r[c]*vnR) + fwt*(fabs(l[c]) + fabs(r[c]));
  for (int c = 1; c < 4; ++c) fsc[c] += (fabs(pL) + fabs(pR))*fabs(d[c-1]);
  fsc[4] += fabs(pL*vnL) + fabs(pR*vnR);
  if (cfg->stab2) {
    double fws = cfg->stab2coef * fw;
    for (int c = 0; c < 5; ++c) f[c] -= fws*(l[c] - r[c]);
  }
  if (ncomp == 5) return;
  muscl_scalar(ncomp, p, q, x, y, z, G, l, r);
  double sw = fabs(vnL) > fabs(vnR) ? fabs(vnL) : fabs(vnR);
  for (int c = 5; c < ncomp; ++c)
    fsc[c] = fabs(l[c]*vnL) + fabs(r[c]*vnR) + sw*(fabs(l[c]) + fabs(r[c]));
  for (int c = 5; c < ncomp; ++c) f[c] = l[c]*vnL + r[c]*vnR + sw*(r[c] - l[c]);
}

/* HLLC flux, Riemann.cpp:480-650 (symL = symR = 0 in the domain) */
static void hllc(const orc_cfg* cfg, int ncomp, const double* x,
                 const double* y, const double* z, const double* G,
                 const double* d, int64_t p, int64_t q, const double* L,
                 const double* R, double* f) {
  const double g = cfg->gamma;
  double l[MAXC], r[MAXC];
  memcpy(l, L, sizeof(double)*(size_t)ncomp);
  memcpy(r, R, sizeof(double)*(size_t)ncomp);
  muscl(ncomp, p, q, x, y, z, G, l, r);
  double nx = -d[0], ny = -d[1], nz = -d[2];
  double len = sqrt(nx*nx + ny*ny + nz*nz);
  nx /= len; ny /= len; nz /= len;
  double qL = l[1]*nx + l[2]*ny + l[3]*nz;
  double qR = r[1]*nx + r[2]*ny + r[3]*nz;
  double pL = (l[0]*l[4]) * (g-1.0);
  double pR = (r[0]*r[4]) * (g-1.0);
  l[4] = (l[4] + 0.5*(l[1]*l[1] + l[2]*l[2] + l[3]*l[3])) * l[0];
  l[1] *= l[0]; l[2] *= l[0]; l[3] *= l[0];
  r[4] = (r[4] + 0.5*(r[1]*r[1] + r[2]*r[2] + r[3]*r[3])) * r[0];
  r[1] *= r[0]; r[2] *= r[0]; r[3] *= r[0];
  double cL = sqrt(g * pL / l[0]);
  double cR = sqrt(g * pR / r[0]);
  double sL = fmin(qL - cL, qR - cR);
  double sR = fmax(qL + cL, qR + cR);
  double tL = sL - qL;
  double tR = sR - qR;
  double sM = (r[0]*qR*tR - l[0]*qL*tL + pL - pR) / (r[0]*tR - l[0]*tL);
  double pS = pL - l[0]*tL*(qL - sM);
  double uL[5], uR[5];
  double s = sL - sM;
  uL[0] = tL*l[0]/s;
  uL[1] = (tL*l[1] + (pS-pL)*nx)/s;
  uL[2] = (tL*l[2] + (pS-pL)*ny)/s;
  uL[3] = (tL*l[3] + (pS-pL)*nz)/s;
  uL[4] = (tL*l[4] - pL*qL + pS*sM)/s;
  s = sR - sM;
  uR[0] = tR*r[0]/s;
  uR[1] = (tR*r[1] + (pS-pR)*nx)/s;
  uR[2] = (tR*r[2] + (pS-pR)*ny)/s;
  uR[3] = (tR*r[3] + (pS-pR)*nz)/s;
  uR[4] = (tR*r[4] - pR*qR + pS*sM)/s;
  double L2 = -2.0*len;
  nx *= L2; ny *= L2; nz *= L2;
  if (sL > 0.0) {
    double qL2 = qL * L2;
    f[0] = l[0]*qL2;
    f[1] = l[1]*qL2 + pL*nx;
    f[2] = l[2]*qL2 + pL*ny;
    f[3] = l[3]*qL2 + pL*nz;
    f[4] = (l[4] + pL)*qL2;
  } else if (sL <= 0.0 && sM > 0.0) {
    double qL2 = qL * L2;
    double sL2 = sL * L2;
    f[0] = l[0]*qL2 + sL2*(uL[0] - l[0]);
    f[1] = l[1]*qL2 + pL*nx + sL2*(uL[1] - l[1]);
    f[2] = l[2]*qL2 + pL*ny + sL2*(uL[2] - l[2]);
    f[3] = l[3]*qL2 + pL*nz + sL2*(uL[3] - l[3]);
    f[4] = (l[4] + pL)*qL2 + sL2*(uL[4] - l[4]);
  } else if (sM <= 0.0 && sR >= 0.0) {
    double qR2 = qR * L2;
    double sR2 = sR * L2;
    f[0] = r[0]*qR2 + sR2*(uR[0] - r[0]);
    f[1] = r[1]*qR2 + pR*nx + sR2*(uR[1] - r[1]);
    f[2] = r[2]*qR2 + pR*ny + sR2*(uR[2] - r[2]);
    f[3] = r[3]*qR2 + pR*nz + sR2*(uR[3] - r[3]);
    f[4] = (r[4] + pR)*qR2 + sR2*(uR[4] - r[4]);
  } else {
    double qR2 = qR * L2;
    f[0] = r[0]*qR2;
    f[1] = r[1]*qR2 + pR*nx;
    f[2] = r[2]*qR2 + pR*ny;
    f[3] = r[3]*qR2 + pR*nz;
    f[4] = (r[4] + pR)*qR2;
  }
  {
    double w = 2.0*len*(fabs(qL) + fabs(qR) + fabs(sL) + fabs(sR) + cL + cR) *
               (1.0 + (cfg->stab2 ? fabs(cfg->stab2coef) : 0.0));
    double pw = 2.0*len*(fabs(pL) + fabs(pR) + fabs(pS)) *
                (1.0 + fabs(qL) + fabs(qR) + fabs(sM));
    for (int c = 0; c < 5; ++c)
      fsc[c] = w*(fabs(l[c]) + fabs(r[c]) + fabs(uL[c]) + fabs(uR[c])) + pw;
  }
  if (cfg->stab2) {
    double sl = fabs(qL) + cL;
    double sr = fabs(qR) + cR;
    double fws = cfg->stab2coef * (sl > sr ? sl : sr) * len;
    for (int c = 0; c < 5; ++c) f[c] -= fws * (l[c] - r[c]);
  }
  if (ncomp == 5) return;
  muscl_scalar(ncomp, p, q, x, y, z, G, l, r);
  double sw = (fabs(qL) > fabs(qR) ? fabs(qL) : fabs(qR)) * len;
  for (int c = 5; c < ncomp; ++c) {
    f[c] = (l[c]*qL + r[c]*qR)*len + sw*(r[c] - l[c]);
    fsc[c] = (fabs(l[c]*qL) + fabs(r[c]*qR))*len + sw*(fabs(r[c]) + fabs(l[c]));
  }
}

typedef void (*fluxfn)(const orc_cfg*, int, const double*, const double*,
                       const double*, const double*, const double*, int64_t,
                       int64_t, const double*, const double*, double*);

int orc_rhs(const orc_cfg* cfg, int64_t npoin, int ncomp, const int64_t* nsup,
            const int64_t* sup0, const double* int0,
            const int64_t* sup1, const double* int1,
            const int64_t* sup2, const double* int2,
            const double* x, const double* y, const double* z,
            int64_t nbtri, const int64_t* triinpoel, const uint8_t* besym,
            const double* G, const double* U, double* R, double* A) {
  fluxfn flux;
  if (cfg->flux == 0) flux = rusanov;
  else if (cfg->flux == 1) flux = hllc;
  else return -1;                      /* "Flux not configured" */
  if (ncomp > MAXC || ncomp < 5) return -2;
  memset(R, 0, sizeof(double) * (size_t)(npoin*ncomp));
  if (A) memset(A, 0, sizeof(double) * (size_t)(npoin*ncomp));
#define RR(n,c) R[(n)*ncomp + (c)]
#define RA(n,c,v) do { if (A) A[(n)*ncomp + (c)] += fabs(v); } while (0)
#define FLUX(k, dd, a, b, ua, ub, ff) \
  do { flux(cfg, ncomp, x, y, z, G, dd, a, b, ua, ub, ff); \
       memcpy(sc[k], fsc, sizeof(double)*(size_t)ncomp); } while (0)
  double sc[6][MAXC];
  /* advdom tet superedges, Riemann.cpp:700-725 */
  for (int64_t e = 0; e < nsup[0]; ++e) {
    const int64_t* N = sup0 + e*4;
    double u[4][MAXC], f[6][MAXC];
    for (int a = 0; a < 4; ++a) primitive(ncomp, N[a], U, u[a]);
    const double* d = int0;
    FLUX(0, d+(e*6+0)*3, N[0], N[1], u[0], u[1], f[0]);
    FLUX(1, d+(e*6+1)*3, N[1], N[2], u[1], u[2], f[1]);
    FLUX(2, d+(e*6+2)*3, N[2], N[0], u[2], u[0], f[2]);
    FLUX(3, d+(e*6+3)*3, N[0], N[3], u[0], u[3], f[3]);
    FLUX(4, d+(e*6+4)*3, N[1], N[3], u[1], u[3], f[4]);
    FLUX(5, d+(e*6+5)*3, N[2], N[3], u[2], u[3], f[5]);
    for (int c = 0; c < ncomp; ++c) {
      RR(N[0],c) = RR(N[0],c) - f[0][c] + f[2][c] - f[3][c];
      RR(N[1],c) = RR(N[1],c) + f[0][c] - f[1][c] - f[4][c];
      RR(N[2],c) = RR(N[2],c) + f[1][c] - f[2][c] - f[5][c];
      RR(N[3],c) = RR(N[3],c) + f[3][c] + f[4][c] + f[5][c];
      RA(N[0],c,sc[0][c]); RA(N[0],c,sc[2][c]); RA(N[0],c,sc[3][c]);
      RA(N[1],c,sc[0][c]); RA(N[1],c,sc[1][c]); RA(N[1],c,sc[4][c]);
      RA(N[2],c,sc[1][c]); RA(N[2],c,sc[2][c]); RA(N[2],c,sc[5][c]);
      RA(N[3],c,sc[3][c]); RA(N[3],c,sc[4][c]); RA(N[3],c,sc[5][c]);
    }
  }
  /* advdom triangle superedges, Riemann.cpp:728-746 */
  for (int64_t e = 0; e < nsup[1]; ++e) {
    const int64_t* N = sup1 + e*3;
    double u[3][MAXC], f[3][MAXC];
    for (int a = 0; a < 3; ++a) primitive(ncomp, N[a], U, u[a]);
    const double* d = int1;
    FLUX(0, d+(e*3+0)*3, N[0], N[1], u[0], u[1], f[0]);
    FLUX(1, d+(e*3+1)*3, N[1], N[2], u[1], u[2], f[1]);
    FLUX(2, d+(e*3+2)*3, N[2], N[0], u[2], u[0], f[2]);
    for (int c = 0; c < ncomp; ++c) {
      RR(N[0],c) = RR(N[0],c) - f[0][c] + f[2][c];
      RR(N[1],c) = RR(N[1],c) + f[0][c] - f[1][c];
      RR(N[2],c) = RR(N[2],c) + f[1][c] - f[2][c];
      RA(N[0],c,sc[0][c]); RA(N[0],c,sc[2][c]);
      RA(N[1],c,sc[0][c]); RA(N[1],c,sc[1][c]);
      RA(N[2],c,sc[1][c]); RA(N[2],c,sc[2][c]);
    }
  }
  /* advdom edges, Riemann.cpp:749-759 */
  for (int64_t e = 0; e < nsup[2]; ++e) {
    const int64_t* N = sup2 + e*2;
    double u[2][MAXC], f[MAXC];
    primitive(ncomp, N[0], U, u[0]);
    primitive(ncomp, N[1], U, u[1]);
    FLUX(0, int2 + e*3, N[0], N[1], u[0], u[1], f);
    for (int c = 0; c < ncomp; ++c) {
      RR(N[0],c) -= f[c];
      RR(N[1],c) += f[c];
      RA(N[0],c,sc[0][c]); RA(N[1],c,sc[0][c]);
    }
  }
  /* advbnd, Riemann.cpp:795-871 */
  const double g = cfg->gamma;
  for (int64_t e = 0; e < nbtri; ++e) {
    const int64_t* N = triinpoel + e*3;
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double n[3];
    cross(ba, ca, n);
    double nx = n[0] / 12.0, ny = n[1] / 12.0, nz = n[2] / 12.0;
    const uint8_t* sym = besym + e*3;
    double f[MAXC][3], fs[MAXC][3];
    for (int a = 0; a < 3; ++a) {
      const double* s = U + N[a]*ncomp;
      double r = s[0], ru = s[1], rv = s[2], rw = s[3], re = s[4];
      double pr = (re - 0.5*(ru*ru + rv*rv + rw*rw)/r) * (g-1.0);
      double vn = sym[a] ? 0.0 : (nx*ru + ny*rv + nz*rw)/r;
      f[0][a] = r*vn;
      f[1][a] = ru*vn + pr*nx;
      f[2][a] = rv*vn + pr*ny;
      f[3][a] = rw*vn + pr*nz;
      f[4][a] = (re + pr)*vn;
      for (int c = 5; c < ncomp; ++c) f[c][a] = s[c]*vn;
      double vs = (fabs(nx*ru) + fabs(ny*rv) + fabs(nz*rw))/fabs(r);
      double ps = fabs(re) + 0.5*(ru*ru + rv*rv + rw*rw)/fabs(r);
      fs[0][a] = fabs(r)*vs;
      fs[1][a] = fabs(ru)*vs + ps*fabs(nx);
      fs[2][a] = fabs(rv)*vs + ps*fabs(ny);
      fs[3][a] = fabs(rw)*vs + ps*fabs(nz);
      fs[4][a] = (fabs(re) + ps)*vs;
      for (int c = 5; c < ncomp; ++c) fs[c][a] = fabs(s[c])*vs;
    }
    for (int c = 0; c < ncomp; ++c) {
      double fab = (f[c][0] + f[c][1])/4.0;
      double fbc = (f[c][1] + f[c][2])/4.0;
      double fca = (f[c][2] + f[c][0])/4.0;
      RR(N[0],c) += fab + fca + f[c][0];
      RR(N[1],c) += fab + fbc + f[c][1];
      RR(N[2],c) += fbc + fca + f[c][2];
      RA(N[0],c,1.5*fs[c][0] + 0.25*(fs[c][1] + fs[c][2]));
      RA(N[1],c,1.5*fs[c][1] + 0.25*(fs[c][0] + fs[c][2]));
      RA(N[2],c,1.5*fs[c][2] + 0.25*(fs[c][0] + fs[c][1]));
    }
  }
#undef RR
#undef RA
#undef FLUX
  return 0;
}

/* -------------------------------------------------------------------- dt -- */

double orc_dt(const orc_cfg* cfg, int64_t npoin, int ncomp, const double* U,
              const double* vol, double cfl) {
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
    double euler_dt = L / den;
    if (euler_dt < mindt) mindt = euler_dt;
  }
  return mindt * cfl;
}

/* ------------------------------------------------------------- time step -- */

int orc_riecg_step(const orc_cfg* cfg, int64_t npoin, int ncomp,
                   const int64_t* nsup, const int64_t* sup0, const double* int0,
                   const int64_t* sup1, const double* int1,
                   const int64_t* sup2, const double* int2,
                   const double* x, const double* y, const double* z,
                   int64_t nbtri, const int64_t* triinpoel,
                   const uint8_t* besym, const double* vol,
                   int64_t nsym, const int64_t* symnodes,
                   const double* symnorms, double dt,
                   double* U, double* Un, double* G, double* R) {
  static const double rkcoef[3] = {1.0/3.0, 1.0/2.0, 1.0};
  const int ng = 3*ncomp;
  for (int stage = 0; stage < 3; ++stage) {
    orc_grad(npoin, ncomp, nsup, sup0, int0, sup1, int1, sup2, int2, x, y, z,
             nbtri, triinpoel, U, G, NULL);
    for (int64_t p = 0; p < npoin; ++p)          /* RieCG.cpp:935-939 */
      for (int c = 0; c < ng; ++c) G[p*ng+c] /= vol[p];
    int rc = orc_rhs(cfg, npoin, ncomp, nsup, sup0, int0, sup1, int1, sup2,
                     int2, x, y, z, nbtri, triinpoel, besym, G, U, R, NULL);
    if (rc) return rc;
    if (stage == 0) memcpy(Un, U, sizeof(double)*(size_t)(npoin*ncomp));
    for (int64_t i = 0; i < npoin; ++i)          /* RieCG.cpp:1016-1021 */
      for (int c = 0; c < ncomp; ++c)
        U[i*ncomp+c] = Un[i*ncomp+c] - rkcoef[stage] * dt * R[i*ncomp+c] / vol[i];
    orc_symbc(npoin, ncomp, U, nsym, symnodes, symnorms);   /* RieCG.cpp:1028 */
  }
  return 0;
}

void orc_diag(int64_t npoin, int ncomp, const double* U, const double* Un,
              const double* v, double* out) {
  memset(out, 0, sizeof(double) * (size_t)(2*ncomp+1));
  for (int64_t i = 0; i < npoin; ++i) {
    for (int c = 0; c < ncomp; ++c)
      out[c] += U[i*ncomp+c] * U[i*ncomp+c] * v[i];
    for (int c = 0; c < ncomp; ++c) {
      double d = U[i*ncomp+c] - Un[i*ncomp+c];
      out[ncomp+c] += d * d * v[i];
    }
    out[2*ncomp] += U[i*ncomp+4] * v[i];
  }
}
