/* CPU oracle: CSR matrix and (Jacobi-)preconditioned conjugate gradients --
 * TEST INFRASTRUCTURE ONLY.
 *
 * Restates src/LinearSolver/CSR.cpp and the solver sequence of
 * src/LinearSolver/ConjugateGradients.cpp (one partition: no halo sums, every
 * node counted once), and the ChoCG pressure matrix assembly
 * (src/Inciter/ChoCG.cpp:146-188).  Row/column indices are 0-based here; the
 * reference stores them 1-based, with the same row structure: every node row
 * holds the node itself and its edge neighbours, columns sorted ascending,
 * and with ncomp > 1 unknowns interleaved node-major (row i*ncomp+k couples
 * only to columns j*ncomp+k).
 *
 * Pinning: the 1- and 3-DOF goldens of
 * tests/unit/LinearSolver/TestConjugateGradients.cpp (normb and final
 * residual), a dense solve of the same systems, and the CSR layout of
 * TestCSR.cpp.
 */
#include "xyst_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* tk::CSR constructor (CSR.cpp:19-84) from the tet connectivity */
int64_t orc_csr_build(int64_t npoin, int64_t ntet, const int64_t* inpoel,
                      int ncomp, int64_t* ia, int64_t* ja, int64_t cap) {
  static const int ed[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};
  /* neighbour lists (psup) */
  int64_t* cnt = (int64_t*)calloc((size_t)npoin + 1, sizeof(int64_t));
  for (int64_t e = 0; e < ntet; ++e)
    for (int k = 0; k < 6; ++k) {
      ++cnt[inpoel[e*4+ed[k][0]] + 1];
      ++cnt[inpoel[e*4+ed[k][1]] + 1];
    }
  for (int64_t p = 0; p < npoin; ++p) cnt[p+1] += cnt[p];
  int64_t* raw = (int64_t*)malloc(sizeof(int64_t) * (size_t)cnt[npoin]);
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)npoin);
  memcpy(pos, cnt, sizeof(int64_t) * (size_t)npoin);
  for (int64_t e = 0; e < ntet; ++e)
    for (int k = 0; k < 6; ++k) {
      int64_t a = inpoel[e*4+ed[k][0]], b = inpoel[e*4+ed[k][1]];
      raw[pos[a]++] = b;
      raw[pos[b]++] = a;
    }
  /* rows: the node plus its unique neighbours, ascending */
  int64_t nnz = 0;
  ia[0] = 0;
  int64_t* row = (int64_t*)malloc(sizeof(int64_t) * 64);
  size_t rowcap = 64;
  for (int64_t p = 0; p < npoin; ++p) {
    size_t n = (size_t)(cnt[p+1] - cnt[p]) + 1;
    if (n > rowcap) { rowcap = n; row = (int64_t*)realloc(row, sizeof(int64_t) * rowcap); }
    row[0] = p;
    memcpy(row + 1, raw + cnt[p], sizeof(int64_t) * (n - 1));
    qsort(row, n, sizeof(int64_t), cmp_i64);
    size_t u = 0;
    for (size_t i = 0; i < n; ++i)
      if (i == 0 || row[i] != row[i-1]) row[u++] = row[i];
    for (int k = 0; k < ncomp; ++k) {
      if (nnz + (int64_t)u > cap) { nnz = -1; goto done; }
      for (size_t i = 0; i < u; ++i) ja[nnz++] = row[i]*ncomp + k;
      ia[p*ncomp + k + 1] = nnz;
    }
  }
done:
  free(cnt); free(raw); free(pos); free(row);
  return nnz;
}

/* CSR::operator() (CSR.cpp:86-104): linear search of the row; -1 if absent */
static int64_t csr_find(int ncomp, const int64_t* ia, const int64_t* ja,
                        int64_t row, int64_t col, int pos) {
  int64_t r = row*ncomp + pos, c = col*ncomp + pos;
  for (int64_t j = ia[r]; j < ia[r+1]; ++j)
    if (ja[j] == c) return j;
  return -1;
}

/* ChoCG::prelhs (ChoCG.cpp:146-188): A = -K, K the P1 stiffness matrix */
int orc_prelhs(int64_t npoin, const double* x, const double* y,
               const double* z, int64_t ntet, const int64_t* inpoel,
               const int64_t* ia, const int64_t* ja, double* a) {
  (void)npoin;
  memset(a, 0, sizeof(double) * (size_t)ia[npoin]);
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
    double g[4][3];
    /* tk::cross, Vector.hpp:36-58 */
    g[1][0] = ca[1]*da[2] - da[1]*ca[2]; g[1][1] = ca[2]*da[0] - da[2]*ca[0];
    g[1][2] = ca[0]*da[1] - da[0]*ca[1];
    g[2][0] = da[1]*ba[2] - ba[1]*da[2]; g[2][1] = da[2]*ba[0] - ba[2]*da[0];
    g[2][2] = da[0]*ba[1] - ba[0]*da[1];
    g[3][0] = ba[1]*ca[2] - ca[1]*ba[2]; g[3][1] = ba[2]*ca[0] - ca[2]*ba[0];
    g[3][2] = ba[0]*ca[1] - ca[0]*ba[1];
    double J = (ba[0]*g[1][0] + ba[1]*g[1][1] + ba[2]*g[1][2]) * 6.0;
    for (int i = 0; i < 3; ++i) g[0][i] = -g[1][i]-g[2][i]-g[3][i];
    for (int p = 0; p < 4; ++p)
      for (int q = 0; q < 4; ++q) {
        int64_t j = csr_find(1, ia, ja, N[p], N[q], 0);
        if (j < 0) return -1;
        a[j] -= (g[p][0]*g[q][0] + g[p][1]*g[q][1] + g[p][2]*g[q][2]) / J;
      }
  }
  return 0;
}

/* The +K Laplacian of TestConjugateGradients.cpp:175-199 (crossdiv by J,
 * J/6 * grad_a . grad_b), added to every component. */
int orc_laplace_test(int64_t npoin, const double* x, const double* y,
                     const double* z, int64_t ntet, const int64_t* inpoel,
                     int ncomp, const int64_t* ia, const int64_t* ja,
                     double* a) {
  memset(a, 0, sizeof(double) * (size_t)ia[npoin*ncomp]);
  for (int64_t e = 0; e < ntet; ++e) {
    const int64_t* N = inpoel + e*4;
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
    double c1[3] = {ca[1]*da[2] - da[1]*ca[2], ca[2]*da[0] - da[2]*ca[0],
                    ca[0]*da[1] - da[0]*ca[1]};
    double J = ba[0]*c1[0] + ba[1]*c1[1] + ba[2]*c1[2];
    if (!(J > 0)) return -1;
    double g[4][3];
    /* tk::crossdiv, Vector.hpp:82-95 */
    g[1][0] = (ca[1]*da[2] - da[1]*ca[2]) / J; g[1][1] = (ca[2]*da[0] - da[2]*ca[0]) / J;
    g[1][2] = (ca[0]*da[1] - da[0]*ca[1]) / J;
    g[2][0] = (da[1]*ba[2] - ba[1]*da[2]) / J; g[2][1] = (da[2]*ba[0] - ba[2]*da[0]) / J;
    g[2][2] = (da[0]*ba[1] - ba[0]*da[1]) / J;
    g[3][0] = (ba[1]*ca[2] - ca[1]*ba[2]) / J; g[3][1] = (ba[2]*ca[0] - ca[2]*ba[0]) / J;
    g[3][2] = (ba[0]*ca[1] - ca[0]*ba[1]) / J;
    for (int i = 0; i < 3; ++i) g[0][i] = -g[1][i]-g[2][i]-g[3][i];
    for (int p = 0; p < 4; ++p)
      for (int k = 0; k < 3; ++k)
        for (int q = 0; q < 4; ++q)
          for (int i = 0; i < ncomp; ++i) {
            int64_t j = csr_find(ncomp, ia, ja, N[p], N[q], i);
            if (j < 0) return -2;
            a[j] += J/6 * g[p][k] * g[q][k];
          }
  }
  return 0;
}

/* CSR::dirichlet (CSR.cpp:106-152), one partition (diagonal set to 1) */
void orc_csr_dirichlet(int64_t nrow, int ncomp, const int64_t* ia,
                       const int64_t* ja, double* a, int64_t i, double val,
                       double* b, int pos) {
  int64_t c = i*ncomp + pos;
  for (int64_t r = 0; r < nrow; ++r)
    for (int64_t j = ia[r]; j < ia[r+1]; ++j)
      if (ja[j] == c) { b[r] += a[j] * val; a[j] = 0.0; break; }
  for (int64_t j = ia[c]; j < ia[c+1]; ++j) a[j] = ja[j] == c ? 1.0 : 0.0;
}

/* CSR::mult (CSR.cpp:154-172) */
void orc_csr_mult(int64_t nrow, const int64_t* ia, const int64_t* ja,
                  const double* a, const double* x, double* r) {
  for (int64_t i = 0; i < nrow; ++i) {
    r[i] = 0.0;
    for (int64_t j = ia[i]; j < ia[i+1]; ++j) r[i] += a[j] * x[ja[j]];
  }
}

/* ConjugateGradients::dot (ConjugateGradients.cpp:130-151), one partition */
static double dot(int64_t n, const double* u, const double* v) {
  double D = 0.0;
  for (int64_t i = 0; i < n; ++i) D += u[i] * v[i];
  return D;
}

int orc_cg(int64_t npoin, int ncomp, const int64_t* ia, const int64_t* ja,
           double* a, double* x, double* b, const double* neubc,
           int64_t ndir, const int64_t* dirnode, const int32_t* dirmask,
           const double* dirval, int apply, int pc, int64_t maxit,
           double tol, double* normb_out, double* normr_out, int64_t* it_out,
           int* converged_out, double* rhist) {
  const int64_t n = npoin * ncomp;
  const int64_t nnz = ia[n];
  double* An = NULL;
  double* r = (double*)calloc((size_t)n, sizeof(double));
  double* z = (double*)calloc((size_t)n, sizeof(double));
  double* d = (double*)calloc((size_t)n, sizeof(double));
  double* p = (double*)calloc((size_t)n, sizeof(double));
  double* q = (double*)calloc((size_t)n, sizeof(double));
  if (apply) {
    /* init + apply + r (ConjugateGradients.cpp:335-556) */
    An = (double*)malloc(sizeof(double) * (size_t)nnz);
    memcpy(An, a, sizeof(double) * (size_t)nnz);
    for (int64_t i = 0; i < n; ++i) b[i] += neubc ? neubc[i] : 0.0;
    memset(r, 0, sizeof(double) * (size_t)n);
    /* Dirichlet rows in decreasing node order (reverse map iteration) */
    for (int64_t k = ndir - 1; k >= 0; --k)
      for (int j = 0; j < ncomp; ++j)
        if (dirmask[k*ncomp + j])
          orc_csr_dirichlet(n, ncomp, ia, ja, a, dirnode[k],
                            dirval[k*ncomp + j], r, j);
    for (int64_t i = 0; i < n; ++i) b[i] -= r[i];
    for (int64_t k = 0; k < ndir; ++k)
      for (int j = 0; j < ncomp; ++j)
        if (dirmask[k*ncomp + j]) b[dirnode[k]*ncomp + j] = dirval[k*ncomp + j];
  }
  /* setup: residual, preconditioner, normb (ConjugateGradients.cpp:105-126,
   * 164-333) */
  orc_csr_mult(n, ia, ja, a, x, r);
  for (int64_t i = 0; i < n; ++i) {
    int64_t row = i;
    double dd = 1.0;                         /* pc "none": 1/count = 1 */
    if (pc) {                                /* "jacobi": diag(A) */
      dd = 0.0;
      for (int64_t j = ia[row]; j < ia[row+1]; ++j)
        if (ja[j] == row) { dd = a[j]; break; }
    }
    d[i] = dd;
  }
  double normb = sqrt(dot(n, b, b));
  for (int64_t i = 0; i < n; ++i) { r[i] *= -1.0; r[i] += b[i]; }
  memcpy(p, r, sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) z[i] = r[i] / d[i];
  double rho = dot(n, r, z), rho0 = 0.0, alpha = 0.0, normr = 0.0;
  int finished = 0, converged = 0;
  int64_t it = 0;
  const double eps = DBL_EPSILON;
  /* solve: next, qAp, q, pq, rz, x (ConjugateGradients.cpp:558-823) */
  for (;;) {
    alpha = it == 0 ? 0.0 : rho / rho0;      /* beta, stored in m_alpha */
    rho0 = rho;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + alpha * p[i];
    orc_csr_mult(n, ia, ja, a, p, q);
    double pq = dot(n, p, q);
    if (fabs(pq) < eps) { finished = 1; alpha = 0.0; }
    else alpha = rho / pq;
    for (int64_t i = 0; i < n; ++i) r[i] -= alpha * q[i];
    for (int64_t i = 0; i < n; ++i) z[i] = r[i] / d[i];
    rho = dot(n, r, z);
    normr = dot(n, r, r);
    for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
    ++it;
    double nb = normb > 1.0e-14 ? normb : 1.0;
    double nr = sqrt(normr);
    if (rhist) rhist[it-1] = nr;
    if (finished || nr < tol*nb || it >= maxit) {
      converged = nr > tol*nb ? 0 : 1;
      normr = nr;
      break;
    }
  }
  if (apply) { memcpy(a, An, sizeof(double) * (size_t)nnz); free(An); }
  *normb_out = normb;
  *normr_out = normr;
  *it_out = it;
  *converged_out = converged;
  free(r); free(z); free(d); free(p); free(q);
  return 0;
}
