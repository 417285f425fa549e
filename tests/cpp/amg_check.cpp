// CPU check of the AMG setup (xg_amg.cpp): the Galerkin coarse operators
// equal P^T B P computed densely (small levels), stay symmetric, and a
// double-precision emulation of the device V-cycle (damped Jacobi, one
// sweep before and after, dense coarsest solve) preconditions CG on S A to
// 1e-10.  Input file: int64 n, nnz; int64 ia[n+1]; int32 ja[nnz]; double
// a[nnz]; double b[n].  Prints key=value statistics.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../xyst_amd/csrc/xg_amg.hpp"

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      std::exit(1);                                                      \
    }                                                                    \
  } while (0)

using xg::Csr;

static void spmv(const Csr& M, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(M.nrow, 0.0);
  for (int64_t i = 0; i < M.nrow; ++i) {
    double s = 0.0;
    for (int64_t p = M.ia[i]; p < M.ia[i + 1]; ++p) s += M.a[p] * x[M.ja[p]];
    y[i] = s;
  }
}

static int g_nu = 1;

static void vcycle(const xg::AmgHierarchy& h, size_t l, const std::vector<double>& b,
                   std::vector<double>& x) {
  const xg::AmgLevel& L = h.lev[l];
  const int64_t n = L.B.nrow;
  if (l + 1 == h.lev.size()) {
    x.assign(n, 0.0);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < n; ++j) x[i] += h.cinv[(size_t)i * n + j] * b[j];
    return;
  }
  x.assign(n, 0.0);
  for (int64_t i = 0; i < n; ++i) x[i] = L.omega * L.dinv[i] * b[i];
  std::vector<double> Bx, r(n), bc, xc, px;
  for (int s = 1; s < g_nu; ++s) {
    spmv(L.B, x, Bx);
    for (int64_t i = 0; i < n; ++i) x[i] += L.omega * L.dinv[i] * (b[i] - Bx[i]);
  }
  spmv(L.B, x, Bx);
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Bx[i];
  spmv(L.R, r, bc);
  vcycle(h, l + 1, bc, xc);
  spmv(L.P, xc, px);
  for (int64_t i = 0; i < n; ++i) x[i] += px[i];
  for (int s = 0; s < g_nu; ++s) {
    spmv(L.B, x, Bx);
    for (int64_t i = 0; i < n; ++i) x[i] += L.omega * L.dinv[i] * (b[i] - Bx[i]);
  }
}

int main(int argc, char** argv) {
  CHECK(argc == 2);
  FILE* f = std::fopen(argv[1], "rb");
  CHECK(f);
  int64_t hdr[2];
  CHECK(std::fread(hdr, 8, 2, f) == 2);
  const int64_t n = hdr[0], nnz = hdr[1];
  std::vector<int64_t> ia(n + 1);
  std::vector<int32_t> ja(nnz);
  std::vector<double> a(nnz), b(n);
  CHECK(std::fread(ia.data(), 8, n + 1, f) == (size_t)(n + 1));
  CHECK(std::fread(ja.data(), 4, nnz, f) == (size_t)nnz);
  CHECK(std::fread(a.data(), 8, nnz, f) == (size_t)nnz);
  CHECK(std::fread(b.data(), 8, n, f) == (size_t)n);
  std::fclose(f);
  xg::AmgHierarchy h;
  xg::AmgOptions opt;
  opt.max_coarse = 200;
  if (getenv("AMG_NU")) g_nu = atoi(getenv("AMG_NU"));
  if (getenv("AMG_THETA")) opt.theta = atof(getenv("AMG_THETA"));
  if (getenv("AMG_COARSE")) opt.max_coarse = atoll(getenv("AMG_COARSE"));
  std::string err;
  CHECK(xg::amg_setup(n, ia.data(), ja.data(), a.data(), opt, h, err) == 0);
  CHECK(h.lev.size() >= 2);
  // Galerkin products and symmetry, checked densely on levels small enough
  for (size_t l = 0; l + 1 < h.lev.size(); ++l) {
    const xg::AmgLevel& L = h.lev[l];
    const Csr& Bc = h.lev[l + 1].B;
    const int64_t m = L.B.nrow, mc = Bc.nrow;
    CHECK(L.P.nrow == m && L.P.ncol == mc && L.R.nrow == mc && L.R.ncol == m);
    double bmax = 0.0;
    for (double v : Bc.a) bmax = std::max(bmax, std::fabs(v));
    std::vector<double> D((size_t)mc * mc, 0.0);
    for (int64_t i = 0; i < mc; ++i)
      for (int64_t p = Bc.ia[i]; p < Bc.ia[i + 1]; ++p) D[(size_t)i * mc + Bc.ja[p]] = Bc.a[p];
    for (int64_t i = 0; i < mc; ++i)
      for (int64_t j = 0; j < i; ++j)
        CHECK(std::fabs(D[(size_t)i * mc + j] - D[(size_t)j * mc + i]) <= 1e-12 * bmax);
    if (m > 3000) continue;
    std::vector<double> Pd((size_t)m * mc, 0.0), Bd((size_t)m * m, 0.0);
    for (int64_t i = 0; i < m; ++i) {
      for (int64_t p = L.P.ia[i]; p < L.P.ia[i + 1]; ++p) Pd[(size_t)i * mc + L.P.ja[p]] = L.P.a[p];
      for (int64_t p = L.B.ia[i]; p < L.B.ia[i + 1]; ++p) Bd[(size_t)i * m + L.B.ja[p]] = L.B.a[p];
    }
    for (int64_t i = 0; i < mc; ++i)
      for (int64_t j = 0; j < mc; ++j) {
        double s = 0.0;
        for (int64_t k = 0; k < m; ++k) {
          if (Pd[(size_t)k * mc + i] == 0.0) continue;
          double t = 0.0;
          for (int64_t q = 0; q < m; ++q) t += Bd[(size_t)k * m + q] * Pd[(size_t)q * mc + j];
          s += Pd[(size_t)k * mc + i] * t;
        }
        CHECK(std::fabs(s - D[(size_t)i * mc + j]) <= 1e-11 * bmax);
      }
  }
  // PCG on B = S A with the V-cycle
  const xg::Csr& B = h.lev[0].B;
  std::vector<double> sg(n, 1.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = ia[i]; p < ia[i + 1]; ++p)
      if (ja[p] == i && a[p] < 0.0) sg[i] = -1.0;
  std::vector<double> bs(n), x(n, 0.0), r(n), z, p(n), q;
  for (int64_t i = 0; i < n; ++i) bs[i] = sg[i] * b[i];
  r = bs;
  double nb = 0.0;
  for (double v : bs) nb += v * v;
  nb = std::sqrt(nb);
  vcycle(h, 0, r, z);
  p = z;
  double rz = 0.0;
  for (int64_t i = 0; i < n; ++i) rz += r[i] * z[i];
  int it = 0;
  double nr = nb;
  for (; it < 200 && nr > 1e-10 * nb; ++it) {
    spmv(B, p, q);
    double pq = 0.0;
    for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
    const double al = rz / pq;
    nr = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += al * p[i];
      r[i] -= al * q[i];
      nr += r[i] * r[i];
    }
    nr = std::sqrt(nr);
    vcycle(h, 0, r, z);
    double rz1 = 0.0;
    for (int64_t i = 0; i < n; ++i) rz1 += r[i] * z[i];
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + rz1 / rz * p[i];
    rz = rz1;
  }
  std::printf("ok levels=%zu complexity=%.4f iterations=%d coarsest=%lld relres=%.3e",
              h.lev.size(), h.op_complexity, it, (long long)h.lev.back().B.nrow, nr / nb);
  for (const auto& L : h.lev) std::printf(" n%lld/nnz%lld/w%.3f", (long long)L.B.nrow, (long long)L.B.ia[L.B.nrow], L.omega);
  std::printf("\n");
  return 0;
}
