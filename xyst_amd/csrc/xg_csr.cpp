// Host side of the pressure-Poisson linear system: the tk::CSR sparsity
// pattern (CSR.cpp:19-84), the ChoCG Laplacian (ChoCG.cpp:146-188) and the
// Dirichlet application of ConjugateGradients::apply / CSR::dirichlet
// (ConjugateGradients.cpp:450-556, CSR.cpp:106-152).
//
// Indices are 0-based (the reference stores 1-based ia/ja); rows keep the
// reference's structure: node row = the node and its edge neighbours, columns
// ascending, ncomp unknowns interleaved node-major with no coupling between
// components.
#include "xg_csr.hpp"

#include <algorithm>
#include <cstring>

namespace xg {

static const int kEdge[6][2] = {{0,1},{1,2},{2,0},{0,3},{1,3},{2,3}};

int csr_pattern(int64_t npoin, int64_t ntet, const size_t* inpoel, int ncomp,
                std::vector<int64_t>& ia, std::vector<int32_t>& ja,
                std::string& err) {
  if ((int64_t)npoin * ncomp >= (1ll << 31)) { err = "csr: too many rows"; return XG_EINVAL; }
  std::vector<int64_t> cnt(npoin + 1, 0);
  for (int64_t e = 0; e < ntet; ++e)
    for (int k = 0; k < 6; ++k) {
      size_t a = inpoel[e * 4 + kEdge[k][0]], b = inpoel[e * 4 + kEdge[k][1]];
      if (a >= (size_t)npoin || b >= (size_t)npoin) { err = "csr: node id out of range"; return XG_EINVAL; }
      ++cnt[a + 1];
      ++cnt[b + 1];
    }
  for (int64_t p = 0; p < npoin; ++p) cnt[p + 1] += cnt[p];
  std::vector<int32_t> raw(cnt[npoin]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ntet; ++e)
      for (int k = 0; k < 6; ++k) {
        size_t a = inpoel[e * 4 + kEdge[k][0]], b = inpoel[e * 4 + kEdge[k][1]];
        raw[pos[a]++] = (int32_t)b;
        raw[pos[b]++] = (int32_t)a;
      }
  }
  ia.assign(npoin * ncomp + 1, 0);
  ja.clear();
  std::vector<int32_t> row;
  for (int64_t p = 0; p < npoin; ++p) {
    row.assign(raw.begin() + cnt[p], raw.begin() + cnt[p + 1]);
    row.push_back((int32_t)p);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    for (int k = 0; k < ncomp; ++k) {
      for (int32_t c : row) ja.push_back(c * ncomp + k);
      ia[p * ncomp + k + 1] = (int64_t)ja.size();
    }
  }
  return XG_OK;
}

static inline int64_t find_col(const int64_t* ia, const int32_t* ja,
                               int64_t row, int32_t col) {
  const int32_t* b = ja + ia[row];
  const int32_t* e = ja + ia[row + 1];
  const int32_t* f = std::lower_bound(b, e, col);
  return (f != e && *f == col) ? (f - ja) : -1;
}

int poisson_matrix(int64_t npoin, const double* x, const double* y,
                   const double* z, int64_t ntet, const size_t* inpoel,
                   const std::vector<int64_t>& ia,
                   const std::vector<int32_t>& ja, std::vector<double>& a,
                   std::string& err) {
  a.assign(ja.size(), 0.0);
  for (int64_t e = 0; e < ntet; ++e) {
    const size_t* N = inpoel + e * 4;
    double ba[3] = {x[N[1]]-x[N[0]], y[N[1]]-y[N[0]], z[N[1]]-z[N[0]]};
    double ca[3] = {x[N[2]]-x[N[0]], y[N[2]]-y[N[0]], z[N[2]]-z[N[0]]};
    double da[3] = {x[N[3]]-x[N[0]], y[N[3]]-y[N[0]], z[N[3]]-z[N[0]]};
    double g[4][3];
    g[1][0] = ca[1]*da[2] - da[1]*ca[2]; g[1][1] = ca[2]*da[0] - da[2]*ca[0];
    g[1][2] = ca[0]*da[1] - da[0]*ca[1];
    g[2][0] = da[1]*ba[2] - ba[1]*da[2]; g[2][1] = da[2]*ba[0] - ba[2]*da[0];
    g[2][2] = da[0]*ba[1] - ba[0]*da[1];
    g[3][0] = ba[1]*ca[2] - ca[1]*ba[2]; g[3][1] = ba[2]*ca[0] - ca[2]*ba[0];
    g[3][2] = ba[0]*ca[1] - ca[0]*ba[1];
    double J = (ba[0]*g[1][0] + ba[1]*g[1][1] + ba[2]*g[1][2]) * 6.0;
    if (!(J > 0)) { err = "Element Jacobian non-positive"; return XG_EINVAL; }
    for (int i = 0; i < 3; ++i) g[0][i] = -g[1][i] - g[2][i] - g[3][i];
    for (int p = 0; p < 4; ++p)
      for (int q = 0; q < 4; ++q) {
        int64_t j = find_col(ia.data(), ja.data(), (int64_t)N[p], (int32_t)N[q]);
        if (j < 0) { err = "poisson matrix: entry outside the pattern"; return XG_EINVAL; }
        a[j] -= (g[p][0]*g[q][0] + g[p][1]*g[q][1] + g[p][2]*g[q][2]) / J;
      }
  }
  return XG_OK;
}

// CSR::dirichlet on one partition with the symmetric pattern: the column
// entries of unknown c are exactly the rows listed in row c, each found by a
// binary search, instead of a scan over all rows.
void csr_dirichlet(int ncomp, const int64_t* ia, const int32_t* ja, double* a,
                   int64_t node, double val, double* b, int pos) {
  (void)ncomp;
  const int64_t c = node * ncomp + pos;
  for (int64_t j = ia[c]; j < ia[c + 1]; ++j) {
    int64_t r = ja[j];                     // rows holding column c
    int64_t k = find_col(ia, ja, r, (int32_t)c);
    if (k >= 0) { b[r] += a[k] * val; a[k] = 0.0; }
  }
  for (int64_t j = ia[c]; j < ia[c + 1]; ++j) a[j] = (ja[j] == c) ? 1.0 : 0.0;
}

}  // namespace xg

extern "C" int xg_csr_pattern(int64_t npoin, int64_t ntet, const size_t* inpoel,
                              int ncomp, int64_t cap, int64_t* nnz, size_t* ia,
                              size_t* ja) {
  if (npoin <= 0 || ntet < 0 || (ntet && !inpoel) || ncomp < 1 || !nnz ||
      cap < 0 || (cap && (!ia || !ja)))
    return XG_EINVAL;
  std::vector<int64_t> i;
  std::vector<int32_t> j;
  std::string err;
  int rc = xg::csr_pattern(npoin, ntet, inpoel, ncomp, i, j, err);
  if (rc) return rc;
  *nnz = (int64_t)j.size();
  if (cap == 0) return XG_OK;
  if (cap < *nnz) return XG_EINVAL;
  for (size_t k = 0; k < i.size(); ++k) ia[k] = (size_t)i[k];
  for (size_t k = 0; k < j.size(); ++k) ja[k] = (size_t)j[k];
  return XG_OK;
}

extern "C" int xg_poisson_matrix(int64_t npoin, const double* x, const double* y,
                                 const double* z, int64_t ntet,
                                 const size_t* inpoel, const size_t* ia,
                                 const size_t* ja, double* a) {
  if (npoin <= 0 || !x || !y || !z || ntet < 0 || (ntet && !inpoel) || !ia ||
      !ja || !a)
    return XG_EINVAL;
  std::vector<int64_t> i(npoin + 1);
  for (int64_t k = 0; k <= npoin; ++k) i[k] = (int64_t)ia[k];
  std::vector<int32_t> j(i[npoin]);
  for (int64_t k = 0; k < i[npoin]; ++k) j[k] = (int32_t)ja[k];
  std::vector<double> v;
  std::string err;
  int rc = xg::poisson_matrix(npoin, x, y, z, ntet, inpoel, i, j, v, err);
  if (rc) return rc;
  std::memcpy(a, v.data(), v.size() * sizeof(double));
  return XG_OK;
}
