// Host setup of the smoothed-aggregation AMG hierarchy (see xg_amg.hpp).
//
// Per level: strength of connection, greedy aggregation (Vanek, Mandel and
// Brezina's three passes), tentative prolongator T (one normalised column
// per aggregate), prolongator smoothing P = (I - w D^-1 B) T with
// w = 4 / (3 rho(D^-1 B)), Galerkin product B_c = P^T B P.  The coarsest
// operator is inverted densely (Cholesky).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "xg_amg.hpp"

namespace xg {
namespace {

template <class F>
void parallel_rows(int64_t n, F&& f) {
  int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  if (nt == 1) { f(0, n, 0); return; }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t r0 = t * chunk, r1 = std::min(n, r0 + chunk);
    if (r0 >= r1) break;
    th.emplace_back([&f, r0, r1, t] { f(r0, r1, t); });
  }
  for (auto& x : th) x.join();
}

int nthreads_for(int64_t n) {
  int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  return n < 4096 ? 1 : nt;
}

// C = A * B (rows of C built independently; each row's entries summed in
// the order (k in row i of A, then the entries of row k of B), columns
// sorted)
void spgemm(const Csr& A, const Csr& B, Csr& C) {
  const int64_t n = A.nrow, m = B.ncol;
  const int nt = nthreads_for(n);
  std::vector<std::vector<int32_t>> cols(n);
  std::vector<std::vector<double>> vals(n);
  parallel_rows(n, [&](int64_t r0, int64_t r1, int) {
    std::vector<double> acc(m, 0.0);
    std::vector<int64_t> mark(m, -1);
    std::vector<int32_t> touched;
    for (int64_t i = r0; i < r1; ++i) {
      touched.clear();
      for (int64_t p = A.ia[i]; p < A.ia[i + 1]; ++p) {
        const int32_t k = A.ja[p];
        const double av = A.a[p];
        for (int64_t q = B.ia[k]; q < B.ia[k + 1]; ++q) {
          const int32_t j = B.ja[q];
          if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; touched.push_back(j); }
          acc[j] += av * B.a[q];
        }
      }
      std::sort(touched.begin(), touched.end());
      cols[i] = touched;
      vals[i].resize(touched.size());
      for (size_t t = 0; t < touched.size(); ++t) vals[i][t] = acc[touched[t]];
    }
  });
  (void)nt;
  C.nrow = n;
  C.ncol = m;
  C.ia.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) C.ia[i + 1] = C.ia[i] + (int64_t)cols[i].size();
  C.ja.resize(C.ia[n]);
  C.a.resize(C.ia[n]);
  parallel_rows(n, [&](int64_t r0, int64_t r1, int) {
    for (int64_t i = r0; i < r1; ++i) {
      std::copy(cols[i].begin(), cols[i].end(), C.ja.begin() + C.ia[i]);
      std::copy(vals[i].begin(), vals[i].end(), C.a.begin() + C.ia[i]);
    }
  });
}

void transpose(const Csr& A, Csr& T) {
  T.nrow = A.ncol;
  T.ncol = A.nrow;
  T.ia.assign(T.nrow + 1, 0);
  for (int64_t p = 0; p < A.ia[A.nrow]; ++p) ++T.ia[A.ja[p] + 1];
  for (int64_t i = 0; i < T.nrow; ++i) T.ia[i + 1] += T.ia[i];
  T.ja.resize(A.ia[A.nrow]);
  T.a.resize(A.ia[A.nrow]);
  std::vector<int64_t> pos(T.ia.begin(), T.ia.end() - 1);
  for (int64_t i = 0; i < A.nrow; ++i)
    for (int64_t p = A.ia[i]; p < A.ia[i + 1]; ++p) {
      const int64_t q = pos[A.ja[p]]++;
      T.ja[q] = (int32_t)i;
      T.a[q] = A.a[p];
    }
}

// spectral radius of D^-1 B: power iteration from a fixed pseudo-random
// start (hash of the row index), 40 steps, times 1.05 -- capped by the
// Gershgorin bound max_i sum_j |b_ij / b_ii|, which it may not exceed
double rho_dinv(const Csr& B, const std::vector<double>& dinv) {
  const int64_t n = B.nrow;
  std::vector<double> x(n), y(n);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    x[i] = (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  double gersh = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p) s += std::fabs(B.a[p]);
    gersh = std::max(gersh, s * std::fabs(dinv[i]));
  }
  double lam = 0.0;
  for (int it = 0; it < 40; ++it) {
    double nx = 0.0;
    for (int64_t i = 0; i < n; ++i) nx += x[i] * x[i];
    nx = std::sqrt(nx);
    if (nx == 0.0) break;
    parallel_rows(n, [&](int64_t r0, int64_t r1, int) {
      for (int64_t i = r0; i < r1; ++i) {
        double s = 0.0;
        for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p) s += B.a[p] * x[B.ja[p]];
        y[i] = dinv[i] * s / nx;
      }
    });
    double ny = 0.0;
    for (int64_t i = 0; i < n; ++i) ny += y[i] * y[i];
    lam = std::sqrt(ny);
    x.swap(y);
  }
  lam *= 1.05;
  if (gersh > 0.0) lam = std::min(lam, gersh);
  return lam > 0.0 ? lam : 1.0;
}

// greedy aggregation on the strength graph; agg[i] = -1 for rows without
// strong connections.  Returns the number of aggregates.
int64_t aggregate(const Csr& B, double theta, std::vector<int64_t>& agg) {
  const int64_t n = B.nrow;
  std::vector<double> diag(n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p)
      if (B.ja[p] == i) diag[i] = B.a[p];
  // strong neighbours (excluding the diagonal)
  std::vector<int64_t> so(n + 1, 0);
  std::vector<int32_t> sj;
  std::vector<double> sv;
  sj.reserve(B.ia[n]);
  sv.reserve(B.ia[n]);
  const double t2 = theta * theta;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p) {
      const int32_t j = B.ja[p];
      if (j == i) continue;
      const double v = B.a[p];
      if (v * v >= t2 * std::fabs(diag[i] * diag[j]) && v != 0.0) {
        sj.push_back(j);
        sv.push_back(std::fabs(v));
      }
    }
    so[i + 1] = (int64_t)sj.size();
  }
  agg.assign(n, -1);
  int64_t na = 0;
  // pass 1: a node whose strong neighbours are all free seeds an aggregate
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0 || so[i] == so[i + 1]) continue;
    bool free = true;
    for (int64_t p = so[i]; p < so[i + 1] && free; ++p) free = agg[sj[p]] < 0;
    if (!free) continue;
    agg[i] = na;
    for (int64_t p = so[i]; p < so[i + 1]; ++p) agg[sj[p]] = na;
    ++na;
  }
  // pass 2: free nodes join the aggregate of their strongest aggregated
  // strong neighbour (aggregates of pass 1 only)
  std::vector<int64_t> agg1(agg);
  for (int64_t i = 0; i < n; ++i) {
    if (agg1[i] >= 0 || so[i] == so[i + 1]) continue;
    double best = -1.0;
    int64_t ba = -1;
    for (int64_t p = so[i]; p < so[i + 1]; ++p)
      if (agg1[sj[p]] >= 0 && sv[p] > best) { best = sv[p]; ba = agg1[sj[p]]; }
    if (ba >= 0) agg[i] = ba;
  }
  // pass 3: what is left forms aggregates with its free strong neighbours
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0 || so[i] == so[i + 1]) continue;
    agg[i] = na;
    for (int64_t p = so[i]; p < so[i + 1]; ++p)
      if (agg[sj[p]] < 0) agg[sj[p]] = na;
    ++na;
  }
  return na;
}

// P = (I - w D^-1 B) T, T(i, agg[i]) = 1 / sqrt(|agg|)
void smoothed_prolongator(const Csr& B, const std::vector<double>& dinv, double w,
                          const std::vector<int64_t>& agg, int64_t na, Csr& P) {
  const int64_t n = B.nrow;
  std::vector<double> size(na, 0.0);
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] >= 0) size[agg[i]] += 1.0;
  std::vector<double> t(n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] >= 0) t[i] = 1.0 / std::sqrt(size[agg[i]]);
  std::vector<std::vector<int32_t>> cols(n);
  std::vector<std::vector<double>> vals(n);
  parallel_rows(n, [&](int64_t r0, int64_t r1, int) {
    std::vector<std::pair<int32_t, double>> e;
    for (int64_t i = r0; i < r1; ++i) {
      e.clear();
      if (agg[i] >= 0) e.push_back({(int32_t)agg[i], t[i]});
      for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p) {
        const int32_t j = B.ja[p];
        if (agg[j] >= 0) e.push_back({(int32_t)agg[j], -w * dinv[i] * B.a[p] * t[j]});
      }
      std::stable_sort(e.begin(), e.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      for (size_t k = 0; k < e.size();) {
        size_t m = k;
        double s = 0.0;
        while (m < e.size() && e[m].first == e[k].first) s += e[m++].second;
        if (s != 0.0) { cols[i].push_back(e[k].first); vals[i].push_back(s); }
        k = m;
      }
    }
  });
  P.nrow = n;
  P.ncol = na;
  P.ia.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) P.ia[i + 1] = P.ia[i] + (int64_t)cols[i].size();
  P.ja.resize(P.ia[n]);
  P.a.resize(P.ia[n]);
  for (int64_t i = 0; i < n; ++i) {
    std::copy(cols[i].begin(), cols[i].end(), P.ja.begin() + P.ia[i]);
    std::copy(vals[i].begin(), vals[i].end(), P.a.begin() + P.ia[i]);
  }
}

void diag_inverse(const Csr& B, std::vector<double>& dinv) {
  dinv.assign(B.nrow, 0.0);
  for (int64_t i = 0; i < B.nrow; ++i)
    for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p)
      if (B.ja[p] == i && B.a[p] != 0.0) dinv[i] = 1.0 / B.a[p];
}

// dense inverse of an SPD matrix by Cholesky; false if not positive definite
bool dense_spd_inverse(const Csr& B, std::vector<double>& inv) {
  const int64_t n = B.nrow;
  std::vector<double> L((size_t)n * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p) L[(size_t)i * n + B.ja[p]] = B.a[p];
  for (int64_t j = 0; j < n; ++j) {
    double d = L[(size_t)j * n + j];
    for (int64_t k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    L[(size_t)j * n + j] = d;
    for (int64_t i = j + 1; i < n; ++i) {
      double s = L[(size_t)i * n + j];
      for (int64_t k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / d;
    }
    for (int64_t i = 0; i < j; ++i) L[(size_t)i * n + j] = 0.0;
  }
  // inv = L^-T L^-1, column by column
  inv.assign((size_t)n * n, 0.0);
  std::vector<double> y(n);
  for (int64_t c = 0; c < n; ++c) {
    for (int64_t i = 0; i < n; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int64_t k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
      y[i] = s / L[(size_t)i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int64_t k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * inv[(size_t)k * n + c];
      inv[(size_t)i * n + c] = s / L[(size_t)i * n + i];
    }
  }
  return true;
}

}  // namespace

int amg_setup(int64_t n, const int64_t* ia, const int32_t* ja, const double* a,
              const AmgOptions& opt, AmgHierarchy& h, std::string& err) {
  const auto t0 = std::chrono::steady_clock::now();
  h = AmgHierarchy{};
  AmgLevel L0;
  Csr& B = L0.B;
  B.nrow = B.ncol = n;
  B.ia.assign(ia, ia + n + 1);
  B.ja.assign(ja, ja + ia[n]);
  B.a.assign(a, a + ia[n]);
  // S A with S = sign(diag A)
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p)
      if (B.ja[p] == i) d = B.a[p];
    if (d < 0.0)
      for (int64_t p = B.ia[i]; p < B.ia[i + 1]; ++p) B.a[p] = -B.a[p];
  }
  h.lev.push_back(std::move(L0));
  const double nnz0 = (double)h.lev[0].B.ia[n];
  double nnz_sum = nnz0;
  for (int l = 0;; ++l) {
    AmgLevel& cur = h.lev[l];
    diag_inverse(cur.B, cur.dinv);
    const double rho = rho_dinv(cur.B, cur.dinv);
    cur.omega = 4.0 / (3.0 * rho);
    if (cur.B.nrow <= opt.max_coarse || l + 1 >= opt.max_levels) break;
    std::vector<int64_t> agg;
    const int64_t na = aggregate(cur.B, opt.theta, agg);
    if (na == 0 || (double)na > 0.8 * (double)cur.B.nrow) break;
    smoothed_prolongator(cur.B, cur.dinv, cur.omega, agg, na, cur.P);
    transpose(cur.P, cur.R);
    Csr BP;
    spgemm(cur.B, cur.P, BP);
    AmgLevel nxt;
    spgemm(cur.R, BP, nxt.B);
    nnz_sum += (double)nxt.B.ia[nxt.B.nrow];
    h.lev.push_back(std::move(nxt));
  }
  AmgLevel& last = h.lev.back();
  if (last.B.nrow > 4 * opt.max_coarse) {
    err = "AMG: coarsening stalled at " + std::to_string(last.B.nrow) + " unknowns";
    return XG_EUNSUPPORTED;
  }
  if (!dense_spd_inverse(last.B, h.cinv)) {
    err = "AMG: coarsest operator not positive definite";
    return XG_EUNSUPPORTED;
  }
  h.op_complexity = nnz_sum / nnz0;
  h.setup_ms = std::chrono::duration<double, std::milli>(
                   std::chrono::steady_clock::now() - t0).count();
  return XG_OK;
}

}  // namespace xg
