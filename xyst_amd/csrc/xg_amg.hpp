// Smoothed-aggregation algebraic multigrid for the pressure-Poisson system
// (no counterpart in the reference): host setup (xg_amg.cpp) and the FP32
// V-cycle on the device (xg_amg.hip), used as a preconditioner of the FP64
// flexible CG (xg_cg.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "xg_internal.hpp"

namespace xg {

// CSR with 0-based int64 row offsets and int32 columns (sorted per row)
struct Csr {
  int64_t nrow = 0, ncol = 0;
  std::vector<int64_t> ia;
  std::vector<int32_t> ja;
  std::vector<double> a;
};

struct AmgLevel {
  Csr B;                      // operator of this level
  std::vector<double> dinv;   // 1 / diag(B) (0 for empty rows)
  double omega = 0.0;         // Jacobi smoother damping, 4 / (3 rho(D^-1 B))
  Csr P, R;                   // prolongator to this level from the next, R = P^T
};

struct AmgHierarchy {
  std::vector<AmgLevel> lev;  // lev[0]: the fine operator
  std::vector<double> cinv;   // dense inverse of the coarsest operator (row major)
  double op_complexity = 0.0; // sum nnz(B_l) / nnz(B_0)
  double setup_ms = 0.0;
};

struct AmgOptions {
  double theta = 0.03;        // strength of connection |b_ij| >= theta sqrt(|b_ii b_jj|)
  int64_t max_coarse = 800;   // stop coarsening at this many unknowns
  int max_levels = 12;
};

// Builds the hierarchy of the SPD matrix S A, S = diag(sign(diag A)) (the
// ChoCG matrix is -K with +1 Dirichlet rows).  Rows without strong
// off-diagonal entries (decoupled Dirichlet rows) stay out of every
// aggregate.  Deterministic (fixed visiting order; rows are built in
// parallel but each row's sums have a fixed order).
int amg_setup(int64_t n, const int64_t* ia, const int32_t* ja, const double* a,
              const AmgOptions& opt, AmgHierarchy& h, std::string& err);

}  // namespace xg

// Device side (xg_amg.hip), as seen by the CG driver.
// (Re)builds and uploads the hierarchy for the matrix (n, ia, ja, a) when
// `version` differs from the one uploaded.
int xg_amg_prepare_internal(xg_ctx* c, int64_t n, const int64_t* ia, const int32_t* ja,
                            const double* a, int64_t version);
// z = V-cycle(S r) (FP32 inside), r and z FP64 device vectors of n
int xg_amg_apply_internal(xg_ctx* c, const double* r, double* z);
void xg_amg_free(xg_ctx* c);
