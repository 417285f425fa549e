/* Multi-threaded riemann::rhs of the oracle, for the CPU baseline: the
 * reference runs one chare per core, each on its own mesh partition, and
 * sums the partitions' nodal contributions across shared nodes (comrhs,
 * RieCG.cpp:955-965).  Here each of nthread OpenMP threads takes a
 * contiguous share of every superedge list and of the boundary triangles
 * (contiguous in a locality-ordered mesh, so a compact partition), evaluates
 * orc_rhs on it into its own nodal array, and the arrays are then summed per
 * node in thread order: deterministic for a given thread count, equal to the
 * single-thread result up to the order of the sums.  TEST INFRASTRUCTURE
 * (the CPU baseline and its check), not part of the product. */
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "xyst_oracle.h"

int orc_rhs_mt(int nthread, const orc_cfg* cfg, int64_t npoin, int ncomp,
               const int64_t* nsup, const int64_t* sup0, const double* int0,
               const int64_t* sup1, const double* int1,
               const int64_t* sup2, const double* int2,
               const double* x, const double* y, const double* z,
               int64_t nbtri, const int64_t* triinpoel, const uint8_t* besym,
               const double* G, const double* U, double* R) {
  if (nthread < 1) return -3;
  const size_t nv = (size_t)(npoin * ncomp);
  double* part = (double*)malloc(sizeof(double) * nv * (size_t)nthread);
  if (!part) return -4;
  int rc = 0;
#pragma omp parallel num_threads(nthread) reduction(| : rc)
  {
    const int t = omp_get_thread_num();
    const int T = omp_get_num_threads();
    int64_t ns[3], b[4];
    for (int k = 0; k < 3; ++k) {
      const int64_t lo = nsup[k] * t / T, hi = nsup[k] * (t + 1) / T;
      ns[k] = hi - lo;
      b[k] = lo;
    }
    b[3] = nbtri * t / T;
    const int64_t nb = nbtri * (t + 1) / T - b[3];
    rc |= orc_rhs(cfg, npoin, ncomp, ns, sup0 + b[0] * 4, int0 + b[0] * 18,
                  sup1 + b[1] * 3, int1 + b[1] * 9, sup2 + b[2] * 2, int2 + b[2] * 3,
                  x, y, z, nb, triinpoel + b[3] * 3, besym + b[3] * 3, G, U,
                  part + nv * (size_t)t, NULL);
#pragma omp barrier
    /* the shared-node sum, in partition order */
    const size_t lo = nv * (size_t)t / (size_t)T, hi = nv * (size_t)(t + 1) / (size_t)T;
    for (size_t i = lo; i < hi; ++i) {
      double s = 0.0;
      for (int k = 0; k < T; ++k) s += part[nv * (size_t)k + i];
      R[i] = s;
    }
  }
  free(part);
  return rc;
}
