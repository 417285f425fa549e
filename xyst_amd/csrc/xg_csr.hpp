// Host-side CSR helpers of the pressure-Poisson solver (xg_csr.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "xg_internal.hpp"

namespace xg {

int csr_pattern(int64_t npoin, int64_t ntet, const size_t* inpoel, int ncomp,
                std::vector<int64_t>& ia, std::vector<int32_t>& ja,
                std::string& err);

int poisson_matrix(int64_t npoin, const double* x, const double* y,
                   const double* z, int64_t ntet, const size_t* inpoel,
                   const std::vector<int64_t>& ia,
                   const std::vector<int32_t>& ja, std::vector<double>& a,
                   std::string& err);

void csr_dirichlet(int ncomp, const int64_t* ia, const int32_t* ja, double* a,
                   int64_t node, double val, double* b, int pos);

}  // namespace xg
