// Internal declarations shared by the host preprocessing and the device code.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xg.h"

namespace xg {

int mesh_vol(int64_t npoin, const double* x, const double* y, const double* z,
             int64_t ntet, const size_t* inpoel, double* vol,
             std::string& err);
int mesh_domint(int64_t npoin, const double* x, const double* y,
                const double* z, int64_t ntet, const size_t* inpoel,
                std::vector<size_t>& edge, std::vector<double>& dint,
                bool count_only, std::string& err, int stride = 3);
int mesh_rcm(int64_t npoin, int64_t ntet, const size_t* inpoel, size_t* perm,
             std::string& err);

}  // namespace xg
