// The declaration of the reference's Kozak.hpp:19-27 that the shim
// implements (same parameter types and order).
#pragma once
#include <array>
#include <vector>

#include "Fields.hpp"

namespace kozak {

void rhs(const std::vector<std::size_t>& inpoel,
         const std::array<std::vector<tk::real>, 3>& coord, tk::real t, tk::real dt,
         const std::vector<tk::real>& tp, const std::vector<tk::real>& dtp,
         const tk::Fields& U, tk::Fields& R);

}  // namespace kozak
