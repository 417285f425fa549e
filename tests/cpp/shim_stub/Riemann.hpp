// The declarations of the reference's Riemann.hpp:19-39 that the shim
// implements (same parameter types and order).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "Fields.hpp"

namespace riemann {

void grad(const std::array<std::vector<std::size_t>, 3>& dsupedge,
          const std::array<std::vector<tk::real>, 3>& dsupint,
          const std::array<std::vector<tk::real>, 3>& coord,
          const std::vector<std::size_t>& triinpoel, const tk::Fields& U, tk::Fields& G);

void rhs(const std::array<std::vector<std::size_t>, 3>& dsupedge,
         const std::array<std::vector<tk::real>, 3>& dsupint,
         const std::array<std::vector<tk::real>, 3>& coord,
         const std::vector<std::size_t>& triinpoel, const std::vector<std::uint8_t>& besym,
         const tk::Fields& G, const tk::Fields& U, const std::vector<tk::real>& v,
         tk::real t, const std::vector<tk::real>& tp, tk::Fields& R);

}  // namespace riemann
