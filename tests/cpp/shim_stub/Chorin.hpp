// The declarations of the reference's Chorin.hpp:19-70 that the shim
// implements (same parameter types and order).
#pragma once
#include <array>
#include <vector>

#include "Fields.hpp"

namespace chorin {

void div(const std::array<std::vector<std::size_t>, 3>& dsupedge,
         const std::array<std::vector<tk::real>, 3>& dsupint,
         const std::array<std::vector<tk::real>, 3>& coord,
         const std::vector<std::size_t>& triinpoel, tk::real dt,
         const std::vector<tk::real>& P, const tk::Fields& G, const tk::Fields& U,
         std::vector<tk::real>& D, bool stab);

void grad(const std::array<std::vector<std::size_t>, 3>& dsupedge,
          const std::array<std::vector<tk::real>, 3>& dsupint,
          const std::array<std::vector<tk::real>, 3>& coord,
          const std::vector<std::size_t>& triinpoel, const std::vector<tk::real>& U,
          tk::Fields& G);

void vgrad(const std::array<std::vector<std::size_t>, 3>& dsupedge,
           const std::array<std::vector<tk::real>, 3>& dsupint,
           const std::array<std::vector<tk::real>, 3>& coord,
           const std::vector<std::size_t>& triinpoel, const tk::Fields& U, tk::Fields& G);

void flux(const std::array<std::vector<std::size_t>, 3>& dsupedge,
          const std::array<std::vector<tk::real>, 3>& dsupint,
          const std::array<std::vector<tk::real>, 3>& coord,
          const std::vector<std::size_t>& triinpoel, const tk::Fields& U,
          const tk::Fields& G, tk::Fields& F);

void rhs(const std::array<std::vector<std::size_t>, 3>& dsupedge,
         const std::array<std::vector<tk::real>, 3>& dsupint,
         const std::array<std::vector<tk::real>, 3>& coord,
         const std::vector<std::size_t>& triinpoel, const std::vector<tk::real>& v,
         tk::real t, const std::vector<tk::real>& P, const tk::Fields& U,
         const tk::Fields& G, tk::Fields& R);

}  // namespace chorin
