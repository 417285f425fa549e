// Stand-in for the reference's Problems.hpp: problems::SRC() with the
// reference's return type; the test harness decides whether it is set.
#pragma once
#include <cstddef>
#include <functional>
#include <vector>

#include "Fields.hpp"

namespace problems {

std::function<std::vector<tk::real>(tk::real, tk::real, tk::real, tk::real, std::size_t)>
SRC();

}  // namespace problems
