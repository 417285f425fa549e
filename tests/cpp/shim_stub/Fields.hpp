// Stand-in for the reference's tk::Fields (Fields.hpp, Data.hpp:420-426) with
// only what the shim uses: unknown-major storage, nunk(), nprop(), vec().
#pragma once
#include <cstddef>
#include <vector>

namespace tk {

using real = double;

class Fields {
 public:
  Fields() = default;
  Fields(std::size_t nunk, std::size_t nprop)
      : m_vec(nunk * nprop, 0.0), m_nunk(nunk), m_nprop(nprop) {}
  std::size_t nunk() const noexcept { return m_nunk; }
  std::size_t nprop() const noexcept { return m_nprop; }
  const std::vector<real>& vec() const { return m_vec; }
  std::vector<real>& vec() { return m_vec; }
  real& operator()(std::size_t i, std::size_t c) { return m_vec[i * m_nprop + c]; }

 private:
  std::vector<real> m_vec;
  std::size_t m_nunk = 0, m_nprop = 0;
};

}  // namespace tk
