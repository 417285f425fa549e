// riemann::grad and riemann::rhs (Riemann.hpp:19-39) on libxg.so.  The
// signatures are the reference's; the mesh arguments were uploaded once per
// mesh by xg_mesh_upload (RieCG::setup, RieCG.cpp:109-245), so the bodies
// pass only the fields.
#include <cstdint>
#include <string>
#include <vector>

#include "Problems.hpp"
#include "Riemann.hpp"
#include "xg_shim.hpp"

namespace xg_shim {

thread_local xg_ctx* t_ctx = nullptr;

void set_problem(xg_ctx* c, const std::string& name, double alpha,
                 double kappa, double p0) {
  xg_problem p{XG_PROBLEM_NONE, alpha, kappa, p0};
  if (name == "vortical_flow")
    p.kind = XG_PROBLEM_VORTICAL_FLOW;
  else if (name == "taylor_green")
    p.kind = XG_PROBLEM_TAYLOR_GREEN;
  else if (name == "slot_cyl")
    p.kind = XG_PROBLEM_SLOT_CYL;
  else if (problems::SRC())
    Throw("xg: the source term of problem '" + name +
          "' is not evaluated on the device");
  check(c, xg_set_problem(c, &p));
}

namespace {

// problems::SRC() set on the host but not on the device: the rhs would
// miss the source term
void check_source(xg_ctx* c) {
  if (!problems::SRC()) return;
  xg_problem p;
  check(c, xg_get_problem(c, &p));
  if (p.kind == XG_PROBLEM_NONE)
    Throw("xg: the problem has a source term but none was set on the device "
          "(xg_shim::set_problem)");
}

}  // namespace

}  // namespace xg_shim

namespace riemann {

void grad(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
          const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
          const std::array<std::vector<tk::real>, 3>& /*coord*/,
          const std::vector<std::size_t>& /*triinpoel*/, const tk::Fields& U,
          tk::Fields& G) {
  xg_ctx* c = xg_shim::t_ctx;
  if (G.nunk() != U.nunk() || G.nprop() != 3 * U.nprop())
    Throw("xg: riemann::grad: G must hold 3 gradients per component of U");
  xg_shim::check(c, xg_riemann_grad(c, static_cast<int>(U.nprop()), U.vec().data(),
                                    G.vec().data()));
}

void rhs(const std::array<std::vector<std::size_t>, 3>& /*dsupedge*/,
         const std::array<std::vector<tk::real>, 3>& /*dsupint*/,
         const std::array<std::vector<tk::real>, 3>& /*coord*/,
         const std::vector<std::size_t>& /*triinpoel*/,
         const std::vector<std::uint8_t>& /*besym*/, const tk::Fields& G,
         const tk::Fields& U, const std::vector<tk::real>& v, tk::real t,
         const std::vector<tk::real>& tp, tk::Fields& R) {
  xg_ctx* c = xg_shim::t_ctx;
  if (R.nunk() != U.nunk() || R.nprop() != U.nprop() || G.nunk() != U.nunk() ||
      G.nprop() != 3 * U.nprop())
    Throw("xg: riemann::rhs: G, U and R sizes disagree");
  if ((!v.empty() && v.size() != U.nunk()) || (!tp.empty() && tp.size() != U.nunk()))
    Throw("xg: riemann::rhs: v and tp need one value per node");
  xg_shim::check_source(c);
  xg_shim::check(c, xg_riemann_rhs(c, static_cast<int>(U.nprop()), G.vec().data(),
                                   U.vec().data(), v.empty() ? nullptr : v.data(), t,
                                   tp.empty() ? nullptr : tp.data(), R.vec().data()));
}

}  // namespace riemann
