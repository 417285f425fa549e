// kozak::rhs (Kozak.hpp:19-27) on libxg.so.  The tet mesh was uploaded once
// per mesh by xg_kozak_upload (KozCG::setup); the body passes the fields, the
// time, the time step and the steady-state arrays.  The library refuses what
// it does not evaluate (more than two transported scalars) with
// XG_EUNSUPPORTED, which becomes a tk::Exception here.
#include <string>
#include <vector>

#include "Kozak.hpp"
#include "Problems.hpp"
#include "xg_shim.hpp"

namespace kozak {

void rhs(const std::vector<std::size_t>& /*inpoel*/,
         const std::array<std::vector<tk::real>, 3>& /*coord*/, tk::real t,
         tk::real dt, const std::vector<tk::real>& tp,
         const std::vector<tk::real>& dtp, const tk::Fields& U, tk::Fields& R) {
  xg_ctx* c = xg_shim::t_ctx;
  if (R.nunk() != U.nunk() || R.nprop() != U.nprop())
    Throw("xg: kozak::rhs: U and R sizes disagree");
  if ((!tp.empty() && tp.size() != U.nunk()) || (!dtp.empty() && dtp.size() != U.nunk()))
    Throw("xg: kozak::rhs: tp and dtp need one value per node");
  if (problems::SRC()) {
    xg_problem p;
    xg_shim::check(c, xg_get_problem(c, &p));
    if (p.kind == XG_PROBLEM_NONE)
      Throw("xg: the problem has a source term but none was set on the device "
            "(xg_shim::set_problem)");
  }
  xg_shim::check(c, xg_kozak_rhs(c, static_cast<int>(U.nprop()), t, dt,
                                 tp.empty() ? nullptr : tp.data(),
                                 dtp.empty() ? nullptr : dtp.data(), U.vec().data(),
                                 R.vec().data()));
}

}  // namespace kozak
