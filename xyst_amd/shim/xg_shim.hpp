// Reference-side binding of libxg.so: what the replaced bodies of
// riemann::grad/rhs (Riemann.hpp:19-39), kozak::rhs (Kozak.hpp:19-27) and
// chorin::div/grad/vgrad/flux/rhs (Chorin.hpp:19-70) share.  These files go into the reference's src/Physics next to the
// headers they implement and are compiled by its own build, linked against
// libxg.so; nothing here is part of libxg.so itself.
#pragma once

#include <string>

#include "Exception.hpp"   // tk::Exception and Throw (Exception.hpp:40-41)
#include "xg.h"

namespace xg_shim {

// The device context of the chare whose entry method runs on this thread.
// A chare sets it before its entry method calls into riemann:: or kozak::
// (one entry method runs at a time on a PE's thread, so a thread-local is
// enough in SMP mode too).
extern thread_local xg_ctx* t_ctx;

// g_cfg's settings of the chorin:: operators (flux "damp2" 0 / "damp4" 1,
// stab, stab2, stab2coef, mat_dyn_viscosity, mat_dyn_diffusivity), which the
// reference's bodies read from g_cfg: set once after the control file is read.
extern xg_chocfg chorin_cfg;

// A failed C-ABI call becomes a tk::Exception with the library's message,
// as the reference's ErrChk/Throw report errors.
inline void check(xg_ctx* c, int rc) {
  if (rc != XG_OK) Throw(std::string("xg: ") + xg_last_error(c));
}

// The device problem of g_cfg's problem name (problem_alpha, problem_kappa,
// problem_p0 for vortical_flow).  Problems with a source term the device
// does not evaluate are refused here, so that a wrong rhs cannot be
// computed silently.
void set_problem(xg_ctx* c, const std::string& name, double alpha,
                 double kappa, double p0);

}  // namespace xg_shim
