// Device evaluation of the source terms of the analytic problems the path
// serves (problems::SRC(), Problems.cpp:1299-1325) and the boundary
// conditions of the explicit solvers (physics::dirbc/farbc/prebc, BC.cpp),
// shared by the RieCG (xg_riemann.hip) and KozCG (xg_kozak.hip) sources.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/xg.h"

namespace xg {

// The problem parameters the kernels read (kind as in xg_problem).
struct Prob {
  int kind;
  double alpha, kappa, p0;
  double g;     // ratio of specific heats
};

__host__ __device__ inline bool has_source(const Prob& P) {
  return P.kind == XG_PROBLEM_VORTICAL_FLOW || P.kind == XG_PROBLEM_TAYLOR_GREEN ||
         P.kind == XG_PROBLEM_SLOT_CYL;
}

// SRC() of the problem at (x, y, z); the sources are independent of time.
//  vortical_flow::src (Problems.cpp:481-507): momentum (a u - k v, k u + a v)
//    of the initial state's velocity, energy s.u + 8 a^3 z^2/(g-1);
//  taylor_green::src (Problems.cpp:433-448): energy only;
//  slot_cyl::src (Problems.cpp:625-670): the momentum (-rho v, rho u) of the
//    rotation u = 0.5 - y, v = x - 0.5 at unit density (the centripetal
//    force); the transported scalars have no source.
__device__ inline void source(const Prob& P, double x, double y, double z,
                              double s[5]) {
  s[0] = s[1] = s[2] = s[3] = s[4] = 0.0;
  if (P.kind == XG_PROBLEM_VORTICAL_FLOW) {
    const double a = P.alpha, k = P.kappa;
    // the initial state (vortical_flow::ic): rho = 1, momentum
    const double ru = a * x - k * y, rv = k * x + a * y;
    s[1] = a * ru - k * rv;
    s[2] = k * ru + a * rv;
    s[4] = (s[1] * ru + s[2] * rv) + 8.0 * a * a * a * z * z / (P.g - 1.0);
  } else if (P.kind == XG_PROBLEM_TAYLOR_GREEN) {
    constexpr double pi = 3.14159265358979323846;
    s[4] = 3.0 * pi / 8.0 * (cos(3.0 * pi * x) * cos(pi * y) -
                             cos(3.0 * pi * y) * cos(pi * x));
  } else if (P.kind == XG_PROBLEM_SLOT_CYL) {
    s[1] = -(x - 0.5);
    s[2] = 0.5 - y;
  }
}

// eos::totalenergy (EOS.hpp:53-56)
__device__ inline double total_energy(double g, double r, double u, double v,
                                      double w, double p) {
  return p / (g - 1.0) + 0.5 * r * (u * u + v * v + w * w);
}

}  // namespace xg
