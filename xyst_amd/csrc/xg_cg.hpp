// Internal access to the context's ConjugateGradients solver (xg_cg.hip) for
// a caller that keeps a second solver on the same context (ChoCG's
// momentum system beside its pressure system) and assembles that solver's
// matrix values on the device.
#pragma once
#include "xg_ctx.hpp"

// makes s the context's solver and returns the previous one (either may be
// null); the CG entry points act on the context's solver
CgSolver* xg_cg_exchange(xg_ctx* c, CgSolver* s);
void xg_cg_free_solver(CgSolver* s);
// the solver's device matrix values (FP64), owned from now on by the caller:
// no Dirichlet rebuild or restore touches them
double* xg_cg_values_dev(xg_ctx* c);
// the solver's device right-hand side; then, after the caller has written
// it, xg_cg_setup / xg_cg_solve as usual with preconditioner pc
double* xg_cg_rhs_dev(xg_ctx* c);
int xg_cg_raw_init(xg_ctx* c, int pc);
