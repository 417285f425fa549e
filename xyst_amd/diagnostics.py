"""Python mirror of the node diagnostics (NodeDiagnostics.cpp:46-264) over
the C-ABI: the volume-weighted sums of the solution, its change over the
step and (density-based solvers) the total energy, computed on the device
from device tensors."""
import numpy as np

from ._xg import lib, ptr

__all__ = ["rhocompute", "errcompute", "precompute"]


def rhocompute(ctx, U, Un, V):
    """Sums of U_c^2 V, (U_c - Un_c)^2 V and U_4 V; U, Un (npoin, ncomp),
    V (npoin,) device tensors, ncomp 1, 3 or 5..7 (5 + transported
    scalars)."""
    n, nc = int(U.shape[0]), int(U.shape[1])
    out = np.empty(2 * nc + 1)
    ctx._check(lib().xg_diag_rho_dev(ctx._h, n, nc, ptr(U), ptr(Un), ptr(V), ptr(out)))
    return out


def errcompute(ctx, U, V, An):
    """The analytic-error sums of rhocompute: (nu_c - An_c)^2 V (ncomp
    values) and |nu_c - An_c| V (ncomp values), nu the primitive form of U
    (rho, u, v, w, specific internal energy, then the scalars) and An the
    analytic solution in that form; U, An (npoin, ncomp), V (npoin,) device
    tensors, ncomp 5..7; or the pressure-based solvers' pressure error
    (NodeDiagnostics.cpp:237-241): U the pressure, An the analytic pressure,
    both (npoin,)."""
    nc = 1 if U.dim() == 1 else int(U.shape[1])
    if tuple(An.shape) != tuple(U.shape):
        raise ValueError("An must have the shape of U")
    out = np.empty(2 * nc)
    ctx._check(lib().xg_diag_err_dev(ctx._h, int(U.shape[0]), nc, ptr(U), ptr(V), ptr(An),
                                     ptr(out)))
    return out


def precompute(ctx, U, Un, P, dP, V):
    """Sums of p^2 V, u_c^2 V, dp^2 V, (u_c - un_c)^2 V for the pressure-based
    solver; U, Un (npoin, ncomp: velocity, and a transported scalar for
    ncomp 4), P, dP, V (npoin,) device tensors."""
    nc = int(U.shape[1])
    out = np.empty(2 * nc + 2)
    ctx._check(lib().xg_diag_pre_nc_dev(ctx._h, int(U.shape[0]), nc, ptr(U), ptr(Un), ptr(P),
                                        ptr(dP), ptr(V), ptr(out)))
    return out
