"""Python mirror of kozak::rhs (Kozak.hpp:19-27) over the C-ABI.

The mesh arguments (inpoel, coord) are uploaded once per context with
`upload`; `rhs(ctx, dt, U)` then has the argument meaning of kozak::rhs for
an unsteady run without a source term (t, tp and dtp unused).
"""
import numpy as np

from ._xg import lib, ptr

__all__ = ["upload", "rhs", "rhs_dev"]


def upload(ctx, coord, inpoel):
    x, y, z = (np.ascontiguousarray(coord[i], dtype=np.float64) for i in range(3))
    inp = np.ascontiguousarray(inpoel, dtype=np.uint64)
    ctx._check(lib().xg_kozak_upload(ctx._h, x.size, ptr(x), ptr(y), ptr(z),
                                     inp.size // 4, ptr(inp)))
    ctx.koz_npoin = x.size


def rhs(ctx, dt, U):
    """R = kozak::rhs(U) for time step dt; host arrays (npoin, 5)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.shape != (ctx.koz_npoin, 5):
        raise ValueError("U must be (%d, 5)" % ctx.koz_npoin)
    R = np.empty_like(U)
    ctx._check(lib().xg_kozak_rhs(ctx._h, 5, float(dt), ptr(U), ptr(R)))
    return R


def rhs_dev(ctx, dt, U, R):
    """Device tensors U, R (npoin, 5) float64; enqueued on the context stream."""
    for a in (U, R):
        if tuple(a.shape) != (ctx.koz_npoin, 5) or not a.is_cuda:
            raise ValueError("device fields must be (%d, 5) on the GPU" % ctx.koz_npoin)
    ctx._check(lib().xg_kozak_rhs_dev(ctx._h, 5, float(dt), ptr(U), ptr(R)))
