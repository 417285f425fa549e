"""Python mirror of kozak::rhs (Kozak.hpp:19-27) and of the KozCG time step
(KozCG.cpp:614-1197) over the C-ABI.

The mesh arguments (inpoel, coord) are uploaded once per context with
`upload`; `rhs(ctx, dt, U)` then has the argument meaning of kozak::rhs for
an unsteady run without a source term (t, tp and dtp unused).  `setup`,
`dt_dev` and `step_dev` advance a GPU-resident KozCG solution: rhs, FCT
(aec, alw, lim), solve and the Dirichlet and symmetry BCs.
"""
import ctypes as C

import numpy as np

from ._xg import FctConfig, lib, ptr

__all__ = ["upload", "rhs", "rhs_dev", "setup", "dt_dev", "step_dev"]


def upload(ctx, coord, inpoel):
    x, y, z = (np.ascontiguousarray(coord[i], dtype=np.float64) for i in range(3))
    inp = np.ascontiguousarray(inpoel, dtype=np.uint64)
    ctx._check(lib().xg_kozak_upload(ctx._h, x.size, ptr(x), ptr(y), ptr(z),
                                     inp.size // 4, ptr(inp)))
    ctx.koz_npoin = x.size


def rhs(ctx, dt, U, t=0.0):
    """R = kozak::rhs(U) for time step dt; host arrays (npoin, 5)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.shape != (ctx.koz_npoin, 5):
        raise ValueError("U must be (%d, 5)" % ctx.koz_npoin)
    R = np.empty_like(U)
    ctx._check(lib().xg_kozak_rhs(ctx._h, 5, float(t), float(dt), None, None, ptr(U),
                                  ptr(R)))
    return R


def rhs_dev(ctx, dt, U, R, t=0.0):
    """Device tensors U, R (npoin, 5) float64; enqueued on the context stream."""
    for a in (U, R):
        if tuple(a.shape) != (ctx.koz_npoin, 5) or not a.is_cuda:
            raise ValueError("device fields must be (%d, 5) on the GPU" % ctx.koz_npoin)
    ctx._check(lib().xg_kozak_rhs_dev(ctx._h, 5, float(t), float(dt), None, None, ptr(U),
                                      ptr(R)))


def setup(ctx, vol, symnodes, symnorms, dirnodes=(), dirmask=None, dirvals=None,
          fct=True, fctdif=1.0, fctclip=False, fctsys=()):
    """KozCG per-mesh data: nodal volumes, symmetry (node, normal) pairs,
    Dirichlet nodes with (ndir, 5) mask and values, FCT settings (fctsys as
    0-based component ids)."""
    v = np.ascontiguousarray(vol, dtype=np.float64)
    sn = np.ascontiguousarray(symnodes, dtype=np.uint64)
    sv = np.ascontiguousarray(symnorms, dtype=np.float64).reshape(-1)
    dn = np.ascontiguousarray(dirnodes, dtype=np.uint64)
    nd = dn.size
    dm = np.ascontiguousarray(np.ones((nd, 5)) if dirmask is None else dirmask,
                              dtype=np.int32).reshape(-1)
    dv = np.ascontiguousarray(np.zeros((nd, 5)) if dirvals is None else dirvals,
                              dtype=np.float64).reshape(-1)
    if v.size != ctx.koz_npoin or sv.size != 3 * sn.size or dm.size != 5 * nd \
            or dv.size != 5 * nd:
        raise ValueError("setup: inconsistent array sizes")
    if len(fctsys) > 5:
        raise ValueError("fctsys holds at most 5 components")
    f = FctConfig(int(bool(fct)), float(fctdif), int(bool(fctclip)), len(fctsys))
    for i, c in enumerate(fctsys):
        f.sys[i] = int(c)
    ctx._check(lib().xg_kozcg_setup(ctx._h, ptr(v), sn.size, ptr(sn), ptr(sv), nd,
                                    ptr(dn), ptr(dm), ptr(dv), C.byref(f)))


def dt_dev(ctx, U, cfl):
    """KozCG::dt, unsteady: cfl * min cbrt(vol)/max(|u|+c, 1e-8)."""
    out = C.c_double()
    ctx._check(lib().xg_kozcg_dt_dev(ctx._h, ptr(U), float(cfl), C.byref(out)))
    return out.value


def step_dev(ctx, dt, U, Un, t=0.0):
    """One KozCG step on device tensors (npoin, 5): U advanced in place, Un
    receives the old solution; enqueued on the context stream."""
    for a in (U, Un):
        if tuple(a.shape) != (ctx.koz_npoin, 5) or not a.is_cuda:
            raise ValueError("device fields must be (%d, 5) on the GPU" % ctx.koz_npoin)
    ctx._check(lib().xg_kozcg_step_dev(ctx._h, float(t), float(dt), ptr(U), ptr(Un)))
