"""Python mirror of kozak::rhs (Kozak.hpp:19-27) and of the KozCG time step
(KozCG.cpp:614-1197) over the C-ABI.

The mesh arguments (inpoel, coord) are uploaded once per context with
`upload`; `rhs(ctx, dt, U)` then has the argument meaning of kozak::rhs
(ncomp = U's columns: 5 + transported scalars; dtp, the nodes' local time
steps of a steady run, optional).  `setup`, `dt_dev`, `dt_steady_dev` and
`step_dev` advance a GPU-resident KozCG solution: rhs, FCT (aec, alw, lim),
solve and the Dirichlet and symmetry BCs, with local time stepping and
freezeflow when asked.
"""
import ctypes as C

import numpy as np

from ._xg import FctConfig, lib, ptr

__all__ = ["upload", "rhs", "rhs_dev", "setup", "dt_dev", "dt_steady_dev", "step_dev"]


def upload(ctx, coord, inpoel):
    x, y, z = (np.ascontiguousarray(coord[i], dtype=np.float64) for i in range(3))
    inp = np.ascontiguousarray(inpoel, dtype=np.uint64)
    ctx._check(lib().xg_kozak_upload(ctx._h, x.size, ptr(x), ptr(y), ptr(z),
                                     inp.size // 4, ptr(inp)))
    ctx.koz_npoin = x.size
    ctx.koz_ncomp = 5


def _shape(ctx, U):
    if U.ndim != 2 or U.shape[0] != ctx.koz_npoin:
        raise ValueError("U must be (%d, ncomp)" % ctx.koz_npoin)
    return U.shape[1]


def rhs(ctx, dt, U, t=0.0, dtp=None):
    """R = kozak::rhs(U) for time step dt (or the local steps dtp); host
    arrays (npoin, ncomp)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    nc = _shape(ctx, U)
    d = None
    if dtp is not None:
        d = np.ascontiguousarray(dtp, dtype=np.float64)
        if d.shape != (ctx.koz_npoin,):
            raise ValueError("dtp must have npoin entries")
    R = np.empty_like(U)
    ctx._check(lib().xg_kozak_rhs(ctx._h, nc, float(t), float(dt), None, ptr(d), ptr(U),
                                  ptr(R)))
    return R


def _dev(ctx, *fields, nc=None):
    for a in fields:
        if a is None:
            continue
        if a.ndim != 2 or a.shape[0] != ctx.koz_npoin or not a.is_cuda:
            raise ValueError("device fields must be (%d, ncomp) on the GPU" % ctx.koz_npoin)
        if nc is not None and a.shape[1] != nc:
            raise ValueError("fields have %d components, the chare %d" % (a.shape[1], nc))


def rhs_dev(ctx, dt, U, R, t=0.0, dtp=None):
    """Device tensors U, R (npoin, ncomp) float64, dtp (npoin) optional;
    enqueued on the context stream."""
    _dev(ctx, U, R, nc=U.shape[1])
    ctx._check(lib().xg_kozak_rhs_dev(ctx._h, int(U.shape[1]), float(t), float(dt), None,
                                      ptr(dtp), ptr(U), ptr(R)))


def setup(ctx, vol, symnodes, symnorms, dirnodes=(), dirmask=None, dirvals=None,
          fct=True, fctdif=1.0, fctclip=False, fctsys=(), ncomp=5):
    """KozCG per-mesh data: nodal volumes, symmetry (node, normal) pairs,
    Dirichlet nodes with (ndir, ncomp) mask and values, FCT settings (fctsys
    as 0-based component ids); ncomp = 5 + transported scalars."""
    ctx._check(lib().xg_kozcg_set_ncomp(ctx._h, int(ncomp)))
    ctx.koz_ncomp = ncomp
    v = np.ascontiguousarray(vol, dtype=np.float64)
    sn = np.ascontiguousarray(symnodes, dtype=np.uint64)
    sv = np.ascontiguousarray(symnorms, dtype=np.float64).reshape(-1)
    dn = np.ascontiguousarray(dirnodes, dtype=np.uint64)
    nd = dn.size
    dm = np.ascontiguousarray(np.ones((nd, ncomp)) if dirmask is None else dirmask,
                              dtype=np.int32).reshape(-1)
    dv = np.ascontiguousarray(np.zeros((nd, ncomp)) if dirvals is None else dirvals,
                              dtype=np.float64).reshape(-1)
    if v.size != ctx.koz_npoin or sv.size != 3 * sn.size or dm.size != ncomp * nd \
            or dv.size != ncomp * nd:
        raise ValueError("setup: inconsistent array sizes")
    if len(fctsys) > 5:
        raise ValueError("fctsys holds at most 5 components")
    f = FctConfig(int(bool(fct)), float(fctdif), int(bool(fctclip)), len(fctsys))
    for i, c in enumerate(fctsys):
        f.sys[i] = int(c)
    ctx._check(lib().xg_kozcg_setup(ctx._h, ptr(v), sn.size, ptr(sn), ptr(sv), nd,
                                    ptr(dn), ptr(dm), ptr(dv), C.byref(f)))


def dt_dev(ctx, U, cfl):
    """KozCG::dt, unsteady: cfl * min cbrt(vol)/max(|u|+c, 1e-8) (the caller
    applies freezeflow)."""
    _dev(ctx, U, nc=ctx.koz_ncomp)
    out = C.c_double()
    ctx._check(lib().xg_kozcg_dt_dev(ctx._h, ptr(U), float(cfl), C.byref(out)))
    return out.value


def dt_steady_dev(ctx, U, cfl, dtp):
    """KozCG::dt, steady: fills dtp (device, npoin), returns its minimum."""
    _dev(ctx, U, nc=ctx.koz_ncomp)
    out = C.c_double()
    ctx._check(lib().xg_kozcg_dt_steady_dev(ctx._h, ptr(U), float(cfl), ptr(dtp),
                                            C.byref(out)))
    return out.value


def step_dev(ctx, dt, U, Un, t=0.0, dtp=None, tp=None, freeze=False):
    """One KozCG step on device tensors (npoin, ncomp): U advanced in place,
    Un receives the old solution; dtp, tp (device, npoin) for local time
    stepping (tp advanced), freeze keeps the flow components (freezeflow);
    enqueued on the context stream."""
    _dev(ctx, U, Un, nc=ctx.koz_ncomp)
    if (dtp is None) != (tp is None):
        raise ValueError("dtp and tp go together")
    ctx._check(lib().xg_kozcg_step2_dev(ctx._h, float(t), float(dt), ptr(dtp), ptr(tp),
                                        int(bool(freeze)), ptr(U), ptr(Un)))
