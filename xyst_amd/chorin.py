"""Python mirror of the chorin:: edge operators (Chorin.hpp:19-70) and of
the ChoCG time step (ChoCG.cpp:817-1695) over the C-ABI.

The mesh arguments (ChoCG superedges with 5 values per edge integral,
coordinates, boundary triangles) are uploaded once per context with
`upload`.  `div`, `grad`, `vgrad` and `flux` add to their output as the
reference's do; `rhs` returns a new R.  Numpy arrays run through the host
entry points; torch device tensors through the device ones (enqueued on the
context stream).
"""
import ctypes as C

import numpy as np

from ._xg import ChoConfig, ChoStep, lib, ptr

__all__ = ["mesh_domint", "upload", "div", "grad", "vgrad", "flux", "rhs", "config", "step_setup",
           "set_ncomp", "initial_dev", "poisson_dev", "neumann", "dt_dev", "step_dev",
           "momentum_setup", "momentum_its"]

FLUX = {"damp2": 0, "damp4": 1}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _idx(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _dev(*arrs):
    kinds = {not isinstance(a, np.ndarray) for a in arrs if a is not None}
    if len(kinds) > 1:
        raise ValueError("mix of host and device arrays")
    return int(kinds.pop()) if kinds else 0


def config(flux="damp2", stab=True, stab2=False, stab2coef=0.0, mu=0.0, diff=0.0):
    """chorin::rhs settings: flux, stab, stab2 (coefficient), viscosity mu
    and the transported scalar's diffusivity diff."""
    if flux not in FLUX:
        raise ValueError("Flux not correctly configured")
    return ChoConfig(FLUX[flux], int(bool(stab)), int(bool(stab2)), stab2coef, mu, diff)


def mesh_domint(coord, inpoel):
    """ChoCG::domint: unique edges (E,2) with p<q and 5 integrals per edge
    (dual-face normal, J/120, P1 stiffness)."""
    x, y, z = (_f64(coord[i]) for i in range(3))
    inp = _idx(inpoel)
    ne = C.c_int64(0)
    L = lib()
    rc = L.xg_mesh_chocg_domint(x.size, ptr(x), ptr(y), ptr(z), inp.size // 4, ptr(inp), 0,
                                C.byref(ne), None, None)
    if rc:
        raise ValueError("xg_mesh_chocg_domint failed (%d)" % rc)
    edge = np.empty((ne.value, 2), np.uint64)
    dint = np.empty((ne.value, 5))
    rc = L.xg_mesh_chocg_domint(x.size, ptr(x), ptr(y), ptr(z), inp.size // 4, ptr(inp),
                                ne.value, C.byref(ne), ptr(edge), ptr(dint))
    if rc:
        raise ValueError("xg_mesh_chocg_domint failed (%d)" % rc)
    return edge, dint


def upload(ctx, coord, dsupedge, dsupint, triinpoel):
    x, y, z = (_f64(coord[i]) for i in range(3))
    s = [_idx(a) for a in dsupedge]
    d = [_f64(a) for a in dsupint]
    nsup = np.array([s[0].size // 4, s[1].size // 3, s[2].size // 2], np.int64)
    tri = _idx(triinpoel)
    ctx._check(lib().xg_chocg_upload(ctx._h, x.size, ptr(x), ptr(y), ptr(z), ptr(nsup),
                                     ptr(s[0]), ptr(d[0]), ptr(s[1]), ptr(d[1]),
                                     ptr(s[2]), ptr(d[2]), tri.size // 3, ptr(tri)))
    ctx.cho_npoin = x.size
    ctx.cho_ncomp = 3


def _cols(a):
    return int(a.shape[1]) if a.ndim == 2 else 1


def div(ctx, dt, P, G, U, D, stab=False):
    """chorin::div: D += div U (+ pressure damping when stab)."""
    dev = _dev(P if stab else None, G if stab else None, U, D)
    ctx._check(lib().xg_chorin_div(ctx._h, dev, float(dt), ptr(P) if stab else None,
                                   ptr(G) if stab else None, _cols(G) if stab else 3,
                                   ptr(U), _cols(U), ptr(D), int(bool(stab))))
    return D


def grad(ctx, U, G):
    """chorin::grad: G(:, 0..2) += gradient of the scalar U."""
    ctx._check(lib().xg_chorin_grad(ctx._h, _dev(U, G), ptr(U), ptr(G), _cols(G)))
    return G


def vgrad(ctx, U, G):
    """chorin::vgrad: G(npoin, 3 ncomp) += gradients of U(npoin, ncomp),
    ncomp 3 or 4."""
    ctx._check(lib().xg_chorin_vgrad(ctx._h, _dev(U, G), _cols(U), ptr(U), ptr(G)))
    return G


def flux(ctx, mu, U, G, F):
    """chorin::flux: F(npoin, 3) += momentum flux divergence."""
    ctx._check(lib().xg_chorin_flux(ctx._h, _dev(U, G, F), float(mu), ptr(U), _cols(U),
                                    ptr(G), _cols(G), ptr(F)))
    return F


def rhs(ctx, cfg, P, U, G, R=None):
    """chorin::rhs: R = advection rhs of U (npoin, ncomp: velocity and a
    transported scalar, ncomp 3 or 4; R allocated for host inputs)."""
    dev = _dev(P, U, G, R)
    if R is None:
        if dev:
            raise ValueError("pass R for device inputs")
        R = np.empty_like(_f64(U))
    ctx._check(lib().xg_chorin_rhs(ctx._h, dev, C.byref(cfg), _cols(U), ptr(P), ptr(U),
                                   ptr(G), ptr(R)))
    return R


def set_ncomp(ctx, ncomp):
    """Components of U for the ChoCG stepping: 3, or 4 with a transported
    scalar; call after `upload`, before `step_setup`."""
    ctx._check(lib().xg_chocg_set_ncomp(ctx._h, int(ncomp)))
    ctx.cho_ncomp = int(ncomp)


def step_setup(ctx, vol, cfg, cfl=0.9, rk=1, p_pc="jacobi", p_maxit=500, p_tol=1e-3,
               symnodes=(), symnorms=(), noslip=(), dirnodes=(), dirmask=None,
               dirvals=None, pdirnodes=(), pdirvals=None):
    """ChoCG per-mesh data (after `upload`, `set_ncomp` and the pressure
    matrix given to the context's ConjugateGradients): cfg from `config`;
    Dirichlet nodes with (ndir, ncomp) masks and values; pressure Dirichlet
    nodes (bc_dir of the pressure block and the hydrostat node) with their
    initial values."""
    st = ChoStep(cfg, float(cfl), int(rk), {"none": 0, "jacobi": 1, "amg": 2}[p_pc], int(p_maxit),
                 float(p_tol))
    nc = getattr(ctx, "cho_ncomp", 3)
    v = _f64(vol)
    sn, sv = _idx(symnodes), _f64(symnorms).reshape(-1)
    ns = _idx(noslip)
    dn = _idx(dirnodes)
    nd = dn.size
    dm = np.ascontiguousarray(np.ones((nd, nc)) if dirmask is None else dirmask,
                              dtype=np.int32).reshape(-1)
    dv = _f64(np.zeros((nd, nc)) if dirvals is None else dirvals).reshape(-1)
    pn = _idx(pdirnodes)
    pv = _f64(np.zeros(pn.size) if pdirvals is None else pdirvals).reshape(-1)
    if v.size != ctx.cho_npoin or sv.size != 3 * sn.size or dm.size != nc * nd or \
            dv.size != nc * nd or pv.size != pn.size:
        raise ValueError("step_setup: inconsistent array sizes")
    ctx._check(lib().xg_chocg_setup(ctx._h, ptr(v), C.byref(st), sn.size, ptr(sn), ptr(sv),
                                    ns.size, ptr(ns), nd, ptr(dn), ptr(dm), ptr(dv),
                                    pn.size, ptr(pn), ptr(pv)))


def _fields(ctx, *arrs):
    n, nc = ctx.cho_npoin, getattr(ctx, "cho_ncomp", 3)
    for a in arrs:
        if tuple(a.shape) != (n, nc) or not a.is_cuda:
            raise ValueError("device fields must be (%d, %d) on the GPU" % (n, nc))


def initial_dev(ctx, U, P):
    """ChoCG start-up on device tensors U (npoin, ncomp), P (npoin,); returns
    the two pressure solves' iteration counts."""
    _fields(ctx, U)
    it = (C.c_int64 * 2)()
    ctx._check(lib().xg_chocg_initial_dev(ctx._h, ptr(U), ptr(P), it))
    return it[0], it[1]


def neumann(coord, tri, besym, pg):
    """The Neumann vector of ChoCG::pinit (ChoCG.cpp:1070-1095): boundary
    faces tri (ntri, 3, outward) flagged in besym (ntri,), pressure gradient
    pg (npoin, 3) at the nodes; host arrays, returns (npoin,)."""
    x, y, z = (_f64(coord[i]) for i in range(3))
    t = _idx(tri)
    b = np.ascontiguousarray(besym, dtype=np.uint8).reshape(-1)
    g = _f64(pg).reshape(-1)
    if t.size != 3 * b.size or g.size != 3 * x.size:
        raise ValueError("neumann: inconsistent array sizes")
    out = np.empty(x.size)
    rc = lib().xg_mesh_neumann(x.size, ptr(x), ptr(y), ptr(z), b.size, ptr(t), ptr(b),
                               ptr(g), ptr(out))
    if rc:
        raise ValueError("xg_mesh_neumann failed (%d)" % rc)
    return out


def poisson_dev(ctx, U, P, pr=None, neubc=None):
    """ChoCG's first pressure solve alone (nstep = 1, the Poisson problems):
    pr (npoin,) the problem's pressure rhs at the nodes, neubc (npoin,) the
    Neumann vector, device tensors or None; U (npoin, ncomp) is projected and
    P (npoin,) receives the solution.  Returns the CG iterations."""
    n = ctx.cho_npoin
    _fields(ctx, U)
    if tuple(P.shape) != (n,):
        raise ValueError("P must be (%d,)" % n)
    for a in (pr, neubc):
        if a is not None and (tuple(a.shape) != (n,) or not a.is_cuda):
            raise ValueError("pr and neubc must be (%d,) device tensors" % n)
    it = C.c_int64()
    ctx._check(lib().xg_chocg_poisson_dev(ctx._h, ptr(pr) if pr is not None else None,
                                          ptr(neubc) if neubc is not None else None,
                                          ptr(U), ptr(P), C.byref(it)))
    return it.value


def dt_dev(ctx, U):
    _fields(ctx, U)
    out = C.c_double()
    ctx._check(lib().xg_chocg_dt_dev(ctx._h, ptr(U), C.byref(out)))
    return out.value


def step_dev(ctx, dt, U, Un, P, dP, freeze=False):
    """One ChoCG step (freeze: the flow frozen, freezeflow > 1); returns the
    pressure solve's iteration count."""
    _fields(ctx, U, Un)
    it = C.c_int64()
    ctx._check(lib().xg_chocg_step2_dev(ctx._h, float(dt), int(bool(freeze)), ptr(U), ptr(Un),
                                        ptr(P), ptr(dP), C.byref(it)))
    return it.value


def momentum_setup(ctx, inpoel, theta, m_pc="jacobi", m_maxit=50, m_tol=1e-3):
    """The theta-implicit momentum system (ChoCG::momlhs/lhs); call after
    step_setup.  inpoel: the tets of the uploaded mesh."""
    inp = _idx(inpoel)
    ctx._check(lib().xg_chocg_momentum_setup(ctx._h, inp.size // 4, ptr(inp), float(theta),
                                             {"none": 0, "jacobi": 1}[m_pc], int(m_maxit),
                                             float(m_tol)))


def momentum_its(ctx):
    out = C.c_int64()
    ctx._check(lib().xg_chocg_momentum_its(ctx._h, C.byref(out)))
    return out.value
