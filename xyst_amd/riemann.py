"""Python mirror of the reference's riemann:: and RieCG preprocessing entry
points over the C-ABI (include/xg.h).

`Context` holds one device-resident partition.  `grad`/`rhs` keep the
argument meaning of riemann::grad / riemann::rhs (Riemann.hpp:19-39): the
mesh arguments (dsupedge, dsupint, coord, triinpoel, besym) are uploaded once
with `upload_mesh`, the per-call arguments are the nodal fields.
"""
import ctypes as C

import numpy as np

from . import _xg
from ._xg import check, lib, ptr

__all__ = ["Context", "mesh_vol", "mesh_domint", "mesh_rcm", "mesh_symbc",
           "mesh_block_order",
           "edges_as_superedges"]

FLUX = {"rusanov": 0, "hllc": 1}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _idx(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def mesh_vol(coord, inpoel):
    """Discretization::vol: nodal volumes."""
    x, y, z = (_f64(coord[i]) for i in range(3))
    inp = _idx(inpoel)
    v = np.empty(x.size)
    check(lib().xg_mesh_vol(x.size, ptr(x), ptr(y), ptr(z), inp.size // 4,
                            ptr(inp), ptr(v)), what="Element Jacobian non-positive")
    return v


def mesh_domint(coord, inpoel):
    """RieCG::domint: unique edges (E,2) with p<q and their integrals (E,3)."""
    x, y, z = (_f64(coord[i]) for i in range(3))
    inp = _idx(inpoel)
    ne = C.c_int64(0)
    L = lib()
    check(L.xg_mesh_domint(x.size, ptr(x), ptr(y), ptr(z), inp.size // 4,
                           ptr(inp), 0, C.byref(ne), None, None), what="domint")
    edge = np.empty((ne.value, 2), np.uint64)
    dint = np.empty((ne.value, 3))
    check(L.xg_mesh_domint(x.size, ptr(x), ptr(y), ptr(z), inp.size // 4,
                           ptr(inp), ne.value, C.byref(ne), ptr(edge), ptr(dint)),
          what="domint")
    return edge, dint


def mesh_rcm(npoin, inpoel):
    """Reverse Cuthill-McKee renumbering; returns perm with perm[old] = new."""
    inp = _idx(inpoel)
    perm = np.empty(npoin, np.uint64)
    check(lib().xg_mesh_rcm(npoin, inp.size // 4, ptr(inp), ptr(perm)), what="rcm")
    return perm.astype(np.int64)


def mesh_block_order(coord, cell_nodes=2048):
    """Recursive-coordinate-bisection cell order (perm[old] = new); nodes
    keep their relative order inside a cell."""
    x, y, z = (_f64(coord[i]) for i in range(3))
    perm = np.empty(x.size, np.uint64)
    check(lib().xg_mesh_block_order(x.size, ptr(x), ptr(y), ptr(z), cell_nodes,
                                    ptr(perm)), what="block order")
    return perm.astype(np.int64)


def mesh_symbc(coord, triinpoel, faceset, symsets):
    """Symmetry BC node list, normals and besym flags."""
    x, y, z = (_f64(coord[i]) for i in range(3))
    tri = _idx(triinpoel)
    fs = np.ascontiguousarray(faceset, dtype=np.int32)
    ss = np.ascontiguousarray(symsets, dtype=np.int32)
    ntri = tri.size // 3
    besym = np.empty(3 * ntri, np.uint8)
    n = C.c_int64(0)
    L = lib()
    args = (x.size, ptr(x), ptr(y), ptr(z), ntri, ptr(tri), ptr(fs), ss.size, ptr(ss))
    check(L.xg_mesh_symbc(*args, 0, C.byref(n), None, None, ptr(besym)), what="symbc")
    nodes = np.empty(n.value, np.uint64)
    norms = np.empty((n.value, 3))
    check(L.xg_mesh_symbc(*args, n.value, C.byref(n), ptr(nodes), ptr(norms),
                          ptr(besym)), what="symbc")
    return nodes.astype(np.int64), norms, besym


def mesh_domsuped(inpoel, edge, dint):
    """RieCG::domsuped / ChoCG::domsuped: the edges of mesh_domint (or the
    5-value ChoCG integrals) grouped into the reference's tet, triangle and
    edge superedges, oriented as the reference orients them; returns the
    (dsupedge, dsupint) triple of lists."""
    inp = _idx(inpoel)
    e = _idx(edge).reshape(-1)
    d = _f64(dint)
    st = d.shape[1] if d.ndim == 2 else 3
    d = d.reshape(-1)
    ns = np.zeros(3, np.int64)
    L = lib()
    args = (inp.size // 4, ptr(inp), e.size // 2, ptr(e), ptr(d), st, ptr(ns))
    check(L.xg_mesh_domsuped(*args, None, None, None, None, None, None), what="domsuped")
    sup = [np.empty(n * k, np.uint64) for n, k in zip(ns, (4, 3, 2))]
    val = [np.empty(n * k * st) for n, k in zip(ns, (6, 3, 1))]
    check(L.xg_mesh_domsuped(*args, ptr(sup[0]), ptr(val[0]), ptr(sup[1]), ptr(val[1]),
                             ptr(sup[2]), ptr(val[2])), what="domsuped")
    return sup, val


def edges_as_superedges(edge, dint):
    """A plain edge list as the reference's (dsupedge, dsupint) triple: every
    edge in the third (single-edge) group."""
    z = np.zeros(0, np.uint64)
    return ([z, z, _idx(edge).reshape(-1)],
            [np.zeros(0), np.zeros(0), _f64(dint).reshape(-1)])


class Context:
    """One GPU-resident partition: device mesh layout + physics config."""

    def __init__(self, device=0, gamma=1.4, flux="rusanov", stab2=False,
                 stab2coef=0.0, stream=None):
        L = lib()
        h = C.c_void_p()
        check(L.xg_ctx_create(device, C.byref(h)), what="xg_ctx_create (no GPU?)")
        self._h = h
        self.npoin = 0
        self.ncomp = 5
        self.set_config(gamma, flux, stab2, stab2coef)
        if stream is not None:
            self.set_stream(stream)

    def close(self):
        if self._h:
            lib().xg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        check(rc, self._h)

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream or a raw hipStream_t integer."""
        handle = getattr(stream, "cuda_stream", stream)
        self._check(lib().xg_set_stream(self._h, C.c_void_p(handle)))

    def set_config(self, gamma=1.4, flux="rusanov", stab2=False, stab2coef=0.0):
        cfg = _xg.Config(gamma, FLUX.get(flux, 99) if isinstance(flux, str) else flux,
                         int(bool(stab2)), stab2coef)
        self._check(lib().xg_set_config(self._h, C.byref(cfg)))

    def upload_mesh(self, coord, dsupedge, dsupint, triinpoel, besym):
        x, y, z = (_f64(coord[i]) for i in range(3))
        s = [_idx(a).reshape(-1) for a in dsupedge]
        d = [_f64(a).reshape(-1) for a in dsupint]
        nsup = np.array([s[0].size // 4, s[1].size // 3, s[2].size // 2], np.int64)
        tri = _idx(triinpoel).reshape(-1)
        bs = np.ascontiguousarray(besym, dtype=np.uint8).reshape(-1)
        self._keep = (x, y, z, s, d, nsup, tri, bs)
        self._check(lib().xg_mesh_upload(
            self._h, x.size, ptr(x), ptr(y), ptr(z), ptr(nsup),
            ptr(s[0]) if s[0].size else None, ptr(d[0]) if d[0].size else None,
            ptr(s[1]) if s[1].size else None, ptr(d[1]) if d[1].size else None,
            ptr(s[2]) if s[2].size else None, ptr(d[2]) if d[2].size else None,
            tri.size // 3, ptr(tri) if tri.size else None,
            ptr(bs) if bs.size else None))
        self._keep = None
        self.npoin = x.size

    def device_bytes(self):
        b = C.c_int64(0)
        self._check(lib().xg_mesh_device_bytes(self._h, C.byref(b)))
        return b.value

    def _fields(self, U, G=None, R=None, dev=False):
        """Shapes must match the uploaded mesh: the C-ABI cannot see them."""
        if U.ndim != 2 or U.shape[0] != self.npoin:
            raise ValueError("U must be (npoin=%d, ncomp)" % self.npoin)
        nc = U.shape[1]
        for a, w in ((G, 3 * nc), (R, nc)):
            if a is not None and tuple(a.shape) != (self.npoin, w):
                raise ValueError("field shape %s, expected (%d, %d)"
                                 % (tuple(a.shape), self.npoin, w))
        if dev:
            import torch
            for a in (U, G, R):
                if a is not None and (not a.is_cuda or a.dtype != torch.float64
                                      or not a.is_contiguous()):
                    raise ValueError("device fields must be contiguous float64 "
                                     "CUDA tensors")

    # riemann::grad / riemann::rhs with host (numpy) fields
    def grad(self, U):
        U = _f64(U)
        if self.npoin:
            self._fields(U)
        G = np.empty((U.shape[0], 3 * U.shape[1]))
        self._check(lib().xg_riemann_grad(self._h, U.shape[1], ptr(U), ptr(G)))
        return G

    def set_problem(self, name="none", alpha=0.1, kappa=1.0, p0=10.0):
        """The problem whose source term riemann::rhs, kozak::rhs and the ChoCG
        step add (problems::SRC())."""
        p = _xg.Problem(_xg.PROBLEMS[name], alpha, kappa, p0)
        self._check(lib().xg_set_problem(self._h, C.byref(p)))

    def rhs(self, G, U, v=None, t=0.0, tp=None):
        """riemann::rhs(..., G, U, v, t, tp, R): v the nodal volumes of this
        partition's elements (needed when the problem has a source)."""
        U, G = _f64(U), _f64(G)
        if self.npoin:
            self._fields(U, G)
        va = _f64(v) if v is not None else None
        ta = _f64(tp) if tp is not None else None
        R = np.empty_like(U)
        self._check(lib().xg_riemann_rhs(self._h, U.shape[1], ptr(G), ptr(U), ptr(va),
                                         t, ptr(ta), ptr(R)))
        return R

    # HBM-resident variants: torch tensors on the GPU
    def grad_dev(self, U, G):
        self._fields(U, G, dev=True)
        self._check(lib().xg_riemann_grad_dev(self._h, U.shape[1], ptr(U), ptr(G)))

    def rhs_dev(self, G, U, R, v=None, t=0.0, tp=None):
        self._fields(U, G, R, dev=True)
        self._check(lib().xg_riemann_rhs_dev(self._h, U.shape[1], ptr(G), ptr(U), ptr(v),
                                             t, ptr(tp), ptr(R)))

    # resident RieCG stepping
    def riecg_setup(self, vol, symnodes, symnorms, ncomp=5):
        """The chare's nodal volumes and symmetry lists; ncomp = 5 + the
        problem's transported scalars (the fields' components per node)."""
        v = _f64(vol)
        if v.shape != (self.npoin,):
            raise ValueError("vol must have npoin entries")
        sn = _idx(symnodes)
        sv = _f64(symnorms).reshape(-1)
        self._check(lib().xg_riecg_setup(self._h, ptr(v), sn.size,
                                         ptr(sn) if sn.size else None,
                                         ptr(sv) if sv.size else None))
        self._check(lib().xg_riecg_set_ncomp(self._h, ncomp))
        self.ncomp = ncomp

    def _stepped(self, *fields):
        """The stepping entries read the chare's ncomp components per node."""
        for a in fields:
            if a is not None and a.shape[1] != self.ncomp:
                raise ValueError("fields have %d components, the chare %d"
                                 % (a.shape[1], self.ncomp))

    def dt_dev(self, U, cfl):
        self._fields(U, dev=True)
        self._stepped(U)
        out = C.c_double(0.0)
        self._check(lib().xg_riecg_dt_dev(self._h, ptr(U), cfl, C.byref(out)))
        return out.value

    def riecg_bc(self, dirnodes=(), dirmask=(), dirvals=(), farnodes=(), farnorms=(),
                 far=(1.0, 0.0, 0.0, 0.0, 1.0), prenodes=(), prevals=()):
        """RieCG's Dirichlet (mask 1: the given initial-condition values),
        farfield and pressure BC lists (RieCG::setupBC)."""
        dn = _idx(dirnodes).reshape(-1)
        dm = np.ascontiguousarray(dirmask, dtype=np.int32).reshape(-1)
        dv = _f64(dirvals).reshape(-1)
        fn = _idx(farnodes).reshape(-1)
        fv = _f64(farnorms).reshape(-1)
        fa = _f64(far).reshape(-1)
        pn = _idx(prenodes).reshape(-1)
        pv = _f64(prevals).reshape(-1)
        nc = self.ncomp
        if dm.size != nc * dn.size or dv.size != nc * dn.size or fv.size != 3 * fn.size \
                or pv.size != 2 * pn.size or fa.size != 5:
            raise ValueError("riecg_bc: list sizes disagree")
        self._check(lib().xg_riecg_bc(self._h, dn.size, ptr(dn) if dn.size else None,
                                      ptr(dm) if dm.size else None,
                                      ptr(dv) if dv.size else None, fn.size,
                                      ptr(fn) if fn.size else None,
                                      ptr(fv) if fv.size else None, ptr(fa), pn.size,
                                      ptr(pn) if pn.size else None,
                                      ptr(pv) if pv.size else None))

    def apply_bc_dev(self, U):
        """RieCG::BC on a device state (dirbc, symbc, farbc, prebc)."""
        self._fields(U, dev=True)
        self._stepped(U)
        self._check(lib().xg_riecg_apply_bc_dev(self._h, ptr(U)))

    def dt_steady_dev(self, U, cfl, dtp):
        """Steady RieCG::dt: fills dtp (device, npoin), returns its minimum."""
        self._fields(U, dev=True)
        self._stepped(U)
        out = C.c_double(0.0)
        self._check(lib().xg_riecg_dt_steady_dev(self._h, ptr(U), cfl, ptr(dtp),
                                                 C.byref(out)))
        return out.value

    def stage_dev(self, stage, dt, U, Un, G, R, t=0.0, dtp=None, tp=None):
        self._fields(U, G, R, dev=True)
        self._fields(Un, dev=True)
        self._stepped(U, Un)
        self._check(lib().xg_riecg_stage_dev(self._h, stage, t, dt, ptr(dtp), ptr(tp),
                                             ptr(U), ptr(Un), ptr(G), ptr(R)))

    def grad_div_dev(self, G):
        """G /= the summed nodal volume (after the partitions' exchange)."""
        if tuple(G.shape) != (self.npoin, 3 * self.ncomp):
            raise ValueError("G must be (npoin, 3 ncomp)")
        self._check(lib().xg_riecg_grad_div_dev(self._h, ptr(G)))

    def solve_dev(self, stage, dt, U, Un, R, t=0.0, dtp=None, tp=None):
        """RieCG::solve with a complete rhs R: update and BCs."""
        self._fields(U, R=R, dev=True)
        self._stepped(U, Un)
        self._check(lib().xg_riecg_solve_dev(self._h, stage, t, dt, ptr(dtp), ptr(tp),
                                             ptr(U), ptr(Un), ptr(R)))
