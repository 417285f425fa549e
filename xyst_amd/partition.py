"""Several partitions of one mesh on one device, the way the reference runs
one RieCG chare per partition (RieCG.cpp, Discretization's NodeCommMap).

Each partition is an `xyst_amd.riemann.Context` holding the partition's own
tets: its edge integrals are RieCG::domint over those tets only, so an edge
on a partition interface carries a partial dual-face normal in every
partition that holds it, exactly as in the chares.  The partial nodal sums of
riemann::grad and riemann::rhs are added at shared nodes by `Comm` (the
C-ABI's xg_comm_sum_dev: the reference's comgrad/comrhs messages,
RieCG.cpp:871-916, 948-990) before the division by the summed nodal volume
and the update.  On several GPUs the same shared-node lists become the
per-neighbour send/receive sets.
"""
import ctypes as C

import numpy as np

from . import riemann as X
from ._xg import XgError, lib, ptr

__all__ = ["Comm", "Part", "split_mesh", "PartitionedRieCG", "PartitionedCG"]


class Comm:
    """The shared nodes of a set of partitions (xg_comm_*)."""

    def __init__(self, gids, device=0):
        L = lib()
        self._gid = [np.ascontiguousarray(g, dtype=np.uint64) for g in gids]
        n = np.array([g.size for g in self._gid], np.int64)
        arr = (C.c_void_p * len(self._gid))(*[g.ctypes.data for g in self._gid])
        h = C.c_void_p()
        rc = L.xg_comm_create(device, len(self._gid), n.ctypes.data, arr, C.byref(h))
        self._h = h
        if rc:
            msg = L.xg_comm_error(h).decode() if h else ""
            self.close()
            raise XgError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            lib().xg_comm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shared(self):
        n = C.c_int64(0)
        lib().xg_comm_shared(self._h, C.byref(n))
        return n.value

    def sum_dev(self, fields, stream=None):
        """Adds the partials of every shared node over the partitions (in
        partition order) and gives every copy the sum; fields[k] is partition
        k's contiguous (nlocal[k], ncomp) float64 device tensor."""
        nc = int(fields[0].shape[1]) if fields[0].dim() == 2 else 1
        arr = (C.c_void_p * len(fields))(*[ptr(f) for f in fields])
        handle = getattr(stream, "cuda_stream", stream)
        rc = lib().xg_comm_sum_dev(self._h, nc, arr, C.c_void_p(handle) if handle else None)
        if rc:
            raise XgError(rc, lib().xg_comm_error(self._h).decode())


class Part:
    """One partition: its tets, its nodes (global ids), the local numbering
    and the boundary faces that belong to its tets."""

    def __init__(self, coord, inpoel, tri, tets):
        ip = inpoel[np.asarray(tets, np.int64)]
        self.gid = np.unique(ip.reshape(-1))
        lid = np.full(coord.shape[1], -1, np.int64)
        lid[self.gid] = np.arange(self.gid.size)
        self.inpoel = lid[ip]
        self.coord = np.ascontiguousarray(coord[:, self.gid])
        keys = set()
        for t in ip:                      # the four faces of each own tet
            for fc in ((t[0], t[1], t[2]), (t[0], t[1], t[3]), (t[0], t[2], t[3]),
                       (t[1], t[2], t[3])):
                keys.add(tuple(sorted(fc)))
        own = np.array([tuple(sorted(f)) in keys for f in tri], bool)
        self.face_index = np.flatnonzero(own)          # into the global face list
        self.tri = lid[tri[own]].reshape(-1, 3)
        self.lid = lid


def split_mesh(coord, inpoel, tri, owner):
    """Partitions from the partition id of every tet (owner[tet])."""
    owner = np.asarray(owner)
    return [Part(coord, inpoel, tri, np.flatnonzero(owner == k))
            for k in range(int(owner.max()) + 1)]


class PartitionedRieCG:
    """RieCG on k partitions of one mesh, all on one device.

    State lives per partition (device tensors, local numbering); `scatter`
    and `gather` move a global host array in and out."""

    def __init__(self, coord, inpoel, tri, owner, device=0, gamma=1.4, flux="rusanov",
                 stab2=False, stab2coef=0.2, problem="none", besym=None, ncomp=5,
                 **problem_args):
        import torch
        self.ncomp = ncomp
        self.torch = torch
        self.dev = torch.device("cuda", device)
        coord = np.ascontiguousarray(coord, dtype=np.float64)
        inpoel = np.asarray(inpoel, np.int64)
        tri = np.asarray(tri, np.int64).reshape(-1, 3)
        self.npoin = coord.shape[1]
        self.parts = split_mesh(coord, inpoel, tri, owner)
        besym = (np.zeros(tri.size, np.uint8) if besym is None
                 else np.asarray(besym, np.uint8).reshape(-1))
        self.ctx, self.v = [], []
        Vol = np.zeros(self.npoin)
        for p in self.parts:
            ctx = X.Context(device, gamma=gamma, flux=flux, stab2=stab2,
                            stab2coef=stab2coef)
            ctx.set_problem(problem, **problem_args)
            edge, dint = X.mesh_domint(p.coord, p.inpoel)
            # the chare's own superedges (RieCG::domsuped), oriented as the
            # reference orients them
            sup, ints = X.mesh_domsuped(p.inpoel, edge, dint)
            bs = besym.reshape(-1, 3)[p.face_index].reshape(-1)
            ctx.upload_mesh(p.coord, sup, ints, p.tri, bs)
            p.vol = X.mesh_vol(p.coord, p.inpoel)          # own elements only
            np.add.at(Vol, p.gid, p.vol)
            self.ctx.append(ctx)
            self.v.append(torch.from_numpy(p.vol).to(self.dev))
        self.Vol = Vol
        self.comm = Comm([p.gid for p in self.parts], device)
        for p, ctx in zip(self.parts, self.ctx):
            ctx.riecg_setup(Vol[p.gid], np.zeros(0, np.int64), np.zeros((0, 3)), ncomp)

    def close(self):
        for c in self.ctx:
            c.close()
        self.comm.close()

    def _local(self, p, nodes, *recs):
        nodes = np.asarray(nodes, np.int64).reshape(-1)
        loc = p.lid[nodes] if nodes.size else nodes
        keep = loc >= 0
        out = [loc[keep]]
        for r in recs:
            r = np.asarray(r)
            out.append(r.reshape(nodes.size, -1)[keep] if nodes.size else r)
        return out

    def setup_bc(self, symnodes=(), symnorms=(), dirnodes=(), dirmask=(), dirvals=(),
                 farnodes=(), farnorms=(), far=(1.0, 0.0, 0.0, 0.0, 1.0), prenodes=(),
                 prevals=()):
        """Global BC lists (RieCG::setupBC), given to every partition that
        holds the node."""
        for p, ctx in zip(self.parts, self.ctx):
            sn, sv = self._local(p, symnodes, np.reshape(symnorms, (-1, 3)))
            ctx.riecg_setup(self.Vol[p.gid], sn, sv, self.ncomp)
            nc = self.ncomp
            dn, dm, dv = self._local(p, dirnodes, np.reshape(dirmask, (-1, nc)),
                                     np.reshape(dirvals, (-1, nc)))
            fn, fv = self._local(p, farnodes, np.reshape(farnorms, (-1, 3)))
            pn, pv = self._local(p, prenodes, np.reshape(prevals, (-1, 2)))
            ctx.riecg_bc(dn, dm, dv, fn, fv, far, pn, pv)

    def scatter(self, U):
        return [self.torch.from_numpy(np.ascontiguousarray(U[p.gid])).to(self.dev)
                for p in self.parts]

    def gather(self, Us):
        U = np.zeros((self.npoin, Us[0].shape[1]))
        for p, u in zip(self.parts, Us):
            U[p.gid] = u.cpu().numpy()
        return U

    def apply_bc(self, Us):
        for ctx, u in zip(self.ctx, Us):
            ctx.apply_bc_dev(u)

    def dt(self, Us, cfl):
        """RieCG::dt with the min_double reduction over partitions."""
        return min(ctx.dt_dev(u, cfl) for ctx, u in zip(self.ctx, Us))

    def step(self, t, dt, Us, Uns):
        """One RK3 step: per stage the partitions' partial gradients, their
        exchange, G/Vol, the partial rhs with the source on each partition's
        own volumes, the rhs exchange, the update and the BCs."""
        torch = self.torch
        G = [torch.empty((u.shape[0], 3 * self.ncomp), dtype=torch.float64,
                         device=self.dev) for u in Us]
        R = [torch.empty_like(u) for u in Us]
        for stage in range(3):
            for ctx, u, g in zip(self.ctx, Us, G):
                ctx.grad_dev(u, g)
            self.comm.sum_dev(G)
            for ctx, g in zip(self.ctx, G):
                ctx.grad_div_dev(g)
            for ctx, u, g, r, v in zip(self.ctx, Us, G, R, self.v):
                ctx.rhs_dev(g, u, r, v=v, t=t)
            self.comm.sum_dev(R)
            for ctx, u, un, r in zip(self.ctx, Us, Uns, R):
                ctx.solve_dev(stage, dt, u, un, r, t=t)

    def rhocompute(self, Us, Uns):
        """NodeDiagnostics::rhocompute summed over the partitions, each with
        its own nodal volumes."""
        from .diagnostics import rhocompute
        return sum(rhocompute(ctx, u, un, v)
                   for ctx, u, un, v in zip(self.ctx, Us, Uns, self.v))

    def errcompute(self, Us, Ans):
        from .diagnostics import errcompute
        return sum(errcompute(ctx, u, v, a)
                   for ctx, u, v, a in zip(self.ctx, Us, self.v, Ans))


class PartitionedCG:
    """The chares of one ConjugateGradients array on one device (xg_pcg_*,
    ConjugateGradients.cpp:105-823): partition k's partial matrix As[k] (an
    object with ia, ja, a, 0-based, as xyst_amd.cg.CSR or the oracle's CSR)
    over its nodes with global ids gids[k]."""

    def __init__(self, As, gids, ncomp=1, device=0):
        L = lib()
        self.ncomp = ncomp
        self._keep = []
        npoin = np.array([len(g) for g in gids], np.int64)
        self.rows = [int(n) * ncomp for n in npoin]

        def arr(lst, dtype):
            a = [np.ascontiguousarray(v, dtype=dtype) for v in lst]
            self._keep.extend(a)
            return (C.c_void_p * len(a))(*[v.ctypes.data for v in a])
        g = arr(gids, np.uint64)
        ia = arr([A.ia for A in As], np.uint64)
        ja = arr([A.ja for A in As], np.uint64)
        av = arr([A.a for A in As], np.float64)
        self._npoin = npoin
        h = C.c_void_p()
        rc = L.xg_pcg_create(device, len(As), ncomp, npoin.ctypes.data, g, ia, ja, av, 0,
                             C.byref(h))
        self._h = h
        if rc:
            msg = L.xg_pcg_error(h).decode() if h else ""
            self.close()
            raise XgError(rc, msg)

    def _check(self, rc):
        if rc:
            raise XgError(rc, lib().xg_pcg_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().xg_pcg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self, xs, bs, pc="none"):
        """ConjugateGradients::setup on every partition; returns normb."""
        bb = [np.ascontiguousarray(b, dtype=np.float64) for b in bs]
        xx = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        for k, b in enumerate(bb):
            if b.size != self.rows[k]:
                raise ValueError("b of partition %d: %d values, need %d" % (k, b.size, self.rows[k]))
        barr = (C.c_void_p * len(bb))(*[b.ctypes.data for b in bb])
        xarr = (C.c_void_p * len(xx))(*[None if x is None else x.ctypes.data for x in xx])
        nb = C.c_double()
        self._check(lib().xg_pcg_setup(self._h, xarr, barr, {"none": 0, "jacobi": 1}[pc],
                                       C.byref(nb)))
        self.normb = nb.value
        return self.normb

    def solve(self, maxit, tol):
        """ConjugateGradients::solve; returns ||r|| of the recurrence."""
        nr, it, cv = C.c_double(), C.c_int64(), C.c_int()
        self._check(lib().xg_pcg_solve(self._h, maxit, tol, C.byref(nr), C.byref(it),
                                       C.byref(cv)))
        self.normr, self.it, self.converged = nr.value, it.value, bool(cv.value)
        return self.normr

    def solution(self):
        out = []
        for k, n in enumerate(self.rows):
            x = np.empty(n)
            self._check(lib().xg_pcg_solution(self._h, k, x.ctypes.data))
            out.append(x)
        return out
