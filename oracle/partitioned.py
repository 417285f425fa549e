"""Partitioned RieCG on the CPU -- TEST INFRASTRUCTURE ONLY.

The reference runs one RieCG chare per mesh partition: each chare integrates
its own tets (RieCG::domint over its elements, so an edge on a partition
interface carries a partial dual-face normal in every chare that holds it),
computes partial gradients and right-hand sides, and adds the neighbours'
contributions at shared nodes (comgrad / comrhs, RieCG.cpp:871-916,
948-990) before dividing by the summed nodal volume and advancing.  This
module restates that data flow with the oracle's per-mesh functions, on a
global copy of the state (the chares' copies of a shared node are equal
after every exchange).
"""
import numpy as np

from . import diag, domint, domsuped, grad, orc_apply_bcs, rhs, src, vol


class Part:
    """One partition: its tets, nodes (global ids), local numbering, the
    boundary faces of its tets, and the per-chare preprocessing."""

    def __init__(self, coord, inpoel, tri, tets):
        self.tets = np.asarray(tets, np.int64)
        ip = inpoel[self.tets]
        self.gid = np.unique(ip.reshape(-1))
        lid = {int(g): i for i, g in enumerate(self.gid)}
        self.inpoel = np.array([[lid[int(g)] for g in t] for t in ip], np.int64)
        self.coord = np.ascontiguousarray(coord[:, self.gid])
        # boundary faces that belong to one of this partition's tets
        faces = {}
        for t in ip:
            for fc in ((t[0], t[2], t[1]), (t[0], t[1], t[3]), (t[0], t[3], t[2]),
                       (t[1], t[2], t[3])):
                faces[tuple(sorted(int(v) for v in fc))] = True
        own = [f for f in tri if tuple(sorted(int(v) for v in f)) in faces]
        self.tri_glob = np.array(own, np.int64).reshape(-1, 3)
        self.tri = np.array([[lid[int(g)] for g in f] for f in self.tri_glob],
                            np.int64).reshape(-1, 3)
        self.vol = vol(self.coord, self.inpoel)          # V: own contributions
        self.edge, self.dint = domint(self.coord, self.inpoel)
        self.se = domsuped(self.inpoel, self.edge, self.dint)
        self.besym = np.zeros(self.tri.size, np.uint8)


def split(coord, inpoel, tri, owner):
    """Partitions from the partition id of every tet."""
    return [Part(coord, inpoel, tri, np.flatnonzero(owner == k))
            for k in range(int(owner.max()) + 1)]


def total_volume(parts, npoin):
    Vol = np.zeros(npoin)
    for p in parts:
        np.add.at(Vol, p.gid, p.vol)
    return Vol


def step(cfg, P, parts, Vol, bc, t, dt, U, coord):
    """One RK3 step of the partitioned RieCG with Dirichlet, symmetry,
    farfield and pressure BCs (bc: a BCLists over global node ids) and the
    problem's source term; U (global) in place, returns Un."""
    rkcoef = (1.0 / 3.0, 1.0 / 2.0, 1.0)
    npoin, ncomp = U.shape
    Un = U.copy()
    for stage in range(3):
        G = np.zeros((npoin, 3 * ncomp))
        for p in parts:                                    # grad + comgrad
            np.add.at(G, p.gid, grad(p.se, p.coord, p.tri, U[p.gid]))
        G /= Vol[:, None]
        R = np.zeros((npoin, ncomp))
        for p in parts:                                    # rhs + comrhs
            Rk = rhs(cfg, p.se, p.coord, p.tri, p.besym, G[p.gid], U[p.gid])
            src(P, p.coord, p.vol, Rk, t)
            np.add.at(R, p.gid, Rk)
        U[:] = Un - rkcoef[stage] * dt * R / Vol[:, None]
        orc_apply_bcs(cfg, P, bc, U, t + rkcoef[stage] * dt, coord)
    return Un


def diagnostics(parts, U, Un):
    """NodeDiagnostics::rhocompute summed over the chares (each with its own
    nodal volumes V, so a shared node counts once in total)."""
    out = 0.0
    for p in parts:
        out = out + diag(U[p.gid], Un[p.gid], p.vol)
    return out
