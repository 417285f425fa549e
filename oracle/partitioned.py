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


class PartitionedCG:
    """The chares of one ConjugateGradients array on the CPU
    (ConjugateGradients.cpp:105-823), one per partition, restated with an
    explicit loop over partitions.  Partition k holds its rows' partial sums
    A_k (from its own elements, Dirichlet rows with 1/count on the diagonal
    as CSR::dirichlet sets them, CSR.cpp:107-152), b_k with the full values
    at shared nodes and x_k; gid_k maps its local nodes to global ids.

    * Partial products at shared nodes are completed by adding every other
      partition's own partial (comres/comq, :167-213, :597-651): each copy
      receives own + the others' in partition order.
    * The Jacobi diagonal is summed the same way (comd, :215-279); without a
      preconditioner each copy contributes 1/count (:225-230), so the sum
      is 1.
    * Dot products skip the nodes a higher partition also holds (tk::slave,
      Reorder.cpp:393-412): each partition's local sum in local order, the
      partitions' sums added in partition order (:129-151).
    * After x += alpha p the copies of a shared node are averaged: the sum
      of all copies in partition order divided by their count (comx/x,
      :721-790).
    Vectors are lists of per-partition arrays (npoin_k * ncomp).
    """

    def __init__(self, As, gids, ncomp):
        self.As = As
        self.gids = [np.asarray(g, np.int64) for g in gids]
        self.ncomp = ncomp
        self.K = len(As)
        holders = {}
        for k, g in enumerate(self.gids):
            for i, v in enumerate(g):
                holders.setdefault(int(v), []).append((k, i))
        # shared nodes: copies in partition order
        self.shared = [h for h in holders.values() if len(h) > 1]
        self.count = [np.array([len(holders[int(v)]) for v in g]) for g in self.gids]
        # 1 where the partition counts the node in dot products
        self.own = [np.array([max(kk for kk, _ in holders[int(v)]) == k for v in g])
                    for k, g in enumerate(self.gids)]

    def _complete(self, vs, avg=False):
        """Every copy of a shared node receives the sum of all copies
        (partition order); avg: divided by their count."""
        nc = self.ncomp
        out = [v.copy() for v in vs]
        for h in self.shared:
            for c in range(nc):
                s = 0.0
                for k, i in h:
                    s += vs[k][i * nc + c]
                if avg:
                    s /= len(h)
                for k, i in h:
                    out[k][i * nc + c] = s
        return out

    def _dot(self, a, b):
        nc = self.ncomp
        D = 0.0
        for k in range(self.K):
            d = 0.0
            for i in range(self.own[k].size):
                if self.own[k][i]:
                    for c in range(nc):
                        d += a[k][i * nc + c] * b[k][i * nc + c]
            D += d
        return D

    def setup(self, xs, bs, pc="jacobi"):
        """ConjugateGradients::setup / initres / rho (:105-329): r = b - A x,
        p = r, the preconditioner, z = r / d, rho = (r, z); returns normb."""
        nc = self.ncomp
        self.x = [np.array(x, np.float64) for x in xs]
        self.b = [np.array(b, np.float64) for b in bs]
        r = self._complete([A.mult(x) for A, x in zip(self.As, self.x)])
        self.r = [-v + b for v, b in zip(r, self.b)]
        self.p = [v.copy() for v in self.r]
        if pc == "jacobi":
            d = []
            for A in self.As:
                n = A.npoin * nc
                dk = np.zeros(n)
                for i in range(n):
                    for j in range(A.ia[i], A.ia[i + 1]):
                        if A.ja[j] == i:
                            dk[i] = A.a[j]
                d.append(dk)
        else:
            d = [np.repeat(1.0 / c, nc) for c in self.count]
        self.d = self._complete(d)
        self.z = [r / d for r, d in zip(self.r, self.d)]
        self.rho = self._dot(self.r, self.z)
        self.normb = np.sqrt(self._dot(self.b, self.b))
        self.q = [np.zeros_like(v) for v in self.r]
        return self.normb

    def solve(self, maxit, tol):
        """ConjugateGradients::solve / next / qAp / q / pq / rz / x
        (:558-823).  Returns (normr, it, converged, history of normr)."""
        eps = np.finfo(np.float64).eps
        it, finished, hist = 0, False, []
        rho0 = 0.0
        while True:
            alpha = 0.0 if it == 0 else self.rho / rho0
            rho0 = self.rho
            self.p = [z + alpha * p for z, p in zip(self.z, self.p)]
            self.q = self._complete([A.mult(p) for A, p in zip(self.As, self.p)])
            pq = self._dot(self.p, self.q)
            if abs(pq) < eps:
                finished, alpha = True, 0.0
            else:
                alpha = self.rho / pq
            self.r = [r - alpha * q for r, q in zip(self.r, self.q)]
            self.z = [r / d for r, d in zip(self.r, self.d)]
            self.rho = self._dot(self.r, self.z)
            normr2 = self._dot(self.r, self.r)
            self.x = self._complete([x + alpha * p for x, p in zip(self.x, self.p)],
                                    avg=True)
            it += 1
            nb = self.normb if self.normb > 1.0e-14 else 1.0
            normr = np.sqrt(normr2)
            hist.append(normr)
            if finished or normr < tol * nb or it >= maxit:
                return normr, it, not normr > tol * nb, np.array(hist)
