"""Synthetic box meshes and initial states for the edge-based flow path.

The reference reads ExodusII meshes; the workloads this package is measured on
are seeded, jittered n^3 hex boxes, each hex split into 6 tetrahedra sharing the
main diagonal (Freudenthal/Kuhn split, conforming across hexes).  Node ids are
lexicographic, i fastest: id = i + (n+1)*(j + (n+1)*k).

Boundary nodes are jittered tangentially only, so the six faces stay planar and
the symmetry normals are exact axis vectors.  Side sets are numbered
1: x=0, 2: x=1, 3: y=0, 4: y=1, 5: z=0, 6: z=1; boundary triangles are ordered
so that cross(b-a, c-a) points out of the box, the convention the reference's
boundary integrals assume (Riemann.cpp:340-346, RieCG.cpp:303-307).
"""
import numpy as np

__all__ = ["box_mesh", "sod_state", "sedov_state", "box_counts", "renumber_mesh"]


def box_counts(n):
    """Closed-form sizes of the n^3 hex-split box: N, T, E, F_b."""
    N = (n + 1) ** 3
    T = 6 * n ** 3
    E = 3 * n * (n + 1) ** 2 + 3 * n ** 2 * (n + 1) + n ** 3
    Fb = 12 * n ** 2
    return N, T, E, Fb


def _tets(n):
    """Connectivity of the 6-tet split, positive orientation, int64 (T,4)."""
    s1, s2 = 1, n + 1
    s3 = (n + 1) ** 2
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # cell order: i fastest, then j, then k
    base = (i + s2 * j + s3 * k).transpose(2, 1, 0).reshape(-1)
    ex, ey, ez = s1, s2, s3
    e111 = ex + ey + ez
    # (a, b) axis pairs of the 6 monotone lattice paths v000 -> v111
    paths = [(ex, ey), (ex, ez), (ey, ex), (ey, ez), (ez, ex), (ez, ey)]
    odd = [False, True, True, False, False, True]   # parity of the permutation
    tets = np.empty((base.size, 6, 4), dtype=np.int64)
    for t, ((a, b), o) in enumerate(zip(paths, odd)):
        v1 = base + a
        v2 = base + a + b
        if o:   # odd permutation: swap the two middle nodes for positive volume
            v1, v2 = v2, v1
        tets[:, t, 0] = base
        tets[:, t, 1] = v1
        tets[:, t, 2] = v2
        tets[:, t, 3] = base + e111
    return tets.reshape(-1, 4)


def _bfaces(n):
    """Boundary triangles (F_b,3) and their side-set ids (F_b,)."""
    s2, s3 = n + 1, (n + 1) ** 2
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = a.T.reshape(-1)
    b = b.T.reshape(-1)
    tris, sets = [], []

    def node(i, j, k):
        return i + s2 * j + s3 * k

    # each entry: side set, fixed axis value, the two face triangles as
    # offsets (dx,dy,dz) relative to the cell corner, outward sign check later
    for sid, (axis, lo) in enumerate([(0, True), (0, False), (1, True),
                                      (1, False), (2, True), (2, False)], 1):
        c = 0 if lo else n - 1
        if axis == 0:
            i, j, k = np.full_like(a, c), a, b
        elif axis == 1:
            i, j, k = a, np.full_like(a, c), b
        else:
            i, j, k = a, b, np.full_like(a, c)
        off = 0 if lo else 1
        # corners of the hex face in the plane axis == off
        if axis == 0:
            q = [(off, 0, 0), (off, 1, 0), (off, 1, 1), (off, 0, 1)]
        elif axis == 1:
            q = [(0, off, 0), (1, off, 0), (1, off, 1), (0, off, 1)]
        else:
            q = [(0, 0, off), (1, 0, off), (1, 1, off), (0, 1, off)]
        # the split's face diagonal joins the face corner nearest v000 with
        # the corner nearest v111: q[0]-q[2] on every face
        for t in ((q[0], q[1], q[2]), (q[0], q[3], q[2])):
            tri = np.stack([node(i + d[0], j + d[1], k + d[2]) for d in t], 1)
            tris.append(tri)
            sets.append(np.full(tri.shape[0], sid, dtype=np.int32))
    return np.concatenate(tris).astype(np.int64), np.concatenate(sets)


def box_mesh(n, jitter=0.1, seed=20261202, length=(1.0, 1.0, 1.0)):
    """Jittered n^3 hex-split box.

    Returns a dict with coord (3,N) float64, inpoel (T,4) int64,
    triinpoel (F_b,3) int64 (outward), faceset (F_b,) int32.
    """
    N = (n + 1) ** 3
    g = np.arange(n + 1, dtype=np.float64) / n
    k, j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1),
                          indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-jitter, jitter, size=(3, N)) / n if jitter else np.zeros((3, N))
    for ax, idx in enumerate((i, j, k)):
        d[ax][(idx == 0) | (idx == n)] = 0.0
    coord = np.stack([g[i] + d[0], g[j] + d[1], g[k] + d[2]])
    for ax in range(3):
        coord[ax] *= length[ax]
    inpoel = _tets(n)
    tri, fset = _bfaces(n)
    # orient outward with the unjittered lattice
    P = np.stack([g[i], g[j], g[k]])
    A, B, C = P[:, tri[:, 0]], P[:, tri[:, 1]], P[:, tri[:, 2]]
    nrm = np.cross((B - A).T, (C - A).T)
    axis = (fset - 1) // 2
    sign = np.where((fset - 1) % 2 == 0, -1.0, 1.0)
    flip = nrm[np.arange(tri.shape[0]), axis] * sign < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    return {"n": n, "coord": np.ascontiguousarray(coord), "inpoel": inpoel,
            "triinpoel": tri, "faceset": fset}


def renumber_mesh(mesh, perm):
    """Apply a node permutation: new id of old node p is perm[p]."""
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    out = dict(mesh)
    out["coord"] = np.ascontiguousarray(mesh["coord"][:, inv])
    out["inpoel"] = perm[mesh["inpoel"]]
    out["triinpoel"] = perm[mesh["triinpoel"]]
    return out


def _totalenergy(r, u, v, w, p, gamma):
    # eos::totalenergy, EOS.hpp:52-56
    return p / (gamma - 1.0) + 0.5 * r * (u * u + v * v + w * w)


def sod_state(coord, gamma=1.4):
    """Sod shock tube along x (Problems.cpp:371-405): (N,5) conserved AoS."""
    x = coord[0]
    r = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    U = np.zeros((x.size, 5))
    U[:, 0] = r
    U[:, 4] = _totalenergy(r, 0.0, 0.0, 0.0, p, gamma)
    return U


def sedov_state(coord, vol, gamma=1.4, energy=1.0, radius=0.02, p0=1e-6,
                centre=(0.5, 0.5, 0.5)):
    """Sedov blast: total energy `energy` spread as rho*E proportional to nodal
    volume over the nodes within `radius` of `centre`; rho=1, u=0, p=p0
    elsewhere.  (N,5) conserved AoS."""
    d2 = sum((coord[a] - centre[a]) ** 2 for a in range(3))
    inside = d2 < radius * radius
    U = np.zeros((coord.shape[1], 5))
    U[:, 0] = 1.0
    U[:, 4] = _totalenergy(1.0, 0.0, 0.0, 0.0, p0, gamma)
    if inside.any():
        U[inside, 4] = energy / vol[inside].sum()
    return U
