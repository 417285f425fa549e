"""Shared test inputs: the reference's Sod fixture mesh and seeded boxes,
with the RieCG preprocessing (volumes, superedges, symmetry BC data) done by
the oracle."""
import os

import numpy as np

import oracle as O
from xyst_amd.meshgen import box_mesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Case:
    def __init__(self, coord, inpoel, tri, faceset, symsets):
        self.coord = np.ascontiguousarray(coord)
        self.inpoel = np.ascontiguousarray(inpoel, dtype=np.int64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int64)
        self.faceset = np.ascontiguousarray(faceset, dtype=np.int32)
        self.symsets = list(symsets)
        self.npoin = self.coord.shape[1]
        self.vol = O.vol(self.coord, self.inpoel)
        self.edge, self.dint = O.domint(self.coord, self.inpoel)
        self.se = O.domsuped(self.inpoel, self.edge, self.dint)
        self.symnodes, self.symnorms, self.besym = O.symbc_setup(
            self.coord, self.tri, self.faceset, self.symsets)


def sod_fixture(name="riecg_sod.npz"):
    d = np.load(os.path.join(GOLDEN, name))
    return Case(d["coord"], d["inpoel"], d["triinpoel"], d["faceset"],
                d["symsets"]), d


def dirichlet_nodes(case, sets):
    """Sorted unique nodes of the boundary faces on the given side sets."""
    on = np.isin(case.faceset, np.asarray(sets))
    return np.unique(case.tri[on].reshape(-1))


def box_case(n, seed=20261202, symsets=(1, 2, 3, 4, 5, 6), jitter=0.1):
    m = box_mesh(n, jitter=jitter, seed=seed)
    return Case(m["coord"], m["inpoel"], m["triinpoel"], m["faceset"], symsets)


def smooth_state(coord, seed=7, gamma=1.4):
    """A smooth, non-uniform, positive state (conserved AoS)."""
    rng = np.random.default_rng(seed)
    x, y, z = coord
    k = rng.uniform(1.0, 3.0, size=(5, 3))
    ph = rng.uniform(0, 2 * np.pi, size=5)
    f = [np.sin(k[i, 0] * x + k[i, 1] * y + k[i, 2] * z + ph[i]) for i in range(5)]
    r = 1.0 + 0.3 * f[0]
    u, v, w = 0.5 * f[1], 0.4 * f[2], 0.3 * f[3]
    p = 1.0 + 0.4 * f[4]
    U = np.empty((x.size, 5))
    U[:, 0] = r
    U[:, 1] = r * u
    U[:, 2] = r * v
    U[:, 3] = r * w
    U[:, 4] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v + w * w)
    return U
