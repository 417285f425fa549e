"""The oracle's partitioned ConjugateGradients (oracle/partitioned.py
PartitionedCG), pinned on the CPU:

* the reference's two-partition unit tests (TestConjugateGradients.cpp,
  "Laplace eq 1DOF 2PE" and "3DOF 2PE"): per partition the mesh, the node
  maps and the Laplacian with a Dirichlet node whose diagonal holds
  1/count (CSR::dirichlet, CSR.cpp:107-152); normb and the final residual
  within the tests' 1e-12;
* a box split into three partitions against the serial oracle CG on the
  assembled system (the sum of the partitions' matrices): same normb, the
  same residual history to round-off, the same solution.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from oracle.partitioned import PartitionedCG
from xyst_amd.meshgen import box_mesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cg_unit.json")


def laplacian(coord, inpoel, ncomp):
    """The unit test's matrix: A(a, b, i) += J/6 grad_a . grad_b per tet
    (TestConjugateGradients.cpp:306-332), on the CSR pattern of the tets."""
    npoin = coord.shape[1]
    A = O.CSR.from_tets(npoin, inpoel, ncomp)
    pos = {}
    for r in range(npoin * ncomp):
        for j in range(A.ia[r], A.ia[r + 1]):
            pos[(r, int(A.ja[j]))] = j
    X, Y, Z = coord
    for N in inpoel:
        ba = np.array([X[N[1]] - X[N[0]], Y[N[1]] - Y[N[0]], Z[N[1]] - Z[N[0]]])
        ca = np.array([X[N[2]] - X[N[0]], Y[N[2]] - Y[N[0]], Z[N[2]] - Z[N[0]]])
        da = np.array([X[N[3]] - X[N[0]], Y[N[3]] - Y[N[0]], Z[N[3]] - Z[N[0]]])
        J = np.dot(ba, np.cross(ca, da))
        assert J > 0
        g = np.zeros((4, 3))
        g[1] = np.cross(ca, da) / J
        g[2] = np.cross(da, ba) / J
        g[3] = np.cross(ba, ca) / J
        g[0] = -g[1] - g[2] - g[3]
        for a in range(4):
            for k in range(3):
                for b in range(4):
                    for i in range(ncomp):
                        A.a[pos[(N[a] * ncomp + i, N[b] * ncomp + i)]] += J / 6 * g[a][k] * g[b][k]
    return A


def dirichlet(A, i, b, count):
    """CSR::dirichlet(i, 0, b, gid, nodecommap, pos) for every component:
    zero the columns (b += a * 0) and the row, 1/count on the diagonal."""
    nc = A.ncomp
    for pos in range(nc):
        col = i * nc + pos
        for r in range(A.npoin * nc):
            for j in range(A.ia[r], A.ia[r + 1]):
                if A.ja[j] == col:
                    b[r] += A.a[j] * 0.0
                    A.a[j] = 0.0
                    break
        for j in range(A.ia[col], A.ia[col + 1]):
            A.a[j] = 1.0 / count if A.ja[j] == col else 0.0


def two_pe_case(ncomp):
    d = json.load(open(GOLD))["two_pe"][str(ncomp)]
    As, bs, xs = [], [], []
    for k in range(2):
        coord = np.array(d["coord"][k])
        inpoel = np.array(d["inpoel"][k], np.int64).reshape(-1, 4)
        A = laplacian(coord, inpoel, ncomp)
        # the pattern is the psup of the test
        p1, p2 = d["psup"][k]
        npoin = len(p2) - 1
        for i in range(npoin):
            nb = sorted([i] + p1[p2[i] + 1:p2[i + 1] + 1])
            row = [int(v) // ncomp for v in A.ja[A.ia[i * ncomp]:A.ia[i * ncomp + 1]]]
            assert sorted(set(row)) == nb
        b = np.ones(npoin * ncomp)
        gid = d["gid"][k]
        if 0 in gid:
            count = 1 + sum(0 in v for m in d["nodecommap"][k].values() for v in [m])
            dirichlet(A, gid.index(0), b, count)
        As.append(A)
        bs.append(b)
        xs.append(np.zeros(npoin * ncomp))
    return d, As, bs, xs


@pytest.mark.parametrize("ncomp", [1, 3])
def test_reference_two_partition_goldens(ncomp):
    d, As, bs, xs = two_pe_case(ncomp)
    cg = PartitionedCG(As, d["gid"], ncomp)
    normb = cg.setup(xs, bs, pc="none")
    assert abs(normb - d["normb"]) < 1e-12
    normr, it, conv, _ = cg.solve(d["maxit"], d["tol"])
    assert conv
    assert abs(normr - d["normres"]) < 1e-12
    # the copies of every shared node agree
    g0, g1 = d["gid"]
    x0, x1 = cg.x
    for i, g in enumerate(g0):
        if g in g1:
            j = g1.index(g)
            assert np.array_equal(x0[i * ncomp:(i + 1) * ncomp], x1[j * ncomp:(j + 1) * ncomp])


def split_box(n, nparts, seed=4):
    """A box split into slabs along x by tet centroid; per partition the
    local mesh, the global ids and the partial Laplacian."""
    m = box_mesh(n, seed=seed)
    coord, inpoel = m["coord"], m["inpoel"]
    cx = coord[0][inpoel].mean(1)
    owner = np.minimum((cx * nparts).astype(int), nparts - 1)
    parts = []
    for k in range(nparts):
        tets = inpoel[owner == k]
        gid = np.unique(tets.reshape(-1))
        lid = {int(g): i for i, g in enumerate(gid)}
        loc = np.array([[lid[int(g)] for g in t] for t in tets], np.int64)
        parts.append((gid, loc, np.ascontiguousarray(coord[:, gid])))
    return m, parts


@pytest.mark.parametrize("pc", ["none", "jacobi"])
def test_three_partitions_equal_the_assembled_serial_solve(pc):
    m, parts = split_box(6, 3)
    coord, inpoel = m["coord"], m["inpoel"]
    N = coord.shape[1]
    count = np.zeros(N, int)
    for gid, _, _ in parts:
        count[gid] += 1
    rng = np.random.default_rng(2)
    # large enough that |(p,q)| stays above the absolute DBL_EPSILON
    # breakdown test (ConjugateGradients.cpp:682-684) down to 1e-10
    bglob = rng.standard_normal(N) * 1e4
    As, bs, xs, gids = [], [], [], []
    for gid, loc, crd in parts:
        A = laplacian(crd, loc, 1)
        b = bglob[gid].copy()
        if 0 in gid:
            dirichlet(A, int(np.flatnonzero(gid == 0)[0]), b, count[0])
        As.append(A)
        bs.append(b)
        xs.append(np.zeros(gid.size))
        gids.append(gid)
    cg = PartitionedCG(As, gids, 1)
    nb = cg.setup(xs, bs, pc=pc)
    normr, it, conv, hist = cg.solve(500, 1e-10)
    assert conv
    # the serial solve of the assembled system
    Ag = laplacian(coord, inpoel, 1)
    bg = bglob.copy()
    Ag.dirichlet(0, 0.0, bg)
    xg = np.zeros(N)
    nbs, nrs, its, cvs, hs = O.cg(Ag, xg, bg, pc=pc, maxit=500, tol=1e-10, history=True)
    assert nb == pytest.approx(nbs, rel=1e-13)
    assert cvs and abs(it - its) <= 1
    k = min(it, its, 30)
    np.testing.assert_allclose(hist[:k], hs[:k], rtol=1e-8)
    x = np.zeros(N)
    for gid, xk in zip(gids, cg.x):
        x[gid] = xk
    np.testing.assert_allclose(x, xg, rtol=0, atol=1e-8 * np.abs(xg).max())
