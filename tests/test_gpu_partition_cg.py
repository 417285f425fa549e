"""Partitions of one linear system on one device (xg_pcg_*, needs an
MI355X): the reference's two-partition ConjugateGradients unit tests
("Laplace eq 1DOF 2PE", "3DOF 2PE": normb and the final residual within the
tests' 1e-12), and a box split into three partitions against the oracle's
partitioned CG (oracle/partitioned.py: normb to 1e-13, the residual history
iterate by iterate to 1e-8 relative, the solution) and against the
single-partition device CG on the assembled system.
"""
import sys

import numpy as np
import pytest

import oracle as O
from oracle.partitioned import PartitionedCG as OraclePCG
from tests.test_oracle_partitioned_cg import dirichlet, laplacian, split_box, two_pe_case

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("ncomp", [1, 3])
def test_reference_two_partition_goldens_on_device(ncomp):
    from xyst_amd.partition import PartitionedCG
    _say("pcg 2PE ncomp=%d" % ncomp)
    d, As, bs, xs = two_pe_case(ncomp)
    cg = PartitionedCG(As, d["gid"], ncomp)
    nb = cg.setup(xs, bs, pc="none")
    assert abs(nb - d["normb"]) < 1e-12
    cg.solve(d["maxit"], d["tol"])
    assert cg.converged
    assert abs(cg.normr - d["normres"]) < 1e-12
    x0, x1 = cg.solution()
    g0, g1 = d["gid"]
    for i, g in enumerate(g0):
        if g in g1:
            j = g1.index(g)
            assert np.array_equal(x0[i * ncomp:(i + 1) * ncomp], x1[j * ncomp:(j + 1) * ncomp])
    cg.close()


@pytest.mark.parametrize("pc", ["none", "jacobi"])
def test_three_partitions_against_the_oracle_and_the_serial_device_solve(pc):
    from xyst_amd import cg as CG
    from xyst_amd.partition import PartitionedCG
    _say("pcg 3 partitions pc=%s" % pc)
    m, parts = split_box(8, 3)
    coord, inpoel = m["coord"], m["inpoel"]
    N = coord.shape[1]
    count = np.zeros(N, int)
    for gid, _, _ in parts:
        count[gid] += 1
    bglob = np.random.default_rng(7).standard_normal(N) * 1e4
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
    # the oracle's partitioned CG, iterate by iterate
    oc = OraclePCG(As, gids, 1)
    nbo = oc.setup(xs, bs, pc=pc)
    nro, ito, cvo, ho = oc.solve(1000, 1e-10)
    hist = []
    dc = PartitionedCG(As, gids, 1)
    nb = dc.setup(xs, bs, pc=pc)
    assert nb == pytest.approx(nbo, rel=1e-13)
    for k in range(1, ito + 1):       # one iteration more per call, from the start
        dc.setup(xs, bs, pc=pc)
        dc.solve(k, 1e-10)
        hist.append(dc.normr)
    _say("  oracle it=%d, device it=%d" % (ito, dc.it))
    assert cvo and dc.converged and abs(dc.it - ito) <= 1
    k = min(len(hist), 40)
    np.testing.assert_allclose(hist[:k], ho[:k], rtol=1e-8)
    xd = dc.solution()
    for a, b in zip(xd, oc.x):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * np.abs(b).max())
    # the single-partition device CG on the assembled system
    Ag = laplacian(coord, inpoel, 1)
    bg = bglob.copy()
    Ag.dirichlet(0, 0.0, bg)
    A1 = CG.CSR(1, Ag.ia, Ag.ja, Ag.a)
    sc = CG.ConjugateGradients(A1)
    sc.init(np.zeros(N), bg, pc=pc)
    sc.solve(1000, 1e-10)
    xs1 = sc.solution()
    x = np.zeros(N)
    for gid, xk in zip(gids, xd):
        x[gid] = xk
    assert abs(sc.it - dc.it) <= 1
    np.testing.assert_allclose(x, xs1, rtol=0, atol=1e-8 * np.abs(xs1).max())
    dc.close()
    sc.ctx.close()
