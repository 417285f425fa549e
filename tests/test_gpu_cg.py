"""Parity of the device ConjugateGradients with the oracle (needs an MI355X).

The device sums dot products in a different order than the oracle's
sequential loops, so iterates agree to round-off that grows with the
iteration count; the checks are: the reference unit-test goldens at their
1e-12 tolerance, iteration counts within one of the oracle's, solutions
within the solve tolerance of each other, and the true residual
||b - A x|| recomputed on the host.
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.test_oracle_cg import _unit_mesh
from xyst_amd.meshgen import box_mesh

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _X():
    from xyst_amd import cg as X
    return X


def _dev_matrix(Ao):
    X = _X()
    return X.CSR(Ao.ncomp, Ao.ia, Ao.ja, Ao.a)


@pytest.mark.parametrize("ncomp", [1, 3])
def test_cg_unit_goldens(ncomp):
    """TestConjugateGradients.cpp tests 1 and 3 on the device."""
    _say("cg goldens ncomp=%d" % ncomp)
    coord, inpoel, d = _unit_mesh()
    g = d["ncomp"][str(ncomp)]
    n = coord.shape[1]
    A = O.CSR.from_tets(n, inpoel, ncomp)
    O.laplace_test(coord, inpoel, A)
    b = np.ones(n * ncomp)
    for i in range(ncomp):
        A.dirichlet(0, 0.0, b, i)
    cg = _X().ConjugateGradients(_dev_matrix(A))
    cg.init(np.zeros(n * ncomp), b)
    normr = cg.solve(g["maxit"], g["tol"])
    assert abs(cg.normb - g["normb"]) < d["tolerance"]
    assert abs(normr - g["normres"]) < d["tolerance"]
    assert cg.converged
    xs = np.linalg.solve(A.dense(), b)
    np.testing.assert_allclose(cg.solution(), xs, rtol=1e-10, atol=1e-12)


def _poisson_case(n, seed=2):
    m = box_mesh(n, seed=seed)
    coord, inpoel = m["coord"], m["inpoel"]
    npoin = coord.shape[1]
    A = O.CSR.from_tets(npoin, inpoel)
    O.prelhs(coord, inpoel, A)
    return coord, inpoel, A


@pytest.mark.parametrize("pc", ["none", "jacobi"])
def test_cg_apply_matches_oracle(pc):
    """init(apply) with Dirichlet and Neumann data on the ChoCG matrix: same
    BC-modified system (normb), iteration count within 2.5 percent (at least
    one: the device sums the dot products in another order, which moves the
    tolerance crossing of the unpreconditioned 1e-8 solve by a couple of
    iterations), solutions within the tolerance, A restored for the next
    solve."""
    _say("cg apply pc=%s" % pc)
    coord, inpoel, A = _poisson_case(8)
    n = coord.shape[1]
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n) * 10.0
    neu = rng.standard_normal(n)
    dirbc = {int(i): [(1, 0.3 * coord[1][i])] for i in np.where(coord[0] < 1e-12)[0]}
    tol = 1e-8
    xo, bo = np.zeros(n), b.copy()
    nbo, nro, ito, cvo = O.cg(A, xo, bo, neubc=neu, dirbc=dirbc, apply=True, pc=pc,
                              maxit=5000, tol=tol)
    assert cvo
    cg = _X().ConjugateGradients(_dev_matrix(A))
    cg.init(np.zeros(n), b, neubc=neu, dirbc=dirbc, apply=True, pc=pc)
    cg.solve(5000, tol)
    assert cg.converged
    assert cg.normb == pytest.approx(nbo, rel=1e-13)
    assert abs(cg.it - ito) <= max(1, ito // 40), (cg.it, ito)
    x = cg.solution()
    scale = np.abs(xo).max()
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-6 * scale)
    # true residual of the BC-modified system (bo holds its rhs)
    # (CSR::dirichlet accumulates A_col*val; ConjugateGradients subtracts it)
    Ab = O.CSR(A.npoin, 1, A.ia, A.ja, A.a.copy())
    rr = np.zeros(n)
    for i in sorted(dirbc, reverse=True):
        Ab.dirichlet(i, dirbc[i][0][1], rr)
    bb = b + neu - rr
    for i in dirbc:
        bb[i] = dirbc[i][0][1]
    np.testing.assert_allclose(bb, bo, rtol=0, atol=1e-12 * np.abs(bo).max())
    assert np.linalg.norm(bb - Ab.mult(x)) < 2 * tol * nbo
    # a solve without BCs runs on the restored matrix: b = A y (consistent),
    # 20 iterations, same residual as the oracle's on the original A
    y = np.sin(3 * coord[0]) * coord[1] * 10.0
    b2 = A.mult(y)
    x2, b2o = np.zeros(n), b2.copy()
    _, nr2o, _, _ = O.cg(A, x2, b2o, pc=pc, maxit=20, tol=1e-30)
    cg.init(np.zeros(n), b2, pc=pc)
    nr2 = cg.solve(20, 1e-30)
    assert cg.it == 20 and nr2 == pytest.approx(nr2o, rel=1e-8)
    np.testing.assert_allclose(cg.solution(), x2, rtol=0, atol=1e-9 * np.abs(x2).max())


@pytest.mark.parametrize("tol", [1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6])
def test_cg_stop_by_tolerance_keeps_the_iterate(tol):
    """A solve that stops on the tolerance returns the iterate of the
    recurrence at that count: the launches queued after the stop (the host
    polls in batches) must not touch x.  Loose tolerances put the stop at
    every position within a batch; x is compared to the oracle's iterate to
    round-off (the recurrence itself agrees to ~1e-11 here)."""
    _say("cg stop by tolerance tol=%g" % tol)
    coord, inpoel, A = _poisson_case(8)
    n = coord.shape[1]
    b = np.cos(2 * coord[0]) + coord[2] * coord[1]
    dirbc = {0: [(1, 0.0)]}
    xo, bo = np.zeros(n), b.copy()
    _, nro, ito, _ = O.cg(A, xo, bo, dirbc=dirbc, apply=True, pc="jacobi", maxit=1000,
                          tol=tol)
    cg = _X().ConjugateGradients(_dev_matrix(A))
    cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve(1000, tol)
    assert cg.it == ito
    x = cg.solution()
    err = float(np.abs(x - xo).max() / np.abs(xo).max())
    _say("  it %d, |x - x_oracle| / max|x| = %.3e" % (ito, err))
    assert err < 1e-9


def test_cg_residual_history_matches_oracle():
    """Iterate by iterate: with maxit = k the device and oracle residual norms
    agree to round-off for the first iterations."""
    _say("cg history")
    coord, inpoel, A = _poisson_case(6)
    n = coord.shape[1]
    b = np.cos(2 * coord[0]) + coord[2]
    dirbc = {int(i): [(1, 0.0)] for i in np.where(coord[0] < 1e-12)[0]}
    xo, bo = np.zeros(n), b.copy()
    _, _, _, _, hist = O.cg(A, xo, bo, dirbc=dirbc, apply=True, pc="jacobi",
                            maxit=30, tol=1e-30, history=True)
    cg = _X().ConjugateGradients(_dev_matrix(A))
    for k in (1, 2, 5, 10, 30):
        cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
        nr = cg.solve(k, 1e-30)
        assert cg.it == k
        assert nr == pytest.approx(hist[k - 1], rel=1e-9), k


def test_cg_breakdown_and_maxit():
    _say("cg breakdown")
    coord, inpoel, _ = _unit_mesh()
    n = coord.shape[1]
    A = O.CSR.from_tets(n, inpoel)
    O.laplace_test(coord, inpoel, A)
    A.dirichlet(0, 0.0, np.zeros(n))
    cg = _X().ConjugateGradients(_dev_matrix(A))
    cg.init(np.zeros(n), np.zeros(n))
    nr = cg.solve(10, 1e-8)
    assert cg.it == 1 and cg.converged and nr == 0.0
    cg.init(np.zeros(n), np.ones(n))
    cg.solve(2, 1e-14)
    assert cg.it == 2 and not cg.converged


def test_cg_absolute_breakdown_as_reference():
    """|(p,q)| < DBL_EPSILON ends the solve as in the reference: same
    iteration count as the oracle, reported not converged."""
    _say("cg absolute breakdown")
    coord, inpoel, A = _poisson_case(4)
    n = coord.shape[1]
    b = np.random.default_rng(1).standard_normal(n)
    dirbc = {int(i): [(1, 0.0)] for i in np.where(coord[0] < 1e-12)[0]}
    xo, bo = np.zeros(n), b.copy()
    _, _, ito, cvo = O.cg(A, xo, bo, dirbc=dirbc, apply=True, pc="jacobi",
                          maxit=2000, tol=1e-13)
    cg = _X().ConjugateGradients(_dev_matrix(A))
    cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve(2000, 1e-13)
    assert not cvo and not cg.converged
    assert abs(cg.it - ito) <= 1


@pytest.mark.parametrize("n", [24, 48])
def test_poisson_solve_to_1e10(n):
    """Pressure-Poisson solve on an n^3 box to 1e-10 relative residual, FP64
    and mixed FP32/FP64: the true residual recomputed on the host."""
    X = _X()
    _say("poisson n=%d" % n)
    m = box_mesh(n, seed=5)
    coord, inpoel = m["coord"], m["inpoel"]
    A = X.CSR.poisson(coord, inpoel)
    npoin = coord.shape[1]
    Ao = O.CSR(npoin, 1, A.ia.astype(np.int64), A.ja.astype(np.int64), A.a.copy())
    x0, y0, z0 = coord
    f = sum(np.sin((k + 1) * np.pi * x0) * np.cos(k * np.pi * y0) * (1 + z0 * k)
            for k in range(4))
    b = Ao.mult(f) * 1e6
    dirbc = {int(i): [(1, 0.0)] for i in np.where(x0 < 1e-12)[0]}
    Ab = O.CSR(npoin, 1, Ao.ia, Ao.ja, Ao.a.copy())
    bb = b.copy()
    for i in sorted(dirbc, reverse=True):
        Ab.dirichlet(i, 0.0, bb)
    for i in dirbc:
        bb[i] = 0.0
    nb = np.linalg.norm(bb)
    cg = X.ConjugateGradients(A)
    cg.init(np.zeros(npoin), b, dirbc=dirbc, apply=True, pc="jacobi")
    assert cg.normb == pytest.approx(nb, rel=1e-12)
    cg.solve(20000, 1e-10)
    assert cg.converged
    x = cg.solution()
    true = np.linalg.norm(bb - Ab.mult(x))
    _say("  fp64 it=%d true/normb=%.3e" % (cg.it, true / nb))
    assert true < 1e-10 * nb
    cg.init(np.zeros(npoin), b, dirbc=dirbc, apply=True, pc="jacobi")
    nr = cg.solve_mixed(20000, 1e-10, inner_tol=1e-4)
    assert cg.converged
    xm = cg.solution()
    truem = np.linalg.norm(bb - Ab.mult(xm))
    _say("  mixed it=%d outer=%d true/normb=%.3e" % (cg.it, cg.nouter, truem / nb))
    assert truem < 1e-10 * nb and nr == pytest.approx(truem, rel=1e-3)
    np.testing.assert_allclose(xm, x, rtol=0, atol=1e-8 * np.abs(x).max())
    # one FP32 recurrence with FP64 reliable updates: a true 1e-10 in about
    # the FP64 iteration count
    it64 = None
    cg.init(np.zeros(npoin), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve(20000, 1e-10)
    it64 = cg.it
    cg.init(np.zeros(npoin), b, dirbc=dirbc, apply=True, pc="jacobi")
    nrr = cg.solve_reliable(20000, 1e-10, delta=0.1)
    assert cg.converged
    xr = cg.solution()
    truer = np.linalg.norm(bb - Ab.mult(xr))
    _say("  reliable it=%d (fp64 %d) updates=%d true/normb=%.3e"
         % (cg.it, it64, cg.nupdates, truer / nb))
    assert truer < 1e-10 * nb and nrr == pytest.approx(truer, rel=1e-3)
    assert cg.it <= 1.3 * it64 + 10
    np.testing.assert_allclose(xr, x, rtol=0, atol=1e-8 * np.abs(x).max())


def test_cg_apply_same_dirichlet_set_new_values():
    """Two init(apply) calls with the same Dirichlet nodes and new values and
    rhs (the ChoCG time-step pattern): the second system equals the oracle's
    fresh application (normb to round-off, same iteration count within one)."""
    _say("cg apply twice")
    coord, inpoel, A = _poisson_case(6)
    n = coord.shape[1]
    rng = np.random.default_rng(4)
    nodes = np.where(coord[0] < 1e-12)[0]
    cg = _X().ConjugateGradients(_dev_matrix(A))
    for it in range(3):
        b = rng.standard_normal(n) * 5.0
        dirbc = {int(i): [(1, float(rng.standard_normal()))] for i in nodes}
        xo, bo = np.zeros(n), b.copy()
        nbo, _, ito, cvo = O.cg(A, xo, bo, dirbc=dirbc, apply=True, pc="jacobi",
                                maxit=3000, tol=1e-9)
        cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
        cg.solve(3000, 1e-9)
        assert cg.normb == pytest.approx(nbo, rel=1e-13), it
        assert cvo and cg.converged and abs(cg.it - ito) <= 1
        np.testing.assert_allclose(cg.solution(), xo, rtol=0,
                                   atol=1e-7 * np.abs(xo).max())


@pytest.mark.parametrize("pc,inner", [("jacobi", 1), ("cg32", 8), ("cg32", 30), ("amg", 1)])
def test_flexible_cg_to_1e10(pc, inner):
    """FP64 flexible CG with Jacobi or FP32 inner-CG preconditioning reaches a
    true relative residual of 1e-10 on the pressure-Poisson system; with
    Jacobi it takes the iterations of Jacobi-PCG to within a few."""
    X = _X()
    _say("fcg %s inner=%d" % (pc, inner))
    m = box_mesh(32, seed=9)
    coord, inpoel = m["coord"], m["inpoel"]
    A = X.CSR.poisson(coord, inpoel)
    npoin = coord.shape[1]
    Ao = O.CSR(npoin, 1, A.ia.astype(np.int64), A.ja.astype(np.int64), A.a.copy())
    x0 = coord[0]
    f = np.sin(np.pi * x0) * np.cos(2 * np.pi * coord[1]) + coord[2] ** 2
    b = Ao.mult(f) * 1e5
    dirbc = {int(i): [(1, 0.0)] for i in np.where(x0 < 1e-12)[0]}
    Ab = O.CSR(npoin, 1, Ao.ia, Ao.ja, Ao.a.copy())
    rr = np.zeros(npoin)
    for i in sorted(dirbc, reverse=True):
        Ab.dirichlet(i, 0.0, rr)
    bb = b - rr
    for i in dirbc:
        bb[i] = 0.0
    nb = np.linalg.norm(bb)
    cg = X.ConjugateGradients(A)
    cg.init(np.zeros(npoin), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve(20000, 1e-10)
    it_pcg = cg.it
    cg.init(np.zeros(npoin), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve_flexible(20000, 1e-10, pc=pc, inner=inner)
    assert cg.converged
    x = cg.solution()
    true = np.linalg.norm(bb - Ab.mult(x))
    _say("  pcg it=%d fcg it=%d inner=%d true/normb=%.3e"
         % (it_pcg, cg.it, cg.inner_total, true / nb))
    assert true < 1e-10 * nb
    if pc == "jacobi":
        assert abs(cg.it - it_pcg) <= max(5, it_pcg // 50)
    elif pc == "amg":
        lv, oc, ms = cg.amg_info()
        _say("  amg levels=%d complexity=%.3f setup=%.0f ms" % (lv, oc, ms))
        assert cg.it <= 30 and lv >= 3 and oc < 2.0
    else:
        assert cg.it < it_pcg


def test_amg_fcg_hydrostat_and_rebuild():
    """The ChoCG situation: a single Dirichlet (hydrostat) node, the
    multigrid-preconditioned FCG to 1e-10; a new Dirichlet set rebuilds the
    hierarchy, the same set reuses it (same setup time reported)."""
    X = _X()
    _say("amg hydrostat")
    m = box_mesh(24, seed=3)
    coord, inpoel = m["coord"], m["inpoel"]
    A = X.CSR.poisson(coord, inpoel)
    n = coord.shape[1]
    Ao = O.CSR(n, 1, A.ia.astype(np.int64), A.ja.astype(np.int64), A.a.copy())
    b = Ao.mult(np.cos(3 * coord[0]) * coord[1]) + 0.01
    cg = X.ConjugateGradients(A)
    out = []
    for dirbc in ({0: [(1, 0.0)]}, {0: [(1, 0.0)]},
                  {int(i): [(1, 0.0)] for i in np.where(coord[2] < 1e-12)[0]}):
        cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
        cg.solve_flexible(200, 1e-10, pc="amg")
        assert cg.converged
        Ab = O.CSR(n, 1, Ao.ia, Ao.ja, Ao.a.copy())
        rr = np.zeros(n)
        for i in sorted(dirbc, reverse=True):
            Ab.dirichlet(i, 0.0, rr)
        bb = b - rr
        for i in dirbc:
            bb[i] = 0.0
        x = cg.solution()
        assert np.linalg.norm(bb - Ab.mult(x)) < 1e-10 * np.linalg.norm(bb)
        out.append((cg.it, cg.amg_info()))
        _say("  it=%d levels/complexity/setup=%s" % out[-1])
    assert out[0][1] == out[1][1]          # same set: hierarchy reused
    assert all(it <= 40 for it, _ in out)
