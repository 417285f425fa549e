"""Pins the CPU oracle of tk::CSR / ConjugateGradients / ChoCG::prelhs."""
import json
import os

import numpy as np
import pytest

import oracle as O
from xyst_amd.meshgen import box_mesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cg_unit.json")


def _unit_mesh():
    d = json.load(open(GOLD))
    inpoel = np.array(d["inpoel_1based"], np.int64).reshape(-1, 4) - 1
    return np.array(d["coord"]), inpoel, d


@pytest.mark.parametrize("ncomp", [1, 3])
def test_cg_unit_goldens(ncomp):
    """TestConjugateGradients.cpp tests 1 and 3: +K Laplacian, Dirichlet 0 at
    node 0 for every component, b = 1, x = 0, pc "none"; normb and the final
    residual at the test's 1e-12 tolerance."""
    coord, inpoel, d = _unit_mesh()
    g = d["ncomp"][str(ncomp)]
    npoin = coord.shape[1]
    A = O.CSR.from_tets(npoin, inpoel, ncomp)
    O.laplace_test(coord, inpoel, A)
    b = np.ones(npoin * ncomp)
    x = np.zeros(npoin * ncomp)
    for i in range(ncomp):
        A.dirichlet(0, 0.0, b, i)
    normb, normr, it, conv = O.cg(A, x, b, maxit=g["maxit"], tol=g["tol"])
    assert abs(normb - g["normb"]) < d["tolerance"]
    assert abs(normr - g["normres"]) < d["tolerance"]
    assert conv
    # and the solution solves the system (dense check)
    M = A.dense()
    xs = np.linalg.solve(M, b)
    np.testing.assert_allclose(x, xs, rtol=1e-10, atol=1e-12)


def test_csr_layout_rows_sorted_with_diagonal():
    coord, inpoel, _ = _unit_mesh()
    A = O.CSR.from_tets(coord.shape[1], inpoel, 3)
    for r in range(A.ia.size - 1):
        cols = A.ja[A.ia[r]:A.ia[r + 1]]
        assert np.all(np.diff(cols) > 0) and r in cols
        assert np.all(cols % 3 == r % 3)        # components do not couple


def test_prelhs_is_minus_stiffness():
    """A = -K: symmetric, zero row sums (constants in the kernel), negative
    definite on functions vanishing at one node, and x^T K x equals the
    integral of |grad u|^2 for a linear u (the exact P1 energy)."""
    m = box_mesh(3, seed=6)
    coord, inpoel = m["coord"], m["inpoel"]
    n = coord.shape[1]
    A = O.CSR.from_tets(n, inpoel)
    O.prelhs(coord, inpoel, A)
    M = A.dense()
    np.testing.assert_allclose(M, M.T, atol=1e-14)
    np.testing.assert_allclose(M.sum(1), 0.0, atol=1e-13)
    u = 2.0 * coord[0] - coord[1] + 0.5 * coord[2]
    vol = O.vol(coord, inpoel).sum()
    assert -u @ M @ u == pytest.approx(vol * (4 + 1 + 0.25), rel=1e-12)


@pytest.mark.parametrize("pc", ["none", "jacobi"])
def test_cg_apply_dirichlet_neumann(pc):
    """init(apply) with Dirichlet values and a Neumann vector on the ChoCG
    matrix: the result solves the BC-modified system, matches a dense solve,
    and A is restored afterwards."""
    m = box_mesh(4, seed=2)
    coord, inpoel = m["coord"], m["inpoel"]
    n = coord.shape[1]
    A = O.CSR.from_tets(n, inpoel)
    O.prelhs(coord, inpoel, A)
    a0 = A.a.copy()
    rng = np.random.default_rng(1)
    b = rng.standard_normal(n)
    neu = rng.standard_normal(n) * 0.1
    xnodes = np.where(coord[0] < 1e-12)[0]
    dirbc = {int(i): [(1, 0.3 * coord[1][i])] for i in xnodes}
    x = np.zeros(n)
    b0 = b.copy()
    normb, normr, it, conv = O.cg(A, x, b, neubc=neu, dirbc=dirbc, apply=True,
                                  pc=pc, maxit=2000, tol=1e-7)
    assert conv and np.array_equal(A.a, a0)
    # dense restatement of the BC application
    M = A.dense()
    rhs = b0 + neu
    for i in sorted(dirbc, reverse=True):
        v = dirbc[i][0][1]
        rhs -= M[:, i] * v
        M[:, i] = 0.0
        M[i, :] = 0.0
        M[i, i] = 1.0
    for i in dirbc:
        rhs[i] = dirbc[i][0][1]
    xs = np.linalg.solve(M, rhs)
    assert np.linalg.norm(rhs - M @ x) <= 1e-7 * normb
    np.testing.assert_allclose(x, xs, rtol=1e-5, atol=1e-6)
    assert normb == pytest.approx(np.linalg.norm(rhs), rel=1e-13)


def test_cg_absolute_breakdown_threshold():
    """The reference stops ('finished', alpha = 0) once |(p,q)| drops below
    machine epsilon in absolute terms (ConjugateGradients.cpp:681-687): on a
    small-scale system a tight tolerance is then reported not converged."""
    m = box_mesh(4, seed=2)
    coord, inpoel = m["coord"], m["inpoel"]
    n = coord.shape[1]
    A = O.CSR.from_tets(n, inpoel)
    O.prelhs(coord, inpoel, A)
    b = np.random.default_rng(1).standard_normal(n)
    dirbc = {int(i): [(1, 0.0)] for i in np.where(coord[0] < 1e-12)[0]}
    x = np.zeros(n)
    normb, normr, it, conv = O.cg(A, x, b, dirbc=dirbc, apply=True, pc="jacobi",
                                  maxit=2000, tol=1e-13)
    assert not conv and it < 2000 and normr > 1e-13 * normb


def test_cg_breakdown_and_maxit():
    coord, inpoel, _ = _unit_mesh()
    n = coord.shape[1]
    A = O.CSR.from_tets(n, inpoel)
    O.laplace_test(coord, inpoel, A)
    A.dirichlet(0, 0.0, np.zeros(n))
    # b = 0: (p,q) = 0 at once -> finished, converged (normr 0 < tol*1)
    x = np.zeros(n)
    nb, nr, it, cv = O.cg(A, x, np.zeros(n), maxit=10, tol=1e-8)
    assert it == 1 and cv and nr == 0.0
    # maxit too small -> not converged
    b = np.ones(n)
    x = np.zeros(n)
    nb, nr, it, cv = O.cg(A, x, b, maxit=2, tol=1e-14)
    assert it == 2 and not cv
