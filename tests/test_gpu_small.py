"""The smallest meshes the device path takes (needs an MI355X): one
tetrahedron (every edge and face on the boundary, one block of 4 nodes)
and two tetrahedra sharing a face.  Each operator against the oracle with
the tolerances of the box tests: FP64 per node and component within
1e-12 x the oracle's round-off scale of that entry.
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.meshes import Case

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _boundary(inpoel):
    """Faces of exactly one tetrahedron, oriented by the partitioner rule
    (Partitioner.cpp:578-587), with side-set ids 1, 2, ... in order."""
    seen = {}
    for t in inpoel:
        for fc in ((t[0], t[2], t[1]), (t[0], t[1], t[3]), (t[0], t[3], t[2]),
                   (t[1], t[2], t[3])):
            k = tuple(sorted(fc))
            seen[k] = None if k in seen else fc
    tri = np.array([f for f in seen.values() if f is not None], np.int64)
    return tri, np.arange(1, tri.shape[0] + 1, dtype=np.int32)


def _mesh(kind):
    X = np.array([[0.0, 1.1, 0.1, 0.05, 0.85],
                  [0.0, 0.05, 0.9, 0.1, 0.8],
                  [0.0, 0.1, 0.05, 1.2, 0.9]])
    if kind == "one":
        inpoel = np.array([[0, 1, 2, 3]], np.int64)
        X = X[:, :4]
    else:
        inpoel = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int64)
    for t in inpoel:   # positive orientation
        if np.linalg.det(X[:, t[1:]].T - X[:, t[0]]) < 0:
            t[[1, 2]] = t[[2, 1]]
        assert np.linalg.det(X[:, t[1:]].T - X[:, t[0]]) > 0
    tri, faceset = _boundary(inpoel)
    return X, inpoel, tri, faceset


def _state(n, seed=3):
    rng = np.random.default_rng(seed)
    r = 1.0 + 0.2 * rng.random(n)
    u = 0.3 * rng.standard_normal((n, 3))
    p = 1.0 + 0.3 * rng.random(n)
    U = np.empty((n, 5))
    U[:, 0] = r
    U[:, 1:4] = r[:, None] * u
    U[:, 4] = p / 0.4 + 0.5 * r * (u ** 2).sum(1)
    return U


def _close(x, ref, A, what):
    err = np.abs(x - ref)
    bad = err > TOL * A + 1e-300
    assert not bad.any(), "%s: %d entries off, worst ratio %.3e" % (
        what, bad.sum(), float(np.max(err / np.maximum(A, 1e-300))))


@pytest.mark.parametrize("kind", ["one", "two"])
@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_riemann_grad_rhs(kind, flux):
    from xyst_amd import riemann as X
    _say("small riemann %s %s" % (kind, flux))
    coord, inpoel, tri, faceset = _mesh(kind)
    c = Case(coord, inpoel, tri, faceset, [1, 2])
    U = _state(c.npoin)
    ctx = X.Context(0, flux=flux)
    ctx.upload_mesh(c.coord, c.se.sup, c.se.int, c.tri, c.besym)
    G = ctx.grad(U)
    Go, Ag = O.grad(c.se, c.coord, c.tri, U, absacc=True)
    _close(G, Go, Ag, "grad")
    Gv = Go / c.vol[:, None]
    Ro, Ar = O.rhs(O.cfg(flux=flux), c.se, c.coord, c.tri, c.besym, Gv, U, absacc=True)
    _close(ctx.rhs(Gv, U), Ro, Ar, "rhs")
    ctx.close()


@pytest.mark.parametrize("kind", ["one", "two"])
def test_kozak_rhs_and_kozcg_step(kind):
    import torch
    from xyst_amd import kozak as K
    from xyst_amd import riemann as X
    _say("small kozak %s" % kind)
    coord, inpoel, tri, faceset = _mesh(kind)
    c = Case(coord, inpoel, tri, faceset, [1])
    U = _state(c.npoin, seed=5)
    dt = 1e-3
    Ro, A = O.kozak_rhs(O.cfg(), coord, inpoel, dt, U, absacc=True)
    ctx = X.Context(0)
    K.upload(ctx, coord, inpoel)
    _close(K.rhs(ctx, dt, U), Ro, A, "kozak rhs")
    # one FCT step, a Dirichlet node, symmetry on face set 1
    dn = np.array([c.npoin - 1], np.int64)
    fct = dict(fct=True, fctdif=1.0, fctclip=False, fctsys=())
    K.setup(ctx, c.vol, c.symnodes, c.symnorms, dn, np.ones((1, 5), np.int32), U[dn], **fct)
    dU = torch.tensor(U, device="cuda")
    dUn = torch.empty_like(dU)
    K.step_dev(ctx, dt, dU, dUn)
    torch.cuda.synchronize()
    Uo, _ = O.kozcg_step(O.cfg(), O.fctcfg(True, 1.0, False, []), coord, inpoel, c.vol, dt,
                         c.symnodes, c.symnorms, dn, U[dn], U.copy())
    scale = np.abs(Uo).max(0)
    assert (np.abs(dU.cpu().numpy() - Uo) / scale).max() <= 1e-11
    assert np.array_equal(dUn.cpu().numpy(), U)
    ctx.close()


@pytest.mark.parametrize("kind", ["one", "two"])
def test_chorin_operators(kind):
    from xyst_amd import chorin as CH
    from xyst_amd import riemann as X
    _say("small chorin %s" % kind)
    coord, inpoel, tri, _ = _mesh(kind)
    e, d = O.chocg_domint(coord, inpoel)
    se = O.domsuped(inpoel, e, d)
    n = coord.shape[1]
    rng = np.random.default_rng(9)
    U = rng.standard_normal((n, 3))
    P = rng.standard_normal(n)
    G9 = rng.standard_normal((n, 9))
    Gp = rng.standard_normal((n, 3))
    ctx = X.Context(0)
    CH.upload(ctx, coord, se.sup, se.int, tri)
    Do, A = O.chorin_div(se, coord, tri, 0.01, P, Gp, U, stab=True, absacc=True)
    _close(CH.div(ctx, 0.01, P, Gp, U, np.zeros(n), stab=True), Do, A, "div")
    Go, A = O.chorin_grad(se, coord, tri, P, absacc=True)
    _close(CH.grad(ctx, P, np.zeros((n, 3))), Go, A, "grad")
    Gvo, A = O.chorin_vgrad(se, coord, tri, U, absacc=True)
    _close(CH.vgrad(ctx, U, np.zeros((n, 9))), Gvo, A, "vgrad")
    Fo, A = O.chorin_flux(se, coord, tri, U, G9, 0.02, absacc=True)
    _close(CH.flux(ctx, 0.02, U, G9, np.zeros((n, 3))), Fo, A, "flux")
    for flux in ("damp2", "damp4"):
        kw = dict(stab=True, stab2=True, stab2coef=0.3, mu=0.02)
        Ro, A = O.chorin_rhs(O.chocfg(flux=flux, **kw), se, coord, tri, P, U, G9, absacc=True)
        _close(CH.rhs(ctx, CH.config(flux, **kw), P, U, G9), Ro, A, "rhs " + flux)
    ctx.close()


@pytest.mark.parametrize("kind", ["one", "two"])
def test_cg_on_the_smallest_poisson_systems(kind):
    """ChoCG's pressure matrix of the mesh with node 0 Dirichlet: the device
    CG reaches the dense solution (at most n iterations in exact arithmetic)."""
    from xyst_amd import cg as X
    _say("small cg %s" % kind)
    coord, inpoel, _, _ = _mesh(kind)
    n = coord.shape[1]
    A = X.CSR.poisson(coord, inpoel)
    b = np.linspace(1.0, 2.0, n)
    cg = X.ConjugateGradients(A)
    cg.init(np.zeros(n), b, dirbc={0: [(1, 0.5)]}, apply=True, pc="jacobi")
    cg.solve(50, 1e-12)
    assert cg.converged and cg.it <= 2 * n
    D = A.dense() if hasattr(A, "dense") else None
    if D is None:
        D = np.zeros((n, n))
        for i in range(n):
            for k in range(A.ia[i], A.ia[i + 1]):
                D[i, A.ja[k]] = A.a[k]
    # the reference's Dirichlet treatment: b -= A[:, 0] * 0.5, row/column 0
    # removed, x[0] = 0.5
    bb = b - D[:, 0] * 0.5
    Dm = D.copy()
    Dm[0, :] = 0.0
    Dm[:, 0] = 0.0
    Dm[0, 0] = 1.0
    bb[0] = 0.5
    x = np.linalg.solve(Dm, bb)
    np.testing.assert_allclose(cg.solution(), x, rtol=1e-10, atol=1e-12)
    cg.ctx.close()
