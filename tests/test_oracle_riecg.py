"""Pins the CPU oracle of the RieCG path (CPU only).

The oracle restates Riemann.cpp / RieCG.cpp; these checks tie it to the
reference's own regression fixture and to what the mathematics fixes, so a
dropped term, a wrong sign or a transposed index fails one of them.
"""
import numpy as np
import pytest

import oracle as O
from tests.meshes import box_case, smooth_state, sod_fixture
from xyst_amd.meshgen import box_counts, sod_state


def _prim_fields(U, gamma=1.4):
    r = U[:, 0]
    u, v, w = U[:, 1] / r, U[:, 2] / r, U[:, 3] / r
    e = U[:, 4] / r
    p = (gamma - 1.0) * r * (e - 0.5 * (u * u + v * v + w * w))
    return np.stack([r, u, v, w, e, p], 1)


def test_sod_regression_fixture():
    """10 RK3 steps of RieCG/Sod against sod.std.exo and diag.std
    (tests/regression/inciter/RieCG/Sod), exodiff/numdiff bands."""
    c, d = sod_fixture()
    U = sod_state(c.coord)
    cfg = O.cfg(float(d["gamma"]))
    t, term = 0.0, float(d["term"])
    V = c.vol.sum()
    for it in range(int(d["nstep"])):
        dt = O.dt(cfg, U, c.vol, float(d["cfl"]))
        if t + dt > term:                 # Discretization::setdt truncation
            dt = term - t
        Un = O.riecg_step(cfg, c.se, c.coord, c.tri, c.besym, c.vol,
                          c.symnodes, c.symnorms, dt, U)
        t += dt
        dg = O.diag(U, Un, c.vol)
        row = d["diag"][it]
        mine = np.concatenate([[it + 1, t, dt], np.sqrt(dg[:10] / V), [dg[10]]])
        # diag.std prints 9 significant digits
        np.testing.assert_allclose(mine, row, rtol=2e-8, atol=1e-14)
    assert abs(t - float(d["t_final"])) < 1e-8          # exodiff TIME STEPS
    mine = _prim_fields(U)
    ref = d["fields"]
    for i, name in enumerate(d["field_names"]):
        diff = np.abs(mine[:, i] - ref[:, i])
        if name in ("xvelocity", "yvelocity", "zvelocity"):
            ok = (diff <= 1e-8) | (np.abs(ref[:, i]) < 1e-7)     # absolute 1e-8 floor 1e-7
        else:
            ok = (diff <= 1e-7 * np.abs(ref[:, i])) | (np.abs(ref[:, i]) < 1e-8)
        assert ok.all(), (name, diff.max())


def _basis_gradients(X):
    """Gradients of the 4 linear basis functions of a tet (brute force)."""
    A = np.ones((4, 4))
    A[:, 1:] = X
    return np.linalg.inv(A)[1:, :].T        # row a = grad N_a


def test_domint_against_basis_integrals():
    """D_pq = (V/8)(grad N_p - grad N_q) summed over tets sharing the edge:
    the dual-face normal (1/2) int (N_q grad N_p - N_p grad N_q)."""
    c = box_case(2, seed=3)
    ref = {}
    for t in c.inpoel:
        X = c.coord[:, t].T
        Gn = _basis_gradients(X)
        V = abs(np.linalg.det(np.c_[X[1:] - X[0]])) / 6.0
        for a in range(4):
            for b in range(a + 1, 4):
                p, q = t[a], t[b]
                d = V / 8.0 * (Gn[a] - Gn[b])
                if p > q:
                    p, q, d = q, p, -d
                ref[(p, q)] = ref.get((p, q), 0.0) + d
    assert len(ref) == c.edge.shape[0] == box_counts(2)[2]
    for (p, q), dint in zip(c.edge, c.dint):
        np.testing.assert_allclose(dint, ref[(p, q)], rtol=1e-12, atol=1e-15)


def test_domint_closure():
    """sum_q D_pq vanishes at interior nodes and equals sum of cross/12 of the
    boundary faces at boundary nodes."""
    c = box_case(3, seed=5)
    S = np.zeros((c.npoin, 3))
    np.add.at(S, c.edge[:, 0], c.dint)
    np.add.at(S, c.edge[:, 1], -c.dint)
    B = np.zeros((c.npoin, 3))
    a, b, cc = (c.coord[:, c.tri[:, i]] for i in range(3))
    n = np.cross((b - a).T, (cc - a).T) / 12.0
    for i in range(3):
        np.add.at(B, c.tri[:, i], n)
    np.testing.assert_allclose(S, B, atol=1e-15)


def test_domsuped_partition_of_edges():
    c = box_case(3, seed=1)
    se = c.se
    seen = {}
    s0, s1, s2 = se.sup[0].reshape(-1, 4), se.sup[1].reshape(-1, 3), se.sup[2].reshape(-1, 2)
    i0, i1, i2 = se.int[0].reshape(-1, 6, 3), se.int[1].reshape(-1, 3, 3), se.int[2].reshape(-1, 3)
    lp = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    for N, D in zip(s0, i0):
        for k, (a, b) in enumerate(lp):
            seen[(N[a], N[b])] = D[k]
    for N, D in zip(s1, i1):
        for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
            seen[(N[a], N[b])] = D[k]
    for N, D in zip(s2, i2):
        seen[(N[0], N[1])] = D
    assert sum(se.nsup * [6, 3, 1]) == c.edge.shape[0] == len(seen)
    ref = {(p, q): d for (p, q), d in zip(c.edge, c.dint)}
    for (p, q), d in seen.items():
        if p < q:
            np.testing.assert_array_equal(d, ref[(p, q)])
        else:
            np.testing.assert_array_equal(d, -ref[(q, p)])
    assert se.nsup[0] > 0 and se.nsup[1] > 0


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_free_stream(flux):
    """A uniform state gives a zero rhs at every node (domain edges cancel
    against the boundary integral)."""
    c = box_case(3, seed=2)
    U = np.zeros((c.npoin, 5))
    U[:, 0] = 1.2
    U[:, 1:4] = 1.2 * np.array([0.3, -0.2, 0.1])
    U[:, 4] = 2.0 / 0.4 + 0.5 * 1.2 * (0.09 + 0.04 + 0.01)
    G = O.grad(c.se, c.coord, c.tri, U)
    assert np.abs(G).max() < 1e-13
    besym = np.zeros_like(c.besym)
    R = O.rhs(O.cfg(flux=flux), c.se, c.coord, c.tri, besym, G, U)
    assert np.abs(R).max() < 1e-13


def test_grad_linear_exact_interior():
    c = box_case(4, seed=9)
    U = np.zeros((c.npoin, 5))
    x, y, z = c.coord
    U[:, 0] = 1.0 + 0.1 * x
    U[:, 1] = U[:, 0] * (2 * x - y + 0.5 * z)
    U[:, 4] = 3.0
    G = O.grad(c.se, c.coord, c.tri, U) / c.vol[:, None]
    interior = np.ones(c.npoin, bool)
    interior[c.tri.reshape(-1)] = False
    np.testing.assert_allclose(G[interior, 0:3] - [0.1, 0, 0], 0.0, atol=1e-12)
    np.testing.assert_allclose(G[interior, 3:6] - [2, -1, 0.5], 0.0, atol=1e-11)


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_conservation_all_symmetry(flux):
    """With every boundary node a symmetry node, mass and energy totals of R
    vanish (antisymmetric edge scatter, zero normal velocity on walls)."""
    c = box_case(4, seed=4)
    U = smooth_state(c.coord)
    G = O.grad(c.se, c.coord, c.tri, U) / c.vol[:, None]
    R = O.rhs(O.cfg(flux=flux), c.se, c.coord, c.tri, c.besym, G, U)
    scale = np.abs(R).sum(0)
    assert abs(R[:, 0].sum()) < 1e-13 * scale[0]
    assert abs(R[:, 4].sum()) < 1e-13 * scale[4]


def test_dt_brute_force():
    c = box_case(3)
    U = smooth_state(c.coord)
    r = U[:, 0]
    vel = np.linalg.norm(U[:, 1:4] / r[:, None], axis=1)
    p = 0.4 * (U[:, 4] - 0.5 * r * vel ** 2)
    cs = np.sqrt(1.4 * np.maximum(p, 0) / r)
    ref = 0.5 * np.min(np.cbrt(c.vol) / np.maximum(vel + cs, 1e-8))
    assert O.dt(O.cfg(), U, c.vol, 0.5) == pytest.approx(ref, rel=1e-14)


def test_flux_not_configured():
    c = box_case(2)
    U = smooth_state(c.coord)
    G = O.grad(c.se, c.coord, c.tri, U)
    with pytest.raises(ValueError, match="Flux not configured"):
        O.rhs(O.cfg(flux="roe"), c.se, c.coord, c.tri, c.besym, G, U)


def test_threaded_rhs_matches_serial():
    """The multi-threaded CPU baseline (per-partition sums added per node in
    partition order) equals the serial rhs to round-off, bit for bit with one
    thread, and repeats bit for bit at a fixed thread count."""
    c = box_case(10, seed=4)
    U = smooth_state(c.coord)
    G = O.grad(c.se, c.coord, c.tri, U) / c.vol[:, None]
    cfg = O.cfg()
    R1, A = O.rhs(cfg, c.se, c.coord, c.tri, c.besym, G, U, absacc=True)
    assert np.array_equal(O.rhs_mt(cfg, c.se, c.coord, c.tri, c.besym, G, U, 1), R1)
    for T in (2, 3, 8):
        Rt = O.rhs_mt(cfg, c.se, c.coord, c.tri, c.besym, G, U, T)
        assert np.all(np.abs(Rt - R1) <= 1e-15 * A + 1e-300), T
        assert np.array_equal(Rt, O.rhs_mt(cfg, c.se, c.coord, c.tri, c.besym, G, U, T))
