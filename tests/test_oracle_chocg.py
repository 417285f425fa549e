"""Pins the CPU oracle of the ChoCG edge operators (Chorin.cpp, ChoCG.cpp).

What fixes them, independent of the oracle's loops:
* ChoCG::domint's 5th value is the P1 stiffness K_pq: it must equal -A(p,q)
  of ChoCG::prelhs (itself pinned by the exact P1 energy identity), and the
  4th sums to 0.3 x the domain volume (6 edges x J/120 per tet).
* The edge-based operators are exact for linear data: grad/vgrad of a linear
  field is V * gradient, div of a linear vector field is V * divergence, the
  viscous momentum flux of a linear gradient field is V * its divergence, and
  the pressure part of the advective rhs is V * grad P, at every node (the
  boundary integral closes the dual cells).
* The 4th-order damping term of div vanishes for a linear pressure with its
  exact gradient (checks its sign and the factor 1/2).
* damp4's MUSCL reconstruction returns the edge midpoint for linear data, so
  the upwind term is inactive: stab on and off agree to round-off.
"""
import numpy as np
import pytest

import oracle as O
from xyst_amd.meshgen import box_mesh


def _case(n=4, seed=3):
    m = box_mesh(n, seed=seed)
    coord, inpoel, tri = m["coord"], m["inpoel"], m["triinpoel"]
    e, d = O.chocg_domint(coord, inpoel)
    se = O.domsuped(inpoel, e, d)
    vol = O.vol(coord, inpoel)
    return coord, inpoel, tri, e, d, se, vol


def test_domint5_is_riecg_normals_plus_mass_and_stiffness():
    coord, inpoel, tri, e, d, se, vol = _case()
    e3, d3 = O.domint(coord, inpoel)
    assert np.array_equal(e, e3)
    assert np.array_equal(d[:, :3], d3)
    assert d[:, 3].sum() == pytest.approx(0.3 * vol.sum(), rel=1e-13)
    A = O.CSR.from_tets(coord.shape[1], inpoel)
    O.prelhs(coord, inpoel, A)
    M = A.dense()
    np.testing.assert_allclose(d[:, 4], -M[e[:, 0], e[:, 1]], rtol=1e-12,
                               atol=1e-15 * np.abs(M).max())
    # every edge is in exactly one superedge group
    assert se.nsup[0] * 6 + se.nsup[1] * 3 + se.nsup[2] == e.shape[0]


def test_grad_vgrad_exact_for_linear_fields():
    coord, inpoel, tri, e, d, se, vol = _case()
    x, y, z = coord
    a = np.array([[0.7, -1.3, 2.1], [1.5, 0.2, -0.4], [-0.3, 0.9, 1.1]])
    U = np.stack([a[i, 0] * x + a[i, 1] * y + a[i, 2] * z + i for i in range(3)], 1)
    G = O.chorin_vgrad(se, coord, tri, U)
    for i in range(3):
        np.testing.assert_allclose(G[:, 3 * i:3 * i + 3] / vol[:, None],
                                   np.tile(a[i], (len(x), 1)), atol=1e-12)
    Gs = O.chorin_grad(se, coord, tri, U[:, 1])
    np.testing.assert_allclose(Gs / vol[:, None], np.tile(a[1], (len(x), 1)), atol=1e-12)


def test_div_exact_for_linear_field_and_stab_vanishes_for_linear_pressure():
    coord, inpoel, tri, e, d, se, vol = _case()
    x, y, z = coord
    U = np.stack([2.0 * x - y, 0.5 * y + z, -1.5 * z + x], 1)
    P = 3.0 * x - 2.0 * y + 0.25 * z
    Gp = np.tile([3.0, -2.0, 0.25], (len(x), 1))
    D = O.chorin_div(se, coord, tri, 0.01, P, Gp, U, stab=False)
    np.testing.assert_allclose(D / vol, 2.0 + 0.5 - 1.5, atol=1e-12)
    Ds = O.chorin_div(se, coord, tri, 0.01, P, Gp, U, stab=True)
    np.testing.assert_allclose(Ds, D, atol=1e-14)
    # a non-linear pressure does switch the damping on
    Ds2 = O.chorin_div(se, coord, tri, 0.01, P + x * x, Gp, U, stab=True)
    assert np.abs(Ds2 - D).max() > 1e-6


def test_flux_viscous_part_exact_for_linear_gradients():
    coord, inpoel, tri, e, d, se, vol = _case()
    x, y, z = coord
    n = len(x)
    rng = np.random.default_rng(5)
    c0 = rng.standard_normal(9)
    c1 = rng.standard_normal((9, 3))
    G = c0[None, :] + (coord.T @ c1.T)            # G_k(x) = c0_k + c1_k . x
    mu = 0.37
    F = O.chorin_flux(se, coord, tri, np.zeros((n, 3)), G, mu)
    # F_i = d_j [ mu (G_ij + G_ji - 2/3 d_ij tr G) ] times V
    exact = np.zeros(3)
    for i in range(3):
        for j in range(3):
            exact[i] += mu * (c1[i * 3 + j, j] + c1[j * 3 + i, j])
        exact[i] -= mu * 2.0 / 3.0 * (c1[0, i] + c1[4, i] + c1[8, i])
    np.testing.assert_allclose(F / vol[:, None], np.tile(exact, (n, 1)), atol=1e-11)
    # constant velocity, mu = 0: the inviscid flux closes at interior nodes
    U = np.tile([0.3, -0.2, 0.5], (n, 1))
    F = O.chorin_flux(se, coord, tri, U, np.zeros((n, 9)), 0.0)
    inner = (x > 1e-9) & (x < 1 - 1e-9) & (y > 1e-9) & (y < 1 - 1e-9) & \
            (z > 1e-9) & (z < 1 - 1e-9)
    assert np.abs(F[inner]).max() < 1e-14
    # with the boundary term: -div(u u) = 0 everywhere for a constant u
    assert np.abs(F).max() < 1e-14


@pytest.mark.parametrize("flux", ["damp2", "damp4"])
def test_rhs_pressure_term_exact_for_linear_pressure(flux):
    coord, inpoel, tri, e, d, se, vol = _case()
    x, y, z = coord
    n = len(x)
    P = 1.0 + 3.0 * x - 2.0 * y + 0.25 * z
    U = np.zeros((n, 3))
    R = O.chorin_rhs(O.chocfg(flux=flux), se, coord, tri, P, U, np.zeros((n, 9)))
    np.testing.assert_allclose(R / vol[:, None], np.tile([3.0, -2.0, 0.25], (n, 1)),
                               atol=1e-12)


def test_damp4_reconstruction_is_midpoint_for_linear_data():
    coord, inpoel, tri, e, d, se, vol = _case()
    x, y, z = coord
    n = len(x)
    a = np.array([[0.7, -1.3, 2.1], [1.5, 0.2, -0.4], [-0.3, 0.9, 1.1]])
    U = np.stack([a[i, 0] * x + a[i, 1] * y + a[i, 2] * z + 2 + i for i in range(3)], 1)
    G = np.tile(a.reshape(-1), (n, 1))             # exact velocity gradients
    P = np.zeros(n)
    R1, A1 = O.chorin_rhs(O.chocfg(flux="damp4", stab=True), se, coord, tri, P, U, G,
                          absacc=True)
    R0 = O.chorin_rhs(O.chocfg(flux="damp4", stab=False), se, coord, tri, P, U, G)
    assert np.all(np.abs(R1 - R0) <= 1e-12 * A1)
    # damp2 has no reconstruction: its upwind term is active for the same data
    R2 = O.chorin_rhs(O.chocfg(flux="damp2", stab=True), se, coord, tri, P, U, G)
    R2o = O.chorin_rhs(O.chocfg(flux="damp2", stab=False), se, coord, tri, P, U, G)
    assert np.abs(R2 - R2o).max() > 1e-6


def test_rhs_flux_not_configured():
    coord, inpoel, tri, e, d, se, vol = _case(2)
    n = coord.shape[1]
    with pytest.raises(ValueError):
        O.chorin_rhs(O.chocfg(flux="bogus"), se, coord, tri, np.zeros(n),
                     np.zeros((n, 3)), np.zeros((n, 9)))


def lid_run():
    """The reference's ChoCG/Lid case on the oracle (ChoCG.cpp control flow,
    chorin operators, ConjugateGradients)."""
    import os
    from oracle.chocg_step import ChoCGRun
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "chocg_lid.npz"))
    r = ChoCGRun(d["coord"], d["inpoel"], d["triinpoel"], d["faceset"], float(d["mu"]),
                 flux=str(d["flux"]), cfl=float(d["cfl"]), rk=int(d["rk"]),
                 noslip=list(d["noslip"]),
                 dirbc={int(d["dirset"]): (list(d["dirmask"]), list(d["dirval"]))},
                 p_hydrostat=int(d["p_hydrostat"]), p_tol=float(d["p_tol"]),
                 p_iter=int(d["p_iter"]))
    return r, d


def test_lid_regression_fixture():
    """10 ChoCG steps of the lid-driven cavity (tests/regression/inciter/
    ChoCG/Lid: damp4 flux with stab, viscosity 0.01, noslip walls, lid
    velocity, Jacobi-PCG pressure projection to 1e-3) against diag_ldc.std
    (numdiff 1e-7) and the fields of ldc.std.exo (exodiff bands)."""
    r, d = lid_run()
    r.initial()
    for it in range(int(d["nstep"])):
        dp = r.step()
        np.testing.assert_allclose(r.diag(dp), d["diag"][it], rtol=1e-8, atol=1e-14)
    assert abs(r.t - float(d["t_final"])) < 1e-8
    mine = {"velocityx": r.U[:, 0], "velocityy": r.U[:, 1], "velocityz": r.U[:, 2],
            "divergence": r.div / r.vol, "pressure": r.P}
    rel = {"velocityx": 3e-7, "velocityy": 2e-6, "velocityz": 3e-7}
    for i, name in enumerate(d["field_names"]):
        ref = d["fields"][:, i]
        diff = np.abs(mine[str(name)] - ref)
        ok = (diff <= rel.get(str(name), 1e-7) * np.abs(ref)) | (np.abs(ref) < 1e-8)
        assert ok.all(), (name, diff.max())


def poiseuille_run(case):
    import os
    from oracle.chocg_step import ChoCGRun
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "chocg_poiseuille.npz"))
    r = ChoCGRun(d["coord"], d["inpoel"], d["triinpoel"], d["faceset"], float(d["mu"]),
                 flux=str(d[case + "_flux"]), cfl=float(d[case + "_cfl"]),
                 rk=int(d[case + "_rk"]), noslip=list(d["noslip"]),
                 dirbc={int(s): (list(d["dirmask"]), [0.0, 0.0, 0.0]) for s in d["dirsets"]},
                 p_dirbc={int(s): (2, float(v)) for s, v in zip(d["p_dirsets"], d["p_dirvals"])},
                 p_tol=float(d["p_tol"]), p_iter=int(d["p_iter"]),
                 theta=float(d[case + "_theta"]))
    return r, d


@pytest.mark.parametrize("case", ["damp2", "damp4", "rk2", "rk3", "rk4", "theta"])
def test_poiseuille_regression_fixtures(case):
    """20 ChoCG steps of the Poiseuille channel (tests/regression/inciter/
    ChoCG/Poiseuille, one PE): damp2 and damp4 fluxes, RK 1-4, a pressure
    drop set by Dirichlet values (2.4 and 0) in the start-up projections,
    mask-1 velocity BCs and noslip walls, and (case theta) the theta = 0.5
    implicit momentum solve with its 3-component CG; diag_poiseuille_*.std
    with the reference's numdiff rule (abs 1e-7 or rel 1e-7) and the exodiff
    bands of velocity and pressure (not for theta, whose fields the
    reference keeps in two partition files)."""
    r, d = poiseuille_run(case)
    r.initial()
    for it in range(int(d["nstep"])):
        dp = r.step()
        row, ref = r.diag(dp), d[case + "_diag"][it]
        diff = np.abs(row - ref)
        assert ((diff <= 1e-7) | (diff <= 1e-7 * np.abs(ref))).all(), (it, diff.max())
    if case == "theta":
        assert r.mit and all(0 < m <= 50 for m in r.mit)
        return
    assert abs(r.t - float(d[case + "_t_final"])) < 1e-8
    mine = {"velocityx": r.U[:, 0], "velocityy": r.U[:, 1], "velocityz": r.U[:, 2],
            "pressure": r.P}
    for i, name in enumerate(d["field_names"]):
        if str(name) not in mine:
            continue                 # divergence is not compared (exodiff.cfg)
        ref = d[case + "_fields"][:, i]
        diff = np.abs(mine[str(name)] - ref)
        assert ((diff <= 1e-7 * np.abs(ref)) | (np.abs(ref) < 1e-8)).all(), (name, diff.max())


@pytest.mark.parametrize("flux", ["damp2", "damp4"])
def test_stab2_weights_from_the_operator_response(flux):
    """chorin's stab2 artificial viscosity (Chorin.cpp:682-687, 797-802) read
    back from the operator: with u_new = u_n - dt R/vol (ChoCG.cpp:1561) the
    term is a graph Laplacian with weights stab2coef*max(|vn_p|, |vn_q|).  A
    transported scalar probes it exactly (its flux is linear in the scalar
    for a fixed velocity): for s = delta_j, R(stab2) - R(no stab2) is
    -w_jq at each neighbour q and +sum_q w_jq at j.  On the velocity
    components the added term conserves (sums to zero) and dissipates
    (u . dR >= 0).  A min for the max, a lost abs, a wrong sign or a swapped
    end fails one of these.  damp4 is probed with zero gradients G: its van
    Leer reconstruction then returns the end values exactly (the limiter is
    0 for an edge difference above muscl_eps and the correction 0 for none),
    so its response is that of damp2 through its own copy of the weight."""
    coord, inpoel, tri, edge, dint, se, _ = _case(4, seed=12)
    n = coord.shape[1]
    rng = np.random.default_rng(5)
    x, y, z = coord
    U = np.zeros((n, 4))
    U[:, 0] = np.sin(2 * x + 1) + 0.3 * rng.standard_normal(n)
    U[:, 1] = np.cos(3 * y) * x + 0.3 * rng.standard_normal(n)
    U[:, 2] = 0.5 - z * y + 0.3 * rng.standard_normal(n)
    P = rng.standard_normal(n)
    G = np.zeros((n, 12))
    coef = 0.7
    on = O.chocfg(flux, stab=False, stab2=True, stab2coef=coef)
    off = O.chocfg(flux, stab=False, stab2=False)
    vn = lambda e, i: abs(U[i, :3] @ dint[e, :3])
    w = {}
    for e, (p, q) in enumerate(edge):
        w[(p, q)] = w[(q, p)] = coef * max(vn(e, p), vn(e, q))
    nbrs = {}
    for p, q in edge:
        nbrs.setdefault(p, []).append(q)
        nbrs.setdefault(q, []).append(p)
    for j in rng.choice(n, 8, replace=False):
        Us = U.copy()
        Us[:, 3] = 0.0
        Us[j, 3] = 1.0
        dR = (O.chorin_rhs(on, se, coord, tri, P, Us, G)
              - O.chorin_rhs(off, se, coord, tri, P, Us, G))[:, 3]
        want = np.zeros(n)
        for q in nbrs[j]:
            want[q] = -w[(j, q)]
        want[j] = sum(w[(j, q)] for q in nbrs[j])
        np.testing.assert_allclose(dR, want, rtol=1e-12, atol=1e-15)
    dR = (O.chorin_rhs(on, se, coord, tri, P, U, G)
          - O.chorin_rhs(off, se, coord, tri, P, U, G))[:, :3]
    assert np.abs(dR.sum(0)).max() < 1e-13 * np.abs(dR).sum()
    diss = (U[:, :3] * dR).sum()
    want = sum(w[(p, q)] * ((U[q, :3] - U[p, :3]) ** 2).sum() for p, q in edge)
    assert diss > 0 and abs(diss - want) < 1e-12 * want
