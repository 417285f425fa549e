"""Parity of the HIP chorin:: operators with the CPU oracle (needs an MI355X).

Tolerance: FP64, per node and component |x_gpu - x_oracle| <= 1e-12 * A, A
being the oracle's sum of the magnitudes of the products each added term is
made of (the device sums the same terms in another order and replaces
divisions by Newton-refined reciprocals).
"""
import sys

import numpy as np
import pytest

import oracle as O
from xyst_amd.meshgen import box_mesh, renumber_mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _close(x, ref, A, what):
    err = np.abs(x - ref)
    bad = err > TOL * A + 1e-300
    assert not bad.any(), "%s: %d entries off, worst ratio %.3e" % (
        what, bad.sum(), float(np.max(err / np.maximum(A, 1e-300))))


def _case(n, order="lex", seed=4):
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    m = box_mesh(n, seed=seed)
    if order != "lex":
        m = renumber_mesh(m, R.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
        m = renumber_mesh(m, R.mesh_block_order(m["coord"], 2048))
    coord, inpoel, tri = m["coord"], m["inpoel"], m["triinpoel"]
    e, d = O.chocg_domint(coord, inpoel)
    se = O.domsuped(inpoel, e, d)
    ctx = R.Context(0)
    X.upload(ctx, coord, se.sup, se.int, tri)
    rng = np.random.default_rng(seed)
    npoin = coord.shape[1]
    x, y, z = coord
    U = np.stack([np.sin(2 * x + y) + 0.1 * rng.standard_normal(npoin),
                  np.cos(x - 3 * z) + 0.1 * rng.standard_normal(npoin),
                  x * y - z + 0.1 * rng.standard_normal(npoin)], 1)
    P = np.cos(3 * x) * y + 0.05 * rng.standard_normal(npoin)
    G9 = rng.standard_normal((npoin, 9))
    Gp = rng.standard_normal((npoin, 3))
    return ctx, se, coord, tri, U, P, G9, Gp


@pytest.mark.parametrize("n,order", [(5, "lex"), (12, "rcm+block")])
def test_chorin_div_grad_vgrad_flux(n, order):
    from xyst_amd import chorin as X
    _say("chorin ops n=%d %s" % (n, order))
    ctx, se, coord, tri, U, P, G9, Gp = _case(n, order)
    npoin = U.shape[0]
    for stab in (False, True):
        Do, A = O.chorin_div(se, coord, tri, 0.013, P, Gp, U, stab=stab, absacc=True)
        D = X.div(ctx, 0.013, P, Gp, U, np.zeros(npoin), stab=stab)
        _close(D, Do, A, "div stab=%s" % stab)
    Go, A = O.chorin_grad(se, coord, tri, P, absacc=True)
    _close(X.grad(ctx, P, np.zeros((npoin, 3))), Go, A, "grad")
    Gvo, A = O.chorin_vgrad(se, coord, tri, U, absacc=True)
    _close(X.vgrad(ctx, U, np.zeros((npoin, 9))), Gvo, A, "vgrad")
    for mu in (0.0, 0.02):
        Fo, A = O.chorin_flux(se, coord, tri, U, G9, mu, absacc=True)
        _close(X.flux(ctx, mu, U, G9, np.zeros((npoin, 3))), Fo, A, "flux mu=%g" % mu)
    ctx.close()


def test_chorin_ops_accumulate_and_pressure_gradient_stride():
    """div/grad add to what the output holds; grad into a 4-column field."""
    from xyst_amd import chorin as X
    _say("chorin accumulate")
    ctx, se, coord, tri, U, P, G9, Gp = _case(6)
    npoin = U.shape[0]
    D0 = np.linspace(-1, 1, npoin)
    Do, A = O.chorin_div(se, coord, tri, 0.0, P, Gp, U, D=D0.copy(), absacc=True)
    _close(X.div(ctx, 0.0, P, Gp, U, D0.copy()), Do, A + np.abs(D0), "div accumulate")
    G4 = np.ones((npoin, 4))
    Go, A = O.chorin_grad(se, coord, tri, P, absacc=True)
    G4 = X.grad(ctx, P, G4)
    _close(G4[:, :3], Go + 1.0, A + 1.0, "grad ldg=4")
    assert np.all(G4[:, 3] == 1.0)
    ctx.close()


@pytest.mark.parametrize("flux", ["damp2", "damp4"])
@pytest.mark.parametrize("stab2,mu", [(False, 0.0), (True, 0.05)])
def test_chorin_rhs(flux, stab2, mu):
    from xyst_amd import chorin as X
    _say("chorin rhs %s stab2=%s mu=%g" % (flux, stab2, mu))
    ctx, se, coord, tri, U, P, G9, Gp = _case(10, "rcm+block", seed=7)
    Ro, A = O.chorin_rhs(O.chocfg(flux=flux, stab=True, stab2=stab2, stab2coef=0.3, mu=mu),
                         se, coord, tri, P, U, G9, absacc=True)
    R = X.rhs(ctx, X.config(flux, stab=True, stab2=stab2, stab2coef=0.3, mu=mu), P, U, G9)
    _close(R, Ro, A, "rhs")
    # bitwise reproducible
    R2 = X.rhs(ctx, X.config(flux, stab=True, stab2=stab2, stab2coef=0.3, mu=mu), P, U, G9)
    assert np.array_equal(R, R2)
    ctx.close()


def test_chorin_device_pointers_match_host_entry():
    import torch
    from xyst_amd import chorin as X
    _say("chorin device entry")
    ctx, se, coord, tri, U, P, G9, Gp = _case(8, "rcm+block", seed=2)
    cfg = X.config("damp4", stab=True, mu=0.01)
    R = X.rhs(ctx, cfg, P, U, G9)
    dR = torch.full((U.shape[0], 3), float("nan"), dtype=torch.float64, device="cuda")
    X.rhs(ctx, cfg, torch.from_numpy(P).cuda(), torch.from_numpy(U).cuda(),
          torch.from_numpy(G9).cuda(), dR)
    D = X.div(ctx, 0.01, P, Gp, U, np.zeros(U.shape[0]), stab=True)
    dD = torch.zeros(U.shape[0], dtype=torch.float64, device="cuda")
    X.div(ctx, 0.01, torch.from_numpy(P).cuda(), torch.from_numpy(Gp).cuda(),
          torch.from_numpy(U).cuda(), dD, stab=True)
    torch.cuda.synchronize()
    assert np.array_equal(dR.cpu().numpy(), R)
    assert np.array_equal(dD.cpu().numpy(), D)
    ctx.close()


# ---------------------------------------------------------------- ChoCG step --

def _device_run(r, cfgkw, problem=None):
    """The device ChoCG on the mesh and BCs of an oracle ChoCGRun r (its
    component count; problem: the source's problem name)."""
    import torch
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    from xyst_amd.cg import CSR, ConjugateGradients
    ctx = R.Context(0)
    X.upload(ctx, r.coord, r.se.sup, r.se.int, r.tri)
    if r.ncomp != 3:
        X.set_ncomp(ctx, r.ncomp)
    if problem is not None:
        ctx.set_problem(problem)
    cg = ConjugateGradients(CSR.poisson(r.coord, r.inpoel), ctx=ctx)
    nc = r.ncomp
    dn = np.array(sorted(r.dir), np.int64)
    mk = np.array([r.dir[p][0] for p in dn], np.int32).reshape(-1, nc)
    cv = np.array([r.dir[p][1] for p in dn]).reshape(-1, nc)
    # the values physics::dirbc sets: mask 1 the initial value, mask 2 the
    # configured one
    dv = np.where(mk == 2, cv, r.U0[dn].reshape(-1, nc))
    dm = (mk != 0).astype(np.int32)
    pd = sorted(set(r.pdir) | ({r.hydrostat} if r.hydrostat is not None else set()))
    X.step_setup(ctx, r.vol, X.config(**cfgkw), cfl=r.cfl, rk=len(r.rkcoef),
                 p_pc=r.p_pc, p_maxit=r.p_iter, p_tol=r.p_tol,
                 symnodes=r.symnodes, symnorms=r.symnorms, noslip=r.noslip,
                 dirnodes=dn, dirmask=dm, dirvals=dv, pdirnodes=np.array(pd, np.int64),
                 pdirvals=np.array([r.pdir.get(p, 0.0) for p in pd]))
    U = torch.tensor(r.U, device="cuda")
    P = torch.tensor(r.P, device="cuda")
    return ctx, cg, U, P


def _lid_oracle():
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


def test_chocg_lid_regression_fixture():
    """The reference's ChoCG/Lid case on the device: start-up projections and
    10 steps, diagnostics against diag_ldc.std (numdiff 1e-7) and the fields
    of ldc.std.exo (exodiff bands); pressure iterations equal the oracle's."""
    import torch
    from xyst_amd import chorin as X
    _say("chocg lid fixture")
    r, d = _lid_oracle()
    ctx, cg, U, P = _device_run(r, dict(flux="damp4", stab=True, mu=float(d["mu"])))
    pit = X.initial_dev(ctx, U, P)
    r.initial()
    assert list(pit) == r.pit[:2], (pit, r.pit)
    from xyst_amd import diagnostics as D
    Un = torch.empty_like(U)
    dP = torch.empty_like(P)
    V = r.vol
    dV = torch.tensor(V, device="cuda")
    t = 0.0
    for it in range(int(d["nstep"])):
        dt = X.dt_dev(ctx, U)
        assert dt == pytest.approx(r.dt_cfl(), rel=1e-14)
        n_it = X.step_dev(ctx, dt, U, Un, P, dP)
        r.step(dt)
        assert n_it == r.pit[-1]
        torch.cuda.synchronize()
        u, un, p, dp = (a.cpu().numpy() for a in (U, Un, P, dP))
        t += dt
        s = [np.sum(p * p * V)] + [np.sum(u[:, c] ** 2 * V) for c in range(3)]
        s += [np.sum(dp * dp * V)] + [np.sum((u[:, c] - un[:, c]) ** 2 * V) for c in range(3)]
        row = np.concatenate([[it + 1, t, dt], np.sqrt(np.array(s) / V.sum())])
        np.testing.assert_allclose(row, d["diag"][it], rtol=1e-7, atol=1e-14)
        np.testing.assert_allclose(D.precompute(ctx, U, Un, P, dP, dV), s, rtol=1e-13)
        err = float(np.max(np.abs(u - r.U)) / np.max(np.abs(r.U)))
        _say("  step %d pit %d |dU|/max|U| vs oracle %.3e" % (it + 1, n_it, err))
        assert err < 1e-9
    mine = {"velocityx": u[:, 0], "velocityy": u[:, 1], "velocityz": u[:, 2],
            "pressure": p}
    rel = {"velocityx": 3e-7, "velocityy": 2e-6, "velocityz": 3e-7}
    for i, name in enumerate(d["field_names"]):
        if str(name) not in mine:
            continue
        ref = d["fields"][:, i]
        diff = np.abs(mine[str(name)] - ref)
        ok = (diff <= rel.get(str(name), 1e-7) * np.abs(ref)) | (np.abs(ref) < 1e-8)
        assert ok.all(), (name, diff.max())
    ctx.close()


@pytest.mark.parametrize("flux,rk", [("damp2", 2), ("damp4", 3)])
def test_chocg_step_parity_box(flux, rk):
    """A 12^3 box (RCM + block order), lid-driven cavity BCs with a
    symmetry wall, damp2/damp4, RK2/RK3: three steps against the oracle's
    ChoCG, fields within 1e-9 of their max (the pressure solves stop at
    1e-6, so round-off moves the iterates only at that order times the
    recurrence's sensitivity)."""
    import torch
    from oracle.chocg_step import ChoCGRun
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    _say("chocg box %s rk%d" % (flux, rk))
    m = box_mesh(12, seed=9)
    m = renumber_mesh(m, R.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    m = renumber_mesh(m, R.mesh_block_order(m["coord"], 2048))
    r = ChoCGRun(m["coord"], m["inpoel"], m["triinpoel"], m["faceset"], 0.01, flux=flux,
                 cfl=0.5, rk=rk, noslip=[1, 2, 3, 4], sym=[5],
                 dirbc={6: ([2, 2, 2], [1.0, 0.0, 0.0])}, p_hydrostat=0, p_tol=1e-6,
                 p_iter=1000)
    ctx, cg, U, P = _device_run(r, dict(flux=flux, stab=True, mu=0.01))
    pit = X.initial_dev(ctx, U, P)
    r.initial()
    assert list(pit) == r.pit[:2]
    Un = torch.empty_like(U)
    dP = torch.empty_like(P)
    for it in range(3):
        dt = X.dt_dev(ctx, U)
        n_it = X.step_dev(ctx, dt, U, Un, P, dP)
        r.step(dt)
        torch.cuda.synchronize()
        u, p = U.cpu().numpy(), P.cpu().numpy()
        eu = float(np.max(np.abs(u - r.U)) / np.max(np.abs(r.U)))
        ep = float(np.max(np.abs(p - r.P)) / np.max(np.abs(r.P)))
        _say("  step %d pit %d/%d |dU| %.3e |dP| %.3e" % (it + 1, n_it, r.pit[-1], eu, ep))
        assert abs(n_it - r.pit[-1]) <= 1
        assert eu < 1e-9 and ep < 1e-9
    ctx.close()


def test_chocg_step_with_multigrid_pressure_solve():
    """The pressure solve by FCG + AMG (p_pc "amg", no reference counterpart)
    gives the oracle's (Jacobi-PCG) fields when both solve to 1e-10: the two
    solutions of the projection differ by the solve tolerance times the
    system's conditioning only."""
    import torch
    from oracle.chocg_step import ChoCGRun
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    from xyst_amd.cg import CSR, ConjugateGradients
    _say("chocg amg pressure")
    m = box_mesh(10, seed=2)
    r = ChoCGRun(m["coord"], m["inpoel"], m["triinpoel"], m["faceset"], 0.01,
                 flux="damp4", cfl=0.5, rk=1, noslip=[1, 2, 3, 4, 5],
                 dirbc={6: ([2, 2, 2], [1.0, 0.0, 0.0])}, p_hydrostat=0, p_tol=1e-10,
                 p_iter=5000)
    ctx = R.Context(0)
    X.upload(ctx, r.coord, r.se.sup, r.se.int, r.tri)
    ConjugateGradients(CSR.poisson(r.coord, r.inpoel), ctx=ctx)
    dn = np.array(sorted(r.dir), np.int64)
    X.step_setup(ctx, r.vol, X.config("damp4", stab=True, mu=0.01), cfl=0.5, rk=1,
                 p_pc="amg", p_maxit=500, p_tol=1e-10, noslip=r.noslip, dirnodes=dn,
                 dirmask=np.ones((dn.size, 3), np.int32),
                 dirvals=np.array([r.dir[p][1] for p in dn]),
                 pdirnodes=np.array([0], np.int64), pdirvals=np.zeros(1))
    U = torch.tensor(r.U, device="cuda")
    P = torch.tensor(r.P, device="cuda")
    pit = X.initial_dev(ctx, U, P)
    r.initial()
    Un, dP = torch.empty_like(U), torch.empty_like(P)
    for it in range(3):
        dt = X.dt_dev(ctx, U)
        n_it = X.step_dev(ctx, dt, U, Un, P, dP)
        r.step(dt)
        torch.cuda.synchronize()
        eu = float(np.max(np.abs(U.cpu().numpy() - r.U)) / np.max(np.abs(r.U)))
        ep = float(np.max(np.abs(P.cpu().numpy() - r.P)) / np.max(np.abs(r.P)))
        _say("  step %d: amg its %d vs pcg %d, |dU| %.2e |dP| %.2e"
             % (it + 1, n_it, r.pit[-1], eu, ep))
        assert n_it < r.pit[-1] and eu < 1e-6 and ep < 1e-6
    ctx.close()


@pytest.mark.parametrize("case", ["damp2", "rk4", "theta"])
def test_chocg_poiseuille_regression_fixture(case):
    """The reference's ChoCG/Poiseuille cases on the device (damp2 with RK1;
    damp4 with RK4 at cfl 1; theta = 0.5 implicit momentum): nonzero
    pressure Dirichlet values in the start-up projections, mask-1 velocity
    BCs, noslip walls; 20 steps against diag_poiseuille_*.std (abs or rel
    1e-7) and each step against the oracle (momentum iterations equal)."""
    import torch
    from tests.test_oracle_chocg import poiseuille_run
    from xyst_amd import chorin as X
    _say("chocg poiseuille %s" % case)
    r, d = poiseuille_run(case)
    ctx, cg, U, P = _device_run(r, dict(flux=str(d[case + "_flux"]), stab=True,
                                        mu=float(d["mu"])))
    if r.theta > 0.0:
        X.momentum_setup(ctx, r.inpoel, r.theta, m_pc=r.m_pc, m_maxit=r.m_iter,
                         m_tol=r.m_tol)
    pit = X.initial_dev(ctx, U, P)
    r.initial()
    assert list(pit) == r.pit[:2]
    Un, dP = torch.empty_like(U), torch.empty_like(P)
    V = r.vol
    t = 0.0
    for it in range(int(d["nstep"])):
        dt = X.dt_dev(ctx, U)
        n_it = X.step_dev(ctx, dt, U, Un, P, dP)
        r.step(dt)
        torch.cuda.synchronize()
        u, un, p, dp = (a.cpu().numpy() for a in (U, Un, P, dP))
        t += dt
        s = [np.sum(p * p * V)] + [np.sum(u[:, c] ** 2 * V) for c in range(3)]
        s += [np.sum(dp * dp * V)] + [np.sum((u[:, c] - un[:, c]) ** 2 * V) for c in range(3)]
        row = np.concatenate([[it + 1, t, dt], np.sqrt(np.array(s) / V.sum())])
        ref = d[case + "_diag"][it]
        diff = np.abs(row - ref)
        assert ((diff <= 1e-7) | (diff <= 1e-7 * np.abs(ref))).all(), (it, diff.max())
        eu = float(np.max(np.abs(u - r.U)) / np.max(np.abs(r.U)))
        if it % 5 == 4:
            _say("  step %d pit %d/%d |dU| vs oracle %.2e" % (it + 1, n_it, r.pit[-1], eu))
        assert eu < 1e-8
        if r.theta > 0.0:
            assert X.momentum_its(ctx) == r.mit[-1]
    ctx.close()
