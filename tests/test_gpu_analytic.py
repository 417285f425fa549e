"""The device path on the reference's analytic-problem regression cases
(needs an MI355X): source terms, Dirichlet / farfield / pressure BCs,
steady-state local time stepping, the analytic-error diagnostics, HLLC and
stab2, and k partitions of one mesh on one device.

Comparison rules (tests/analytic_cases.py): the reference's diag*.std rows
within diag.ndiff.cfg and, tighter, to the printed digits (conserved norms
2e-8, error norms 5e-7, residual norms 1e-5 relative); the final fields in
the exodiff.cfg bands.  Per call, the device against the oracle at
1e-12 of the round-off scale (as in test_gpu_riecg.py).
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.analytic_cases import (GAMMA_VF, PIPE_BANDS, TG_BANDS, VF_BANDS,
                                  all_dirichlet, check_diag, check_fields,
                                  diag_row, fixture, pipe_case,
                                  primitive_fields, unitcube)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _riecg_device(c, flux, stab2, problem, besym, sn, sv):
    from xyst_amd import riemann as X
    ctx = X.Context(0, gamma=GAMMA_VF if problem != "none" else 1.4, flux=flux,
                    stab2=stab2, stab2coef=0.2)
    ctx.set_problem(problem)
    edge, dint = X.mesh_domint(c.coord, c.inpoel)
    # the reference's superedges: the MUSCL epsilon makes the edge flux
    # depend on the orientation they give each edge (to ~1e-10 of the
    # round-off scale), so the device takes the same grouping as the oracle
    sup, ints = X.mesh_domsuped(c.inpoel, edge, dint)
    ctx.upload_mesh(c.coord, sup, ints, c.tri, besym)
    ctx.riecg_setup(c.vol, sn, sv)
    return ctx


@pytest.mark.parametrize("case", ["rusanov", "stab2", "steady"])
def test_vortical_flow_on_device(case):
    """RieCG/VorticalFlow (serial reference runs) on the device: source,
    Dirichlet BCs on all sides, device diagnostics incl. error norms."""
    import torch
    from xyst_amd import diagnostics as D
    _say("vortical flow %s" % case)
    d = fixture()
    c = unitcube(d)
    P = O.problem("vortical_flow", gamma=GAMMA_VF)
    besym = np.zeros(c.tri.size, np.uint8)
    ctx = _riecg_device(c, "rusanov", case == "stab2", "vortical_flow", besym,
                        np.zeros(0, np.int64), np.zeros((0, 3)))
    U0 = O.problem_ic(P, c.coord)
    dn, dm = all_dirichlet(c)
    ctx.riecg_bc(dn, dm, U0[dn])
    dev = torch.device("cuda")
    U = torch.from_numpy(U0.copy()).to(dev)
    Un, R = torch.empty_like(U), torch.empty_like(U)
    G = torch.empty((c.npoin, 15), dtype=torch.float64, device=dev)
    V = torch.from_numpy(c.vol).to(dev)
    An = torch.from_numpy(O.analytic_prim(P, c.coord)).to(dev)
    dtp = torch.empty(c.npoin, dtype=torch.float64, device=dev)
    tp = torch.zeros(c.npoin, dtype=torch.float64, device=dev)
    ctx.apply_bc_dev(U)
    ref = d["vf_%s_diag" % case]
    Vt = c.vol.sum()
    t = 0.0
    for it in range(ref.shape[0]):
        if case == "steady":
            dt = ctx.dt_steady_dev(U, 0.8, dtp)
        else:
            dt = ctx.dt_dev(U, 0.8)
        if t + dt > 1.0:
            dt = 1.0 - t
        for s in range(3):
            if case == "steady":
                ctx.stage_dev(s, dt, U, Un, G, R, t=t, dtp=dtp, tp=tp)
            else:
                ctx.stage_dev(s, dt, U, Un, G, R, t=t)
        t += dt
        mine = diag_row(it + 1, t, dt, D.rhocompute(ctx, U, Un, V), Vt,
                        D.errcompute(ctx, U, V, An))
        check_diag(mine, ref[it])
    check_fields(primitive_fields(U.cpu().numpy(), GAMMA_VF),
                 d["vf_%s_fields" % case], VF_BANDS)
    if case == "steady":   # tp advanced by the local steps, as RieCG.cpp:1047-1050
        assert float(tp.min()) > 0.0
    ctx.close()


@pytest.mark.parametrize("case", ["hllc", "hllc_stab2"])
def test_vortical_flow_hllc_eight_partitions_on_one_device(case):
    """The reference's 8-chare HLLC runs reproduced by 8 partitions of the
    mesh on one device (the chare of every tet read from the reference's
    partition files), with the shared-node exchange on the device."""
    import torch
    from xyst_amd.partition import PartitionedRieCG
    _say("partitioned vortical flow %s" % case)
    d = fixture()
    c = unitcube(d)
    P = O.problem("vortical_flow", gamma=GAMMA_VF)
    run = PartitionedRieCG(c.coord, c.inpoel, c.tri, d["vf_chare_of_tet"], gamma=GAMMA_VF,
                           flux="hllc", stab2="stab2" in case, problem="vortical_flow")
    assert len(run.parts) == 8 and run.comm.shared > 0
    np.testing.assert_allclose(run.Vol, c.vol, rtol=1e-14)
    U0 = O.problem_ic(P, c.coord)
    dn, dm = all_dirichlet(c)
    run.setup_bc(dirnodes=dn, dirmask=dm, dirvals=U0[dn])
    Us = run.scatter(U0)
    Uns = [torch.empty_like(u) for u in Us]
    An = O.analytic_prim(P, c.coord)
    Ans = run.scatter(An)
    ref = d["vf_%s_diag" % case]
    Vt = run.Vol.sum()
    t = 0.0
    for it in range(ref.shape[0]):
        dt = run.dt(Us, 0.8)
        if t + dt > 1.0:
            dt = 1.0 - t
        run.step(t, dt, Us, Uns)
        t += dt
        mine = diag_row(it + 1, t, dt, run.rhocompute(Us, Uns), Vt,
                        run.errcompute(Us, Ans))
        check_diag(mine, ref[it])
    # every copy of a shared node holds the same state
    U = run.gather(Us)
    for p, u in zip(run.parts, Us):
        assert np.array_equal(u.cpu().numpy(), U[p.gid])
    check_fields(primitive_fields(U, GAMMA_VF), d["vf_%s_fields" % case], VF_BANDS)
    run.close()


def test_partitions_match_the_partitioned_oracle_per_step():
    """One partitioned RK3 step on the device against the partitioned oracle
    (oracle/partitioned.py) node by node, and k = 1 against the serial run."""
    import torch
    from oracle import partitioned as PP
    from xyst_amd.partition import PartitionedRieCG
    d = fixture()
    c = unitcube(d)
    P = O.problem("vortical_flow", gamma=GAMMA_VF)
    cfg = O.cfg(GAMMA_VF, flux="hllc")
    U0 = O.problem_ic(P, c.coord)
    U0[:, 1] += 0.05 * np.sin(4 * c.coord[2])
    dn, dm = all_dirichlet(c)
    # per-node bound: the rhs of each stage agrees to 1e-12 of its round-off
    # scale A (the oracle's sum of term magnitudes), which the update
    # multiplies by rkcoef dt / vol; three stages, a factor 2 for the change
    # of A over the step
    Go = O.grad(c.se, c.coord, c.tri, U0) / c.vol[:, None]
    _, A = O.rhs(cfg, c.se, c.coord, c.tri, np.zeros(c.tri.size, np.uint8), Go, U0,
                 absacc=True)
    A = A + np.abs(O.problem_src(P, c.coord) * c.vol[:, None])
    for owner in (d["vf_chare_of_tet"], np.zeros(c.inpoel.shape[0], np.int64)):
        _say("partitioned step, %d partitions" % (owner.max() + 1))
        run = PartitionedRieCG(c.coord, c.inpoel, c.tri, owner, gamma=GAMMA_VF,
                               flux="hllc", problem="vortical_flow")
        run.setup_bc(dirnodes=dn, dirmask=dm, dirvals=O.problem_ic(P, c.coord)[dn])
        Us = run.scatter(U0)
        Uns = [torch.empty_like(u) for u in Us]
        dt = 0.01
        run.step(0.0, dt, Us, Uns)
        parts = PP.split(c.coord, c.inpoel, c.tri, owner)
        Uo = U0.copy()
        bc = O.BCLists(5, dirnodes=dn, dirmask=dm)
        PP.step(cfg, P, parts, PP.total_volume(parts, c.npoin), bc, 0.0, dt, Uo, c.coord)
        Ud = run.gather(Us)
        tol = 2e-12 * dt * A / c.vol[:, None] + 1e-15 * np.abs(Uo)
        err = np.abs(Ud - Uo)
        _say("  max err/tol %.3f" % (err / tol).max())
        assert np.all(err <= tol)
        run.close()


def test_taylor_green_on_device():
    import torch
    from xyst_amd import diagnostics as D
    d = fixture()
    c = unitcube(d)
    P = O.problem("taylor_green", gamma=GAMMA_VF)
    besym = np.zeros(c.tri.size, np.uint8)
    ctx = _riecg_device(c, "rusanov", False, "taylor_green", besym,
                        np.zeros(0, np.int64), np.zeros((0, 3)))
    U0 = O.problem_ic(P, c.coord)
    dn, dm = all_dirichlet(c)
    ctx.riecg_bc(dn, dm, U0[dn])
    dev = torch.device("cuda")
    U = torch.from_numpy(U0.copy()).to(dev)
    Un, R = torch.empty_like(U), torch.empty_like(U)
    G = torch.empty((c.npoin, 15), dtype=torch.float64, device=dev)
    V = torch.from_numpy(c.vol).to(dev)
    An = torch.from_numpy(O.analytic_prim(P, c.coord)).to(dev)
    ref = d["tg_diag"]
    Vt, t, it, row = c.vol.sum(), 0.0, 0, 0
    while row < ref.shape[0]:
        dt = min(ctx.dt_dev(U, 0.8), 1.0 - t)
        for s in range(3):
            ctx.stage_dev(s, dt, U, Un, G, R, t=t)
        t += dt
        it += 1
        if it % 2:
            continue
        check_diag(diag_row(it, t, dt, D.rhocompute(ctx, U, Un, V), Vt,
                            D.errcompute(ctx, U, V, An)), ref[row])
        row += 1
    check_fields(primitive_fields(U.cpu().numpy(), GAMMA_VF), d["tg_fields"], TG_BANDS)
    ctx.close()


@pytest.mark.parametrize("name,key,every", [("vortical_flow", "koz_vf", 1),
                                            ("taylor_green", "koz_tg", 2)])
def test_kozcg_source_on_device(name, key, every):
    """KozCG (no FCT) with the source in kozak::rhs against the reference's
    KozCG/VorticalFlow and KozCG/TaylorGreen."""
    import torch
    from xyst_amd import diagnostics as D
    from xyst_amd import kozak as K
    from xyst_amd import riemann as X
    d = fixture()
    c = unitcube(d)
    P = O.problem(name, gamma=GAMMA_VF)
    ctx = X.Context(0, gamma=GAMMA_VF)
    ctx.set_problem(name)
    K.upload(ctx, c.coord, c.inpoel)
    U0 = O.problem_ic(P, c.coord)
    dn, _ = all_dirichlet(c)
    K.setup(ctx, c.vol, np.zeros(0, np.int64), np.zeros((0, 3)), dirnodes=dn,
            dirmask=np.ones((dn.size, 5), np.int32), dirvals=U0[dn],
            fct=False)
    dev = torch.device("cuda")
    U = torch.from_numpy(U0.copy()).to(dev)
    Un = torch.empty_like(U)
    V = torch.from_numpy(c.vol).to(dev)
    An = torch.from_numpy(O.analytic_prim(P, c.coord)).to(dev)
    ref = d[key + "_diag"]
    Vt, t, it, row = c.vol.sum(), 0.0, 0, 0
    while row < ref.shape[0]:
        dt = min(K.dt_dev(ctx, U, 0.8), 1.0 - t)
        K.step_dev(ctx, dt, U, Un, t=t)
        t += dt
        it += 1
        if it % every:
            continue
        check_diag(diag_row(it, t, dt, D.rhocompute(ctx, U, Un, V), Vt,
                            D.errcompute(ctx, U, V, An)), ref[row])
        row += 1
    bands = list(VF_BANDS if key == "koz_vf" else TG_BANDS)
    if key == "koz_vf":
        bands[3] = ("rel", 7e-6, 1e-8)
    check_fields(primitive_fields(U.cpu().numpy(), GAMMA_VF), d[key + "_fields"], bands)
    ctx.close()


def test_pipe_pressure_bc_on_device():
    """RieCG/Pipe: symmetry walls and pressure BCs on inlet and outlet."""
    import torch
    from xyst_amd import diagnostics as D
    from xyst_amd import riemann as X
    d = fixture()
    c = pipe_case(d)
    g = 1.4
    sn, sv, besym = X.mesh_symbc(c.coord, c.tri, c.faceset, [2, 4, 5, 6])
    ctx = _riecg_device(c, "rusanov", False, "none", besym, sn, sv)
    pn, pv = O.prebc_lists(c.tri, c.faceset, [([1], 1.0, 100000.0), ([3], 1.0, 99900.0)])
    ctx.riecg_bc(prenodes=pn, prevals=pv)
    U0 = np.zeros((c.npoin, 5))
    U0[:, 0] = 1.0
    U0[:, 4] = 99950.0 / (g - 1.0)
    dev = torch.device("cuda")
    U = torch.from_numpy(U0).to(dev)
    ctx.apply_bc_dev(U)
    Un, R = torch.empty_like(U), torch.empty_like(U)
    G = torch.empty((c.npoin, 15), dtype=torch.float64, device=dev)
    V = torch.from_numpy(c.vol).to(dev)
    ref = d["pipe_diag"]
    Vt, t = c.vol.sum(), 0.0
    for it in range(20):
        dt = ctx.dt_dev(U, 0.5)
        for s in range(3):
            ctx.stage_dev(s, dt, U, Un, G, R, t=t)
        t += dt
        mine = diag_row(it + 1, t, dt, D.rhocompute(ctx, U, Un, V), Vt)
        r = ref[it]
        ab, rel = np.abs(mine - r), np.abs(mine - r) / np.abs(r)
        assert ((ab[1:8] <= 2e-7) | (rel[1:8] <= 1e-7)).all(), rel
        assert ((ab[8:13] <= 5e-7) | (rel[8:13] <= 6e-5)).all(), rel
        assert (rel[[1, 2, 3, 7, 13]] <= 1e-10).all(), rel
    check_fields(primitive_fields(U.cpu().numpy(), g), d["pipe_fields"], PIPE_BANDS)
    ctx.close()


def test_rhs_with_source_and_farbc_against_oracle():
    """riemann::rhs with the vortical-flow source, per node at 1e-12 of the
    round-off scale, and farbc on every characteristic branch bit for bit
    against the oracle."""
    import torch
    from xyst_amd import riemann as X
    d = fixture()
    c = unitcube(d)
    P = O.problem("vortical_flow", gamma=GAMMA_VF)
    U = O.problem_ic(P, c.coord)
    U[:, 2] += 0.1 * np.cos(3 * c.coord[0])
    besym = np.zeros(c.tri.size, np.uint8)
    ctx = _riecg_device(c, "hllc", True, "vortical_flow", besym, np.zeros(0, np.int64),
                        np.zeros((0, 3)))
    cfg = O.cfg(GAMMA_VF, flux="hllc", stab2=True, stab2coef=0.2)
    Go = O.grad(c.se, c.coord, c.tri, U) / c.vol[:, None]
    Ro, A = O.rhs(cfg, c.se, c.coord, c.tri, besym, Go, U, absacc=True)
    S = O.problem_src(P, c.coord)
    O.src(P, c.coord, c.vol, Ro)
    A = A + np.abs(S * c.vol[:, None])
    R = ctx.rhs(Go, U, v=c.vol, t=0.0)
    assert np.all(np.abs(R - Ro) <= 1e-12 * A + 1e-300)
    with pytest.raises(Exception, match="nodal volumes"):
        ctx.rhs(Go, U)
    # farbc: random states and normals, Mach numbers on both sides of -1, 0, 1
    rng = np.random.default_rng(5)
    n = 64
    nodes = np.arange(n)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    far = (1.1, 300.0, -120.0, 40.0, 0.9e5)
    X0 = np.zeros((c.npoin, 5))
    X0[:, 0] = rng.uniform(0.8, 1.2, c.npoin)
    X0[:, 1:4] = X0[:, :1] * rng.normal(scale=100.0, size=(c.npoin, 3))
    X0[:, 4] = 1e5 / (GAMMA_VF - 1) + 0.5 * (X0[:, 1:4] ** 2).sum(1) / X0[:, 0]
    ctx.riecg_bc(farnodes=nodes, farnorms=nrm, far=far)
    dX = torch.from_numpy(X0.copy()).cuda()
    ctx.apply_bc_dev(dX)
    Xo = X0.copy()
    O.orc_apply_bcs(O.cfg(GAMMA_VF), P, O.BCLists(5, farnodes=nodes, farnorms=nrm, far=far),
                    Xo, 0.0)
    Xd = dX.cpu().numpy()
    assert not np.array_equal(Xd[nodes], X0[nodes])          # the BC did act
    assert np.array_equal(Xd[n:], X0[n:])                    # and only on its nodes
    assert np.all(np.abs(Xd - Xo) <= 1e-14 * np.abs(Xo))
    ctx.close()
