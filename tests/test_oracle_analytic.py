"""Pins the oracle's source terms, boundary conditions, steady-state local
time stepping, analytic-error diagnostics, the HLLC flux and stab2 to the
reference's own regression results (CPU only).

* RieCG/VorticalFlow: Rusanov, Rusanov + stab2 and steady state are serial
  runs (diag.std, diag_stab2.std, diag_steady.std, vortical_flow{,_stab2,
  _steady}.std.exo); HLLC and HLLC + stab2 are the reference's 8-chare
  partitioned runs (diag_hllc*.std, vortical_flow_hllc*_pe4_u0.5.std.exo.*),
  reproduced here by the partitioned restatement (oracle/partitioned.py) on
  the chare of every tet read from those files.  Serial HLLC differs from
  them by ~1e-4 (partial dual-face normals on chare interfaces), so matching
  them to 1e-9 pins the HLLC star states and the partitioned data flow.
* RieCG/TaylorGreen, KozCG/VorticalFlow, KozCG/TaylorGreen: the source terms
  of riemann::rhs and kozak::rhs.
* RieCG/Pipe: pressure BCs (prebc) with symmetry walls.
* farbc: the characteristic branches against their closed forms.
"""
import numpy as np
import pytest

import oracle as O
from oracle import partitioned as PP
from tests.analytic_cases import (GAMMA_VF, PIPE_BANDS, TG_BANDS, VF_BANDS,
                                  all_dirichlet, check_diag, check_fields,
                                  diag_row, fixture, pipe_case,
                                  primitive_fields, unitcube)


def _riecg_run(name, flux="rusanov", stab2=False, steady=False, diag_every=1,
               key=None):
    d = fixture()
    c = unitcube(d)
    P = O.problem(name, gamma=GAMMA_VF)
    cfg = O.cfg(GAMMA_VF, flux=flux, stab2=stab2)
    U = O.problem_ic(P, c.coord)
    dn, dm = all_dirichlet(c)
    bc = O.BCLists(5, dirnodes=dn, dirmask=dm)
    O.orc_apply_bcs(cfg, P, bc, U, 0.0, c.coord)
    besym = np.zeros(c.tri.size, np.uint8)
    V = c.vol.sum()
    An = O.analytic_prim(P, c.coord)
    ref = d[key + "_diag"]
    t, it, row = 0.0, 0, 0
    tp = np.zeros(c.npoin)
    while row < ref.shape[0]:
        if steady:
            dt, dtp = O.dt_steady(cfg, U, c.vol, 0.8)
        else:
            dt, dtp = O.dt(cfg, U, c.vol, 0.8), None
        if t + dt > 1.0:
            dt = 1.0 - t
        Un = O.riecg_step2(cfg, P, c.se, c.coord, c.tri, besym, c.vol, c.vol, bc, t,
                           dt, U, dtp=dtp, tp=tp)
        t += dt
        it += 1
        if it % diag_every:
            continue
        mine = diag_row(it, t, dt, O.diag(U, Un, c.vol), V, O.diag_err(U, c.vol, An))
        check_diag(mine, ref[row])
        row += 1
    return U, t, d


@pytest.mark.parametrize("case", ["rusanov", "stab2", "steady"])
def test_vortical_flow_serial(case):
    U, t, d = _riecg_run("vortical_flow", stab2=case == "stab2",
                         steady=case == "steady", key="vf_" + case)
    assert abs(t - float(d["vf_%s_t_final" % case])) < 1e-8
    check_fields(primitive_fields(U, GAMMA_VF), d["vf_%s_fields" % case], VF_BANDS)


@pytest.mark.parametrize("case", ["hllc", "hllc_stab2"])
def test_vortical_flow_partitioned_hllc(case):
    d = fixture()
    coord = d["coord"]
    inpoel = d["inpoel"].astype(np.int64)
    tri = d["triinpoel"].astype(np.int64)
    c = unitcube(d)
    P = O.problem("vortical_flow", gamma=GAMMA_VF)
    cfg = O.cfg(GAMMA_VF, flux="hllc", stab2="stab2" in case)
    parts = PP.split(coord, inpoel, tri, d["vf_chare_of_tet"])
    assert len(parts) == 8
    Vol = PP.total_volume(parts, c.npoin)
    np.testing.assert_allclose(Vol, c.vol, rtol=1e-14)
    U = O.problem_ic(P, coord)
    dn, dm = all_dirichlet(c)
    bc = O.BCLists(5, dirnodes=dn, dirmask=dm)
    An = O.analytic_prim(P, coord)
    ref = d["vf_%s_diag" % case]
    V = Vol.sum()
    t = 0.0
    for it in range(ref.shape[0]):
        dt = O.dt(cfg, U, Vol, 0.8)
        if t + dt > 1.0:
            dt = 1.0 - t
        Un = PP.step(cfg, P, parts, Vol, bc, t, dt, U, coord)
        t += dt
        mine = diag_row(it + 1, t, dt, PP.diagnostics(parts, U, Un), V,
                        O.diag_err(U, Vol, An))
        check_diag(mine, ref[it])
    check_fields(primitive_fields(U, GAMMA_VF), d["vf_%s_fields" % case], VF_BANDS)


def test_serial_hllc_is_not_the_partitioned_result():
    """The partitioned and serial HLLC runs differ beyond the fixture's own
    bands, so the fixture does discriminate the partitioned data flow."""
    with pytest.raises(AssertionError):
        _riecg_run("vortical_flow", flux="hllc", key="vf_hllc")


def test_taylor_green_riecg():
    U, t, d = _riecg_run("taylor_green", diag_every=2, key="tg")
    check_fields(primitive_fields(U, GAMMA_VF), d["tg_fields"], TG_BANDS)


@pytest.mark.parametrize("name,key,every", [("vortical_flow", "koz_vf", 1),
                                            ("taylor_green", "koz_tg", 2)])
def test_kozcg_source(name, key, every):
    """KozCG (no FCT) with the problem's source in kozak::rhs."""
    d = fixture()
    c = unitcube(d)
    P = O.problem(name, gamma=GAMMA_VF)
    cfg = O.cfg(GAMMA_VF)
    U = O.problem_ic(P, c.coord)
    dn, _ = all_dirichlet(c)
    dirvals = U[dn].copy()
    An = O.analytic_prim(P, c.coord)
    V = c.vol.sum()
    ref = d[key + "_diag"]
    fct = O.fctcfg(fct=False)
    t, it, row = 0.0, 0, 0
    while row < ref.shape[0]:
        dt = O.dt(cfg, U, c.vol, 0.8)
        if t + dt > 1.0:
            dt = 1.0 - t
        U, Un = O.kozcg_step(cfg, fct, c.coord, c.inpoel, c.vol, dt, [], [], dn, dirvals,
                             U, P=P, t=t)
        t += dt
        it += 1
        if it % every:
            continue
        mine = diag_row(it, t, dt, O.diag(U, Un, c.vol), V, O.diag_err(U, c.vol, An))
        check_diag(mine, ref[row])
        row += 1
    bands = list(VF_BANDS if key == "koz_vf" else TG_BANDS)
    if key == "koz_vf":
        bands[3] = ("rel", 7e-6, 1e-8)
    check_fields(primitive_fields(U, GAMMA_VF), d[key + "_fields"], bands)


def test_pipe_pressure_bc():
    """RieCG/Pipe: userdef IC (rho 1, u 0, p 99950), symmetry walls, pressure
    BCs 100000 / 99900 on the inlet / outlet side sets; the BCs are applied
    to the initial state too (RieCG::merge, RieCG.cpp:754)."""
    d = fixture()
    c = pipe_case(d)
    g = 1.4
    cfg, P = O.cfg(g), O.problem("none", gamma=g)
    U = np.zeros((c.npoin, 5))
    U[:, 0] = 1.0
    U[:, 4] = 99950.0 / (g - 1.0)
    pn, pv = O.prebc_lists(c.tri, c.faceset, [([1], 1.0, 100000.0), ([3], 1.0, 99900.0)])
    sn, sv, _, _, besym = O.riecg_bc_setup(c.coord, c.tri, c.faceset, [2, 4, 5, 6])
    bc = O.BCLists(5, symnodes=sn, symnorms=sv, prenodes=pn, prevals=pv)
    O.orc_apply_bcs(cfg, P, bc, U, 0.0)
    ref = d["pipe_diag"]
    V, t = c.vol.sum(), 0.0
    for it in range(20):
        dt = O.dt(cfg, U, c.vol, 0.5)
        Un = O.riecg_step2(cfg, P, c.se, c.coord, c.tri, besym, c.vol, c.vol, bc, t, dt, U)
        t += dt
        mine = diag_row(it + 1, t, dt, O.diag(U, Un, c.vol), V)
        r = ref[it]
        ab, rel = np.abs(mine - r), np.abs(mine - r) / np.abs(r)
        assert ((ab[1:8] <= 2e-7) | (rel[1:8] <= 1e-7)).all(), rel     # diag.ndiff.cfg
        assert ((ab[8:13] <= 5e-7) | (rel[8:13] <= 6e-5)).all(), rel
        assert (rel[[1, 2, 3, 7, 13]] <= 1e-10).all(), rel               # 12 digits printed
    check_fields(primitive_fields(U, g), d["pipe_fields"], PIPE_BANDS)


def test_farbc_characteristic_branches():
    """physics::farbc against its closed forms: supersonic inflow takes the
    free stream, subsonic inflow the free stream with the inside pressure,
    subsonic outflow keeps rho and momentum and takes the free-stream
    pressure, supersonic outflow leaves the node alone."""
    g = 1.4
    fr, fp = 1.2, 1.0e5
    a = np.sqrt(g * fp / fr)
    U0 = np.array([0.9, 0.9 * 30.0, -0.9 * 5.0, 0.9 * 2.0, 0.0])
    p_in = 0.8e5
    U0[4] = p_in / (g - 1) + 0.5 * (U0[1] ** 2 + U0[2] ** 2 + U0[3] ** 2) / U0[0]
    n = np.array([1.0, 0.0, 0.0])
    for M in (-2.0, -0.5, 0.5, 2.0):
        far = (fr, M * a, 10.0, -3.0, fp)
        U = U0[None, :].copy()
        bc = O.BCLists(5, farnodes=[0], farnorms=n, far=far)
        O.orc_apply_bcs(O.cfg(g), O.problem("none", gamma=g), bc, U, 0.0)
        r, ru, rv, rw, re = U[0]
        pr = (g - 1) * (re - 0.5 * (ru * ru + rv * rv + rw * rw) / r)
        if M <= -1.0:
            np.testing.assert_allclose([r, ru / r, rv / r, rw / r, pr],
                                       [fr, far[1], far[2], far[3], fp], rtol=1e-13)
        elif M < 0.0:
            np.testing.assert_allclose([r, ru / r, rv / r, rw / r, pr],
                                       [fr, far[1], far[2], far[3], p_in], rtol=1e-12)
        elif M < 1.0:
            np.testing.assert_array_equal(U[0, :4], U0[:4])
            assert pr == pytest.approx(fp, rel=1e-12)
        else:
            np.testing.assert_array_equal(U[0], U0)


def test_source_and_ic_closed_forms():
    """The vortical flow IC is a steady Euler solution under its source: at
    any point the source equals the divergence of the Euler flux of the IC
    (central differences), i.e. d/dt U = -div F + S = 0."""
    P = O.problem("vortical_flow", gamma=GAMMA_VF)
    g = GAMMA_VF
    rng = np.random.default_rng(4)
    X = rng.uniform(0.1, 0.9, size=(3, 5))
    h = 1e-5

    def flux(x):
        s = O.problem_ic(P, x.reshape(3, 1))[0]
        r, ru, rv, rw, re = s
        p = (g - 1) * (re - 0.5 * (ru * ru + rv * rv + rw * rw) / r)
        vel = np.array([ru, rv, rw]) / r
        F = np.zeros((3, 5))
        for j in range(3):
            F[j] = [s[j + 1], ru * vel[j], rv * vel[j], rw * vel[j], (re + p) * vel[j]]
            F[j, j + 1] += p
        return F

    S = O.problem_src(P, X)
    for k in range(X.shape[1]):
        div = np.zeros(5)
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            div += (flux(X[:, k] + e)[j] - flux(X[:, k] - e)[j]) / (2 * h)
        np.testing.assert_allclose(div, S[k], atol=1e-6)
