"""Pins the oracle's transported-scalar path (scalar MUSCL, the scalar
Rusanov and HLLC fluxes, scalar boundary fluxes and gradients, per-component
Dirichlet masks) to the reference's RieCG/SlotCyl results (CPU only):
one scalar c0 (cone, hump, slotted cylinder) carried by a rigid rotation
with its momentum source, 20 steps at cfl 0.9 on unitsquare_01_3.6k.

Fixture: tests/golden/riecg_slotcyl.npz (make_regression_fixtures.py):
diag.std (one file for both runs) and slot_cyl{,_hllc}.std.exo.
"""
import numpy as np
import pytest

import oracle as O
from tests.analytic_cases import (SC_BANDS, check_fields, check_slotcyl_diag,
                                  slotcyl_case, slotcyl_diag_row, slotcyl_dirichlet,
                                  slotcyl_fields, slotcyl_fixture)


def run_slotcyl(flux, check=True):
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    g = float(d["gamma"])
    P = O.problem("slot_cyl", gamma=g)
    cfg = O.cfg(g, flux=flux)
    U = O.problem_ic(P, c.coord)
    assert U.shape == (c.npoin, 6)
    dn, dm = slotcyl_dirichlet(c)
    bc = O.BCLists(6, dirnodes=dn, dirmask=dm)
    O.orc_apply_bcs(cfg, P, bc, U, 0.0, c.coord)
    besym = np.zeros(c.tri.size, np.uint8)
    V = c.vol.sum()
    ref = d["diag"]
    t = 0.0
    rows = []
    for it in range(int(d["nstep"])):
        dt = O.dt(cfg, U, c.vol, float(d["cfl"]))
        Un = O.riecg_step2(cfg, P, c.se, c.coord, c.tri, besym, c.vol, c.vol, bc, t, dt, U)
        t += dt
        # NodeDiagnostics evaluates the analytic solution at the new time
        An = O.analytic_prim(P, c.coord, t)
        mine = slotcyl_diag_row(it + 1, t, dt, O.diag(U, Un, c.vol), V,
                                O.diag_err(U, c.vol, An))
        rows.append(mine)
        if check:
            check_slotcyl_diag(mine, ref[it], tight=flux == "rusanov")
    return U, t, d, np.array(rows)


def test_slot_cyl_initial_condition_is_the_reference_analytic_field():
    """c0_analytic of the output file is IC() at the final time: the rotated
    shapes, including the slot's edges."""
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    P = O.problem("slot_cyl", gamma=float(d["gamma"]))
    tf = float(d["rusanov_t_final"])
    A = O.analytic_prim(P, c.coord, tf)
    names = list(d["field_names"])
    ref = d["rusanov_fields"]
    assert np.abs(A[:, 5] - ref[:, names.index("c0_analytic")]).max() < 1e-14
    assert np.abs(A[:, 0] - ref[:, names.index("density_analytic")]).max() < 1e-14
    assert np.abs(A[:, 1] - ref[:, names.index("xvelocity_analytic")]).max() < 1e-14
    # the three shapes are all present
    c0 = A[:, 5]
    assert np.isclose(c0.max(), 0.6) and (c0 > 0).sum() > 50


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_slot_cyl(flux):
    U, t, d, _ = run_slotcyl(flux)
    assert abs(t - float(d["%s_t_final" % flux])) < 1e-8
    check_fields(slotcyl_fields(U, float(d["gamma"])), d["%s_fields" % flux][:, :7],
                 SC_BANDS)


def test_scalar_flux_is_discriminated():
    """The Rusanov and HLLC scalar fluxes differ beyond the c0 band, so each
    output file pins its own flux."""
    d = slotcyl_fixture()
    a, b = d["rusanov_fields"][:, 6], d["hllc_fields"][:, 6]
    assert np.abs(a - b).max() > 1e-4


def run_slotcyl_partitioned(owner):
    """The Rusanov run as the reference's chares (oracle/partitioned.py):
    every chare integrates its own tets, so the dual-face normals of edges on
    chare interfaces are partial and the nonlinear dissipation differs from
    the serial run's."""
    from oracle import partitioned as PP
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    g = float(d["gamma"])
    P = O.problem("slot_cyl", gamma=g)
    cfg = O.cfg(g)
    parts = PP.split(c.coord, c.inpoel, c.tri, owner)
    Vol = PP.total_volume(parts, c.npoin)
    U = O.problem_ic(P, c.coord)
    dn, dm = slotcyl_dirichlet(c)
    bc = O.BCLists(6, dirnodes=dn, dirmask=dm)
    O.orc_apply_bcs(cfg, P, bc, U, 0.0, c.coord)
    t = 0.0
    rows = []
    for it in range(int(d["nstep"])):
        dt = O.dt(cfg, U, Vol, float(d["cfl"]))
        Un = PP.step(cfg, P, parts, Vol, bc, t, dt, U, c.coord)
        t += dt
        An = O.analytic_prim(P, c.coord, t)
        rows.append(slotcyl_diag_row(it + 1, t, dt, PP.diagnostics(parts, U, Un),
                                     Vol.sum(), O.diag_err(U, Vol, An)))
    return U, t, np.array(rows)


def test_slot_cyl_eight_chares():
    """The reference's 8-chare Rusanov run (slot_cyl_pe4_u0.5.std.exo.*):
    the partitioned restatement on the chare of every tet read from those
    files reproduces its fields, and its diagnostics stay within the bands
    of the (serial) diag.std."""
    d = slotcyl_fixture()
    U, t, rows = run_slotcyl_partitioned(d["chare_of_tet"])
    assert abs(t - float(d["pe8_t_final"])) < 1e-8
    check_fields(slotcyl_fields(U, float(d["gamma"])), d["pe8_fields"][:, :7], SC_BANDS)
    for mine, ref in zip(rows, d["diag"]):
        check_slotcyl_diag(mine, ref, tight=False)
    # and the serial fields do not pass for it: the fixture discriminates
    with pytest.raises(AssertionError):
        check_fields(d["rusanov_fields"][:, :7], d["pe8_fields"][:, :7], SC_BANDS)


def run_kozcg_slotcyl(check=True):
    """KozCG/SlotCyl: kozak::rhs with the scalar's advection and the
    momentum source, FCT (reference defaults) over six components, Dirichlet
    masks, and freezeflow 3 from freezetime 0: KozCG::dt multiplies the time
    step once the time exceeds freezetime (from the second step on), and
    while it does the flow components keep their values (KozCG.cpp:669-674,
    1141, 1168-1175)."""
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    g = float(d["gamma"])
    P = O.problem("slot_cyl", gamma=g)
    cfg = O.cfg(g)
    fct = O.fctcfg(fct=True, fctdif=1.0, fctclip=False, fctsys=())
    U = O.problem_ic(P, c.coord)
    dn, dm = slotcyl_dirichlet(c)
    dirvals = U[dn].copy()
    V = c.vol.sum()
    ref = d["koz_diag"]
    ff, ft = float(d["koz_freezeflow"]), float(d["koz_freezetime"])
    t = 0.0
    rows = []
    for it in range(int(d["nstep"])):
        dt = O.dt(cfg, U, c.vol, float(d["koz_cfl"]))
        freeze = t > ft
        if freeze:
            dt *= ff
        U, Un = O.kozcg_step(cfg, fct, c.coord, c.inpoel, c.vol, dt, [], [], dn, dirvals, U,
                             P=P, t=t, freeze=freeze, dirmask=dm)
        t += dt
        An = O.analytic_prim(P, c.coord, t)
        mine = slotcyl_diag_row(it + 1, t, dt, O.diag(U, Un, c.vol), V,
                                O.diag_err(U, c.vol, An))
        rows.append(mine)
        if check:
            check_slotcyl_diag(mine, ref[it], tight=True)
    return U, t, d, np.array(rows)


def test_kozcg_slot_cyl():
    U, t, d, _ = run_kozcg_slotcyl()
    assert abs(t - float(d["koz_t_final"])) < 1e-8
    check_fields(slotcyl_fields(U, float(d["gamma"])), d["koz_fields"][:, :7], SC_BANDS)
