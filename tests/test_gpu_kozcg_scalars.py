"""KozCG with transported scalars, local time steps and freezeflow on the
device (needs an MI355X): kozak::rhs with ncomp = 6 and 7 and with steady
local time steps against the oracle (1e-12 of the oracle's round-off scale,
as in test_gpu_kozcg.py), KozCG steps with scalars and with local time
stepping against the oracle's steps (1e-11 of each component's magnitude),
and the reference's KozCG/SlotCyl run (one scalar, FCT over six components,
freezeflow 3) within its diag.ndiff.cfg (and the printed digits) and
exodiff.cfg.
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.analytic_cases import (SC_BANDS, check_fields, check_slotcyl_diag,
                                  slotcyl_case, slotcyl_diag_row, slotcyl_dirichlet,
                                  slotcyl_fields, slotcyl_fixture)
from tests.meshes import smooth_state
from tests.test_gpu_kozcg import STEP_TOL, TOL, _case, _check_step, _ctx, _kozcg_case
from tests.test_oracle_slotcyl import run_kozcg_slotcyl
from xyst_amd.meshgen import sod_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scalars(U, coord, ns, seed=5):
    rng = np.random.default_rng(seed)
    x, y, z = coord
    S = []
    for _ in range(ns):
        k = rng.uniform(1.0, 4.0, 3)
        S.append(0.5 + 0.3 * np.sin(k[0] * x + k[1] * y + k[2] * z) + 0.4 * (x > 0.45))
    return np.concatenate([U, np.stack(S, 1)], 1)


def _assert_rhs(R, Ro, A, what):
    err = np.abs(R - Ro)
    bad = err > TOL * A
    assert not bad.any(), "%s: %d entries off, worst ratio %.3e" % (
        what, bad.sum(), float(np.max(err / np.maximum(A, 1e-300))))


@pytest.mark.parametrize("ns", [1, 2])
def test_kozak_rhs_scalars(ns):
    from xyst_amd import kozak as K
    _say("kozak rhs + %d scalars" % ns)
    m = _case(12, "rcm+block", seed=4)
    coord, inpoel = m["coord"], m["inpoel"]
    U = _scalars(smooth_state(coord, seed=3), coord, ns)
    Ro, A = O.kozak_rhs(O.cfg(), coord, inpoel, 2e-3, U, absacc=True)
    ctx = _ctx()
    K.upload(ctx, coord, inpoel)
    R = K.rhs(ctx, 2e-3, U)
    _assert_rhs(R, Ro, A, "rhs")
    assert np.all(np.abs(R.sum(0)) <= 1e-13 * A.sum(0))
    ctx.close()


def test_kozak_rhs_local_time_steps():
    """Steady runs: each tet takes the mean of its corners' dtp."""
    from xyst_amd import kozak as K
    _say("kozak rhs steady")
    m = _case(12, "rcm+block", seed=6)
    coord, inpoel = m["coord"], m["inpoel"]
    U = smooth_state(coord, seed=4)
    dtp = 2e-3 * (1.0 + 0.5 * np.sin(3 * coord[0]) * np.cos(2 * coord[2]))
    Ro = O.kozak_rhs(O.cfg(), coord, inpoel, 7.0, U, dtp=dtp)
    ctx = _ctx()
    K.upload(ctx, coord, inpoel)
    R = K.rhs(ctx, 7.0, U, dtp=dtp)          # dt unused when dtp is given
    scale = np.abs(Ro).max(0)
    assert (np.abs(R - Ro) <= 1e-12 * scale).all(), np.abs(R - Ro).max(0) / scale
    # uniform local steps are the time-accurate rhs
    Ru = K.rhs(ctx, 2e-3, U)
    R2 = K.rhs(ctx, 7.0, U, dtp=np.full(coord.shape[1], 2e-3))
    assert (np.abs(R2 - Ru) <= 1e-14 * np.abs(Ru).max(0)).all()
    ctx.close()


@pytest.mark.parametrize("ns,fct", [(1, dict(fct=True, fctclip=False, fctsys=())),
                                    (2, dict(fct=True, fctclip=True, fctsys=(0, 5))),
                                    (1, dict(fct=False))])
def test_kozcg_steps_with_scalars(ns, fct):
    """Five device steps with scalars, each against the oracle's step from
    the same state (Dirichlet x faces on every component, symmetry on the
    others)."""
    import torch
    from xyst_amd import kozak as K
    _say("kozcg + %d scalars %s" % (ns, fct))
    c, dn = _kozcg_case(12, "rcm+block")
    U0 = sod_state(c.coord)
    U0[:, 1] = 0.1 * U0[:, 0]
    U0 = _scalars(U0, c.coord, ns)
    nc = U0.shape[1]
    ctx = _ctx()
    K.upload(ctx, c.coord, c.inpoel)
    K.setup(ctx, c.vol, c.symnodes, c.symnorms, dn, np.ones((dn.size, nc), np.int32),
            U0[dn], ncomp=nc, **fct)
    U = torch.tensor(U0, device="cuda")
    Un = torch.empty_like(U)
    Uh = U0.copy()
    f = O.fctcfg(fct.get("fct", True), fct.get("fctdif", 1.0), fct.get("fctclip", False),
                 list(fct.get("fctsys", ())))
    for step in range(5):
        dt = K.dt_dev(ctx, U, 0.5)
        assert dt == pytest.approx(O.dt(O.cfg(), Uh, c.vol, 0.5), rel=1e-14)
        K.step_dev(ctx, dt, U, Un)
        torch.cuda.synchronize()
        Ug = U.cpu().numpy()
        Uo, _ = O.kozcg_step(O.cfg(), f, c.coord, c.inpoel, c.vol, dt, c.symnodes,
                             c.symnorms, dn, U0[dn], Uh)
        _check_step(Ug, Uo, "step %d" % step)
        Uh = Ug
    ctx.close()


@pytest.mark.parametrize("fct", [True, False])
def test_kozcg_local_time_stepping(fct):
    """Steady KozCG: device dtp against the oracle's, five steps with local
    time steps (rhs, low-order solution or update per node, tp += dtp)
    against the oracle's steps."""
    import torch
    from xyst_amd import kozak as K
    _say("kozcg steady fct=%s" % fct)
    c, dn = _kozcg_case(12, "rcm+block")
    U0 = sod_state(c.coord)
    U0[:, 1] = 0.1 * U0[:, 0]
    ctx = _ctx()
    K.upload(ctx, c.coord, c.inpoel)
    K.setup(ctx, c.vol, c.symnodes, c.symnorms, dn, np.ones((dn.size, 5), np.int32),
            U0[dn], fct=fct)
    U = torch.tensor(U0, device="cuda")
    Un = torch.empty_like(U)
    dtp = torch.empty(c.npoin, dtype=torch.float64, device="cuda")
    tp = torch.zeros(c.npoin, dtype=torch.float64, device="cuda")
    Uh = U0.copy()
    tph = np.zeros(c.npoin)
    f = O.fctcfg(fct, 1.0, False, [])
    for step in range(5):
        dt = K.dt_steady_dev(ctx, U, 0.5, dtp)
        dto, dtpo = O.dt_steady(O.cfg(), Uh, c.vol, 0.5)
        assert dt == pytest.approx(dto, rel=1e-14)
        np.testing.assert_allclose(dtp.cpu().numpy(), dtpo, rtol=1e-14)
        K.step_dev(ctx, dt, U, Un, dtp=dtp, tp=tp)
        torch.cuda.synchronize()
        Ug = U.cpu().numpy()
        Uo, _ = O.kozcg_step(O.cfg(), f, c.coord, c.inpoel, c.vol, dt, c.symnodes,
                             c.symnorms, dn, U0[dn], Uh, dtp=dtp.cpu().numpy(), tp=tph)
        _check_step(Ug, Uo, "step %d" % step)
        Uh = Ug
    np.testing.assert_allclose(tp.cpu().numpy(), tph, rtol=1e-14)
    ctx.close()


def test_kozcg_slot_cyl_on_device():
    """KozCG/SlotCyl: 20 steps, cfl 0.5, freezeflow 3 from time 0 (dt x 3 and
    the flow frozen from the second step on), FCT over the six components."""
    import torch
    from xyst_amd import diagnostics as D
    from xyst_amd import kozak as K
    from xyst_amd import riemann as X
    _say("kozcg slot_cyl")
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    g = float(d["gamma"])
    P = O.problem("slot_cyl", gamma=g)
    ctx = X.Context(0, gamma=g)
    ctx.set_problem("slot_cyl")
    K.upload(ctx, c.coord, c.inpoel)
    U0 = O.problem_ic(P, c.coord)
    dn, dm = slotcyl_dirichlet(c)
    K.setup(ctx, c.vol, np.zeros(0, np.int64), np.zeros((0, 3)), dn, dm, U0[dn],
            fct=True, fctdif=1.0, fctclip=False, fctsys=(), ncomp=6)
    dev = torch.device("cuda")
    U = torch.from_numpy(U0.copy()).to(dev)
    Un = torch.empty_like(U)
    V = torch.from_numpy(c.vol).to(dev)
    ref = d["koz_diag"]
    ff, ft = float(d["koz_freezeflow"]), float(d["koz_freezetime"])
    t = 0.0
    for it in range(int(d["nstep"])):
        dt = K.dt_dev(ctx, U, float(d["koz_cfl"]))
        freeze = t > ft
        if freeze:
            dt *= ff
        K.step_dev(ctx, dt, U, Un, t=t, freeze=freeze)
        t += dt
        An = torch.from_numpy(O.analytic_prim(P, c.coord, t)).to(dev)
        mine = slotcyl_diag_row(it + 1, t, dt, D.rhocompute(ctx, U, Un, V), c.vol.sum(),
                                D.errcompute(ctx, U, V, An))
        check_slotcyl_diag(mine, ref[it], tight=True)
    Ug = U.cpu().numpy()
    ctx.close()
    assert abs(t - float(d["koz_t_final"])) < 1e-8
    check_fields(slotcyl_fields(Ug, g), d["koz_fields"][:, :7], SC_BANDS)
    Uo, _, _, _ = run_kozcg_slotcyl(check=False)
    scale = np.abs(Uo).max(0)
    assert (np.abs(Ug - Uo) <= 1e-10 * scale).all(), np.abs(Ug - Uo).max(0) / scale
