"""Transported scalars on the device (needs an MI355X): riemann::grad/rhs
with ncomp = 6 and 7 against the oracle, and the reference's RieCG/SlotCyl
runs (one scalar carried by a rigid rotation, Rusanov and HLLC, 20 steps)
reproduced on the device, serially and as the reference's eight chares on
one device.

Tolerances: per call the device against the oracle at 1e-12 of the oracle's
round-off scale (as in test_gpu_riecg.py); the SlotCyl runs against the
reference's diag.std within diag.ndiff.cfg (Rusanov also to the printed 12
digits: 1e-9 relative) and the final fields within exodiff.cfg
(tests/analytic_cases.py); the device run against the oracle's run of the
same case to 1e-10 of each field's magnitude.
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.analytic_cases import (SC_BANDS, check_fields, check_slotcyl_diag,
                                  slotcyl_case, slotcyl_diag_row, slotcyl_dirichlet,
                                  slotcyl_fields, slotcyl_fixture)
from tests.meshes import box_case, smooth_state
from tests.test_gpu_riecg import _parity
from tests.test_oracle_slotcyl import run_slotcyl, run_slotcyl_partitioned

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _with_scalars(U, coord, ns, seed=3):
    """A flow state with ns smooth scalars, each with a jump (the limiter's
    branches all taken)."""
    rng = np.random.default_rng(seed)
    x, y, z = coord
    S = []
    for _ in range(ns):
        k = rng.uniform(1.0, 4.0, 3)
        s = 0.5 + 0.3 * np.sin(k[0] * x + k[1] * y + k[2] * z)
        s += 0.4 * (x + 0.3 * y > 0.55)
        S.append(s)
    return np.concatenate([U, np.stack(S, 1)], 1)


@pytest.mark.parametrize("ns", [1, 2])
@pytest.mark.parametrize("flux,stab2", [("rusanov", False), ("rusanov", True), ("hllc", False)])
def test_scalar_parity_box(ns, flux, stab2):
    c = box_case(8, seed=31)
    U = _with_scalars(smooth_state(c.coord), c.coord, ns)
    _parity(c, U, flux=flux, stab2=stab2, label="box8 +%d scalars" % ns)


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_scalar_parity_slot_cyl_mesh(flux):
    """The slot_cyl initial state (steep scalar edges) on its own mesh, no
    symmetry faces."""
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    P = O.problem("slot_cyl", gamma=float(d["gamma"]))
    U = O.problem_ic(P, c.coord)
    _parity(c, U, flux=flux, besym=np.zeros(c.tri.size, np.uint8), label="slot_cyl")


def _device_run(flux, check=True):
    import torch
    from xyst_amd import diagnostics as D
    from xyst_amd import riemann as X
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    g = float(d["gamma"])
    P = O.problem("slot_cyl", gamma=g)
    ctx = X.Context(0, gamma=g, flux=flux)
    ctx.set_problem("slot_cyl")
    edge, dint = X.mesh_domint(c.coord, c.inpoel)
    sup, ints = X.mesh_domsuped(c.inpoel, edge, dint)
    ctx.upload_mesh(c.coord, sup, ints, c.tri, np.zeros(c.tri.size, np.uint8))
    ctx.riecg_setup(c.vol, np.zeros(0, np.int64), np.zeros((0, 3)), ncomp=6)
    U0 = O.problem_ic(P, c.coord)
    dn, dm = slotcyl_dirichlet(c)
    ctx.riecg_bc(dn, dm, U0[dn])
    dev = torch.device("cuda")
    U = torch.from_numpy(U0.copy()).to(dev)
    Un, R = torch.empty_like(U), torch.empty_like(U)
    G = torch.empty((c.npoin, 18), dtype=torch.float64, device=dev)
    V = torch.from_numpy(c.vol).to(dev)
    ctx.apply_bc_dev(U)
    ref = d["diag"]
    Vt = c.vol.sum()
    t = 0.0
    for it in range(int(d["nstep"])):
        dt = ctx.dt_dev(U, float(d["cfl"]))
        for s in range(3):
            ctx.stage_dev(s, dt, U, Un, G, R, t=t)
        t += dt
        An = torch.from_numpy(O.analytic_prim(P, c.coord, t)).to(dev)
        mine = slotcyl_diag_row(it + 1, t, dt, D.rhocompute(ctx, U, Un, V), Vt,
                                D.errcompute(ctx, U, V, An))
        if check:
            check_slotcyl_diag(mine, ref[it], tight=flux == "rusanov")
    out = U.cpu().numpy()
    ctx.close()
    return out, t, d


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_slot_cyl_on_device(flux):
    _say("slot_cyl on device, %s" % flux)
    U, t, d = _device_run(flux)
    assert abs(t - float(d["%s_t_final" % flux])) < 1e-8
    check_fields(slotcyl_fields(U, float(d["gamma"])), d["%s_fields" % flux][:, :7],
                 SC_BANDS)
    Uo, _, _, _ = run_slotcyl(flux, check=False)
    scale = np.abs(Uo).max(0)
    assert (np.abs(U - Uo) <= 1e-10 * scale).all(), np.abs(U - Uo).max(0) / scale


def test_slot_cyl_eight_partitions_on_one_device():
    """The reference's 8-chare Rusanov run (the chare of every tet read from
    its partition files) as eight partitions on one device, the shared-node
    sums of G and R exchanged for all six components: its fields within the
    exodiff bands, and the oracle's partitioned run to 1e-10."""
    import torch
    from xyst_amd.partition import PartitionedRieCG
    _say("slot_cyl, 8 partitions")
    d = slotcyl_fixture()
    c = slotcyl_case(d)
    g = float(d["gamma"])
    P = O.problem("slot_cyl", gamma=g)
    run = PartitionedRieCG(c.coord, c.inpoel, c.tri, d["chare_of_tet"], gamma=g,
                           flux="rusanov", problem="slot_cyl", ncomp=6)
    assert len(run.parts) == 8 and run.comm.shared > 0
    U0 = O.problem_ic(P, c.coord)
    dn, dm = slotcyl_dirichlet(c)
    run.setup_bc(dirnodes=dn, dirmask=dm, dirvals=U0[dn])
    Us = run.scatter(U0)
    Uns = [torch.empty_like(u) for u in Us]
    run.apply_bc(Us)
    ref = d["diag"]
    Vt = run.Vol.sum()
    t = 0.0
    for it in range(int(d["nstep"])):
        dt = run.dt(Us, float(d["cfl"]))
        run.step(t, dt, Us, Uns)
        t += dt
        Ans = run.scatter(O.analytic_prim(P, c.coord, t))
        mine = slotcyl_diag_row(it + 1, t, dt, run.rhocompute(Us, Uns), Vt,
                                run.errcompute(Us, Ans))
        check_slotcyl_diag(mine, ref[it], tight=False)
    U = run.gather(Us)
    run.close()
    assert abs(t - float(d["pe8_t_final"])) < 1e-8
    check_fields(slotcyl_fields(U, g), d["pe8_fields"][:, :7], SC_BANDS)
    Uo, _, _ = run_slotcyl_partitioned(d["chare_of_tet"])
    scale = np.abs(Uo).max(0)
    assert (np.abs(U - Uo) <= 1e-10 * scale).all(), np.abs(U - Uo).max(0) / scale


def test_two_scalars_steps_against_the_oracle():
    """ncomp = 7 through the resident stepping: five RK3 steps on a box with
    symmetry walls and two scalars, device against oracle."""
    import torch
    from xyst_amd import riemann as X
    _say("two scalars, 5 steps")
    c = box_case(10, seed=8)
    U0 = _with_scalars(smooth_state(c.coord), c.coord, 2, seed=9)
    cfg = O.cfg(1.4)
    P = O.problem("none")
    ctx = X.Context(0, gamma=1.4)
    ctx.upload_mesh(c.coord, c.se.sup, c.se.int, c.tri, c.besym)
    ctx.riecg_setup(c.vol, c.symnodes, c.symnorms, ncomp=7)
    dev = torch.device("cuda")
    U = torch.from_numpy(U0.copy()).to(dev)
    Un, R = torch.empty_like(U), torch.empty_like(U)
    G = torch.empty((c.npoin, 21), dtype=torch.float64, device=dev)
    Uo = U0.copy()
    bc = O.BCLists(7, symnodes=c.symnodes, symnorms=c.symnorms)
    t = 0.0
    for _ in range(5):
        dt = ctx.dt_dev(U, 0.5)
        dto = O.dt(cfg, Uo, c.vol, 0.5)
        assert abs(dt - dto) <= 1e-13 * dto
        for s in range(3):
            ctx.stage_dev(s, dt, U, Un, G, R, t=t)
        O.riecg_step2(cfg, P, c.se, c.coord, c.tri, c.besym, c.vol, c.vol, bc, t, dto, Uo)
        t += dt
    out = U.cpu().numpy()
    ctx.close()
    scale = np.abs(Uo).max(0)
    assert (np.abs(out - Uo) <= 1e-11 * scale).all(), np.abs(out - Uo).max(0) / scale
    assert np.abs(out[:, 5:] - U0[:, 5:]).max() > 1e-4      # the scalars moved


def test_more_scalars_than_served_are_refused():
    from xyst_amd import riemann as X
    from xyst_amd._xg import XgError
    c = box_case(4, seed=1)
    ctx = X.Context(0)
    ctx.upload_mesh(c.coord, c.se.sup, c.se.int, c.tri, c.besym)
    U = np.ones((c.npoin, 8))
    with pytest.raises(XgError, match="transported scalars"):
        ctx.grad(U)
    with pytest.raises(XgError):
        ctx.riecg_setup(c.vol, c.symnodes, c.symnorms, ncomp=8)
    ctx.close()
