"""ChoCG with a transported scalar on the device (needs an MI355X).

* chorin::vgrad and chorin::rhs with ncomp 4 (velocity + one scalar; damp2
  and damp4, stab2, viscosity and diffusivity) against the oracle, per entry
  within 1e-12 of the oracle's round-off magnitude A (the tolerance of
  tests/test_gpu_chocg.py).
* The reference's ChoCG/SlotCyl runs (damp2 RK3, damp4 RK4, damp4 RK4 with
  the flow frozen after t = 0.1) on the device, step by step against the
  oracle's run (dt to 1e-12, c0 to 1e-12 of max |c0| after each step for
  damp2 and 1e-8 for the damp4 runs, whose limiter amplifies round-off) and
  at the end against the reference's fields and diagnostics in the bands of
  tests/test_oracle_chocg_slotcyl.py.  The pressure solves' iteration
  counts are printed, not compared: Jacobi-PCG stopped at 1e-2 amplifies
  round-off (154 vs 156 iterations in one step), and the pressure feeds
  back into neither c0 nor the velocity the BCs hold.
* The refusals: five components, the momentum system with a scalar, a
  problem without a ChoCG source, slot_cyl with three components.
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.test_gpu_chocg import _close, _device_run
from tests.test_oracle_chocg_slotcyl import (CASES, check_fields, check_row, fixture,
                                             run_setup)
from xyst_amd.meshgen import box_mesh, renumber_mesh

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case4(n, seed=5):
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    m = box_mesh(n, seed=seed)
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
    U = np.stack([np.sin(2 * x + y), np.cos(x - 3 * z), x * y - z,
                  np.exp(-((x - 0.4) ** 2 + (y - 0.6) ** 2) * 8.0)], 1)
    U += 0.05 * rng.standard_normal(U.shape)
    P = np.cos(3 * x) * y + 0.05 * rng.standard_normal(npoin)
    G = rng.standard_normal((npoin, 12))
    return ctx, se, coord, tri, U, P, G


def test_chorin_vgrad_rhs_with_a_scalar():
    from xyst_amd import chorin as X
    _say("chorin vgrad/rhs ncomp 4")
    ctx, se, coord, tri, U, P, G = _case4(11)
    npoin = U.shape[0]
    Go, A = O.chorin_vgrad(se, coord, tri, U, absacc=True)
    _close(X.vgrad(ctx, U, np.zeros((npoin, 12))), Go, A, "vgrad ncomp 4")
    for flux in ("damp2", "damp4"):
        for stab2, mu, diff in ((False, 0.0, 0.0), (True, 0.03, 0.02)):
            ko = dict(flux=flux, stab=True, stab2=stab2, stab2coef=0.3, mu=mu, diff=diff)
            Ro, A = O.chorin_rhs(O.chocfg(**ko), se, coord, tri, P, U, G, absacc=True)
            R = X.rhs(ctx, X.config(**ko), P, U, G)
            _close(R, Ro, A, "rhs %s stab2=%s diff=%g" % (flux, stab2, diff))
    # the diffusivity acts on the scalar only
    cfg0 = X.config("damp4", stab=True, mu=0.01)
    cfg1 = X.config("damp4", stab=True, mu=0.01, diff=0.05)
    R0, R1 = X.rhs(ctx, cfg0, P, U, G), X.rhs(ctx, cfg1, P, U, G)
    assert np.array_equal(R0[:, :3], R1[:, :3]) and not np.array_equal(R0[:, 3], R1[:, 3])
    ctx.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_chocg_slotcyl_device(case):
    import torch
    from xyst_amd import chorin as X
    from xyst_amd import diagnostics as D
    _say("chocg slotcyl %s" % case)
    d = fixture()
    r = run_setup(d, case)
    flux = CASES[case][0]
    ctx, cg, U, P = _device_run(r, dict(flux=flux, stab=True), problem="slot_cyl")
    pit = X.initial_dev(ctx, U, P)
    r.initial()
    _say("  start-up pressure iterations %s, oracle %s" % (list(pit), r.pit[:2]))
    Un = torch.empty_like(U)
    dP = torch.empty_like(P)
    dV = torch.tensor(r.vol, device="cuda")
    V = r.vol.sum()
    ff, t = 1.0, 0.0
    ctol = 1e-12 if case == "damp2" else 1e-8
    for it in range(int(d["nstep"])):
        dt = X.dt_dev(ctx, U)
        if t > float(d["freezetime"]):
            ff = CASES[case][2]
        dt *= ff
        r.step()
        assert dt == pytest.approx(r.dt, rel=1e-12)
        n_it = X.step_dev(ctx, dt, U, Un, P, dP, freeze=ff > 1.0)
        t += dt
        u = U.cpu().numpy()
        err = float(np.abs(u[:, 3] - r.U[:, 3]).max() / np.abs(r.U[:, 3]).max())
        _say("  step %d pit %d (oracle %d) |dc0|/max|c0| vs oracle %.3e"
             % (it + 1, n_it, r.pit[-1], err))
        assert err <= ctol
        assert np.array_equal(u[:, :3], r.U[:, :3])     # the velocity BCs
        s = D.precompute(ctx, U, Un, P, dP, dV)
        row = np.concatenate([[it + 1, t, dt], np.sqrt(s / V)])
        check_row(case, it, row, d[case + "_diag"][it])
    assert abs(t - float(d[case + "_t_final"])) < 1e-12
    err = check_fields(d, case, U.cpu().numpy())
    _say("  c0 vs the reference's field: %.3e" % err)
    ctx.close()


def test_chocg_scalar_refusals():
    import torch
    from xyst_amd import chorin as X
    from xyst_amd._xg import XG_EINVAL, XG_EUNSUPPORTED, lib
    _say("chocg scalar refusals")
    d = fixture()
    r = run_setup(d, "damp2")
    assert lib().xg_chocg_set_ncomp(None, 4) == XG_EINVAL
    ctx, cg, U, P = _device_run(r, dict(flux="damp2", stab=True), problem="slot_cyl")
    h = ctx._h
    assert lib().xg_chocg_set_ncomp(h, 5) == XG_EUNSUPPORTED
    assert lib().xg_chocg_momentum_setup(h, r.inpoel.shape[0],
                                         np.ascontiguousarray(r.inpoel, np.uint64).ctypes.data,
                                         0.5, 1, 10, 1e-3) == XG_EUNSUPPORTED
    Un, dP = torch.empty_like(U), torch.empty_like(P)
    ctx.set_problem("vortical_flow")
    with pytest.raises(Exception):
        X.step_dev(ctx, 1e-3, U, Un, P, dP)
    ctx.close()
    # slot_cyl with three components
    from xyst_amd import riemann as R
    ctx = R.Context(0)
    X.upload(ctx, r.coord, r.se.sup, r.se.int, r.tri)
    from xyst_amd.cg import CSR, ConjugateGradients
    cg = ConjugateGradients(CSR.poisson(r.coord, r.inpoel), ctx=ctx)
    X.step_setup(ctx, r.vol, X.config("damp2"))
    ctx.set_problem("slot_cyl")
    U3 = torch.zeros((r.npoin, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(Exception):
        X.step_dev(ctx, 1e-3, U3, torch.empty_like(U3), P, dP)
    del cg
    ctx.close()
