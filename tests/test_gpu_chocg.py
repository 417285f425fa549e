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
