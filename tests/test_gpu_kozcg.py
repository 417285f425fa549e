"""Parity of the HIP kozak::rhs with the CPU oracle (needs an MI355X).

Tolerance: FP64, per node and component |R_gpu - R_oracle| <= 1e-12 * A,
where A is the oracle's first-order round-off magnitude of that entry (the
device evaluates the same Taylor-Galerkin terms in another order, with
reciprocals refined by Newton steps instead of divisions).
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.meshes import smooth_state
from xyst_amd.meshgen import box_mesh, renumber_mesh, sedov_state, sod_state

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx():
    from xyst_amd import riemann as X
    return X.Context(0)


def _case(n, order="lex", seed=3):
    from xyst_amd import riemann as X
    m = box_mesh(n, seed=seed)
    if order != "lex":
        m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
        m = renumber_mesh(m, X.mesh_block_order(m["coord"], 2048))
    return m


@pytest.mark.parametrize("n,order,state", [(6, "lex", "smooth"), (12, "rcm+block", "smooth"),
                                           (16, "rcm+block", "sod"), (10, "lex", "sedov")])
def test_kozak_rhs_parity(n, order, state):
    from xyst_amd import kozak as K
    _say("kozak parity n=%d %s %s" % (n, order, state))
    m = _case(n, order)
    coord, inpoel = m["coord"], m["inpoel"]
    if state == "smooth":
        U = smooth_state(coord, seed=n)
    elif state == "sod":
        U = sod_state(coord)
    else:
        U = sedov_state(coord, O.vol(coord, inpoel))
    dt = 2e-3
    Ro, A = O.kozak_rhs(O.cfg(), coord, inpoel, dt, U, absacc=True)
    ctx = _ctx()
    K.upload(ctx, coord, inpoel)
    R = K.rhs(ctx, dt, U)
    err = np.abs(R - Ro)
    bad = err > TOL * A
    assert not bad.any(), "%d entries off, worst ratio %.3e" % (
        bad.sum(), float(np.max(err / np.maximum(A, 1e-300))))
    _say("  worst |dR|/A = %.3e" % float(np.max(err / np.maximum(A, 1e-300))))
    # total of the rhs vanishes (sum over corners of grad N = 0)
    assert np.all(np.abs(R.sum(0)) <= 1e-13 * A.sum(0))
    # bitwise reproducible
    assert np.array_equal(R, K.rhs(ctx, dt, U))
    ctx.close()


def test_kozak_rhs_dev_matches_host_entry():
    import torch
    from xyst_amd import kozak as K
    _say("kozak dev entry")
    m = _case(14, "rcm+block", seed=8)
    U = smooth_state(m["coord"], seed=2)
    ctx = _ctx()
    K.upload(ctx, m["coord"], m["inpoel"])
    R = K.rhs(ctx, 1e-3, U)
    dU = torch.from_numpy(U).cuda()
    dR = torch.full_like(dU, np.nan)
    K.rhs_dev(ctx, 1e-3, dU, dR)
    torch.cuda.synchronize()
    assert np.array_equal(dR.cpu().numpy(), R)
    ctx.close()


def test_kozak_errors():
    """No tet mesh uploaded -> XG_ENOMESH; ncomp other than 5 -> unsupported."""
    from xyst_amd._xg import XG_ENOMESH, XG_EUNSUPPORTED, lib
    from xyst_amd import kozak as K
    _say("kozak errors")
    ctx = _ctx()
    U = np.zeros((8, 5))
    assert lib().xg_kozak_rhs(ctx._h, 5, 1e-3, U.ctypes.data, U.ctypes.data) == XG_ENOMESH
    m = _case(2)
    K.upload(ctx, m["coord"], m["inpoel"])
    U = np.zeros((m["coord"].shape[1], 6))
    assert lib().xg_kozak_rhs(ctx._h, 6, 1e-3, U.ctypes.data, U.ctypes.data) == XG_EUNSUPPORTED
    ctx.close()
