"""The reference's ChoCG/Poisson cases on the device (needs an MI355X): the
pressure matrix, the Neumann vector of the host C-ABI and the start-up's
first pressure solve (xg_chocg_poisson_dev), with the diagnostics from the
device sums (precompute, and the pressure error of xg_diag_err_dev).

Compared with the oracle (same iteration count; the solution to 1e-11 of
max |p| where the solve does not depend on the summation order, 1e-6 for
the unpreconditioned Neumann case, which amplifies round-off by ~1e8) and
with the reference's diag_poisson_*.std and .exo fields in the bands of
tests/test_oracle_chocg_poisson.py.
"""
import sys

import numpy as np
import pytest

from tests.test_oracle_chocg_poisson import (CASES, TIGHT, check_diag, check_fields,
                                             fixture, mesh, setup)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _device(d, case):
    import torch
    from xyst_amd import chorin as X
    from xyst_amd import riemann as R
    from xyst_amd.cg import CSR, ConjugateGradients
    r, pr, neubc_o, ic = setup(d, case)
    key, problem, pc, dirsets, neusets = CASES[case]
    coord, inpoel, tri, faceset = mesh(d, key)
    ctx = R.Context(0)
    X.upload(ctx, coord, r.se.sup, r.se.int, tri)
    cg = ConjugateGradients(CSR.poisson(coord, inpoel), ctx=ctx)
    pd = np.array(sorted(r.pdir), np.int64)
    X.step_setup(ctx, r.vol, X.config(flux="damp2", stab=False), p_pc=pc,
                 p_maxit=int(d["p_iter"]), p_tol=float(d["p_tol"]), pdirnodes=pd,
                 pdirvals=ic[pd])
    neubc = None
    if neubc_o is not None:
        neubc = X.neumann(coord, tri, np.isin(faceset, neusets), _pg(problem, coord))
        assert np.array_equal(neubc, neubc_o)
        neubc = torch.tensor(neubc, device="cuda")
    U = torch.zeros((coord.shape[1], 3), dtype=torch.float64, device="cuda")
    P = torch.zeros(coord.shape[1], dtype=torch.float64, device="cuda")
    dpr = torch.tensor(pr, device="cuda")
    return ctx, cg, r, pr, neubc_o, ic, U, P, dpr, neubc


def _pg(problem, coord):
    from oracle.chocg_step import pressure_problem
    return pressure_problem(problem, coord)[2]


@pytest.mark.parametrize("case", sorted(CASES))
def test_chocg_poisson_device(case):
    import torch
    from xyst_amd import chorin as X
    from xyst_amd import diagnostics as D
    _say("chocg poisson %s" % case)
    d = fixture()
    ctx, cg, r, pr, neubc_o, ic, U, P, dpr, neubc = _device(d, case)
    it = X.poisson_dev(ctx, U, P, dpr, neubc)
    sol = r.poisson(pr, neubc_o)
    assert it == r.pit[0], (it, r.pit)
    torch.cuda.synchronize()
    p, u = P.cpu().numpy(), U.cpu().numpy()
    tol = 1e-11 if case in TIGHT else 1e-6
    err = float(np.abs(p - sol).max() / np.abs(sol).max())
    _say("  it %d |dp|/max|p| vs oracle %.3e" % (it, err))
    assert err <= tol
    assert float(np.abs(u - r.U).max() / np.abs(r.U).max()) <= 1e3 * tol
    # diagnostics from the device sums
    V = torch.tensor(r.vol, device="cuda")
    s = D.precompute(ctx, U, U, P, P, V)
    e = D.errcompute(ctx, P, V, torch.tensor(ic, device="cuda"))
    tot = r.vol.sum()
    row = np.concatenate([[1, float(d["dt"]), float(d["dt"])],
                          np.sqrt([s[0] / tot, s[4] / tot, e[0] / tot]), [e[1] / tot]])
    np.testing.assert_allclose(row[3:], r.pressure_diag(sol, ic), rtol=1e3 * tol)
    check_diag(d, case, row)
    check_fields(d, case, p, ic)
    ctx.close()


def test_chocg_poisson_errors():
    """Null fields and wrongly shaped tensors are refused; without a
    pressure rhs the solve takes the divergence of U (zero here: p = the
    Dirichlet lift, the harmonic extension of the boundary values)."""
    import torch
    from xyst_amd import chorin as X
    from xyst_amd._xg import XG_EINVAL, lib
    _say("chocg poisson errors")
    d = fixture()
    ctx, cg, r, pr, neubc_o, ic, U, P, dpr, neubc = _device(d, "const")
    assert lib().xg_chocg_poisson_dev(ctx._h, None, None, None, None, None) == XG_EINVAL
    with pytest.raises(ValueError):
        X.poisson_dev(ctx, U[:-1], P, dpr)
    with pytest.raises(ValueError):
        X.poisson_dev(ctx, U, P, dpr[:-1])
    it = X.poisson_dev(ctx, U, P)
    sol = r.poisson(None, None)
    assert it == r.pit[0]
    torch.cuda.synchronize()
    err = float(np.abs(P.cpu().numpy() - sol).max() / np.abs(sol).max())
    assert err <= 1e-11, err
    ctx.close()
