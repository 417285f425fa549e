"""xg_diag_rho_dev_noalloc (the diagnostics sums without any device
allocation, for callers that cannot allocate) against xg_diag_rho_dev on the
same inputs: bit for bit, on two calls in a row with the second on a
smaller input (fewer workgroup partials than the first left in the scratch
array), for every supported ncomp; and the device's free memory is the same
before and after the calls.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _p(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("ncomp", [5, 3, 1])
def test_noalloc_diag_matches_the_allocating_entry(ncomp):
    import torch
    from xyst_amd import riemann as X
    from xyst_amd._xg import lib
    _say("noalloc diag ncomp=%d" % ncomp)
    ctx = X.Context(0)
    rng = np.random.default_rng(ncomp)
    ns = 2 * ncomp + 1
    cases = []
    for n in (700_001, 4_999):       # 1024 workgroup partials, then 20
        U = torch.from_numpy(rng.standard_normal((n, ncomp))).cuda()
        Un = torch.from_numpy(rng.standard_normal((n, ncomp))).cuda()
        V = torch.from_numpy(rng.uniform(0.5, 1.5, n)).cuda()
        cases.append((n, U, Un, V))
    outs = [torch.full((ns,), np.nan, dtype=torch.float64, device="cuda") for _ in cases]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for (n, U, Un, V), out in zip(cases, outs):
        rc = lib().xg_diag_rho_dev_noalloc(ctx._h, n, ncomp, _p(U), _p(Un), _p(V), _p(out))
        assert rc == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    for (n, U, Un, V), out in zip(cases, outs):
        ref = np.empty(ns)
        rc = lib().xg_diag_rho_dev(ctx._h, n, ncomp, _p(U), _p(Un), _p(V),
                                   ref.ctypes.data_as(C.c_void_p))
        assert rc == 0
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (n, got, ref)
        # and the sums are the sums (first component, float64 on the host)
        u, un, v = U.cpu().numpy(), Un.cpu().numpy(), V.cpu().numpy()
        np.testing.assert_allclose(got[0], (u[:, 0] ** 2 * v).sum(), rtol=1e-12)
        np.testing.assert_allclose(got[ncomp], ((u[:, 0] - un[:, 0]) ** 2 * v).sum(),
                                   rtol=1e-12)
    ctx.close()


def test_noalloc_diag_rejects_bad_arguments():
    import torch
    from xyst_amd import riemann as X
    from xyst_amd._xg import XG_EINVAL, XG_EUNSUPPORTED, lib
    ctx = X.Context(0)
    U = torch.zeros((8, 6), dtype=torch.float64, device="cuda")
    out = torch.zeros(13, dtype=torch.float64, device="cuda")
    assert lib().xg_diag_rho_dev_noalloc(ctx._h, 8, 6, _p(U), _p(U), _p(U),
                                         _p(out)) == XG_EUNSUPPORTED
    assert lib().xg_diag_rho_dev_noalloc(ctx._h, 8, 5, _p(U), None, _p(U),
                                         _p(out)) == XG_EINVAL
    ctx.close()
