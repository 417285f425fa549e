"""The device node diagnostics against the oracle's (needs an MI355X).

Sums of non-negative terms (and of U_4 V) in another order: relative
agreement to 1e-13."""
import sys

import numpy as np
import pytest

import oracle as O
from tests.meshes import smooth_state
from xyst_amd.meshgen import box_mesh

pytestmark = pytest.mark.gpu


def test_rhocompute_and_precompute_match_oracle():
    import torch
    from xyst_amd import diagnostics as D
    from xyst_amd import riemann as R
    print("diagnostics", flush=True)
    sys.stdout.flush()
    m = box_mesh(20, seed=5)
    coord, inpoel = m["coord"], m["inpoel"]
    V = O.vol(coord, inpoel)
    U = smooth_state(coord, seed=1)
    Un = smooth_state(coord, seed=2)
    ctx = R.Context(0)
    dU, dUn, dV = (torch.tensor(a, device="cuda") for a in (U, Un, V))
    got = D.rhocompute(ctx, dU, dUn, dV)
    ref = O.diag(U, Un, V)
    np.testing.assert_allclose(got, ref, rtol=1e-13)
    # the three-component form has no energy term
    got3 = D.rhocompute(ctx, dU[:, :3].contiguous(), dUn[:, :3].contiguous(), dV)
    np.testing.assert_allclose(got3[:6], np.concatenate([(U[:, :3] ** 2 * V[:, None]).sum(0),
                               ((U[:, :3] - Un[:, :3]) ** 2 * V[:, None]).sum(0)]), rtol=1e-13)
    assert got3[6] == 0.0
    P, dP = U[:, 4].copy(), Un[:, 0].copy()
    pre = D.precompute(ctx, dU[:, 1:4].contiguous(), dUn[:, 1:4].contiguous(),
                       torch.tensor(P, device="cuda"), torch.tensor(dP, device="cuda"), dV)
    u, un = U[:, 1:4], Un[:, 1:4]
    want = np.concatenate([[np.sum(P * P * V)], (u ** 2 * V[:, None]).sum(0),
                           [np.sum(dP * dP * V)], ((u - un) ** 2 * V[:, None]).sum(0)])
    np.testing.assert_allclose(pre, want, rtol=1e-13)
    # deterministic
    assert np.array_equal(got, D.rhocompute(ctx, dU, dUn, dV))
    ctx.close()
