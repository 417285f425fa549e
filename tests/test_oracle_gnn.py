"""Pins the numpy oracle of the learned preconditioner: its gradients against
central finite differences of its loss, the loss identity against the
energy-norm definition, and the scale equivariance of the map r -> z."""
import numpy as np
import pytest

import oracle as O
from oracle import gnn as G
from xyst_amd.meshgen import box_mesh


def _system(n=3, seed=4):
    m = box_mesh(n, seed=seed)
    coord, inpoel = m["coord"], m["inpoel"]
    A = O.CSR.from_tets(coord.shape[1], inpoel)
    O.prelhs(coord, inpoel, A)
    A.a = -A.a                       # K: symmetric positive semi-definite
    M = A.dense()
    M += np.eye(M.shape[0]) * 1e-2 * np.abs(np.diag(M)).mean()   # make it definite
    d = np.diag(M).copy()
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(M.shape[0])
    r = M @ e
    return M, d, r, e


def test_loss_is_relative_energy_error():
    M, d, r, e = _system()
    p = G.init_params(1)
    z, _ = G.forward(p, M, d, r)
    L = G.loss(M, r, e, z)
    assert L == pytest.approx((z - e) @ M @ (z - e) / (e @ M @ e), rel=1e-10)
    assert G.loss(M, r, e, e) == pytest.approx(0.0, abs=1e-12)
    assert G.loss(M, r, e, np.zeros_like(e)) == pytest.approx(1.0, rel=1e-12)


def test_gradients_match_finite_differences():
    M, d, r, e = _system()
    p = G.init_params(3)
    for k in p:
        p[k] = p[k] + 0.05          # biases away from relu kinks
    L, g, _ = G.loss_grad(p, M, d, r, e)
    rng = np.random.default_rng(7)
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            h = 1e-6 * max(1.0, abs(flat[idx]))
            old = flat[idx]
            flat[idx] = old + h
            Lp, _, _ = G.loss_grad(p, M, d, r, e)
            flat[idx] = old - h
            Lm, _, _ = G.loss_grad(p, M, d, r, e)
            flat[idx] = old
            fd = (Lp - Lm) / (2 * h)
            an = g[k].reshape(-1)[idx]
            assert an == pytest.approx(fd, rel=1e-4, abs=1e-8), (k, idx)


def test_positive_scale_equivariance():
    M, d, r, e = _system()
    p = G.init_params(5)
    z1, _ = G.forward(p, M, d, r)
    z2, _ = G.forward(p, M, d, 37.0 * r)
    np.testing.assert_allclose(z2, 37.0 * z1, rtol=1e-12)
