"""The learned preconditioner on the device (needs an MI355X): forward and
gradients against the numpy oracle (oracle/gnn.py), training lowers the pair
loss, and flexible CG with it reaches 1e-10.

Tolerances: activations are stored in BF16 (8 significant bits) with FP32
sums, so the device output matches the float64 oracle to a few 1e-3
relative; gradients are compared by relative norm error and cosine.
"""
import sys

import numpy as np
import pytest

import oracle as O
from oracle import gnn as OG
from xyst_amd.meshgen import box_mesh

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _system(n, seed=3):
    """Poisson system with p = 0 on x = 0, on the device and densely for the
    oracle."""
    from xyst_amd import cg as X
    m = box_mesh(n, seed=seed)
    coord, inpoel = m["coord"], m["inpoel"]
    A = X.CSR.poisson(coord, inpoel)
    npoin = coord.shape[1]
    dirbc = {int(i): [(1, 0.0)] for i in np.where(coord[0] < 1e-12)[0]}
    Ao = O.CSR(npoin, 1, A.ia.astype(np.int64), A.ja.astype(np.int64), A.a.copy())
    rr = np.zeros(npoin)
    for i in sorted(dirbc, reverse=True):
        Ao.dirichlet(i, 0.0, rr)
    M = Ao.dense()
    cg = X.ConjugateGradients(A)
    cg.init(np.zeros(npoin), np.ones(npoin), dirbc=dirbc, apply=True, pc="jacobi")
    return cg, M, coord, dirbc


def _pair(M, coord, seed):
    rng = np.random.default_rng(seed)
    x, y, z = coord
    e = np.sin(2 * x + rng.uniform()) * np.cos(3 * y) + 0.3 * rng.standard_normal(x.size)
    e[np.abs(np.diag(M) - 1.0) < 1e-15] = 0.0     # Dirichlet rows (diagonal 1)
    return M @ e, e


def test_forward_and_gradients_match_oracle():
    import torch
    from xyst_amd import gnn
    _say("gnn forward/backward parity")
    cg, M, coord, dirbc = _system(6)
    p = gnn.init_params(11)
    for k in p:
        p[k] = p[k] + 0.02
    net = gnn.LearnedPreconditioner(cg, params=p)
    net.prepare()
    r, e = _pair(M, coord, 1)
    d = np.diag(M).copy()
    zo, cache = OG.forward(p, M, d, r)
    Lo, go, _ = OG.loss_grad(p, M, d, r, e)
    dr = torch.from_numpy(r).cuda()
    de = torch.from_numpy(e).cuda()
    dz = torch.empty_like(dr)
    net.apply_dev(dr, dz)
    torch.cuda.synchronize()
    z = dz.cpu().numpy()
    rel = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    _say("  forward rel err %.2e" % rel)
    assert rel < 2e-2
    L = net.loss_grad_dev(dr, de, dz)
    _say("  loss dev %.6f oracle %.6f" % (L, Lo))
    assert L == pytest.approx(Lo, rel=2e-2, abs=2e-3)
    g = gnn.unflatten(net.grad())
    for k in go:
        a, b = np.asarray(g[k]).reshape(-1), np.asarray(go[k]).reshape(-1)
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.linalg.norm(a) < 1e-6
            continue
        cos = a @ b / (np.linalg.norm(a) * nb)
        err = np.linalg.norm(a - b) / nb
        _say("  grad %-3s cos %.5f rel %.2e" % (k, cos, err))
        assert cos > 0.995 and err < 0.1, k


def test_training_lowers_loss_and_fcg_converges():
    """Training on solves of a 16^3 system lowers the loss of a held-out
    residual/error pair; flexible CG with the trained network reaches 1e-10
    (true residual)."""
    import torch
    from xyst_amd import gnn
    _say("gnn training + fcg")
    cg, M, coord, dirbc = _system(16, seed=5)
    n = coord.shape[1]
    net = gnn.LearnedPreconditioner(cg, seed=2)
    x, y, z = coord
    # held-out pair from the training distribution: the residual and error
    # of 12 Jacobi-PCG iterations on a right-hand side not trained on
    bh = (np.cos(3 * x) * y + z * z) * 1e4
    bbh = bh.copy()
    for i in dirbc:
        bbh[i] = 0.0
    cg.init(np.zeros(n), bh, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve(12, 1e-30)
    x12 = cg.solution()
    xs = np.linalg.solve(M, bbh)
    e = xs - x12
    r = M @ e
    dr, de = torch.from_numpy(r).cuda(), torch.from_numpy(e).cuda()
    dz = torch.empty_like(dr)

    def held_out_loss():
        net.prepare()
        L = net.loss_grad_dev(dr, de, dz)
        net.set_params(net.params())        # drop the accumulated gradient
        return L

    rng = np.random.default_rng(0)
    cg.init(np.zeros(n), np.ones(n), dirbc=dirbc, apply=True, pc="jacobi")
    L0 = held_out_loss()
    for s in range(6):
        b = sum(np.sin((k + 1) * np.pi * x + rng.uniform(0, 6)) *
                np.cos(rng.integers(1, 4) * np.pi * y) for k in range(3)) * 1e4
        cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
        L, npairs = net.train_system(stride=4, max_pairs=60, batch=4, lr=2e-3)
        _say("  system %d: mean loss %.4f over %d pairs" % (s, L, npairs))
        assert npairs > 5 and L > -1e-6    # an energy-norm ratio: never negative
    cg.init(np.zeros(n), np.ones(n), dirbc=dirbc, apply=True, pc="jacobi")
    L1 = held_out_loss()
    _say("  held-out loss %.4f -> %.4f" % (L0, L1))
    assert L1 < L0
    b = (np.sin(2 * np.pi * x) * z + y) * 1e4
    cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve(20000, 1e-10)
    it_pcg = cg.it
    cg.init(np.zeros(n), b, dirbc=dirbc, apply=True, pc="jacobi")
    cg.solve_flexible(5000, 1e-10, pc="gnn")
    _say("  pcg %d it, fcg+gnn %d it, %d fallbacks" % (it_pcg, cg.it, cg.inner_total))
    assert cg.converged
    xs = cg.solution()
    rr = np.zeros(n)
    bb = b.copy()
    Ab = O.CSR(n, 1, cg.A.ia.astype(np.int64), cg.A.ja.astype(np.int64), cg.A.a.copy())
    for i in sorted(dirbc, reverse=True):
        Ab.dirichlet(i, 0.0, rr)
    bb -= rr
    for i in dirbc:
        bb[i] = 0.0
    assert np.linalg.norm(bb - Ab.mult(xs)) < 2e-10 * np.linalg.norm(bb)
