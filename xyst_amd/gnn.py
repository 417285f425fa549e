"""Python mirror of the learned graph preconditioner over the C-ABI (no
counterpart in the reference; see include/xg.h, xg_gnn_*).

`LearnedPreconditioner(cg)` attaches a network to a device
ConjugateGradients system; parameters travel as a flat float64 array of
NPARAM values (we[32] be[32], per layer W[32][64] b[32], wo[32] bo[1]).
"""
import ctypes as C

import numpy as np

from ._xg import lib, ptr

NPARAM = 6337
CH = 32
NLAYER = 3

__all__ = ["LearnedPreconditioner", "NPARAM", "init_params", "flatten", "unflatten"]


def init_params(seed=0):
    """Initial weights: encoder N(0, 0.5), layers N(0, 0.5/sqrt(64)) with zero
    biases, a small decoder N(0, 0.05/sqrt(32)), so the untrained
    preconditioner is close to Jacobi."""
    rng = np.random.default_rng(seed)
    p = {"we": rng.standard_normal(CH) * 0.5, "be": rng.standard_normal(CH) * 0.1}
    for l in range(NLAYER):
        p["W%d" % l] = rng.standard_normal((CH, 2 * CH)) * (0.5 / np.sqrt(2 * CH))
        p["b%d" % l] = np.zeros(CH)
    p["wo"] = rng.standard_normal(CH) * (0.05 / np.sqrt(CH))
    p["bo"] = np.zeros(1)
    return p


def flatten(p):
    parts = [p["we"], p["be"]]
    for l in range(NLAYER):
        parts += [np.asarray(p["W%d" % l]).reshape(-1), p["b%d" % l]]
    parts += [p["wo"], np.asarray(p["bo"]).reshape(-1)]
    out = np.concatenate([np.asarray(a, np.float64).reshape(-1) for a in parts])
    assert out.size == NPARAM
    return out


def unflatten(v):
    v = np.asarray(v, np.float64)
    p = {"we": v[0:32].copy(), "be": v[32:64].copy()}
    o = 64
    for l in range(NLAYER):
        p["W%d" % l] = v[o:o + 2048].reshape(CH, 2 * CH).copy()
        p["b%d" % l] = v[o + 2048:o + 2080].copy()
        o += 2080
    p["wo"] = v[o:o + 32].copy()
    p["bo"] = v[o + 32:o + 33].copy()
    return p


class LearnedPreconditioner:
    """A network bound to the system of a xyst_amd.cg.ConjugateGradients."""

    def __init__(self, cg, params=None, seed=0):
        self.cg = cg
        self.ctx = cg.ctx
        self.ctx._check(lib().xg_gnn_create(self.ctx._h, cg.n, cg.A.ja.size))
        self.set_params(init_params(seed) if params is None else params)

    def set_params(self, p):
        v = flatten(p) if isinstance(p, dict) else np.ascontiguousarray(p, np.float64)
        self.ctx._check(lib().xg_gnn_set_params(self.ctx._h, ptr(v), v.size))

    def params(self):
        v = np.empty(NPARAM)
        self.ctx._check(lib().xg_gnn_get_params(self.ctx._h, ptr(v), None, NPARAM))
        return v

    def grad(self):
        g = np.empty(NPARAM)
        self.ctx._check(lib().xg_gnn_get_params(self.ctx._h, None, ptr(g), NPARAM))
        return g

    def prepare(self):
        """Normalised matrix of the system as currently initialised."""
        self.ctx._check(lib().xg_gnn_prepare(self.ctx._h))

    def apply_dev(self, r, z):
        """z = network(r) for device tensors (float64, n)."""
        self.ctx._check(lib().xg_gnn_apply_dev(self.ctx._h, ptr(r), ptr(z)))

    def loss_grad_dev(self, r, e, z):
        loss = C.c_double()
        self.ctx._check(lib().xg_gnn_loss_grad_dev(self.ctx._h, ptr(r), ptr(e), ptr(z),
                                                   C.byref(loss)))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx._check(lib().xg_gnn_adam(self.ctx._h, lr, beta1, beta2, eps))

    def train_system(self, stride=10, max_pairs=100000, batch=8, lr=1e-3):
        """Train on the solves of the system as currently initialised;
        returns (mean pair loss, number of pairs)."""
        L, n = C.c_double(), C.c_int64()
        self.ctx._check(lib().xg_gnn_train_system(self.ctx._h, stride, max_pairs, batch,
                                                  lr, C.byref(L), C.byref(n)))
        return L.value, n.value
