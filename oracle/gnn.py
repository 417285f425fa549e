"""CPU oracle of the learned graph preconditioner -- TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference; this numpy restatement
(float64) is the checker of the device forward and backward kernels.

Network (per node i, 32 channels, A the system matrix with diagonal d; |d|
is used where a square root is taken, as the ChoCG matrix is -K, negative
definite):
  s    = ||D^-1/2 r|| / sqrt(N)                     (scale, one scalar)
  x_i  = r_i / (sqrt|d_i| s)
  h0_i = relu(we x_i + be)                          (encoder, 1 -> 32)
  for l = 1..3:
    m_i  = sum_j Ahat_ij h_{l-1,j},  Ahat = |D|^-1/2 A |D|^-1/2
    a_i  = W_l [h_{l-1,i}; m_i] + b_l               (64 -> 32)
    h_l  = h_{l-1} + relu(a_i)
  y_i  = wo . h3_i + bo                             (decoder, 32 -> 1)
  z_i  = r_i / d_i + s y_i / sqrt|d_i|   (a learned correction of Jacobi),
         r_i / d_i alone on decoupled rows (a single nonzero, the diagonal:
         the Dirichlet rows, where that is exact)

Loss of a pair (r, e) with A e = r, the energy-norm error of z relative to
that of e:  L = (z - e)^T A (z - e) / (e^T r) = (z^T A z - 2 z^T r) / (e^T r) + 1.
dL/dz = 2 (A z - r) / (e^T r).
"""
import numpy as np

C = 32
NL = 3


def init_params(seed=0, c=C):
    rng = np.random.default_rng(seed)
    p = {"we": rng.standard_normal(c) * 0.5, "be": rng.standard_normal(c) * 0.1}
    for l in range(NL):
        p["W%d" % l] = rng.standard_normal((c, 2 * c)) * (0.5 / np.sqrt(2 * c))
        p["b%d" % l] = np.zeros(c)
    p["wo"] = rng.standard_normal(c) * (0.5 / np.sqrt(c))
    p["bo"] = np.zeros(1)
    return p


def _ahat(A_dense, d):
    s = 1.0 / np.sqrt(np.abs(d))
    return A_dense * s[:, None] * s[None, :]


def forward(p, A_dense, d, r):
    """Returns z and the cache needed by backward."""
    n = r.size
    sq = np.sqrt(np.abs(d))
    s = np.sqrt(np.sum(r * r / np.abs(d)) / n)
    x = r / (sq * s)
    Ah = _ahat(A_dense, d)
    pre0 = np.outer(x, p["we"]) + p["be"]
    h = np.maximum(pre0, 0.0)
    cache = {"x": x, "s": s, "sq": sq, "Ah": Ah, "pre0": pre0, "h": [h], "m": [], "a": []}
    for l in range(NL):
        m = Ah @ h
        a = np.concatenate([h, m], 1) @ p["W%d" % l].T + p["b%d" % l]
        h = h + np.maximum(a, 0.0)
        cache["m"].append(m)
        cache["a"].append(a)
        cache["h"].append(h)
    y = h @ p["wo"] + p["bo"][0]
    dec = decoupled(A_dense)
    z = np.where(dec, r / d, r / d + s * y / sq)
    cache["y"] = y
    cache["dec"] = dec
    return z, cache


def decoupled(A_dense):
    """Rows whose only nonzero is the diagonal."""
    nz = (A_dense != 0.0).sum(1)
    return (nz == 1) & (np.diag(A_dense) != 0.0)


def loss(A_dense, r, e, z):
    etr = e @ r
    return float((z @ (A_dense @ z) - 2.0 * z @ r) / etr + 1.0)


def backward(p, cache, dz):
    """Gradients of the parameters given dL/dz (s is treated as a constant of
    the input, as the device does: it depends on r only)."""
    g = {}
    dy = np.where(cache["dec"], 0.0, dz * cache["s"] / cache["sq"])
    h3 = cache["h"][-1]
    g["wo"] = h3.T @ dy
    g["bo"] = np.array([dy.sum()])
    dh = np.outer(dy, p["wo"])
    Ah = cache["Ah"]
    for l in reversed(range(NL)):
        hprev = cache["h"][l]
        m = cache["m"][l]
        a = cache["a"][l]
        da = dh * (a > 0.0)
        g["W%d" % l] = da.T @ np.concatenate([hprev, m], 1)
        g["b%d" % l] = da.sum(0)
        dhm = da @ p["W%d" % l]
        dm = dhm[:, C:]
        dh = dh + dhm[:, :C] + Ah.T @ dm
    da0 = dh * (cache["pre0"] > 0.0)
    g["we"] = da0.T @ cache["x"]
    g["be"] = da0.sum(0)
    return g


def loss_grad(p, A_dense, d, r, e):
    z, cache = forward(p, A_dense, d, r)
    L = loss(A_dense, r, e, z)
    dz = 2.0 * (A_dense @ z - r) / (e @ r)
    return L, backward(p, cache, dz), z
