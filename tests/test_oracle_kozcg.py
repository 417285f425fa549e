"""Pins the CPU oracle of kozak::rhs (Kozak.cpp:29-180).

The checks do not restate the oracle's loops: the Taylor-Galerkin rhs is
rebuilt from its definition as a P1 Galerkin integral, with the shape
function gradients of every tet obtained by inverting its 4x4 barycentric
system (not by the cross products the reference uses); free-stream and
conservation invariants; and the round-off-magnitude variant must reproduce
the plain evaluation bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from tests.meshes import smooth_state
from xyst_amd.meshgen import box_mesh

GAMMA = 1.4


def _euler_flux(U, g=GAMMA):
    """F[j] (npoin, 5) for j = x, y, z of the Euler equations."""
    r, ru, rv, rw, rE = U.T
    p = (g - 1.0) * (rE - 0.5 * (ru * ru + rv * rv + rw * rw) / r)
    vel = np.stack([ru / r, rv / r, rw / r])
    F = []
    for j in range(3):
        f = np.stack([U[:, j + 1], ru * vel[j], rv * vel[j], rw * vel[j],
                      (rE + p) * vel[j]], axis=1)
        f[:, j + 1] += p
        F.append(f)
    return F


def _galerkin_tg(coord, inpoel, U, dt, g=GAMMA):
    """Two-step Taylor-Galerkin rhs from its definition: per tet e with
    volume V and P1 basis gradients dN_a (from the inverse of the barycentric
    matrix), U_e = mean(U_a) - dt/2 sum_a dN_a . F(U_a), and
    R_a += V dN_a . F(U_e) = int_e grad N_a . F(U_e)."""
    R = np.zeros_like(U)
    X = coord.T
    for e in range(inpoel.shape[0]):
        N = inpoel[e]
        M = np.ones((4, 4))
        M[:, 1:] = X[N]
        Minv = np.linalg.inv(M)            # columns: barycentric coefficients
        dN = Minv[1:, :].T                 # (4, 3): gradient of N_a
        V = abs(np.linalg.det(M)) / 6.0
        Fa = _euler_flux(U[N], g)
        ue = U[N].mean(0)
        for j in range(3):
            ue = ue - 0.5 * dt * (dN[:, j:j + 1] * Fa[j]).sum(0)
        Fe = _euler_flux(ue[None, :], g)
        for j in range(3):
            R[N] += V * dN[:, j:j + 1] * Fe[j][0][None, :]
    return R


@pytest.mark.parametrize("dt", [0.0, 2e-3])
def test_kozak_rhs_is_the_galerkin_tg_integral(dt):
    m = box_mesh(3, seed=11)
    coord, inpoel = m["coord"], m["inpoel"]
    U = smooth_state(coord, seed=3)
    R = O.kozak_rhs(O.cfg(), coord, inpoel, dt, U)
    Rg = _galerkin_tg(coord, inpoel, U, dt)
    scale = np.abs(Rg).max()
    np.testing.assert_allclose(R, Rg, rtol=0, atol=1e-13 * scale)


def test_kozak_rhs_mag_reproduces_values():
    m = box_mesh(4, seed=5)
    coord, inpoel = m["coord"], m["inpoel"]
    U = smooth_state(coord, seed=8)
    R = O.kozak_rhs(O.cfg(), coord, inpoel, 1e-3, U)
    R2, A = O.kozak_rhs(O.cfg(), coord, inpoel, 1e-3, U, absacc=True)
    assert np.array_equal(R, R2)
    assert np.all(A >= np.abs(R) - 1e-300)


def test_kozak_free_stream_and_conservation():
    m = box_mesh(5, seed=2)
    coord, inpoel = m["coord"], m["inpoel"]
    n = coord.shape[1]
    U = np.tile([1.2, 0.3, -0.2, 0.1, 2.5], (n, 1))
    R, A = O.kozak_rhs(O.cfg(), coord, inpoel, 1e-3, U, absacc=True)
    x, y, z = coord
    interior = (x > 1e-9) & (x < 1 - 1e-9) & (y > 1e-9) & (y < 1 - 1e-9) & \
               (z > 1e-9) & (z < 1 - 1e-9)
    assert np.all(np.abs(R[interior]) <= 1e-13 * A[interior])
    # sum_a grad N_a = 0 per tet: the total rhs vanishes for any state
    U = smooth_state(coord, seed=4)
    R, A = O.kozak_rhs(O.cfg(), coord, inpoel, 3e-3, U, absacc=True)
    assert np.all(np.abs(R.sum(0)) <= 1e-13 * A.sum(0))


def test_kozak_steady_dtp_equals_uniform_dt():
    m = box_mesh(3, seed=1)
    coord, inpoel = m["coord"], m["inpoel"]
    U = smooth_state(coord, seed=6)
    dtp = np.full(coord.shape[1], 2.5e-3)
    R1 = O.kozak_rhs(O.cfg(), coord, inpoel, 2.5e-3, U)
    R2 = O.kozak_rhs(O.cfg(), coord, inpoel, 99.0, U, dtp=dtp)
    assert np.array_equal(R1, R2)


def _prim_fields(U, gamma=GAMMA):
    r = U[:, 0]
    u, v, w = U[:, 1] / r, U[:, 2] / r, U[:, 3] / r
    e = U[:, 4] / r
    p = (gamma - 1.0) * r * (e - 0.5 * (u * u + v * v + w * w))
    return np.stack([r, u, v, w, e, p], 1)


def kozcg_sod_run(step=None):
    """10 KozCG steps of the reference's KozCG/Sod case; `step` defaults to
    the oracle's.  Yields (it, t, dt, U_new, U_old) per step."""
    from tests.meshes import dirichlet_nodes, sod_fixture
    from xyst_amd.meshgen import sod_state
    c, d = sod_fixture("kozcg_sod.npz")
    U = sod_state(c.coord)
    cfg = O.cfg(float(d["gamma"]))
    f = O.fctcfg(True, float(d["fctdif"]), bool(d["fctclip"]), list(d["fctsys"]))
    dn = dirichlet_nodes(c, d["dirsets"])
    dv = U[dn].copy()                  # Sod's initial state is time-independent
    step = step or (lambda dt, U: O.kozcg_step(cfg, f, c.coord, c.inpoel, c.vol, dt,
                                               c.symnodes, c.symnorms, dn, dv, U))
    t = 0.0
    for it in range(int(d["nstep"])):
        dt = O.dt(cfg, U, c.vol, float(d["cfl"]))
        U, Un = step(dt, U)
        t += dt
        yield it, t, dt, U, Un, c, d


def test_sod_regression_fixture():
    """10 KozCG steps (Taylor-Galerkin rhs, FCT with fctclip and fctsys
    {1,2,5}, Dirichlet on side sets 1 and 3, symmetry on 2,4,5,6) against
    diag.std and sod.std.exo of tests/regression/inciter/KozCG/Sod."""
    for it, t, dt, U, Un, c, d in kozcg_sod_run():
        V = c.vol.sum()
        dg = O.diag(U, Un, c.vol)
        mine = np.concatenate([[it + 1, t, dt], np.sqrt(dg[:10] / V), [dg[10]]])
        np.testing.assert_allclose(mine, d["diag"][it], rtol=2e-8, atol=1e-14)
    assert abs(t - float(d["t_final"])) < 1e-8
    mine = _prim_fields(U)
    ref = d["fields"]
    for i, name in enumerate(d["field_names"]):
        diff = np.abs(mine[:, i] - ref[:, i])
        if name.startswith("velocity"):
            ok = (diff <= 1e-8) | (np.abs(ref[:, i]) < 1e-7)
        else:
            ok = (diff <= 1e-7 * np.abs(ref[:, i])) | (np.abs(ref[:, i]) < 1e-8)
        assert ok.all(), (name, diff.max())


def test_fct_off_is_forward_euler_and_limited_step_is_bounded():
    """Without FCT the step is U + dt R/vol; with FCT and fctclip each new
    nodal value lies within the low-order solution's bounds over the tets
    around the node (the defining property of the limiter), checked on a
    jittered box with a discontinuous state."""
    m = box_mesh(3, jitter=0.1, seed=11)
    coord, inpoel = m["coord"], m["inpoel"]
    vol = O.vol(coord, inpoel)
    U = smooth_state(coord)
    U[coord[0] > 0.5, 0] *= 0.3
    U[coord[0] > 0.5, 4] *= 0.2
    cfg = O.cfg(GAMMA)
    dt = O.dt(cfg, U, vol, 0.5)
    none = np.zeros(0, np.int64)
    off = O.fctcfg(False)
    U1, _ = O.kozcg_step(cfg, off, coord, inpoel, vol, dt, none, np.zeros(0), none,
                         np.zeros(0), U)
    R = O.kozak_rhs(cfg, coord, inpoel, dt, U)
    np.testing.assert_allclose(U1, U + dt * R / vol[:, None], rtol=0, atol=1e-15)
    on = O.fctcfg(True, 1.0, True, [])
    U2, _ = O.kozcg_step(cfg, on, coord, inpoel, vol, dt, none, np.zeros(0), none,
                         np.zeros(0), U)
    # low-order solution, restated from its definition: mass-lumped minus the
    # consistent-minus-lumped mass difference of U
    ML = np.zeros_like(U)
    for t in inpoel:
        J = np.linalg.det(np.c_[coord[:, t[1:]].T - coord[:, t[0]]])
        Mdiff = J / 120.0 * (4.0 * np.eye(4) - np.ones((4, 4)))
        ML[t] += Mdiff @ U[t]
    ul = U + dt * R / vol[:, None] - ML / vol[:, None]
    hi = np.full_like(U, -np.inf)
    lo = np.full_like(U, np.inf)
    for t in inpoel:
        hi[t] = np.maximum(hi[t], ul[t].max(0))
        lo[t] = np.minimum(lo[t], ul[t].min(0))
    tol = 1e-12 * np.abs(U).max(0)
    assert (U2 <= hi + tol).all() and (U2 >= lo - tol).all()
    assert not np.allclose(U2, ul)        # the limiter did add antidiffusion
