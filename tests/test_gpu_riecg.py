"""Parity of the HIP RieCG path with the CPU oracle (needs an MI355X).

Tolerance: FP64, per node and component |x_gpu - x_oracle| <= 1e-12 * A,
where A is the oracle's round-off scale of that entry: the summed magnitudes
of the products every added term is made of (the GPU evaluates the same
terms with fused multiply-adds and sums them in a different order).
"""
import sys

import numpy as np
import pytest

import oracle as O
from tests.meshes import Case, box_case, smooth_state, sod_fixture
from xyst_amd.meshgen import box_mesh, sod_state

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(**kw):
    from xyst_amd import riemann as X
    return X.Context(0, **kw)


def _assert_close(x, ref, A, what):
    err = np.abs(x - ref)
    bound = TOL * A + 1e-300
    bad = err > bound
    assert not bad.any(), "%s: %d entries off, worst ratio %.3e" % (
        what, bad.sum(), float(np.max(err / np.maximum(A, 1e-300))))


def _parity(c, U, flux="rusanov", stab2=False, besym=None, label=""):
    besym = c.besym if besym is None else besym
    _say("parity %s flux=%s stab2=%s npoin=%d" % (label, flux, stab2, c.npoin))
    ctx = _ctx(flux=flux, stab2=stab2, stab2coef=0.2)
    ctx.upload_mesh(c.coord, c.se.sup, c.se.int, c.tri, besym)
    G = ctx.grad(U)
    Go, Ag = O.grad(c.se, c.coord, c.tri, U, absacc=True)
    _assert_close(G, Go, Ag, "grad")
    Gv = Go / c.vol[:, None]
    R = ctx.rhs(Gv, U)
    Ro, Ar = O.rhs(O.cfg(flux=flux, stab2=stab2, stab2coef=0.2), c.se, c.coord,
                   c.tri, besym, Gv, U, absacc=True)
    _assert_close(R, Ro, Ar, "rhs")
    ctx.close()


@pytest.mark.parametrize("flux", ["rusanov", "hllc"])
def test_parity_sod_fixture_mesh(flux):
    c, _ = sod_fixture()
    U = sod_state(c.coord)
    U[:, 2] = 0.05 * U[:, 0] * np.cos(7 * c.coord[0])
    _parity(c, U, flux=flux, label="sod-fixture")


@pytest.mark.parametrize("flux,stab2", [("rusanov", False), ("rusanov", True),
                                        ("hllc", False), ("hllc", True)])
def test_parity_box_smooth(flux, stab2):
    c = box_case(8, seed=21)
    _parity(c, smooth_state(c.coord), flux=flux, stab2=stab2, label="box8-smooth")


def test_parity_box_shock_no_symmetry():
    c = box_case(10, seed=5)
    U = sod_state(c.coord)
    _parity(c, U, besym=np.zeros_like(c.besym), label="box10-sod")


def test_deterministic_bitwise():
    import torch
    c = box_case(12, seed=6)
    U = smooth_state(c.coord)
    ctx = _ctx()
    ctx.upload_mesh(c.coord, c.se.sup, c.se.int, c.tri, c.besym)
    dU = torch.from_numpy(U).cuda()
    G1 = torch.empty((c.npoin, 15), dtype=torch.float64, device="cuda")
    G2 = torch.empty_like(G1)
    R1 = torch.empty_like(dU)
    R2 = torch.empty_like(dU)
    for G, R in ((G1, R1), (G2, R2)):
        ctx.grad_dev(dU, G)
        ctx.rhs_dev(G, dU, R)
    torch.cuda.synchronize()
    assert torch.equal(G1, G2) and torch.equal(R1, R2)
    ctx.close()


def _resident_case(n, seed):
    """Mesh preprocessing done by the product's host code (domint edges as a
    single-edge superedge group), as the resident RieCG path uses it."""
    from xyst_amd import riemann as X
    m = box_mesh(n, seed=seed)
    edge, dint = X.mesh_domint(m["coord"], m["inpoel"])
    sup, ints = X.edges_as_superedges(edge, dint)
    vol = X.mesh_vol(m["coord"], m["inpoel"])
    sn, sv, besym = X.mesh_symbc(m["coord"], m["triinpoel"], m["faceset"],
                                 [1, 2, 3, 4, 5, 6])
    se = O.Superedges([s.astype(np.int64) for s in sup], ints)
    return m, se, vol, sn, sv, besym


def test_resident_riecg_sod_20_steps():
    """Config: 32^3 jittered box, Sod along x, symmetry on all faces, 20 RK3
    steps at CFL 0.5 -- every dt, the state and conserved totals."""
    import torch
    n, nstep = 32, 20
    m, se, vol, sn, sv, besym = _resident_case(n, 20261202)
    coord, tri = m["coord"], m["triinpoel"]
    _say("resident sod n=%d" % n)
    ctx = _ctx()
    ctx.upload_mesh(coord, se.sup, se.int, tri, besym)
    ctx.riecg_setup(vol, sn, sv)
    U0 = sod_state(coord)
    Uo = U0.copy()
    dU = torch.from_numpy(U0.copy()).cuda()
    dUn, dR = torch.empty_like(dU), torch.empty_like(dU)
    dG = torch.empty((dU.shape[0], 15), dtype=torch.float64, device="cuda")
    cfg = O.cfg()
    tot0 = (U0 * vol[:, None]).sum(0)
    for it in range(nstep):
        dt = ctx.dt_dev(dU, 0.5)
        dto = O.dt(cfg, Uo, vol, 0.5)
        assert dt == pytest.approx(dto, rel=1e-11), it
        for s in range(3):
            ctx.stage_dev(s, dto, dU, dUn, dG, dR)
        O.riecg_step(cfg, se, coord, tri, besym, vol, sn, sv, dto, Uo)
        Ug = dU.cpu().numpy()
        scale = np.abs(Uo).max(0)
        err = np.abs(Ug - Uo).max(0) / scale
        assert err.max() < 1e-10, (it, err)
    Ug = dU.cpu().numpy()
    tot = (Ug * vol[:, None]).sum(0)
    # mass and total energy are conserved with every wall a symmetry wall
    assert abs(tot[0] - tot0[0]) < 1e-13 * tot0[0]
    assert abs(tot[4] - tot0[4]) < 1e-13 * tot0[4]
    ctx.close()


def test_errors_surface_as_codes():
    from xyst_amd._xg import XgError
    ctx = _ctx()
    with pytest.raises(XgError, match="ENOMESH"):
        ctx.grad(np.ones((4, 5)))
    c = box_case(2)
    ctx.upload_mesh(c.coord, c.se.sup, c.se.int, c.tri, c.besym)
    with pytest.raises(XgError, match="EUNSUPPORTED"):
        ctx.grad(np.ones((c.npoin, 8)))   # more than two transported scalars
    with pytest.raises(XgError, match="Flux not configured"):
        ctx.set_config(flux=7)
    bad = [s.copy() for s in c.se.sup]
    bad[2][0] = c.npoin + 3
    with pytest.raises(XgError, match="EINVAL"):
        ctx.upload_mesh(c.coord, bad, c.se.int, c.tri, c.besym)
    ctx.close()


@pytest.mark.parametrize("n", [160])
def test_full_size_properties(n):
    """160^3 (24.6M tets, 28.9M edges): free-stream preservation and exact
    mass/energy conservation of the rhs with all-symmetry walls, and
    bitwise repeatability -- size-independent properties."""
    import torch
    from xyst_amd import riemann as X
    _say("full size n=%d setup" % n)
    m, se, vol, sn, sv, besym = _resident_case(n, 7)
    ctx = _ctx()
    ctx.upload_mesh(m["coord"], se.sup, se.int, m["triinpoel"], besym)
    N = m["coord"].shape[1]
    dU = torch.zeros((N, 5), dtype=torch.float64, device="cuda")
    dU[:, 0] = 1.0
    dU[:, 4] = 2.5
    dG = torch.empty((N, 15), dtype=torch.float64, device="cuda")
    dR = torch.empty_like(dU)
    ctx.grad_dev(dU, dG)
    ctx.rhs_dev(dG, dU, dR)
    torch.cuda.synchronize()
    assert float(dG.abs().max()) < 1e-12
    assert float(dR.abs().max()) < 1e-12
    _say("full size: free stream ok")
    x = torch.from_numpy(m["coord"]).cuda()
    r = 1.0 + 0.2 * torch.sin(6.0 * x[0]) * torch.cos(5.0 * x[1])
    dU[:, 0] = r
    dU[:, 1] = r * 0.3 * torch.sin(4.0 * x[2])
    dU[:, 4] = 2.5 + 0.5 * torch.cos(3.0 * x[0] + x[1])
    ctx.grad_dev(dU, dG)
    dGv = dG / torch.from_numpy(vol).cuda()[:, None]
    ctx.rhs_dev(dGv, dU, dR)
    R1 = dR.clone()
    ctx.rhs_dev(dGv, dU, dR)
    torch.cuda.synchronize()
    assert torch.equal(R1, dR)
    tot = R1.sum(0).cpu().numpy()
    scale = R1.abs().sum(0).cpu().numpy()
    assert abs(tot[0]) < 1e-12 * scale[0] and abs(tot[4]) < 1e-12 * scale[4]
    _say("full size: conservation and repeatability ok")
    ctx.close()


def _exodiff(mine, ref, vel_names=("xvelocity", "yvelocity", "zvelocity")):
    """The reference's RieCG/Sod exodiff.cfg on two ExodusII files: nodes
    matched by coordinates (COORDINATES absolute 1e-6), TIME STEPS absolute
    1e-8, NODAL VARIABLES relative 1e-7 floor 1e-8, velocities absolute 1e-8
    floor 1e-7; every step of `ref`."""
    from xyst_amd import exodus
    A = exodus.read_mesh(mine, orient=False)["coord"]
    B = exodus.read_mesh(ref, orient=False)["coord"]
    key = lambda P: np.lexsort(np.round(P, 9)[::-1])
    ia, ib = key(A), key(B)
    assert A.shape == B.shape and np.abs(A[:, ia] - B[:, ib]).max() < 1e-6
    tref = exodus.times(ref)
    assert np.abs(exodus.times(mine) - tref).max() < 1e-8
    for step in range(tref.size):
        nm, fm, _ = exodus.read_fields(mine, step)
        nr, fr, _ = exodus.read_fields(ref, step)
        assert nm == nr
        for i, name in enumerate(nr):
            a, b = fm[ia, i], fr[ib, i]
            diff = np.abs(a - b)
            if name in vel_names:
                ok = (diff <= 1e-8) | (np.abs(b) < 1e-7)
            else:
                ok = (diff <= 1e-7 * np.abs(b)) | (np.abs(b) < 1e-8)
            assert ok.all(), (step, name, float(diff.max()))


def test_riecg_sod_exodus_end_to_end(tmp_path):
    """The reference's RieCG/Sod case from its files to an ExodusII result:
    the mesh read from rectangle_01_1.5k.exo (xg_exo), preprocessing by the
    host C-ABI, 10 RK3 steps on the device with the diagnostics summed on
    the device against diag.std, the initial and final fields written with
    xg_exo and compared with the reference's sod.std.exo in its exodiff
    bands."""
    import os
    import torch
    from xyst_amd import diagnostics, exodus
    from xyst_amd import riemann as X
    _say("riecg sod exodus end to end")
    here = os.path.dirname(os.path.abspath(__file__))
    exo = os.path.join(here, "golden", "exo")
    d = np.load(os.path.join(here, "golden", "riecg_sod.npz"))
    m = exodus.read_mesh(os.path.join(exo, "rectangle_01_1.5k.exo"))
    coord, inpoel, tri = m["coord"], m["inpoel"], m["triinpoel"]
    edge, dint = X.mesh_domint(coord, inpoel)
    sup, ints = X.edges_as_superedges(edge, dint)
    vol = X.mesh_vol(coord, inpoel)
    sn, sv, besym = X.mesh_symbc(coord, tri, m["faceset"], [int(s) for s in d["symsets"]])
    ctx = _ctx(gamma=float(d["gamma"]), flux="rusanov")
    ctx.upload_mesh(coord, sup, ints, tri, besym)
    ctx.riecg_setup(vol, sn, sv)
    U = torch.tensor(sod_state(coord), device="cuda")
    Un, Uold = torch.empty_like(U), torch.empty_like(U)
    G = torch.empty((U.shape[0], 15), dtype=torch.float64, device="cuda")
    R = torch.empty_like(U)
    dvol = torch.tensor(vol, device="cuda")
    names = ["density", "xvelocity", "yvelocity", "zvelocity", "energy", "pressure"]

    def fields(Uh, gamma=float(d["gamma"])):
        r = Uh[:, 0]
        vel = Uh[:, 1:4] / r[:, None]
        e = Uh[:, 4] / r
        return np.column_stack([r, vel, e, (gamma - 1.0) * r * (e - 0.5 * (vel ** 2).sum(1))])

    out = str(tmp_path / "sod.exo")
    w = exodus.Writer(out, m, names=names)
    w.write_step(0.0, fields(U.cpu().numpy()))
    t, term, V = 0.0, float(d["term"]), vol.sum()
    for it in range(int(d["nstep"])):
        dt = ctx.dt_dev(U, float(d["cfl"]))
        if t + dt > term:
            dt = term - t
        Uold.copy_(U)
        for s in range(3):
            ctx.stage_dev(s, dt, U, Un, G, R)
        t += dt
        dg = diagnostics.rhocompute(ctx, U, Uold, dvol)
        mine = np.concatenate([[it + 1, t, dt], np.sqrt(dg[:10] / V), [dg[10]]])
        # the tightest band of diag.ndiff.cfg (rel 1e-7) on every column: the
        # single-edge superedges sum in another order than the reference's
        np.testing.assert_allclose(mine, d["diag"][it], rtol=1e-7, atol=1e-14)
    w.write_step(t, fields(U.cpu().numpy()))
    w.close()
    ctx.close()
    _exodiff(out, os.path.join(exo, "sod.std.exo"))
